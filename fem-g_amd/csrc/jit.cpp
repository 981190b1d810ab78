// Run-time specialisation of coefficient programs.
//
// The reference evaluates coefficients by walking an expression tree at every
// Gauss point (fem-master/src/function/executionTree/ExecutionTree.h:58-153).
// Here a coefficient program (the GFEM_OP_* bytecode of a term) is turned
// into straight-line device code, compiled once per program shape with
// hipRTC for the device's architecture and cached per context; the program's
// constants stay kernel arguments, so a new source position or wavenumber
// reuses the compiled kernel.  Without a compiled kernel (compilation
// failure, or GFEM_JIT=0) the interpreter kernels run instead.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"

namespace gfem {

namespace {

const char *kPrelude = R"(
typedef long long i64;
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 cneg(double2 a) { return make_double2(-a.x, -a.y); }
__device__ __forceinline__ double2 cexp_(double2 a) {
  const double m = exp(a.x); double s, c; sincos(a.y, &s, &c); return make_double2(m * c, m * s);
}
__device__ __forceinline__ double2 clog_(double2 a) { return make_double2(log(hypot(a.x, a.y)), atan2(a.y, a.x)); }
__device__ __forceinline__ double2 csqrt_(double2 a) {
  const double r = hypot(a.x, a.y);
  if(r == 0.) return make_double2(0., 0.);
  const double re = sqrt(0.5 * (r + fabs(a.x))); const double im = 0.5 * a.y / re;
  return (a.x >= 0.) ? make_double2(re, im) : make_double2(fabs(im), a.y >= 0. ? re : -re);
}
__device__ __forceinline__ double2 csin_(double2 a) { double s, c; sincos(a.x, &s, &c); return make_double2(s * cosh(a.y), c * sinh(a.y)); }
__device__ __forceinline__ double2 ccos_(double2 a) { double s, c; sincos(a.x, &s, &c); return make_double2(c * cosh(a.y), -s * sinh(a.y)); }
__device__ __forceinline__ double2 cabs_(double2 a) { return make_double2(hypot(a.x, a.y), 0.); }
__device__ __forceinline__ double2 cpow_(double2 a, double2 b) {
  if(a.y == 0. && b.y == 0. && (a.x > 0. || b.x == floor(b.x))) return make_double2(pow(a.x, b.x), 0.);
  return (a.x == 0. && a.y == 0.) ? make_double2(0., 0.) : cexp_(cmul(b, clog_(a)));
}
struct JitLinArgs {
  double consts[2 * 32];
  double fac_re, fac_im;
  int Q, dim, nv, N, accumulate, EB;
};
)";

// A block takes EB consecutive elements: their vertex coordinates and
// determinants are staged in LDS once, then one thread per (element, Gauss
// point) evaluates c_q = w_q det fac f(x_q) into LDS, and one thread per
// (element, local dof) reduces b_e(i) = sum_q phi_i(q) c_q in q order.
const char *kLinearBody = R"(
extern "C" __global__ void __launch_bounds__(256)
gfem_jit_linear(JitLinArgs a, const double *__restrict__ w, const double *__restrict__ pts,
                const double *__restrict__ bf, i64 ne, const int *__restrict__ vtx,
                const double *__restrict__ xyz, const double *__restrict__ F64, double2 *__restrict__ be)
{
  extern __shared__ double s_dyn[];
  const int Q = a.Q, N = a.N, EB = a.EB, nv = a.nv;
  double *s_bf = s_dyn;             // [Q][N]
  double *s_u = s_bf + Q * N;       // [Q][3] reference points
  double *s_w = s_u + 3 * Q;        // [Q]
  double *s_X = s_w + Q;            // [EB][4][3] vertex coordinates
  double *s_det = s_X + 12 * EB;    // [EB]
  double2 *s_c = reinterpret_cast<double2 *>(s_det + EB + ((Q * N + 4 * Q + 13 * EB) & 1));  // [EB][Q]
  const int t = threadIdx.x;
  const i64 e0 = (i64)blockIdx.x * EB;
  const int nb = (int)(ne - e0 < (i64)EB ? ne - e0 : (i64)EB);
  for(int k = t; k < Q * N; k += 256) s_bf[k] = bf[k];
  for(int k = t; k < 3 * Q; k += 256) s_u[k] = pts[k];
  for(int k = t; k < Q; k += 256) s_w[k] = w[k];
  for(int k = t; k < nb * nv * 3; k += 256) {
    const int le = k / (3 * nv), r = k - le * 3 * nv, v = r / 3, c = r - 3 * v;
    s_X[12 * le + r] = xyz[3 * (i64)vtx[(e0 + le) * nv + v] + c];
  }
  for(int k = t; k < nb; k += 256) s_det[k] = F64[(e0 + k) * 8 + 6];
  __syncthreads();
  for(int p = t; p < nb * Q; p += 256) {
    const int le = p / Q, q = p - le * Q;
    const double *X = &s_X[12 * le];
    const double *u = &s_u[3 * q];
    double xp[3];
    if(a.dim == 1) {
      const double s = 0.5 * (u[0] + 1.);
      for(int i = 0; i < 3; ++i) xp[i] = X[i] + s * (X[3 + i] - X[i]);
    } else {
      for(int i = 0; i < 3; ++i) {
        double s = X[i];
        for(int j = 0; j < a.dim; ++j) s += u[j] * (X[3 * (j + 1) + i] - X[i]);
        xp[i] = s;
      }
    }
    const double x = xp[0], y = xp[1], z = xp[2];
    (void)x; (void)y; (void)z;
    @EXPR@
    const double wd = s_w[q] * s_det[le];
    s_c[p] = cmul(make_double2(a.fac_re * wd, a.fac_im * wd), f);
  }
  __syncthreads();
  for(int k = t; k < nb * N; k += 256) {
    const int le = k / N, i = k - le * N;
    const i64 o = (e0 + le) * N + i;
    double2 acc = a.accumulate ? be[o] : make_double2(0., 0.);
    const double2 *c = &s_c[le * Q];
    for(int q = 0; q < Q; ++q) {
      const double b = s_bf[q * N + i];
      acc.x += c[q].x * b;
      acc.y += c[q].y * b;
    }
    be[o] = acc;
  }
}
)";

// straight-line code computing `double2 f` from the program
std::string expr_code(const DevExpr &ex)
{
  std::ostringstream o;
  std::vector<std::string> st;
  int n = 0;
  const bool real = ex.real_only != 0;
  const char *T = real ? "double" : "double2";
  auto fresh = [&]() { return "v" + std::to_string(n++); };
  for(int k = 0; k < ex.n_ops; ++k) {
    const int op = ex.ops[2 * k], arg = ex.ops[2 * k + 1];
    const std::string v = fresh();
    if(op == GFEM_OP_CONST) {
      if(real)
        o << "const double " << v << " = a.consts[" << 2 * arg << "];\n";
      else
        o << "const double2 " << v << " = make_double2(a.consts[" << 2 * arg << "], a.consts[" << 2 * arg + 1 << "]);\n";
      st.push_back(v);
    }
    else if(op >= GFEM_OP_X && op <= GFEM_OP_Z) {
      const char *c = op == GFEM_OP_X ? "x" : (op == GFEM_OP_Y ? "y" : "z");
      if(real)
        o << "const double " << v << " = " << c << ";\n";
      else
        o << "const double2 " << v << " = make_double2(" << c << ", 0.);\n";
      st.push_back(v);
    }
    else if((op >= GFEM_OP_ADD && op <= GFEM_OP_DIV) || op == GFEM_OP_POW) {
      const std::string b = st.back();
      st.pop_back();
      const std::string a = st.back();
      st.pop_back();
      o << "const " << T << " " << v << " = ";
      if(real) {
        static const char *sym[] = {"+", "-", "*", "/"};
        if(op == GFEM_OP_POW)
          o << "pow(" << a << ", " << b << ");\n";
        else
          o << a << " " << sym[op - GFEM_OP_ADD] << " " << b << ";\n";
      }
      else {
        static const char *fn[] = {"cadd", "csub", "cmul", "cdiv"};
        o << (op == GFEM_OP_POW ? "cpow_" : fn[op - GFEM_OP_ADD]) << "(" << a << ", " << b << ");\n";
      }
      st.push_back(v);
    }
    else {
      const std::string a = st.back();
      st.pop_back();
      o << "const " << T << " " << v << " = ";
      if(real) {
        switch(op) {
        case GFEM_OP_NEG: o << "-" << a; break;
        case GFEM_OP_EXP: o << "exp(" << a << ")"; break;
        case GFEM_OP_SIN: o << "sin(" << a << ")"; break;
        case GFEM_OP_COS: o << "cos(" << a << ")"; break;
        case GFEM_OP_ABS: o << "fabs(" << a << ")"; break;
        default: o << a; break;
        }
      }
      else {
        switch(op) {
        case GFEM_OP_NEG: o << "cneg(" << a << ")"; break;
        case GFEM_OP_EXP: o << "cexp_(" << a << ")"; break;
        case GFEM_OP_SQRT: o << "csqrt_(" << a << ")"; break;
        case GFEM_OP_SIN: o << "csin_(" << a << ")"; break;
        case GFEM_OP_COS: o << "ccos_(" << a << ")"; break;
        case GFEM_OP_ABS: o << "cabs_(" << a << ")"; break;
        default: o << a; break;
        }
      }
      o << ";\n";
      st.push_back(v);
    }
  }
  const std::string r = st.empty() ? (real ? "0." : "make_double2(0., 0.)") : st.back();
  if(real)
    o << "const double2 f = make_double2(" << r << ", 0.);\n";
  else
    o << "const double2 f = " << r << ";\n";
  return o.str();
}

struct Cache {
  std::mutex m;
  std::map<std::pair<int, std::string>, hipFunction_t> fn;  // (device, source) -> kernel
  std::vector<hipModule_t> modules;
};

Cache &cache()
{
  static Cache c;
  return c;
}

bool jit_enabled()
{
  const char *e = std::getenv("GFEM_JIT");
  return !(e && e[0] == '0');
}

}  // namespace

hipFunction_t jit_linear_kernel(int device, const DevExpr &ex, std::string *log)
{
  if(!jit_enabled()) return nullptr;
  std::string src = std::string(kPrelude) + kLinearBody;
  const std::string code = expr_code(ex);
  src.replace(src.find("@EXPR@"), 6, code);
  Cache &C = cache();
  std::lock_guard<std::mutex> lk(C.m);
  auto key = std::make_pair(device, src);
  auto it = C.fn.find(key);
  if(it != C.fn.end()) return it->second;
  hipDeviceProp_t prop;
  if(hipGetDeviceProperties(&prop, device) != hipSuccess) return nullptr;
  std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  const size_t colon = arch.find(':');
  if(colon != std::string::npos) arch = arch.substr(0, colon);
  hiprtcProgram prog;
  if(hiprtcCreateProgram(&prog, src.c_str(), "gfem_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) return nullptr;
  const char *opts[] = {arch.c_str(), "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, 6, opts);
  if(r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string l(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &l[0]);
    if(log) *log = l;
    hiprtcDestroyProgram(&prog);
    C.fn[key] = nullptr;
    return nullptr;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> bin(cs);
  hiprtcGetCode(prog, bin.data());
  hiprtcDestroyProgram(&prog);
  hipModule_t mod;
  hipFunction_t fn = nullptr;
  if(hipModuleLoadData(&mod, bin.data()) != hipSuccess) return nullptr;
  if(hipModuleGetFunction(&fn, mod, "gfem_jit_linear") != hipSuccess) return nullptr;
  C.modules.push_back(mod);
  C.fn[key] = fn;
  return fn;
}

}  // namespace gfem
