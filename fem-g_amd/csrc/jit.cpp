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
#include <cstring>
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
  int Q, dim, nv, N, accumulate, EB, real_out;
};
)";

// One thread per element, the group's shape fixed at compile time (Q Gauss
// points, N dofs, NV vertices, dimension DIM): the element's vertices and
// determinant in registers, x_q, f(x_q) and c_q = w_q det fac f(x_q) per
// point, b_e(i) = sum_q phi_i(q) c_q accumulated in registers in q order
// (onto the element load already there for a second term on the same
// right-hand side).  Rule and basis are read with uniform (scalar) loads;
// no LDS, no barrier.
const char *kLinearBody = R"(
__device__ __forceinline__ void lin_element(const JitLinArgs &a, const double *__restrict__ w,
                                            const double *__restrict__ pts, const double *__restrict__ bf, i64 e,
                                            const int *__restrict__ vtx, const double *__restrict__ xyz,
                                            const double *__restrict__ F64, void *__restrict__ be)
{
  constexpr int Q = @Q@, N = @N@, NV = @NV@, DIM = @DIM@;
  double X[NV][3];
#pragma unroll
  for(int v = 0; v < NV; ++v) {
    const i64 n = vtx[e * NV + v];
#pragma unroll
    for(int i = 0; i < 3; ++i) X[v][i] = xyz[3 * n + i];
  }
  const double det = F64[e * 8 + 6];
  double2 *o = reinterpret_cast<double2 *>(be) + e * N;
  double2 acc[N];
  const bool ro = a.real_out != 0;  // element loads stored as double (imaginary part 0)
  const double *orr = reinterpret_cast<const double *>(be) + e * N;
#pragma unroll
  for(int i = 0; i < N; ++i)
    acc[i] = a.accumulate ? (ro ? make_double2(orr[i], 0.) : o[i]) : make_double2(0., 0.);
  for(int q = 0; q < Q; ++q) {
    double xp[3];
    if(DIM == 1) {
      const double s = 0.5 * (pts[3 * q] + 1.);
#pragma unroll
      for(int i = 0; i < 3; ++i) xp[i] = X[0][i] + s * (X[1][i] - X[0][i]);
    } else {
#pragma unroll
      for(int i = 0; i < 3; ++i) {
        double s = X[0][i];
#pragma unroll
        for(int j = 0; j < DIM; ++j) s += pts[3 * q + j] * (X[j + 1][i] - X[0][i]);
        xp[i] = s;
      }
    }
    const double x = xp[0], y = xp[1], z = xp[2];
    (void)x; (void)y; (void)z;
    @EXPR@
    const double wd = w[q] * det;
    const double2 c = cmul(make_double2(a.fac_re * wd, a.fac_im * wd), f);
    if(ro) {
#pragma unroll
      for(int i = 0; i < N; ++i) acc[i].x += c.x * bf[q * N + i];
    }
    else {
#pragma unroll
      for(int i = 0; i < N; ++i) {
        const double b = bf[q * N + i];
        acc[i].x += c.x * b;
        acc[i].y += c.y * b;
      }
    }
  }
  if(ro) {
    double *ow = reinterpret_cast<double *>(be) + e * N;
#pragma unroll
    for(int i = 0; i < N; ++i) ow[i] = acc[i].x;
  }
  else {
#pragma unroll
    for(int i = 0; i < N; ++i) o[i] = acc[i];
  }
}

extern "C" __global__ void __launch_bounds__(512)
gfem_jit_linear(JitLinArgs a, const double *__restrict__ w, const double *__restrict__ pts,
                const double *__restrict__ bf, i64 ne, const int *__restrict__ vtx,
                const double *__restrict__ xyz, const double *__restrict__ F64, void *__restrict__ be)
{
  const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= ne) return;
  lin_element(a, w, pts, bf, e, vtx, xyz, F64, be);
}

// the same program over the groups of several systems: item blockIdx.y
struct JitLinItem {
  JitLinArgs a;
  const double *w, *pts, *bf;
  i64 ne;
  const int *vtx;
  const double *xyz, *F64;
  void *be;
};
extern "C" __global__ void __launch_bounds__(512) gfem_jit_linear_many(const JitLinItem *__restrict__ items)
{
  const JitLinItem &it = items[blockIdx.y];
  const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= it.ne) return;
  lin_element(it.a, it.w, it.pts, it.bf, e, it.vtx, it.xyz, it.F64, it.be);
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
  std::map<hipFunction_t, hipFunction_t> many;              // kernel -> its entry over several groups
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

hipFunction_t jit_linear_kernel(int device, const DevExpr &ex, int Q, int N, int nv, int dim, std::string *log)
{
  if(!jit_enabled()) return nullptr;
  std::string src = std::string(kPrelude) + kLinearBody;
  const std::string code = expr_code(ex);
  src.replace(src.find("@EXPR@"), 6, code);
  const std::pair<const char *, int> shape[] = {{"@Q@", Q}, {"@N@", N}, {"@NV@", nv}, {"@DIM@", dim}};
  for(const auto &kv : shape) src.replace(src.find(kv.first), std::strlen(kv.first), std::to_string(kv.second));
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
  hipFunction_t fm = nullptr;
  if(hipModuleGetFunction(&fm, mod, "gfem_jit_linear_many") != hipSuccess) fm = nullptr;
  C.modules.push_back(mod);
  C.fn[key] = fn;
  C.many[fn] = fm;
  return fn;
}

hipFunction_t jit_linear_many(hipFunction_t single)
{
  if(!single) return nullptr;
  Cache &C = cache();
  std::lock_guard<std::mutex> lk(C.m);
  auto it = C.many.find(single);
  return it == C.many.end() ? nullptr : it->second;
}

}  // namespace gfem
