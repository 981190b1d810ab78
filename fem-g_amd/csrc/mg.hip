// Shifted-Laplacian multigrid hierarchy (layout and method in mg.h): DOF
// positions, trilinear transfers, Galerkin products of the unit term
// matrices (host, once per system), and the per-request level operators
// (device, at submission).  The V-cycle itself runs in batch.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "internal.h"
#include "mg.h"

namespace gfem {
namespace {

constexpr int kTB = 256;
inline unsigned blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kTB - 1) / kTB); }
__device__ inline int64_t gid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline bool row_type(int32_t t) { return t == GFEM_AT_UNKNOWN || t == GFEM_AT_UNKNOWN_BUBBLE; }

// ------------------------------------------------------------ DOF positions

// the element of smallest global index that holds each row DOF
__global__ void k_mg_owner(const int32_t *__restrict__ num, const int32_t *__restrict__ type, int64_t ne, int N,
                           int64_t eoff, unsigned long long *__restrict__ owner)
{
  const int64_t i = gid();
  if(i >= ne * N) return;
  if(!row_type(type[i])) return;
  atomicMin(&owner[num[i]], (unsigned long long)(eoff + i / N));
}

// the owner writes the DOF's node position: curved groups from their node
// list, affine groups from the vertices and the node's lattice multi-index
__global__ void k_mg_dof_xyz(const int32_t *__restrict__ num, const int32_t *__restrict__ type, int64_t ne, int N,
                             int nv, int p, int64_t eoff, const int32_t *__restrict__ vtx,
                             const int32_t *__restrict__ conn, const int *__restrict__ nodes,
                             const double *__restrict__ xyz, const unsigned long long *__restrict__ owner,
                             double *__restrict__ out)
{
  const int64_t i = gid();
  if(i >= ne * N) return;
  if(!row_type(type[i])) return;
  const int64_t e = i / N;
  const int a = (int)(i - e * N);
  const int32_t d = num[i];
  if(owner[d] != (unsigned long long)(eoff + e)) return;
  double X[3] = {0., 0., 0.};
  if(conn) {
    const int32_t v = conn[e * N + a];
    for(int c = 0; c < 3; ++c) X[c] = xyz[3 * (int64_t)v + c];
  }
  else {
    for(int k = 0; k < nv; ++k) {
      const double lam = (double)nodes[4 * a + k] / (double)p;
      const int32_t v = vtx[e * nv + k];
      for(int c = 0; c < 3; ++c) X[c] += lam * xyz[3 * (int64_t)v + c];
    }
  }
  for(int c = 0; c < 3; ++c) out[3 * (int64_t)d + c] = X[c];
}

// FP64 complex values to complex64
__global__ void k_mg_to_f2(int64_t n, const double2 *__restrict__ a, float2 *__restrict__ out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) out[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// ------------------------------------------------------------ host sparse algebra

unsigned host_threads()
{
  return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

// fn(begin, end) over contiguous ranges of [0, n), one per host thread
template <class Fn>
void host_parallel(int64_t n, Fn fn)
{
  const unsigned nth = (unsigned)std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / 65536 + 1));
  if(nth == 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  for(unsigned w = 0; w < nth; ++w) th.emplace_back(fn, n * w / nth, n * (w + 1) / nth);
  for(auto &t : th) t.join();
}

struct Csr {
  int64_t rows = 0, cols = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
  std::vector<double> w;  // transfers: real weights
};

// 1-D linear interpolation of coordinate x between the nodes of grid g
inline void interp1(const std::vector<double> &g, double x, int &j, double &t)
{
  const int G = (int)g.size();
  if(G == 1) {
    j = 0;
    t = 0.;
    return;
  }
  j = (int)(std::upper_bound(g.begin(), g.end(), x) - g.begin()) - 1;
  j = std::min(std::max(j, 0), G - 2);
  t = (x - g[j]) / (g[j + 1] - g[j]);
  t = std::min(std::max(t, 0.), 1.);
}

// trilinear interpolation of grid values (x fastest) at points pts
Csr trilinear(const std::vector<double> &pts, const std::vector<double> g[3])
{
  const int64_t n = (int64_t)pts.size() / 3;
  const int64_t G0 = (int64_t)g[0].size(), G1 = (int64_t)g[1].size(), G2 = (int64_t)g[2].size();
  Csr P;
  P.rows = n;
  P.cols = G0 * G1 * G2;
  P.ptr.assign((size_t)n + 1, 0);
  for(int64_t i = 0; i < n; ++i) {
    int j[3];
    double t[3];
    for(int a = 0; a < 3; ++a) interp1(g[a], pts[3 * i + a], j[a], t[a]);
    for(int dz = 0; dz < 2; ++dz)
      for(int dy = 0; dy < 2; ++dy)
        for(int dx = 0; dx < 2; ++dx) {
          if((dx && G0 == 1) || (dy && G1 == 1) || (dz && G2 == 1)) continue;
          const double w = (dx ? t[0] : 1. - t[0]) * (dy ? t[1] : 1. - t[1]) * (dz ? t[2] : 1. - t[2]);
          if(w < 1e-12) continue;
          P.idx.push_back((int32_t)((j[0] + dx) + G0 * ((j[1] + dy) + G1 * (j[2] + dz))));
          P.w.push_back(w);
        }
    P.ptr[i + 1] = (int64_t)P.idx.size();
  }
  return P;
}

Csr transpose(const Csr &A)
{
  Csr T;
  T.rows = A.cols;
  T.cols = A.rows;
  T.ptr.assign((size_t)T.rows + 1, 0);
  for(int32_t c : A.idx) T.ptr[(size_t)c + 1]++;
  for(int64_t r = 0; r < T.rows; ++r) T.ptr[r + 1] += T.ptr[r];
  T.idx.resize(A.idx.size());
  T.w.resize(A.idx.size());
  std::vector<int64_t> pos(T.ptr.begin(), T.ptr.end() - 1);
  for(int64_t r = 0; r < A.rows; ++r)
    for(int64_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
      const int64_t q = pos[A.idx[k]]++;
      T.idx[q] = (int32_t)r;
      T.w[q] = A.w[k];
    }
  return T;
}

// complex FP32 value of the host products (the coarse operators are complex64)
struct CF {
  float re, im;
};

// a pattern with nt complex values per entry, [nnz][nt], left uninitialised
// at allocation (the unit matrices of large meshes take gigabytes)
struct MultiCsr {
  int64_t rows = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> idx;
  std::unique_ptr<CF[]> v;
};

// C = P^T X P for every value plane of X, rows of C in parallel; each row's
// columns ascending
MultiCsr galerkin(const MultiCsr &X, const Csr &P, const Csr &Pt, int nt)
{
  const int64_t nc = P.cols;
  unsigned nth = host_threads();
  std::vector<std::vector<int32_t>> cols(nth);
  std::vector<std::vector<CF>> vals(nth);
  std::vector<std::vector<int64_t>> rowlen(nth);
  std::vector<std::thread> th;
  auto work = [&](unsigned w) {
    const int64_t r0 = nc * w / nth, r1 = nc * (w + 1) / nth;
    std::vector<int64_t> mark((size_t)nc, -1);
    std::vector<CF> acc((size_t)nc * nt);
    std::vector<int32_t> list;
    for(int64_t I = r0; I < r1; ++I) {
      list.clear();
      for(int64_t a = Pt.ptr[I]; a < Pt.ptr[I + 1]; ++a) {
        const int64_t i = Pt.idx[a];
        const float wi = (float)Pt.w[a];
        for(int64_t k = X.ptr[i]; k < X.ptr[i + 1]; ++k) {
          const int64_t j = X.idx[k];
          const CF *xv = &X.v[(size_t)k * nt];
          for(int64_t b = P.ptr[j]; b < P.ptr[j + 1]; ++b) {
            const int32_t J = P.idx[b];
            const float f = wi * (float)P.w[b];
            CF *ac = &acc[(size_t)J * nt];
            if(mark[J] != I) {
              mark[J] = I;
              list.push_back(J);
              for(int t = 0; t < nt; ++t) ac[t] = CF{0.f, 0.f};
            }
            for(int t = 0; t < nt; ++t) {
              ac[t].re += f * xv[t].re;
              ac[t].im += f * xv[t].im;
            }
          }
        }
      }
      std::sort(list.begin(), list.end());
      rowlen[w].push_back((int64_t)list.size());
      for(int32_t J : list) {
        cols[w].push_back(J);
        for(int t = 0; t < nt; ++t) vals[w].push_back(acc[(size_t)J * nt + t]);
      }
    }
  };
  for(unsigned w = 0; w < nth; ++w) th.emplace_back(work, w);
  for(auto &t : th) t.join();
  MultiCsr C;
  C.rows = nc;
  C.ptr.assign((size_t)nc + 1, 0);
  int64_t r = 0;
  size_t nvals = 0;
  for(unsigned w = 0; w < nth; ++w) nvals += vals[w].size();
  C.v.reset(new CF[std::max<size_t>(nvals, 1)]);
  size_t at = 0;
  for(unsigned w = 0; w < nth; ++w) {
    for(int64_t len : rowlen[w]) {
      C.ptr[r + 1] = C.ptr[r] + len;
      ++r;
    }
    C.idx.insert(C.idx.end(), cols[w].begin(), cols[w].end());
    std::copy(vals[w].begin(), vals[w].end(), C.v.get() + at);
    at += vals[w].size();
  }
  return C;
}

template <typename T>
hipError_t up(T **d, const T *h, size_t n)
{
  return dupload(d, h, n);
}

// ------------------------------------------------------------ per-request operators

// level 0: S = A + i beta sum_s c_s X_s (the system's pattern)
__global__ void k_mg_level0(int64_t nnz, const double2 *__restrict__ a64, const float2 *__restrict__ xs, float2 cs,
                            float2 *__restrict__ out)
{
  const int64_t k = gid();
  if(k >= nnz) return;
  const float2 x = xs[k];
  out[k] = make_float2((float)a64[k].x + (cs.x * x.x - cs.y * x.y), (float)a64[k].y + (cs.x * x.y + cs.y * x.x));
}

// coarse level values: sum_t c_t G_t
__global__ void k_mg_combine(int64_t nnz, const float2 *__restrict__ g, Coefs cf, float2 *__restrict__ out)
{
  const int64_t k = gid();
  if(k >= nnz) return;
  float re = 0.f, im = 0.f;
  for(int t = 0; t < cf.nt; ++t) {
    const float2 v = g[(int64_t)t * nnz + k], c = cf.c[t];
    re += c.x * v.x - c.y * v.y;
    im += c.x * v.y + c.y * v.x;
  }
  out[k] = make_float2(re, im);
}

// omega / diag(S); 64-bit or 32-bit row pointers
template <typename P>
__global__ void k_mg_dinv(int64_t n, const P *__restrict__ ai, const int32_t *__restrict__ aj,
                          const float2 *__restrict__ a, float2 *__restrict__ dinv)
{
  const int64_t r = gid();
  if(r >= n) return;
  float2 d = make_float2(1.f, 0.f);
  for(int64_t k = ai[r]; k < ai[r + 1]; ++k)
    if(aj[k] == (int32_t)r) {
      d = a[k];
      break;
    }
  const float m = d.x * d.x + d.y * d.y;
  dinv[r] = m > 0.f ? make_float2(kMgOmega * d.x / m, -kMgOmega * d.y / m) : make_float2(0.f, 0.f);
}

size_t a256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

void mg_free(MgSystem *m)
{
  if(!m) return;
  for(int l = 0; l < kMgMaxLevels; ++l) {
    dfree(m->ai[l]);
    dfree(m->aj[l]);
    dfree(m->g[l]);
    dfree(m->pi[l]);
    dfree(m->pj[l]);
    dfree(m->pw[l]);
    dfree(m->ri[l]);
    dfree(m->rj[l]);
    dfree(m->rw[l]);
  }
  dfree(m->gdense);
  dfree(m->xs);
  delete m;
}

size_t mg_request_bytes(const gfem_system *sys)
{
  const MgSystem *m = sys->mg;
  if(!m) return 0;
  size_t b = a256(sizeof(MgReq));
  for(int l = 0; l < m->nlev; ++l) {
    const size_t n = (size_t)m->n[l];
    b += 5 * a256(8 * n);  // dinv, b, x, r, y
    if(l < m->nlev - 1) b += a256(8 * (size_t)m->nnz[l]);  // level values (level 0: S, the system pattern)
  }
  const size_t nc = (size_t)m->n[m->nlev - 1];
  b += a256(8 * nc * nc);  // coarsest inverse, column-major
  return b;
}

void mg_bind(const gfem_system *sys, char *base, MgReq *req)
{
  const MgSystem *m = sys->mg;
  char *p = base;
  auto take = [&](size_t bytes) {
    char *q = p;
    p += a256(bytes);
    return q;
  };
  take(sizeof(MgReq));
  std::memset(req, 0, sizeof(MgReq));
  req->nlev = m->nlev;
  req->lsmall = m->lsmall;
  for(int l = 0; l < m->nlev; ++l) {
    MgLevelReq &L = req->lev[l];
    const size_t n = (size_t)m->n[l];
    L.n = (int)n;
    if(l == 0) {
      L.ai64 = sys->d_ai;
      L.aj = sys->d_aj;
    }
    else {
      L.ai = m->ai[l];
      L.aj = m->aj[l];
    }
    L.dinv = reinterpret_cast<float2 *>(take(8 * n));
    L.b = reinterpret_cast<float2 *>(take(8 * n));
    L.x = reinterpret_cast<float2 *>(take(8 * n));
    L.r = reinterpret_cast<float2 *>(take(8 * n));
    L.y = reinterpret_cast<float2 *>(take(8 * n));
    if(l < m->nlev - 1) {
      L.a = reinterpret_cast<float2 *>(take(8 * (size_t)m->nnz[l]));
      L.pi = m->pi[l];
      L.pj = m->pj[l];
      L.pw = m->pw[l];
      L.ri = m->ri[l];
      L.rj = m->rj[l];
      L.rw = m->rw[l];
    }
  }
  const size_t nc = (size_t)m->n[m->nlev - 1];
  req->ainv = reinterpret_cast<float2 *>(take(8 * nc * nc));
}

hipError_t mg_prepare(const gfem_system *sys, MgReq &req, const double2 *a64, hipStream_t s)
{
  const MgSystem *m = sys->mg;
  Coefs cf{};
  cf.nt = m->nterms;
  float2 cshift = make_float2(0.f, 0.f);
  for(int t = 0; t < m->nterms; ++t) {
    const DevTerm &d = sys->groups[m->term_group[t]].dterms[m->term_index[t]];
    std::complex<double> c(d.c_re, d.c_im);
    if(t == m->shift) {
      const std::complex<double> ib(0., m->beta);
      cshift = make_float2((float)(ib * c).real(), (float)(ib * c).imag());
      c *= 1. + ib;
    }
    cf.c[t] = make_float2((float)c.real(), (float)c.imag());
  }
  // level 0: S from the request's matrix and the shifted term's unit matrix
  const MgLevelReq &L0 = req.lev[0];
  hipLaunchKernelGGL(k_mg_level0, dim3(blocks(sys->nnz)), dim3(kTB), 0, s, (int64_t)sys->nnz, a64, m->xs, cshift,
                     const_cast<float2 *>(L0.a));
  hipLaunchKernelGGL(k_mg_dinv<int64_t>, dim3(blocks(L0.n)), dim3(kTB), 0, s, (int64_t)L0.n, sys->d_ai, sys->d_aj,
                     L0.a, const_cast<float2 *>(L0.dinv));
  for(int l = 1; l < m->nlev - 1; ++l) {
    const MgLevelReq &L = req.lev[l];
    hipLaunchKernelGGL(k_mg_combine, dim3(blocks(m->nnz[l])), dim3(kTB), 0, s, (int64_t)m->nnz[l], m->g[l], cf,
                       const_cast<float2 *>(L.a));
    hipLaunchKernelGGL(k_mg_dinv<int32_t>, dim3(blocks(L.n)), dim3(kTB), 0, s, (int64_t)L.n, m->ai[l], m->aj[l], L.a,
                       const_cast<float2 *>(L.dinv));
  }
  req.gdense = m->gdense;
  req.cf = cf;
  return hipGetLastError();
}

}  // namespace gfem

using namespace gfem;

extern "C" gfem_status gfem_mg_setup(gfem_system *sys, int n_shift, const int *shift_group, const int *shift_term,
                                     double beta)
{
  if(!sys) return fail(nullptr, GFEM_ERR_INVALID_ARG, "mg_setup: system is null");
  gfem_ctx *ctx = sys->ctx;
  if(!sys->finalized) return fail(ctx, GFEM_ERR_STATE, "mg_setup: system not finalized");
  if(!sys->cplx) return fail(ctx, GFEM_ERR_UNSUPPORTED, "mg_setup: complex systems only");
  if(sys->symmetric) return fail(ctx, GFEM_ERR_UNSUPPORTED, "mg_setup: symmetric storage");
  if(n_shift < 1 || !shift_group || !shift_term)
    return fail(ctx, GFEM_ERR_INVALID_ARG, "mg_setup: at least one shifted (group, term)");
  if(!(beta > 0.)) return fail(ctx, GFEM_ERR_INVALID_ARG, "mg_setup: beta > 0");
  if(sys->batch_refs > 0)  // their cycles read the hierarchy this call replaces
    return fail(ctx, GFEM_ERR_STATE, "mg_setup: requests of this system are in a batch");
  if(sys->n >= (int64_t(1) << 31) || sys->nnz >= (int64_t(1) << 31))
    return fail(ctx, GFEM_ERR_TOO_LARGE, "mg_setup: system too large");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  // the constant-coefficient bilinear terms, in classes of equal
  // coefficient, form and test-function factor (one integral() over several
  // element groups: its terms keep equal coefficients, so the class has one
  // unit matrix, the sum of theirs); the shifted slots must form one class
  std::vector<std::pair<int, int>> all;  // every member
  std::vector<std::vector<std::pair<int, int>>> cls;
  auto same = [&](const DevTerm &a, const DevTerm &b) {
    return a.c_re == b.c_re && a.c_im == b.c_im && a.form_dof == b.form_dof && a.form_tf == b.form_tf &&
           a.fac_re == b.fac_re && a.fac_im == b.fac_im;
  };
  auto term = [&](const std::pair<int, int> &s) -> const DevTerm & { return sys->groups[s.first].dterms[s.second]; };
  for(int g = 0; g < (int)sys->groups.size(); ++g) {
    const Group &G = sys->groups[g];
    if(G.family != 0) return fail(ctx, GFEM_ERR_UNSUPPORTED, "mg_setup: Lagrange fields only");
    for(int t = 0; t < (int)G.dterms.size(); ++t) {
      const DevTerm &d = G.dterms[t];
      if(d.kind != 0) continue;
      if(d.per_point) return fail(ctx, GFEM_ERR_UNSUPPORTED, "mg_setup: expression coefficients of bilinear terms");
      all.push_back({g, t});
      size_t c = 0;
      while(c < cls.size() && !same(term(cls[c][0]), d)) ++c;
      if(c == cls.size()) cls.emplace_back();
      cls[c].push_back({g, t});
    }
  }
  int shift = -1;
  for(int i = 0; i < n_shift; ++i) {
    int c = -1;
    for(int u = 0; u < (int)cls.size() && c < 0; ++u)
      for(const auto &s : cls[u])
        if(s.first == shift_group[i] && s.second == shift_term[i]) c = u;
    if(c < 0) return fail(ctx, GFEM_ERR_INVALID_ARG, "mg_setup: a shifted (group, term) is not a bilinear term");
    if(shift >= 0 && c != shift)
      return fail(ctx, GFEM_ERR_INVALID_ARG, "mg_setup: the shifted terms have different coefficients or forms");
    shift = c;
  }
  if((int)cls.size() > kMgMaxTerms) return fail(ctx, GFEM_ERR_UNSUPPORTED, "mg_setup: too many bilinear terms");
  const int nt = (int)cls.size();
  const int64_t n = sys->n, nnz = sys->nnz;
  if(n == 0) return fail(ctx, GFEM_ERR_INVALID_ARG, "mg_setup: no unknowns");

  SetupTimer timer;
  // the shifted term's unit matrix on level 0, complex64, handed to the
  // hierarchy at the end (freed here on every earlier return)
  struct DevBuf {
    float2 *p = nullptr;
    ~DevBuf() { dfree(p); }
  } xs_dev;
  // DOF positions
  std::vector<double> xyz((size_t)(3 * n), 0.);
  {
    unsigned long long *owner = nullptr;
    double *dx = nullptr;
    hipError_t e = dalloc(&owner, (size_t)n);
    if(e == hipSuccess) e = dalloc(&dx, (size_t)(3 * n));
    if(e == hipSuccess) e = hipMemset(owner, 0xFF, sizeof(unsigned long long) * (size_t)n);
    if(e == hipSuccess) e = hipMemset(dx, 0, sizeof(double) * (size_t)(3 * n));
    for(const Group &G : sys->groups) {
      if(e != hipSuccess || G.ne == 0) continue;
      hipLaunchKernelGGL(k_mg_owner, dim3(blocks(G.ne * G.N)), dim3(kTB), 0, nullptr, G.d_num, G.d_type, G.ne, G.N,
                         G.e_off, owner);
    }
    for(const Group &G : sys->groups) {
      if(e != hipSuccess || G.ne == 0) continue;
      hipLaunchKernelGGL(k_mg_dof_xyz, dim3(blocks(G.ne * G.N)), dim3(kTB), 0, nullptr, G.d_num, G.d_type, G.ne, G.N,
                         G.nv, std::max(G.order, 1), G.e_off, G.d_vtx, G.d_conn, G.d_nodes, sys->d_xyz, owner, dx);
    }
    if(e == hipSuccess) e = hipGetLastError();
    if(e == hipSuccess) e = hipMemcpy(xyz.data(), dx, sizeof(double) * (size_t)(3 * n), hipMemcpyDeviceToHost);
    dfree(owner);
    dfree(dx);
    if(e != hipSuccess) return hip_fail(ctx, e, "mg_setup: DOF positions");
  }

  timer.mark("mg: DOF positions");
  // unit matrices of the terms: each assembled alone with coefficient 1
  MultiCsr X;
  X.rows = n;
  X.ptr.resize((size_t)n + 1);
  X.idx.resize((size_t)nnz);
  X.v.reset(new CF[(size_t)std::max<int64_t>(nnz, 1) * nt]);
  {
    std::vector<int64_t> ai((size_t)n + 1);
    GFEM_HIP(ctx, hipMemcpy(ai.data(), sys->d_ai, sizeof(int64_t) * ai.size(), hipMemcpyDeviceToHost));
    GFEM_HIP(ctx, hipMemcpy(X.idx.data(), sys->d_aj, sizeof(int32_t) * (size_t)nnz, hipMemcpyDeviceToHost));
    for(size_t i = 0; i < ai.size(); ++i) X.ptr[i] = ai[i];
    std::vector<std::complex<double>> saved(all.size());
    for(size_t i = 0; i < all.size(); ++i) saved[i] = {term(all[i]).c_re, term(all[i]).c_im};
    double2 *vals = nullptr, *rhs = nullptr, *h = nullptr;  // h: pinned staging of one unit matrix
    GFEM_HIP(ctx, dalloc(&vals, (size_t)nnz));
    hipError_t e = dalloc(&rhs, (size_t)(n * sys->nrhs));
    if(e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&h), sizeof(double2) * (size_t)nnz);
    if(e == hipSuccess) e = dalloc(&xs_dev.p, (size_t)std::max<int64_t>(nnz, 1));
    gfem_status st = e == hipSuccess ? GFEM_OK : hip_fail(ctx, e, "mg_setup: buffers");
    for(int t = 0; t < nt && st == GFEM_OK; ++t) {
      for(int u = 0; u < nt; ++u)
        for(const auto &s : cls[u])
          if(st == GFEM_OK) st = gfem_term_update(sys, s.first, s.second, u == t ? 1. : 0., 0., 0, nullptr, nullptr);
      if(st == GFEM_OK) st = gfem_assemble_into(sys, GFEM_PREC_FP64, vals, rhs, nullptr);
      if(st == GFEM_OK) {
        e = hipMemcpy(h, vals, sizeof(double2) * (size_t)nnz, hipMemcpyDeviceToHost);
        if(e == hipSuccess && t == shift) {
          // the shifted term's unit matrix stays on the device (level 0 of each request)
          hipLaunchKernelGGL(k_mg_to_f2, dim3(blocks(nnz)), dim3(kTB), 0, nullptr, nnz, vals, xs_dev.p);
          e = hipGetLastError();
        }
        if(e != hipSuccess) st = hip_fail(ctx, e, "mg_setup: unit matrix");
      }
      timer.mark("mg: unit matrix assembled and copied");
      if(st == GFEM_OK)
        host_parallel(nnz, [&](int64_t b, int64_t end) {
          for(int64_t k = b; k < end; ++k) X.v[(size_t)k * nt + t] = CF{(float)h[k].x, (float)h[k].y};
        });
      timer.mark("mg: unit matrix converted");
    }
    for(size_t i = 0; i < all.size(); ++i) {
      const gfem_status r =
        gfem_term_update(sys, all[i].first, all[i].second, saved[i].real(), saved[i].imag(), 0, nullptr, nullptr);
      if(st == GFEM_OK) st = r;
    }
    dfree(vals);
    dfree(rhs);
    if(h) (void)hipHostFree(h);
    if(st != GFEM_OK) return st;
  }

  timer.mark("mg: unit matrices");
  // grids: twice the mean DOF spacing, halved per level down to kMgCoarsest
  double lo[3], hi[3];
  for(int a = 0; a < 3; ++a) {
    lo[a] = 1e300;
    hi[a] = -1e300;
  }
  for(int64_t i = 0; i < n; ++i)
    for(int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], xyz[3 * i + a]);
      hi[a] = std::max(hi[a], xyz[3 * i + a]);
    }
  double vol = 1.;
  int dims = 0;
  for(int a = 0; a < 3; ++a)
    if(hi[a] - lo[a] > 1e-12 * (1. + std::fabs(hi[a]))) {
      vol *= hi[a] - lo[a];
      ++dims;
    }
  if(dims == 0) return fail(ctx, GFEM_ERR_UNSUPPORTED, "mg_setup: DOFs at one point");
  const double h0 = std::pow(vol / (double)n, 1. / dims);
  std::vector<double> grid[3];
  for(int a = 0; a < 3; ++a) {
    const double ext = hi[a] - lo[a];
    const int G = ext > 1e-12 * (1. + std::fabs(hi[a])) ? std::max(2, (int)std::lround(ext / (2. * h0)) + 1) : 1;
    grid[a].resize(G);
    for(int i = 0; i < G; ++i) grid[a][i] = G == 1 ? lo[a] : lo[a] + ext * i / (G - 1);
  }
  // P_0 interpolates at the DOF positions snapped to the lattice of half the
  // level-1 spacing: weights 1, 1/2, 1/4 or 1/8 (on average 3.4 per row
  // instead of 8), which keeps P_0^T X P_0 and the level-1 stencil small
  std::vector<double> snap(xyz);
  for(int64_t i = 0; i < n; ++i)
    for(int a = 0; a < 3; ++a) {
      const int G = (int)grid[a].size();
      if(G < 2) continue;
      const double half = 0.5 * (grid[a][G - 1] - grid[a][0]) / (G - 1);
      const double q = std::round((xyz[3 * i + a] - grid[a][0]) / half);
      snap[3 * i + a] = std::min(std::max(grid[a][0] + q * half, grid[a][0]), grid[a][G - 1]);
    }
  std::vector<Csr> Ps;
  std::vector<MultiCsr> Ls;
  Ls.push_back(std::move(X));
  Ps.push_back(trilinear(snap, grid));
  auto coarser = [](const std::vector<double> &g) {
    std::vector<double> c;
    for(size_t i = 0; i < g.size(); i += 2) c.push_back(g[i]);
    if(c.back() != g.back()) c.push_back(g.back());
    return c;
  };
  while(true) {
    const Csr Pt = transpose(Ps.back());
    Ls.push_back(galerkin(Ls.back(), Ps.back(), Pt, nt));
    int gmax = 0;
    for(int a = 0; a < 3; ++a) gmax = std::max(gmax, (int)grid[a].size());
    if(gmax <= kMgCoarsest || (int)Ls.size() >= kMgMaxLevels) break;
    std::vector<double> cg[3];
    for(int a = 0; a < 3; ++a) cg[a] = grid[a].size() > (size_t)kMgCoarsest ? coarser(grid[a]) : grid[a];
    std::vector<double> pts;
    pts.reserve(3 * grid[0].size() * grid[1].size() * grid[2].size());
    for(double z : grid[2])
      for(double y : grid[1])
        for(double x : grid[0]) {
          pts.push_back(x);
          pts.push_back(y);
          pts.push_back(z);
        }
    Ps.push_back(trilinear(pts, cg));
    for(int a = 0; a < 3; ++a) grid[a] = cg[a];
  }
  timer.mark("mg: transfers and Galerkin products");
  const int nlev = (int)Ls.size();
  const int64_t ncoarse = Ls.back().rows;
  if(ncoarse > kMgInvMax) return fail(ctx, GFEM_ERR_UNSUPPORTED, "mg_setup: coarsest level too large");

  MgSystem *m = new MgSystem();
  m->nlev = nlev;
  m->nterms = nt;
  m->shift = shift;
  m->beta = beta;
  for(int t = 0; t < nt; ++t) {  // a class's coefficient: its first member's
    m->term_group[t] = cls[t][0].first;
    m->term_index[t] = cls[t][0].second;
  }
  m->lsmall = nlev - 1;
  for(int l = nlev - 2; l >= 0; --l)
    if(Ls[l].rows <= kMgSmall) m->lsmall = l;
  hipError_t e = hipSuccess;
  for(int l = 0; l < nlev && e == hipSuccess; ++l) {
    const MultiCsr &L = Ls[l];
    m->n[l] = (int)L.rows;
    m->nnz[l] = (int64_t)L.idx.size();
    if(l > 0 && l < nlev - 1) {
      std::vector<int32_t> ai(L.ptr.begin(), L.ptr.end());
      std::vector<float2> g((size_t)nt * L.idx.size());
      const int64_t lnnz = (int64_t)L.idx.size();
      host_parallel(lnnz, [&](int64_t b, int64_t end) {
        for(int t = 0; t < nt; ++t)
          for(int64_t k = b; k < end; ++k)
            g[(size_t)t * lnnz + k] = make_float2(L.v[k * nt + t].re, L.v[k * nt + t].im);
      });
      e = up(&m->ai[l], ai.data(), ai.size());
      if(e == hipSuccess) e = up(&m->aj[l], L.idx.data(), L.idx.size());
      if(e == hipSuccess) e = up(&m->g[l], g.data(), g.size());
    }
    if(l < nlev - 1 && e == hipSuccess) {
      const Csr &P = Ps[l];
      const Csr Pt = transpose(P);
      std::vector<int32_t> pi(P.ptr.begin(), P.ptr.end()), ri(Pt.ptr.begin(), Pt.ptr.end());
      std::vector<float> pw(P.w.begin(), P.w.end()), rw(Pt.w.begin(), Pt.w.end());
      e = up(&m->pi[l], pi.data(), pi.size());
      if(e == hipSuccess) e = up(&m->pj[l], P.idx.data(), P.idx.size());
      if(e == hipSuccess) e = up(&m->pw[l], pw.data(), pw.size());
      if(e == hipSuccess) e = up(&m->ri[l], ri.data(), ri.size());
      if(e == hipSuccess) e = up(&m->rj[l], Pt.idx.data(), Pt.idx.size());
      if(e == hipSuccess) e = up(&m->rw[l], rw.data(), rw.size());
    }
  }
  if(e == hipSuccess) {
    // coarsest level dense, per term
    const MultiCsr &L = Ls.back();
    const int64_t nc = L.rows;
    std::vector<float2> gd((size_t)(nt * nc * nc), make_float2(0.f, 0.f));
    for(int64_t r = 0; r < nc; ++r)
      for(int64_t k = L.ptr[r]; k < L.ptr[r + 1]; ++k)
        for(int t = 0; t < nt; ++t)
          gd[((size_t)t * nc + r) * nc + L.idx[k]] =
            make_float2(L.v[k * nt + t].re, L.v[k * nt + t].im);
    e = up(&m->gdense, gd.data(), gd.size());
  }
  if(e == hipSuccess) {
    // the shifted term's unit matrix on the fine level
    m->xs = xs_dev.p;
    xs_dev.p = nullptr;
  }
  if(e != hipSuccess) {
    mg_free(m);
    return hip_fail(ctx, e, "mg_setup: upload");
  }
  mg_free(sys->mg);
  sys->mg = m;
  timer.mark("mg: upload");
  return GFEM_OK;
}

extern "C" gfem_status gfem_mg_info(const gfem_system *sys, int *levels, int *rows, int64_t *nnz)
{
  if(!sys || !levels) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "mg_info");
  *levels = sys->mg ? sys->mg->nlev : 0;
  if(sys->mg)
    for(int l = 0; l < sys->mg->nlev; ++l) {
      if(rows) rows[l] = sys->mg->n[l];
      if(nnz) nnz[l] = sys->mg->nnz[l];
    }
  return GFEM_OK;
}
