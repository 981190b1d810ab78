// Batched solves: many assembled systems (requests) iterate together.
//
// Each request owns its matrix values (assembled with its own parameters into
// its own buffers by gfem_assemble_into), right-hand side, FP64 iterate and
// residual, FP32 Krylov basis and Hessenberg state, all carved from one
// device arena whose size is the batch's HBM budget.  A step runs one
// GMRES(m) restart cycle for every live request: each kernel launch covers
// all requests (row-parallel kernels over a block -> (slot, first row) map,
// reductions over a (block, slot) grid), requests advance in lock-step
// through the cycle and drop out when their residual estimate is reached.
// After the cycle each request takes its FP64 residual; the converged ones
// complete and free their slot.  Requests join between steps (continuous
// batching).  A request's arithmetic does not depend on the others in the
// batch: its result is bitwise the same alone or in a mix.
//
// The algorithm per request is the one of solver.hip (GMRES in FP32 with FP64
// iterative refinement for A x = -b, fem-master/src/system/Solver.cpp:119-123).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "internal.h"
#include "krylov.h"
#include "mg.h"

namespace gfem {

namespace {

constexpr int kLanes = 8;                       // lanes per CSR row
constexpr int kTerms = 4;                       // terms of a lane in flight together (row kernels)
constexpr int kSmallTerms = 8;                  // the same in the one-block V-cycle kernel (same sums)
constexpr int kXcds = 8;                        // XCDs of an MI355X
constexpr int kRowsPerBlock = 256 / kLanes;      // 32 rows per 256-thread block
constexpr int kElemsPerBlock = 256;              // vector kernels: one element per thread
constexpr int kDotSpan = 1024;                   // rows per reduction block (at least)
constexpr int kPartMax = 1024;                   // reduction blocks per request (at most)
constexpr int kCheck = 5;                        // iterations between convergence checks

// device view of one request slot
struct BItem {
  int64_t n, nnz;
  const int64_t *ai;
  const int32_t *aj;
  const double2 *a64;
  const float2 *a32;
  const double2 *b;
  double2 *x, *r;
  float2 *w, *V;     // V: (m+1) rows of n
  double2 *st;       // H[(m+1) m] | cs[m] | sn[m] | g[m+1] | est | y[m+1] | nrm[2]
  double2 *part;     // (m+2) rows of kPartMax partial sums
  const double2 *dinv;  // right preconditioner: inverse diagonal of A (ones without it)
  int nparts;        // reduction blocks of this request (a function of n only)
  double2 *dc;       // delayed CGS2 state (krylov.h layout, its dot products at the end)
  double2 *part2;    // 2(m+1) rows of nparts2 partial sums of the delayed-CGS2 sweep
  int nparts2;       // blocks of kD2Span rows of this request
  // shifted-Laplacian multigrid (mg.h), or null: the Krylov vector goes
  // through one V-cycle into z before the product, and a cycle's correction
  // V y (in u) through one more
  const MgReq *mg;
  float2 *z, *u;
};

// rows per reduction block of a request of n rows: kDotSpan, or more so that
// a request has at most kPartMax blocks
inline int64_t dot_span(int64_t n) { return std::max<int64_t>(kDotSpan, (n + kPartMax - 1) / kPartMax); }

// the delayed-CGS2 sweep: one block per kD2Span rows of a request
constexpr int kD2Span = 256;
inline int64_t d2_parts(int64_t n) { return std::max<int64_t>(1, (n + kD2Span - 1) / kD2Span); }

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconjmul(double2 a, double2 b)  // conj(a) b
{
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 f2d(float2 v) { return make_double2(v.x, v.y); }

struct St {  // offsets into BItem::st
  int m;
  __device__ int64_t H(int j, int i) const { return (int64_t)j * (m + 1) + i; }
  __device__ int64_t cs() const { return (int64_t)(m + 1) * m; }
  __device__ int64_t sn() const { return cs() + m; }
  __device__ int64_t g() const { return sn() + m; }
  __device__ int64_t est() const { return g() + m + 1; }
  __device__ int64_t y() const { return est() + 1; }
  __device__ int64_t nrm() const { return y() + m + 1; }
};

__device__ __forceinline__ double2 lanes_sum(double2 v)
{
#pragma unroll
  for(int d = 1; d < kLanes; d <<= 1) {
    v.x += __shfl_xor(v.x, d, kLanes);
    v.y += __shfl_xor(v.y, d, kLanes);
  }
  return v;
}

// w = A32 V_j (FP32 products, lanes combined in a fixed butterfly)
__global__ void __launch_bounds__(256) bk_spmv32(const BItem *__restrict__ items, const int *__restrict__ act,
                                                 const int2 *__restrict__ rblk, int nrblk, int j,
                                                 const int *__restrict__ dj)
{
  if(dj) j = *dj;
  // each XCD takes a contiguous run of the row blocks (the grid is a
  // multiple of kXcds): its L2 holds one segment of each request's x
  const int per = (int)(gridDim.x / kXcds);
  const int b = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if(b >= nrblk) return;
  const int2 bb = rblk[b];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const int64_t row = (int64_t)bb.y + threadIdx.x / kLanes;
  const int l = threadIdx.x % kLanes;
  if(row >= I.n) return;
  const float2 *x = I.mg ? I.z : I.V + (int64_t)j * I.n;
  float sr = 0.f, si = 0.f;
  // kU terms of the lane at a time (columns and values, then the x entries,
  // in flight together), summed in term order
  constexpr int kU = kTerms;
  const int64_t k1 = I.ai[row + 1];
  for(int64_t k0 = I.ai[row] + l; k0 < k1; k0 += kU * kLanes) {
    int32_t c[kU];
    float2 a[kU], v[kU];
#pragma unroll
    for(int u = 0; u < kU; ++u) {
      const bool in = k0 + u * kLanes < k1;
      c[u] = in ? I.aj[k0 + u * kLanes] : 0;
      a[u] = in ? I.a32[k0 + u * kLanes] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for(int u = 0; u < kU; ++u) v[u] = k0 + u * kLanes < k1 ? x[c[u]] : make_float2(0.f, 0.f);
#pragma unroll
    for(int u = 0; u < kU; ++u)
      if(k0 + u * kLanes < k1) {
        sr += a[u].x * v[u].x - a[u].y * v[u].y;
        si += a[u].x * v[u].y + a[u].y * v[u].x;
      }
  }
  const double2 s = lanes_sum(make_double2(sr, si));
  if(l == 0) I.w[row] = make_float2((float)s.x, (float)s.y);
}

// r = -b - A64 x (FP64)
__global__ void __launch_bounds__(256) bk_residual(const BItem *__restrict__ items, const int *__restrict__ act,
                                                   const int2 *__restrict__ rblk)
{
  const int2 bb = rblk[blockIdx.x];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const int64_t row = (int64_t)bb.y + threadIdx.x / kLanes;
  const int l = threadIdx.x % kLanes;
  if(row >= I.n) return;
  double2 s = make_double2(0., 0.);
  // kU terms of the lane at a time, summed in term order
  constexpr int kU = kTerms;
  const int64_t k1 = I.ai[row + 1];
  for(int64_t k0 = I.ai[row] + l; k0 < k1; k0 += kU * kLanes) {
    int32_t c[kU];
    double2 a[kU], v[kU];
#pragma unroll
    for(int u = 0; u < kU; ++u) {
      const bool in = k0 + u * kLanes < k1;
      c[u] = in ? I.aj[k0 + u * kLanes] : 0;
      a[u] = in ? I.a64[k0 + u * kLanes] : make_double2(0., 0.);
    }
#pragma unroll
    for(int u = 0; u < kU; ++u) v[u] = k0 + u * kLanes < k1 ? I.x[c[u]] : make_double2(0., 0.);
#pragma unroll
    for(int u = 0; u < kU; ++u)
      if(k0 + u * kLanes < k1) {
        const double2 p = zmul(a[u], v[u]);
        s.x += p.x;
        s.y += p.y;
      }
  }
  s = lanes_sum(s);
  if(l == 0) I.r[row] = make_double2(-I.b[row].x - s.x, -I.b[row].y - s.y);
}

// Partial inner products, one block per (request, row range) of the dot map
// (slot, part, first row, end row): each request has nparts blocks in
// proportion to its size, so large and small requests finish together.
// mode:
//   0: conj(V_k) . w, k <= j     1: w . w     2: r . r     3: b . b
// Mode 0 takes the basis vectors kGroup at a time, so w is read once per
// group.
constexpr int kGroup = 8;

__global__ void __launch_bounds__(256) bk_dots(const BItem *__restrict__ items, const int *__restrict__ act,
                                               const int4 *__restrict__ dblk, int mode, int j,
                                               const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int4 db = dblk[blockIdx.x];
  if(!act[db.x]) return;
  const BItem &I = items[db.x];
  const int nk = mode == 0 ? j + 1 : 1;
  __shared__ double2 s_red[4][kGroup];
  for(int k0 = 0; k0 < nk; k0 += kGroup) {
    const int kg = min(kGroup, nk - k0);
    double2 acc[kGroup];
#pragma unroll
    for(int q = 0; q < kGroup; ++q) acc[q] = make_double2(0., 0.);
    for(int64_t i = (int64_t)db.z + threadIdx.x; i < (int64_t)db.w; i += blockDim.x) {
      if(mode == 0) {
        const double2 wv = f2d(I.w[i]);
#pragma unroll
        for(int q = 0; q < kGroup; ++q)
          if(q < kg) {
            const double2 p = zconjmul(f2d(I.V[(int64_t)(k0 + q) * I.n + i]), wv);
            acc[q].x += p.x;
            acc[q].y += p.y;
          }
      }
      else {
        const double2 v = mode == 1 ? f2d(I.w[i]) : (mode == 2 ? I.r[i] : I.b[i]);
        acc[0].x += v.x * v.x + v.y * v.y;
      }
    }
#pragma unroll
    for(int q = 0; q < kGroup; ++q) {
      if(q >= kg) break;
#pragma unroll
      for(int d = 32; d >= 1; d >>= 1) {
        acc[q].x += __shfl_xor(acc[q].x, d, 64);
        acc[q].y += __shfl_xor(acc[q].y, d, 64);
      }
      if((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6][q] = acc[q];
    }
    __syncthreads();
    if((int)threadIdx.x < kg) {
      double2 t = make_double2(0., 0.);
      for(int w = 0; w < 4; ++w) {
        t.x += s_red[w][threadIdx.x].x;
        t.y += s_red[w][threadIdx.x].y;
      }
      I.part[(int64_t)(k0 + threadIdx.x) * kPartMax + db.y] = t;
    }
    __syncthreads();
  }
}

// Sum the partials (fixed order), grid (m+2, slots).  Destination by mode:
//   0: H column j (first pass, pass = 0) or y (second pass, pass = 1)
//   1: H[j+1, j] = |w|^2     2: nrm[0] = |r|^2     3: nrm[1] = |b|^2
__global__ void __launch_bounds__(64) bk_dots_final(const BItem *__restrict__ items, const int *__restrict__ act,
                                                    int m, int mode, int j, int pass, const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int it = blockIdx.y, k = blockIdx.x;
  if(!act[it]) return;
  const int nk = mode == 0 ? j + 1 : 1;
  if(k >= nk) return;
  const BItem &I = items[it];
  double2 s = make_double2(0., 0.);
  for(int p = threadIdx.x; p < I.nparts; p += 64) {
    const double2 v = I.part[(int64_t)k * kPartMax + p];
    s.x += v.x;
    s.y += v.y;
  }
#pragma unroll
  for(int d = 32; d >= 1; d >>= 1) {
    s.x += __shfl_xor(s.x, d, 64);
    s.y += __shfl_xor(s.y, d, 64);
  }
  if(threadIdx.x != 0) return;
  const St o{m};
  double2 *st = I.st;
  if(mode == 0)
    st[pass == 0 ? o.H(j, k) : o.y() + k] = s;
  else if(mode == 1)
    st[o.H(j, j + 1)] = s;
  else
    st[o.nrm() + (mode == 2 ? 0 : 1)] = s;
}

// w -= sum_{k<=j} V_k h_k with h = H column j (pass 0) or y (pass 1)
__global__ void __launch_bounds__(256) bk_orth(const BItem *__restrict__ items, const int *__restrict__ act,
                                               const int2 *__restrict__ vblk, int m, int j, int pass,
                                               const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int2 bb = vblk[blockIdx.x];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const int64_t i = (int64_t)bb.y + threadIdx.x;
  if(i >= I.n) return;
  const St o{m};
  const double2 *h = I.st + (pass == 0 ? o.H(j, 0) : o.y());
  double2 s = f2d(I.w[i]);
  for(int k = 0; k <= j; ++k) {
    const double2 p = zmul(f2d(I.V[(int64_t)k * I.n + i]), h[k]);
    s.x -= p.x;
    s.y -= p.y;
  }
  I.w[i] = make_float2((float)s.x, (float)s.y);
}

// Delayed CGS2 (krylov.h), one global reduction per Krylov step.
// Block partials of conj(V_k) . u and conj(V_k) . w, k <= j, u = V_j, over
// the block's kD2Span rows of its request (map: slot, part, first row, end
// row); the four waves take the basis vectors round robin, u and w stay in
// L1.  tail: w = u (the cycle's last column needs V_k^H u only).
__global__ void __launch_bounds__(256) bk_dots2(const BItem *__restrict__ items, const int *__restrict__ act,
                                                const int4 *__restrict__ d2blk, int j, const int *__restrict__ dj,
                                                int tail)
{
  if(dj) j = *dj;
  const int4 db = d2blk[blockIdx.x];
  if(!act[db.x]) return;
  const BItem &I = items[db.x];
  const float2 *u = I.V + (int64_t)j * I.n;
  const float2 *w = tail ? u : I.w;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for(int k = wv; k <= j; k += 4) {
    const float2 *vk = I.V + (int64_t)k * I.n;
    double2 au = make_double2(0., 0.), aw = make_double2(0., 0.);
#pragma unroll 4
    for(int64_t i = (int64_t)db.z + lane; i < (int64_t)db.w; i += 64) {
      const double2 v = f2d(vk[i]);
      const double2 pa = zconjmul(v, f2d(u[i])), pb = zconjmul(v, f2d(w[i]));
      au.x += pa.x;
      au.y += pa.y;
      aw.x += pb.x;
      aw.y += pb.y;
    }
#pragma unroll
    for(int d = 32; d >= 1; d >>= 1) {
      au.x += __shfl_xor(au.x, d, 64);
      au.y += __shfl_xor(au.y, d, 64);
      aw.x += __shfl_xor(aw.x, d, 64);
      aw.y += __shfl_xor(aw.y, d, 64);
    }
    if(lane == 0) {
      I.part2[(int64_t)(2 * k) * I.nparts2 + db.y] = au;
      I.part2[(int64_t)(2 * k + 1) * I.nparts2 + db.y] = aw;
    }
  }
}

// the sweep's 2(j+1) sums of each request in a fixed order; grid (2(m+1), slots)
__global__ void __launch_bounds__(64) bk_dots2_final(const BItem *__restrict__ items, const int *__restrict__ act,
                                                     int m, int j, const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int it = blockIdx.y, o = blockIdx.x;
  if(!act[it] || o >= 2 * (j + 1)) return;
  const BItem &I = items[it];
  double2 s = make_double2(0., 0.);
  for(int p = threadIdx.x; p < I.nparts2; p += 64) {
    const double2 v = I.part2[(int64_t)o * I.nparts2 + p];
    s.x += v.x;
    s.y += v.y;
  }
#pragma unroll
  for(int d = 32; d >= 1; d >>= 1) {
    s.x += __shfl_xor(s.x, d, 64);
    s.y += __shfl_xor(s.y, d, 64);
  }
  if(threadIdx.x == 0) I.dc[krylov::dcgs_dots_offset(m) + o] = s;
}

// the bookkeeping of step j per request (krylov::dcgs_coef): columns finished
// -> kcnt; grid (slots), 64 threads
__global__ void __launch_bounds__(64) bk_dcgs_coef(const BItem *__restrict__ items, const int *__restrict__ act,
                                                   int *__restrict__ kcnt, int m, int j, const int *__restrict__ dj,
                                                   int tail)
{
  if(dj) j = *dj;
  const int it = blockIdx.x;
  if(!act[it]) return;
  const BItem &I = items[it];
  krylov::dcgs_coef(m, j, tail, I.dc, I.st);
  if(threadIdx.x == 0) kcnt[it] = tail ? m : j;
}

// q_j = (u - sum_{k<j} V_k s_k) / r into V_j, u_{j+1} = (w - sum_{k<j} V_k c_k
// - q_j c_j) / r into V_{j+1}: one read of the basis (map: slot, first row)
__global__ void __launch_bounds__(256) bk_dcgs_update(const BItem *__restrict__ items, const int *__restrict__ act,
                                                      const int2 *__restrict__ vblk, int m, int j,
                                                      const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int2 bb = vblk[blockIdx.x];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const int64_t i = (int64_t)bb.y + threadIdx.x;
  if(i >= I.n) return;
  const double2 *sv = I.dc + krylov::dcgs_s_offset(m), *cv = I.dc + krylov::dcgs_c_offset(m);
  const double inv = I.dc[krylov::dcgs_r_offset(m)].x;
  double2 q = f2d(I.V[(int64_t)j * I.n + i]), un = f2d(I.w[i]);
  for(int k = 0; k < j; ++k) {
    const double2 vk = f2d(I.V[(int64_t)k * I.n + i]);
    const double2 a = zmul(vk, sv[k]), b = zmul(vk, cv[k]);
    q.x -= a.x;
    q.y -= a.y;
    un.x -= b.x;
    un.y -= b.y;
  }
  q = make_double2(q.x * inv, q.y * inv);
  const double2 qc = zmul(q, cv[j]);
  un = make_double2((un.x - qc.x) * inv, (un.y - qc.y) * inv);
  I.V[(int64_t)j * I.n + i] = make_float2((float)q.x, (float)q.y);
  I.V[(int64_t)(j + 1) * I.n + i] = make_float2((float)un.x, (float)un.y);
}

// H column j += y (second Gram-Schmidt pass), grid (slots)
__global__ void bk_add_y(const BItem *__restrict__ items, const int *__restrict__ act, int m, int j,
                         const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int it = blockIdx.x;
  if(!act[it]) return;
  const St o{m};
  double2 *st = items[it].st;
  for(int k = threadIdx.x; k <= j; k += blockDim.x) {
    st[o.H(j, k)].x += st[o.y() + k].x;
    st[o.H(j, k)].y += st[o.y() + k].y;
  }
}

// V_{j+1} = w / |w| (pass j >= 0) or V_0 = r / |r| (j < 0)
__global__ void __launch_bounds__(256) bk_scale(const BItem *__restrict__ items, const int *__restrict__ act,
                                                const int2 *__restrict__ vblk, int m, int j, const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int2 bb = vblk[blockIdx.x];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const int64_t i = (int64_t)bb.y + threadIdx.x;
  if(i >= I.n) return;
  const St o{m};
  const double n2 = j < 0 ? I.st[o.nrm()].x : I.st[o.H(j, j + 1)].x;
  const double nv = sqrt(n2);
  const double inv = nv > 0. ? 1. / nv : 0.;
  const double2 v = j < 0 ? I.r[i] : f2d(I.w[i]);
  I.V[(int64_t)(j + 1) * I.n + i] = make_float2((float)(v.x * inv), (float)(v.y * inv));
}

// g = [|r|, 0, ...], cycle iteration count 0; grid (slots)
__global__ void bk_cycle_start(const BItem *__restrict__ items, const int *__restrict__ act, int *__restrict__ kcnt,
                               int m)
{
  const int it = blockIdx.x;
  if(!act[it]) return;
  const St o{m};
  double2 *st = items[it].st;
  for(int i = threadIdx.x; i <= m; i += blockDim.x)
    st[o.g() + i] = make_double2(i == 0 ? sqrt(st[o.nrm()].x) : 0., 0.);
  if(threadIdx.x == 0) kcnt[it] = 0;
}

// Givens rotations of column j, residual estimate |g_{j+1}|; grid (slots)
__global__ void bk_givens(const BItem *__restrict__ items, const int *__restrict__ act, int *__restrict__ kcnt, int m,
                          int j, const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int it = blockIdx.x;
  if(!act[it] || threadIdx.x != 0) return;
  const St o{m};
  double2 *st = items[it].st;
  double2 *col = st + o.H(j, 0);
  double2 *cs = st + o.cs(), *sn = st + o.sn(), *g = st + o.g();
  col[j + 1] = make_double2(sqrt(col[j + 1].x), 0.);
  for(int i = 0; i < j; ++i) {
    const double c = cs[i].x;
    const double2 s = sn[i];
    const double2 a = col[i], b = col[i + 1];
    const double2 sb = zmul(s, b), csa = zconjmul(s, a);
    col[i] = make_double2(c * a.x + sb.x, c * a.y + sb.y);
    col[i + 1] = make_double2(-csa.x + c * b.x, -csa.y + c * b.y);
  }
  const double2 a = col[j], b = col[j + 1];
  const double na = hypot(a.x, a.y), nb = hypot(b.x, b.y);
  double c;
  double2 s;
  if(nb == 0.) {
    c = 1.;
    s = make_double2(0., 0.);
  }
  else if(na == 0.) {
    c = 0.;
    s = make_double2(b.x / nb, -b.y / nb);
  }
  else {
    const double t = hypot(na, nb);
    c = na / t;
    const double2 ua = make_double2(a.x / na, a.y / na);
    s = make_double2((ua.x * b.x + ua.y * b.y) / t, (ua.y * b.x - ua.x * b.y) / t);
  }
  cs[j] = make_double2(c, 0.);
  sn[j] = s;
  const double2 sb = zmul(s, b);
  col[j] = make_double2(c * a.x + sb.x, c * a.y + sb.y);
  col[j + 1] = make_double2(0., 0.);
  const double2 gj = g[j], csg = zconjmul(s, gj);
  g[j] = make_double2(c * gj.x, c * gj.y);
  g[j + 1] = make_double2(-csg.x, -csg.y);
  st[o.est()] = make_double2(hypot(g[j + 1].x, g[j + 1].y), 0.);
  kcnt[it] = j + 1;
}

// y = R^-1 g over the request's iterations of this cycle; grid (slots)
__global__ void bk_backsolve(const BItem *__restrict__ items, const int *__restrict__ live,
                             const int *__restrict__ kcnt, int m)
{
  const int it = blockIdx.x;
  if(!live[it] || threadIdx.x != 0) return;
  const St o{m};
  double2 *st = items[it].st;
  const int k = kcnt[it];
  for(int i = k - 1; i >= 0; --i) {
    double2 s = st[o.g() + i];
    for(int l = i + 1; l < k; ++l) {
      const double2 p = zmul(st[o.H(l, i)], st[o.y() + l]);
      s.x -= p.x;
      s.y -= p.y;
    }
    const double2 d = st[o.H(i, i)];
    const double dd = d.x * d.x + d.y * d.y;
    st[o.y() + i] = dd > 0. ? make_double2((s.x * d.x + s.y * d.y) / dd, (s.y * d.x - s.x * d.y) / dd)
                            : make_double2(0., 0.);
  }
}

// x += sum_k V_k y_k (FP64)
__global__ void __launch_bounds__(256) bk_update_x(const BItem *__restrict__ items, const int *__restrict__ live,
                                                   const int *__restrict__ kcnt, const int2 *__restrict__ vblk, int m)
{
  const int2 bb = vblk[blockIdx.x];
  if(!live[bb.x]) return;
  const BItem &I = items[bb.x];
  const int64_t i = (int64_t)bb.y + threadIdx.x;
  if(i >= I.n) return;
  const St o{m};
  const int k = kcnt[bb.x];
  double2 s = make_double2(0., 0.);
  for(int l = 0; l < k; ++l) {
    const double2 p = zmul(f2d(I.V[(int64_t)l * I.n + i]), I.st[o.y() + l]);
    s.x += p.x;
    s.y += p.y;
  }
  if(I.mg) {  // u = V y, then x += M^-1 u (bk_mg_addz)
    I.u[i] = make_float2((float)s.x, (float)s.y);
    return;
  }
  const double2 d = zmul(I.dinv[i], s);  // x += D^-1 V y (right preconditioning)
  I.x[i] = make_double2(I.x[i].x + d.x, I.x[i].y + d.y);
}

// ------------------------------------------------------------ multigrid V-cycle
// One V-cycle of every active request with a hierarchy (mg.h): levels below
// lsmall row-parallel over per-level block maps (slot, first row), the rest
// of the cycle in one block per request.  Level 0's right-hand side is the
// Krylov vector V_j (src 0) or the cycle's correction u (src 1); its result
// goes to z.  The same operations in the same order whatever else is in the
// batch.

__device__ __forceinline__ float2 cmulf(float2 a, float2 b)
{
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ const float2 *mg_rhs(const BItem &I, int l, int src, int j)
{
  return l > 0 ? I.mg->lev[l].b : (src ? I.u : I.V + (int64_t)j * I.n);
}

__device__ __forceinline__ int64_t mg_row0(const MgLevelReq &L, int64_t r)
{
  return L.ai64 ? L.ai64[r] : (int64_t)L.ai[r];
}

// sum over row r of S times v (v = D b when DB: the pre-smoothed iterate
// computed on the fly), lanes of kLanes terms apart starting at `lane`; kU
// terms of the lane at a time (columns and values, then the v and D entries,
// in flight together), summed in term order
template <bool DB, int kU = kTerms>
__device__ __forceinline__ float2 mg_row_dot(const MgLevelReq &L, int64_t r, const float2 *v, int lane)
{
  float sr = 0.f, si = 0.f;
  const int64_t k1 = mg_row0(L, r + 1);
  for(int64_t k0 = mg_row0(L, r) + lane; k0 < k1; k0 += kU * kLanes) {
    int32_t c[kU];
    float2 a[kU], x[kU], d[kU];
#pragma unroll
    for(int u = 0; u < kU; ++u) {
      const bool in = k0 + u * kLanes < k1;
      c[u] = in ? L.aj[k0 + u * kLanes] : 0;
      a[u] = in ? L.a[k0 + u * kLanes] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for(int u = 0; u < kU; ++u) {
      const bool in = k0 + u * kLanes < k1;
      x[u] = in ? v[c[u]] : make_float2(0.f, 0.f);
      if(DB) d[u] = in ? L.dinv[c[u]] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for(int u = 0; u < kU; ++u)
      if(k0 + u * kLanes < k1) {
        const float2 p = cmulf(a[u], DB ? cmulf(d[u], x[u]) : x[u]);
        sr += p.x;
        si += p.y;
      }
  }
  return make_float2(sr, si);
}

// sum over row r of a transfer operator (P or its transpose: pointers, columns,
// real weights) times v, lanes of kLanes terms apart starting at `lane`,
// kU terms of the lane at a time, summed in term order (the sum does not
// depend on kU)
template <int kU = kTerms>
__device__ __forceinline__ float2 mg_transfer_dot(const int32_t *ptr, const int32_t *idx, const float *w,
                                                  const float2 *v, int64_t r, int lane)
{
  float sr = 0.f, si = 0.f;
  const int32_t k1 = ptr[r + 1];
  for(int32_t k0 = ptr[r] + lane; k0 < k1; k0 += kU * kLanes) {
    int32_t c[kU];
    float wt[kU];
    float2 x[kU];
#pragma unroll
    for(int u = 0; u < kU; ++u) {
      const bool in = k0 + u * kLanes < k1;
      c[u] = in ? idx[k0 + u * kLanes] : 0;
      wt[u] = in ? w[k0 + u * kLanes] : 0.f;
    }
#pragma unroll
    for(int u = 0; u < kU; ++u) x[u] = k0 + u * kLanes < k1 ? v[c[u]] : make_float2(0.f, 0.f);
#pragma unroll
    for(int u = 0; u < kU; ++u)
      if(k0 + u * kLanes < k1) {
        sr += wt[u] * x[u].x;
        si += wt[u] * x[u].y;
      }
  }
  return make_float2(sr, si);
}

// block b of a per-level map: each XCD takes a contiguous run of the blocks
// (the grid is a multiple of kXcds), so its L2 holds one segment of each
// request's vectors; false past the map
__device__ __forceinline__ bool mg_block(int nblk, int &b)
{
  const int per = (int)(gridDim.x / kXcds);
  b = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  return b < nblk;
}

__device__ __forceinline__ float2 lanes_sumf(float2 v)
{
  const double2 d = lanes_sum(make_double2(v.x, v.y));
  return make_float2((float)d.x, (float)d.y);
}

// level 0's pre-smoothed iterate x = D b, one thread per row (requests that
// run level 0 row-parallel), so that bk_mg_down gathers x alone
__global__ void __launch_bounds__(256) bk_mg_x0(const BItem *__restrict__ items, const int *__restrict__ act,
                                                const int2 *__restrict__ vblk, int src, int j,
                                                const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int2 bb = vblk[blockIdx.x];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  if(!I.mg || I.mg->lsmall == 0) return;
  const int64_t i = (int64_t)bb.y + threadIdx.x;
  if(i >= I.n) return;
  const MgLevelReq &L = I.mg->lev[0];
  L.x[i] = cmulf(L.dinv[i], mg_rhs(I, 0, src, j)[i]);
}

// x = D b, r = b - S x (level 0: x formed by bk_mg_x0)
__global__ void __launch_bounds__(256) bk_mg_down(const BItem *__restrict__ items, const int *__restrict__ act,
                                                  const int2 *__restrict__ blk, int nblk, int l, int src, int j,
                                                  const int *__restrict__ dj)
{
  if(dj) j = *dj;
  int bi;
  if(!mg_block(nblk, bi)) return;
  const int2 bb = blk[bi];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const MgLevelReq &L = I.mg->lev[l];
  const int64_t row = (int64_t)bb.y + threadIdx.x / kLanes;
  const int lane = threadIdx.x % kLanes;
  if(row >= L.n) return;
  const float2 *b = mg_rhs(I, l, src, j);
  if(l == 0) {  // one gather per term
    const float2 s = lanes_sumf(mg_row_dot<false>(L, row, L.x, lane));
    if(lane == 0) L.r[row] = make_float2(b[row].x - s.x, b[row].y - s.y);
    return;
  }
  const float2 s = lanes_sumf(mg_row_dot<true>(L, row, b, lane));
  if(lane == 0) {
    const float2 br = b[row];
    L.x[row] = cmulf(L.dinv[row], br);
    L.r[row] = make_float2(br.x - s.x, br.y - s.y);
  }
}

// b_{l+1} = P^T r_l, rows of level l + 1
__global__ void __launch_bounds__(256) bk_mg_restrict(const BItem *__restrict__ items, const int *__restrict__ act,
                                                      const int2 *__restrict__ blk, int nblk, int l)
{
  int bi;
  if(!mg_block(nblk, bi)) return;
  const int2 bb = blk[bi];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const MgLevelReq &L = I.mg->lev[l], &C = I.mg->lev[l + 1];
  const int64_t row = (int64_t)bb.y + threadIdx.x / kLanes;
  const int lane = threadIdx.x % kLanes;
  if(row >= C.n) return;
  const float2 s = lanes_sumf(mg_transfer_dot(L.ri, L.rj, L.rw, L.r, row, lane));
  if(lane == 0) C.b[row] = s;
}

__device__ __forceinline__ const float2 *mg_coarse_x(const MgReq &M, int l)
{
  return l + 1 == M.nlev - 1 ? M.lev[l + 1].x : M.lev[l + 1].y;
}

// x_l += P xhat_{l+1}
__global__ void __launch_bounds__(256) bk_mg_prolong(const BItem *__restrict__ items, const int *__restrict__ act,
                                                     const int2 *__restrict__ blk, int nblk, int l)
{
  int bi;
  if(!mg_block(nblk, bi)) return;
  const int2 bb = blk[bi];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const MgLevelReq &L = I.mg->lev[l];
  const int64_t row = (int64_t)bb.y + threadIdx.x / kLanes;
  const int lane = threadIdx.x % kLanes;
  if(row >= L.n) return;
  const float2 *xc = mg_coarse_x(*I.mg, l);
  const float2 s = lanes_sumf(mg_transfer_dot(L.pi, L.pj, L.pw, xc, row, lane));
  if(lane == 0) L.x[row] = make_float2(L.x[row].x + s.x, L.x[row].y + s.y);
}

// y_l = x_l + D (b_l - S x_l); level 0 writes z
__global__ void __launch_bounds__(256) bk_mg_post(const BItem *__restrict__ items, const int *__restrict__ act,
                                                  const int2 *__restrict__ blk, int nblk, int l, int src, int j,
                                                  const int *__restrict__ dj)
{
  if(dj) j = *dj;
  int bi;
  if(!mg_block(nblk, bi)) return;
  const int2 bb = blk[bi];
  if(!act[bb.x]) return;
  const BItem &I = items[bb.x];
  const MgLevelReq &L = I.mg->lev[l];
  const int64_t row = (int64_t)bb.y + threadIdx.x / kLanes;
  const int lane = threadIdx.x % kLanes;
  if(row >= L.n) return;
  const float2 *b = mg_rhs(I, l, src, j);
  const float2 s = lanes_sumf(mg_row_dot<false>(L, row, L.x, lane));
  if(lane == 0) {
    const float2 br = b[row], xi = L.x[row];
    const float2 d = cmulf(L.dinv[row], make_float2(br.x - s.x, br.y - s.y));
    float2 *out = l == 0 ? I.z : L.y;
    out[row] = make_float2(xi.x + d.x, xi.y + d.y);
  }
}

// the levels from lsmall down, the coarsest exactly, and back up, in one
// block per request: rows kLanes lanes each, as the row-parallel kernels
// (the same sums in the same order), stages separated by block barriers
__global__ void __launch_bounds__(kMgSmallThreads) bk_mg_small(const BItem *__restrict__ items,
                                                              const int *__restrict__ act, int src, int j,
                                                              const int *__restrict__ dj)
{
  if(dj) j = *dj;
  const int it = blockIdx.x;
  if(!act[it] || !items[it].mg) return;
  const BItem &I = items[it];
  const MgReq &M = *I.mg;
  const int lane = threadIdx.x % kLanes;
  const int r0 = threadIdx.x / kLanes;
  constexpr int kRows = kMgSmallThreads / kLanes;  // rows per pass
  for(int l = M.lsmall; l < M.nlev - 1; ++l) {
    const MgLevelReq &L = M.lev[l], &C = M.lev[l + 1];
    const float2 *b = mg_rhs(I, l, src, j);
    for(int64_t base = 0; base < L.n; base += kRows) {
      const int64_t r = base + r0;
      const float2 s = lanes_sumf(r < L.n ? mg_row_dot<true, kSmallTerms>(L, r, b, lane) : make_float2(0.f, 0.f));
      if(r < L.n && lane == 0) {
        L.x[r] = cmulf(L.dinv[r], b[r]);
        L.r[r] = make_float2(b[r].x - s.x, b[r].y - s.y);
      }
    }
    __syncthreads();
    for(int64_t base = 0; base < C.n; base += kRows) {
      const int64_t r = base + r0;
      const float2 s = lanes_sumf(r < C.n ? mg_transfer_dot<kSmallTerms>(L.ri, L.rj, L.rw, L.r, r, lane)
                                          : make_float2(0.f, 0.f));
      if(r < C.n && lane == 0) C.b[r] = s;
    }
    __syncthreads();
  }
  {
    // coarsest: x = S^-1 b with the inverse column-major; each row's sum in
    // kCg column groups (columns g, g + kCg, ...) whose loads are in flight
    // together, the groups' partial sums added in group order
    const MgLevelReq &C = M.lev[M.nlev - 1];
    const int n = C.n;
    constexpr int kRowSlots = 128, kCg = kMgSmallThreads / kRowSlots;
    static_assert(kMgCoarsest * kMgCoarsest * kMgCoarsest <= kRowSlots, "coarsest rows per column group");
    __shared__ float2 s_part[kMgSmallThreads];
    const int r = threadIdx.x % kRowSlots, g = threadIdx.x / kRowSlots;
    float sr = 0.f, si = 0.f;
    if(r < n) {
#pragma unroll 4
      for(int c = g; c < n; c += kCg) {
        const float2 p = cmulf(M.ainv[(int64_t)c * n + r], C.b[c]);
        sr += p.x;
        si += p.y;
      }
    }
    s_part[threadIdx.x] = make_float2(sr, si);
    __syncthreads();
    if(g == 0 && r < n) {
      float2 t = s_part[r];
      for(int q = 1; q < kCg; ++q) {
        t.x += s_part[q * kRowSlots + r].x;
        t.y += s_part[q * kRowSlots + r].y;
      }
      C.x[r] = t;
    }
    __syncthreads();
  }
  for(int l = M.nlev - 2; l >= M.lsmall; --l) {
    const MgLevelReq &L = M.lev[l];
    const float2 *xc = mg_coarse_x(M, l);
    for(int64_t base = 0; base < L.n; base += kRows) {
      const int64_t r = base + r0;
      const float2 s = lanes_sumf(r < L.n ? mg_transfer_dot<kSmallTerms>(L.pi, L.pj, L.pw, xc, r, lane)
                                          : make_float2(0.f, 0.f));
      if(r < L.n && lane == 0) L.x[r] = make_float2(L.x[r].x + s.x, L.x[r].y + s.y);
    }
    __syncthreads();
    const float2 *b = mg_rhs(I, l, src, j);
    float2 *out = l == 0 ? I.z : L.y;
    for(int64_t base = 0; base < L.n; base += kRows) {
      const int64_t r = base + r0;
      const float2 s = lanes_sumf(r < L.n ? mg_row_dot<false, kSmallTerms>(L, r, L.x, lane) : make_float2(0.f, 0.f));
      if(r < L.n && lane == 0) {
        const float2 d = cmulf(L.dinv[r], make_float2(b[r].x - s.x, b[r].y - s.y));
        out[r] = make_float2(L.x[r].x + d.x, L.x[r].y + d.y);
      }
    }
    __syncthreads();
  }
}

// Inverse of the coarsest level's S = sum_t c_t G_t (dense, n <= kMgInvMax),
// one block per request admitted since the last step (act): in-place
// Gauss-Jordan with partial pivoting on the complex64 matrix held in LDS (row
// swaps recorded, columns swapped back at the end), written column-major so
// the cycle's coarse product reads it with consecutive lanes on consecutive
// rows.
__global__ void __launch_bounds__(kMgInvThreads) bk_mg_invert(const BItem *__restrict__ items,
                                                              const int *__restrict__ act)
{
  const int it = blockIdx.x;
  if(!act[it] || !items[it].mg) return;
  const MgReq &M = *items[it].mg;
  const int n = M.lev[M.nlev - 1].n;
  const float2 *gd = M.gdense;
  float2 *ainv = M.ainv;
  __shared__ float2 A[kMgInvMax * kMgInvMax];
  __shared__ float2 s_f[kMgInvMax];  // the multipliers of the column being eliminated
  __shared__ int s_perm[kMgInvMax];
  const int t = threadIdx.x;
  for(int e = t; e < n * n; e += kMgInvThreads) {
    float re = 0.f, im = 0.f;
    for(int k = 0; k < M.cf.nt; ++k) {
      const float2 g = gd[(int64_t)k * n * n + e], q = M.cf.c[k];
      re += q.x * g.x - q.y * g.y;
      im += q.x * g.y + q.y * g.x;
    }
    A[e] = make_float2(re, im);
  }
  __syncthreads();
  for(int c = 0; c < n; ++c) {
    // pivot: largest |A[r][c]|, r >= c (one wave, smallest index on ties)
    if(t < 64) {
      float best = -1.f;
      int bi = c;
      for(int r = c + t; r < n; r += 64) {
        const float2 v = A[r * n + c];
        const float m = v.x * v.x + v.y * v.y;
        if(m > best) {
          best = m;
          bi = r;
        }
      }
      for(int d = 32; d > 0; d >>= 1) {
        const float ob = __shfl_down(best, d, 64);
        const int oi = __shfl_down(bi, d, 64);
        if(t + d < 64 && (ob > best || (ob == best && oi < bi))) {
          best = ob;
          bi = oi;
        }
      }
      if(t == 0) s_perm[c] = bi;
    }
    __syncthreads();
    const int p = s_perm[c];
    if(p != c)
      for(int k = t; k < n; k += kMgInvThreads) {
        const float2 x = A[c * n + k];
        A[c * n + k] = A[p * n + k];
        A[p * n + k] = x;
      }
    __syncthreads();
    const float2 d = A[c * n + c];
    const float dm = d.x * d.x + d.y * d.y;
    const float2 inv = dm > 0.f ? make_float2(d.x / dm, -d.y / dm) : make_float2(0.f, 0.f);
    // the other rows' multipliers, read before any of those rows changes
    for(int r = t; r < n; r += kMgInvThreads) s_f[r] = A[r * n + c];
    __syncthreads();
    // pivot row: A[c][c] = 1, then scaled by 1 / d
    for(int k = t; k < n; k += kMgInvThreads) {
      const float2 x = k == c ? make_float2(1.f, 0.f) : A[c * n + k];
      A[c * n + k] = make_float2(x.x * inv.x - x.y * inv.y, x.x * inv.y + x.y * inv.x);
    }
    __syncthreads();
    // every other row: A[r][c] = 0, A[r][:] -= f_r A[c][:]
    for(int e = t; e < n * n; e += kMgInvThreads) {
      const int r = e / n, k = e - r * n;
      if(r == c) continue;
      const float2 f = s_f[r], y = A[c * n + k];
      const float2 x = k == c ? make_float2(0.f, 0.f) : A[e];
      A[e] = make_float2(x.x - (f.x * y.x - f.y * y.y), x.y - (f.x * y.y + f.y * y.x));
    }
    __syncthreads();
  }
  // undo the row swaps as column swaps, last first, and write column-major
  for(int c = n - 1; c >= 0; --c) {
    const int p = s_perm[c];
    if(p != c)
      for(int r = t; r < n; r += kMgInvThreads) {
        const float2 x = A[r * n + c];
        A[r * n + c] = A[r * n + p];
        A[r * n + p] = x;
      }
    __syncthreads();
  }
  for(int e = t; e < n * n; e += kMgInvThreads) {
    const int r = e / n, c = e - r * n;
    ainv[(int64_t)c * n + r] = A[e];
  }
}

// x += z (the V-cycle of the cycle's correction)
__global__ void __launch_bounds__(256) bk_mg_addz(const BItem *__restrict__ items, const int *__restrict__ live,
                                                  const int2 *__restrict__ vblk)
{
  const int2 bb = vblk[blockIdx.x];
  if(!live[bb.x]) return;
  const BItem &I = items[bb.x];
  if(!I.mg) return;
  const int64_t i = (int64_t)bb.y + threadIdx.x;
  if(i >= I.n) return;
  I.x[i] = make_double2(I.x[i].x + (double)I.z[i].x, I.x[i].y + (double)I.z[i].y);
}

// per-slot scalars to one array: est, |r|^2, |b|^2
__global__ void bk_collect(const BItem *__restrict__ items, int nslots, int m, double *__restrict__ out)
{
  const int it = blockIdx.x * blockDim.x + threadIdx.x;
  if(it >= nslots || !items[it].st) return;
  const St o{m};
  out[3 * it] = items[it].st[o.est()].x;
  out[3 * it + 1] = items[it].st[o.nrm()].x;
  out[3 * it + 2] = items[it].st[o.nrm() + 1].x;
}

// Right scaling: the Krylov iteration runs on A D^-1 with D the 2-norms of
// the rows of A (for the symmetric patterns assembled here, those of its
// columns), so the FP32 copy of the matrix is A with its columns scaled by
// D^-1 and the correction of a cycle is D^-1 V y.  Unlike the diagonal, the
// norms stay away from zero on indefinite (Helmholtz) matrices.  The FP64
// residual -b - A x of each cycle is unchanged: the solution and its
// tolerance are the unscaled ones.
__global__ void bk_diag_inv(int64_t n, const int64_t *__restrict__ ai, const int32_t *__restrict__ aj,
                            const double2 *__restrict__ a, int scale, double2 *__restrict__ dinv)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= n) return;
  double d = 1.;
  if(scale) {
    double q = 0.;
    for(int64_t k = ai[r]; k < ai[r + 1]; ++k) q += a[k].x * a[k].x + a[k].y * a[k].y;
    if(q > 0.) d = 1. / sqrt(q);
  }
  (void)aj;
  dinv[r] = make_double2(d, 0.);
}

// FP32 copy of A D^-1
__global__ void bk_round_cols(int64_t nnz, const double2 *__restrict__ a, const int32_t *__restrict__ aj,
                              const double2 *__restrict__ dinv, float2 *__restrict__ b)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= nnz) return;
  const double2 v = zmul(a[i], dinv[aj[i]]);
  b[i] = make_float2((float)v.x, (float)v.y);
}

// Krylov index of a captured iteration, kept on the device
__global__ void bk_set_j(int *__restrict__ dj, int v)
{
  if(threadIdx.x == 0) *dj = v;
}

__global__ void bk_next_j(int *__restrict__ dj)
{
  if(threadIdx.x == 0) *dj += 1;
}

__global__ void bk_zero(int64_t n, double2 *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) x[i] = make_double2(0., 0.);
}

// first-fit allocator over one device allocation
class Arena {
 public:
  bool init(size_t bytes)
  {
    if(hipMalloc(&base_, bytes) != hipSuccess) return false;
    free_[0] = bytes;
    size_ = bytes;
    return true;
  }
  ~Arena()
  {
    if(base_) (void)hipFree(base_);
  }
  // offset of a block of `bytes`, or -1
  int64_t alloc(size_t bytes)
  {
    bytes = (bytes + 255) / 256 * 256;
    for(auto it = free_.begin(); it != free_.end(); ++it)
      if(it->second >= bytes) {
        const size_t off = it->first, sz = it->second;
        free_.erase(it);
        if(sz > bytes) free_[off + bytes] = sz - bytes;
        used_[off] = bytes;
        return (int64_t)off;
      }
    return -1;
  }
  void release(int64_t off)
  {
    auto u = used_.find((size_t)off);
    if(u == used_.end()) return;
    size_t o = u->first, sz = u->second;
    used_.erase(u);
    auto nx = free_.lower_bound(o);
    if(nx != free_.end() && o + sz == nx->first) {
      sz += nx->second;
      nx = free_.erase(nx);
    }
    if(nx != free_.begin()) {
      auto pv = std::prev(nx);
      if(pv->first + pv->second == o) {
        pv->second += sz;
        return;
      }
    }
    free_[o] = sz;
  }
  char *at(int64_t off) const { return static_cast<char *>(base_) + off; }
  size_t size() const { return size_; }

 private:
  void *base_ = nullptr;
  size_t size_ = 0;
  std::map<size_t, size_t> free_, used_;
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// bytes of one request on `sys` (layout of Slot::bind)
size_t request_bytes(const gfem_system *sys, int m)
{
  const size_t n = (size_t)sys->n, nnz = (size_t)std::max<int64_t>(sys->nnz, 1);
  const size_t st = (size_t)(m + 1) * m + 3 * (size_t)m + 2 + (size_t)(m + 1) + 2;
  size_t b = align256(16 * nnz) + align256(8 * nnz) + align256(16 * n * (size_t)sys->nrhs) + 3 * align256(16 * n) +
             align256(8 * n) + align256(8 * n * (size_t)(m + 1)) + align256(16 * st) +
             align256(16 * (size_t)(m + 2) * kPartMax) + align256(16 * krylov::dcgs_size(m)) +
             align256(16 * 2 * (size_t)(m + 1) * (size_t)d2_parts(sys->n));
  if(sys->mg) b += 2 * align256(8 * n) + mg_request_bytes(sys);  // z, u, the cycle's state
  return b;
}

}  // namespace
}  // namespace gfem

using namespace gfem;

struct gfem_batch {
  struct Slot {
    bool used = false, live = false;
    int64_t id = -1;
    gfem_system *sys = nullptr;
    int64_t off = -1;
    BItem item{};
    double tol = 0., bnorm = 0., beta = 0., stop = 0., rel = 0.;
    int max_iter = 0, iters = 0, cycles = 0;
    bool converged = false;
    bool pending = false;  // submitted; initial residual at the next step
  };
  gfem_ctx *ctx = nullptr;
  int m = 50;
  int max_items = 0;
  Arena arena;
  std::vector<Slot> slots;
  BItem *d_items = nullptr;
  int *d_act = nullptr, *d_live = nullptr, *d_kcnt = nullptr;
  int2 *d_rblk = nullptr, *d_vblk = nullptr;
  int4 *d_dblk = nullptr, *d_d2blk = nullptr;
  int64_t cap_rblk = 0, cap_vblk = 0, cap_dblk = 0, cap_d2blk = 0;
  double *d_scal = nullptr;
  int *h_flags = nullptr;   // pinned: act | live
  double *h_scal = nullptr; // pinned: 3 per slot
  int64_t next_id = 1;
  std::vector<int64_t> done;          // completed, not yet fetched
  std::map<int64_t, int> where;       // id -> slot
  bool maps_dirty = true;
  int64_t n_rblk = 0, n_vblk = 0, n_dblk = 0, n_d2blk = 0;
  // Krylov iterations as HIP graphs (1 and kCheck iterations), captured on
  // cap_stream with the Krylov index read from d_j; valid until the block
  // maps change.  GFEM_BATCH_GRAPH=0 launches the kernels one by one.
  bool use_graph = true;
  bool dcgs = true;    // delayed CGS2 (GFEM_BATCH_CGS2=1: classical CGS2)
  bool jacobi = true;  // right scaling by row norms (GFEM_KRYLOV_SCALE=0: none)
  int *d_j = nullptr;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t g1 = nullptr, g5 = nullptr;
  // multigrid: per level l, blocks over its rows (d_mgblk) and over the rows
  // of level l + 1 (d_mgcblk) of the live requests that run level l row-
  // parallel (l < lsmall); mg_big = the most such levels, n_mg = live
  // requests with a hierarchy
  int2 *d_mgblk[kMgMaxLevels] = {}, *d_mgcblk[kMgMaxLevels] = {};
  int64_t n_mgblk[kMgMaxLevels] = {}, n_mgcblk[kMgMaxLevels] = {};
  int64_t cap_mgblk[kMgMaxLevels] = {}, cap_mgcblk[kMgMaxLevels] = {};
  int mg_big = 0, n_mg = 0;
  std::vector<MgReq> mg_host;  // per slot: the request's cycle layout
};

namespace {

void drop_graphs(gfem_batch *B)
{
  if(B->g1) (void)hipGraphExecDestroy(B->g1);
  if(B->g5) (void)hipGraphExecDestroy(B->g5);
  B->g1 = B->g5 = nullptr;
}

// block maps over the live slots
gfem_status build_maps(gfem_batch *B, hipStream_t s)
{
  std::vector<int2> rb, vb;
  std::vector<int4> db, d2;
  for(size_t i = 0; i < B->slots.size(); ++i) {
    const auto &S = B->slots[i];
    if(!S.live) continue;
    for(int64_t r = 0; r < S.item.n; r += kRowsPerBlock) rb.push_back(make_int2((int)i, (int)r));
    for(int64_t r = 0; r < S.item.n; r += kElemsPerBlock) vb.push_back(make_int2((int)i, (int)r));
    const int64_t span = dot_span(S.item.n);
    for(int64_t r = 0, p = 0; r < S.item.n; r += span, ++p)
      db.push_back(make_int4((int)i, (int)p, (int)r, (int)std::min<int64_t>(S.item.n, r + span)));
    for(int64_t r = 0, p = 0; r < S.item.n; r += kD2Span, ++p)
      d2.push_back(make_int4((int)i, (int)p, (int)r, (int)std::min<int64_t>(S.item.n, r + kD2Span)));
  }
  if(rb.empty()) rb.push_back(make_int2(0, 0));
  if(vb.empty()) vb.push_back(make_int2(0, 0));
  if(db.empty()) db.push_back(make_int4(0, 0, 0, 0));
  if(d2.empty()) d2.push_back(make_int4(0, 0, 0, 0));
  gfem_ctx *ctx = B->ctx;
  if((int64_t)d2.size() > B->cap_d2blk) {
    dfree(B->d_d2blk);
    B->cap_d2blk = (int64_t)d2.size() * 2;
    GFEM_HIP(ctx, dalloc(&B->d_d2blk, (size_t)B->cap_d2blk));
  }
  GFEM_HIP(ctx, hipMemcpyAsync(B->d_d2blk, d2.data(), sizeof(int4) * d2.size(), hipMemcpyHostToDevice, s));
  if((int64_t)db.size() > B->cap_dblk) {
    dfree(B->d_dblk);
    B->cap_dblk = (int64_t)db.size() * 2;
    GFEM_HIP(ctx, dalloc(&B->d_dblk, (size_t)B->cap_dblk));
  }
  GFEM_HIP(ctx, hipMemcpyAsync(B->d_dblk, db.data(), sizeof(int4) * db.size(), hipMemcpyHostToDevice, s));
  if((int64_t)rb.size() > B->cap_rblk) {
    dfree(B->d_rblk);
    B->cap_rblk = (int64_t)rb.size() * 2;
    GFEM_HIP(ctx, dalloc(&B->d_rblk, (size_t)B->cap_rblk));
  }
  if((int64_t)vb.size() > B->cap_vblk) {
    dfree(B->d_vblk);
    B->cap_vblk = (int64_t)vb.size() * 2;
    GFEM_HIP(ctx, dalloc(&B->d_vblk, (size_t)B->cap_vblk));
  }
  GFEM_HIP(ctx, hipMemcpyAsync(B->d_rblk, rb.data(), sizeof(int2) * rb.size(), hipMemcpyHostToDevice, s));
  GFEM_HIP(ctx, hipMemcpyAsync(B->d_vblk, vb.data(), sizeof(int2) * vb.size(), hipMemcpyHostToDevice, s));
  // multigrid level maps
  std::vector<int2> mb[kMgMaxLevels], mc[kMgMaxLevels];
  B->mg_big = 0;
  B->n_mg = 0;
  for(size_t i = 0; i < B->slots.size(); ++i) {
    const auto &S = B->slots[i];
    if(!S.live || !S.item.mg) continue;
    const MgReq &M = B->mg_host[i];
    ++B->n_mg;
    B->mg_big = std::max(B->mg_big, M.lsmall);
    for(int l = 0; l < M.lsmall; ++l) {
      for(int r = 0; r < M.lev[l].n; r += kRowsPerBlock) mb[l].push_back(make_int2((int)i, r));
      for(int r = 0; r < M.lev[l + 1].n; r += kRowsPerBlock) mc[l].push_back(make_int2((int)i, r));
    }
  }
  for(int l = 0; l < kMgMaxLevels; ++l) {
    for(int q = 0; q < 2; ++q) {
      std::vector<int2> &v = q ? mc[l] : mb[l];
      int2 *&d = q ? B->d_mgcblk[l] : B->d_mgblk[l];
      int64_t &cap = q ? B->cap_mgcblk[l] : B->cap_mgblk[l];
      (q ? B->n_mgcblk[l] : B->n_mgblk[l]) = (int64_t)v.size();
      if(v.empty()) continue;
      if((int64_t)v.size() > cap) {
        dfree(d);
        cap = (int64_t)v.size() * 2;
        GFEM_HIP(ctx, dalloc(&d, (size_t)cap));
      }
      GFEM_HIP(ctx, hipMemcpyAsync(d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice, s));
    }
  }
  GFEM_HIP(ctx, hipStreamSynchronize(s));  // the sources are local vectors
  B->n_rblk = (int64_t)rb.size();
  B->n_vblk = (int64_t)vb.size();
  B->n_dblk = (int64_t)db.size();
  B->n_d2blk = (int64_t)d2.size();
  B->maps_dirty = false;
  drop_graphs(B);
  return GFEM_OK;
}

// push act (per slot) and live flags through the pinned buffer; the stream
// is drained first so no earlier copy from that buffer is still pending
gfem_status push_flags(gfem_batch *B, const std::vector<int> &act, hipStream_t s)
{
  GFEM_HIP(B->ctx, hipStreamSynchronize(s));
  const size_t ns = B->slots.size();
  for(size_t i = 0; i < ns; ++i) {
    B->h_flags[i] = act[i];
    B->h_flags[ns + i] = B->slots[i].live ? 1 : 0;
  }
  GFEM_HIP(B->ctx, hipMemcpyAsync(B->d_act, B->h_flags, sizeof(int) * ns, hipMemcpyHostToDevice, s));
  GFEM_HIP(B->ctx, hipMemcpyAsync(B->d_live, B->h_flags + ns, sizeof(int) * ns, hipMemcpyHostToDevice, s));
  return GFEM_OK;
}

// One V-cycle of the requests with a hierarchy flagged in act: right-hand
// side V_j (src 0, j or dj) or u (src 1), result in z
// blocks of a launch over nblk map entries: a multiple of kXcds
unsigned xcd_grid(int64_t nblk)
{
  return (unsigned)((nblk + kXcds - 1) / kXcds * kXcds);
}

void launch_mg(gfem_batch *B, hipStream_t s, const int *act, int src, int j, const int *dj)
{
  if(B->n_mg == 0) return;
  const int ns = (int)B->slots.size();
  if(B->mg_big > 0)
    hipLaunchKernelGGL(bk_mg_x0, dim3((unsigned)B->n_vblk), dim3(256), 0, s, B->d_items, act, B->d_vblk, src, j, dj);
  for(int l = 0; l < B->mg_big; ++l) {
    if(B->n_mgblk[l])
      hipLaunchKernelGGL(bk_mg_down, dim3(xcd_grid(B->n_mgblk[l])), dim3(256), 0, s, B->d_items, act, B->d_mgblk[l],
                         (int)B->n_mgblk[l], l, src, j, dj);
    if(B->n_mgcblk[l])
      hipLaunchKernelGGL(bk_mg_restrict, dim3(xcd_grid(B->n_mgcblk[l])), dim3(256), 0, s, B->d_items, act,
                         B->d_mgcblk[l], (int)B->n_mgcblk[l], l);
  }
  hipLaunchKernelGGL(bk_mg_small, dim3((unsigned)ns), dim3(kMgSmallThreads), 0, s, B->d_items, act, src, j, dj);
  for(int l = B->mg_big - 1; l >= 0; --l) {
    if(!B->n_mgblk[l]) continue;
    hipLaunchKernelGGL(bk_mg_prolong, dim3(xcd_grid(B->n_mgblk[l])), dim3(256), 0, s, B->d_items, act, B->d_mgblk[l],
                       (int)B->n_mgblk[l], l);
    hipLaunchKernelGGL(bk_mg_post, dim3(xcd_grid(B->n_mgblk[l])), dim3(256), 0, s, B->d_items, act, B->d_mgblk[l],
                       (int)B->n_mgblk[l], l, src, j, dj);
  }
}

// One Krylov iteration j of every active request: w = A V_j, two
// Gram-Schmidt passes, |w|, V_{j+1}, Givens rotation.  With dj the index is
// read from device memory (captured graphs), else j is used.
// Delayed CGS2 step j of every active request (krylov.h): w = A V_j, one
// sweep for the dot products, their sums, the bookkeeping (column j-1 and
// its rotation), one sweep writing q_j and u_{j+1}.  tail (j = m): column m-1
// only, without a product.
void launch_dcgs(gfem_batch *B, hipStream_t s, int j, const int *dj, bool tail)
{
  const int ns = (int)B->slots.size(), m = B->m;
  if(!tail) launch_mg(B, s, B->d_act, 0, j, dj);
  if(!tail)
    hipLaunchKernelGGL(bk_spmv32, dim3((unsigned)((B->n_rblk + kXcds - 1) / kXcds * kXcds)), dim3(256), 0, s, B->d_items,
                       B->d_act, B->d_rblk, (int)B->n_rblk, j, dj);
  hipLaunchKernelGGL(bk_dots2, dim3((unsigned)B->n_d2blk), dim3(256), 0, s, B->d_items, B->d_act, B->d_d2blk, j, dj,
                     tail ? 1 : 0);
  hipLaunchKernelGGL(bk_dots2_final, dim3((unsigned)(2 * (m + 1)), (unsigned)ns), dim3(64), 0, s, B->d_items, B->d_act,
                     m, j, dj);
  hipLaunchKernelGGL(bk_dcgs_coef, dim3((unsigned)ns), dim3(64), 0, s, B->d_items, B->d_act, B->d_kcnt, m, j, dj,
                     tail ? 1 : 0);
  if(!tail)
    hipLaunchKernelGGL(bk_dcgs_update, dim3((unsigned)B->n_vblk), dim3(256), 0, s, B->d_items, B->d_act, B->d_vblk, m,
                       j, dj);
}

void launch_iteration(gfem_batch *B, hipStream_t s, int j, const int *dj)
{
  if(B->dcgs) {
    launch_dcgs(B, s, j, dj, false);
    return;
  }
  const int ns = (int)B->slots.size(), m = B->m;
  const dim3 red((unsigned)B->n_dblk), fin((unsigned)(m + 2), (unsigned)ns);
  launch_mg(B, s, B->d_act, 0, j, dj);
  hipLaunchKernelGGL(bk_spmv32, dim3((unsigned)((B->n_rblk + kXcds - 1) / kXcds * kXcds)), dim3(256), 0, s, B->d_items,
                       B->d_act, B->d_rblk, (int)B->n_rblk, j, dj);
  for(int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(bk_dots, red, dim3(256), 0, s, B->d_items, B->d_act, B->d_dblk, 0, j, dj);
    hipLaunchKernelGGL(bk_dots_final, fin, dim3(64), 0, s, B->d_items, B->d_act, m, 0, j, pass, dj);
    hipLaunchKernelGGL(bk_orth, dim3((unsigned)B->n_vblk), dim3(256), 0, s, B->d_items, B->d_act, B->d_vblk, m, j,
                       pass, dj);
  }
  hipLaunchKernelGGL(bk_add_y, dim3((unsigned)ns), dim3(64), 0, s, B->d_items, B->d_act, m, j, dj);
  hipLaunchKernelGGL(bk_dots, red, dim3(256), 0, s, B->d_items, B->d_act, B->d_dblk, 1, j, dj);
  hipLaunchKernelGGL(bk_dots_final, fin, dim3(64), 0, s, B->d_items, B->d_act, m, 1, j, 0, dj);
  hipLaunchKernelGGL(bk_scale, dim3((unsigned)B->n_vblk), dim3(256), 0, s, B->d_items, B->d_act, B->d_vblk, m, j, dj);
  hipLaunchKernelGGL(bk_givens, dim3((unsigned)ns), dim3(64), 0, s, B->d_items, B->d_act, B->d_kcnt, m, j, dj);
}

// a graph of `iters` iterations, each followed by d_j += 1 (kernels only)
hipError_t capture(gfem_batch *B, int iters, hipGraphExec_t *out)
{
  hipStream_t cs = B->cap_stream;
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if(e != hipSuccess) return e;
  for(int it = 0; it < iters; ++it) {
    launch_iteration(B, cs, 0, B->d_j);
    hipLaunchKernelGGL(bk_next_j, dim3(1), dim3(64), 0, cs, B->d_j);
  }
  hipGraph_t g = nullptr;
  const hipError_t le = hipGetLastError();
  e = hipStreamEndCapture(cs, &g);
  if(le != hipSuccess) e = le;
  if(e == hipSuccess) e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  if(g) (void)hipGraphDestroy(g);
  return e;
}

gfem_status collect(gfem_batch *B, hipStream_t s)
{
  const int ns = (int)B->slots.size();
  hipLaunchKernelGGL(bk_collect, dim3((unsigned)((ns + 63) / 64)), dim3(64), 0, s, B->d_items, ns, B->m, B->d_scal);
  GFEM_HIP(B->ctx, hipMemcpyAsync(B->h_scal, B->d_scal, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, s));
  GFEM_HIP(B->ctx, hipStreamSynchronize(s));
  return GFEM_OK;
}

// residuals and norms of the slots flagged in act; then for each: done when
// converged / out of iterations, else prepared for the next cycle
gfem_status residual_and_prepare(gfem_batch *B, std::vector<int> &act, hipStream_t s, bool first)
{
  gfem_ctx *ctx = B->ctx;
  const int ns = (int)B->slots.size(), m = B->m;
  gfem_status st = push_flags(B, act, s);
  if(st != GFEM_OK) return st;
  hipLaunchKernelGGL(bk_residual, dim3((unsigned)B->n_rblk), dim3(256), 0, s, B->d_items, B->d_act, B->d_rblk);
  hipLaunchKernelGGL(bk_dots, dim3((unsigned)B->n_dblk), dim3(256), 0, s, B->d_items, B->d_act, B->d_dblk, 2, 0,
                     nullptr);
  hipLaunchKernelGGL(bk_dots_final, dim3(1, (unsigned)ns), dim3(64), 0, s, B->d_items, B->d_act, m, 2, 0, 0, nullptr);
  if(first) {
    hipLaunchKernelGGL(bk_dots, dim3((unsigned)B->n_dblk), dim3(256), 0, s, B->d_items, B->d_act, B->d_dblk, 3, 0,
                       nullptr);
    hipLaunchKernelGGL(bk_dots_final, dim3(1, (unsigned)ns), dim3(64), 0, s, B->d_items, B->d_act, m, 3, 0, 0,
                       nullptr);
  }
  GFEM_HIP(ctx, hipGetLastError());
  st = collect(B, s);
  if(st != GFEM_OK) return st;
  static const bool debug = std::getenv("GFEM_BATCH_DEBUG") != nullptr;
  for(int i = 0; i < ns; ++i) {
    if(!act[i]) continue;
    auto &S = B->slots[i];
    if(first) S.bnorm = std::sqrt(B->h_scal[3 * i + 2]);
    S.beta = std::sqrt(B->h_scal[3 * i + 1]);
    if(debug)
      std::fprintf(stderr, "gfem batch: slot %d id %lld n %lld |b| %.6e |r| %.6e iters %d\n", i, (long long)S.id,
                   (long long)S.item.n, S.bnorm, S.beta, S.iters);
    S.rel = S.bnorm > 0. ? S.beta / S.bnorm : 0.;
    if(S.rel <= S.tol || S.iters >= S.max_iter || !(S.rel == S.rel)) {
      S.converged = S.rel <= S.tol;
      S.live = false;
      act[i] = 0;
      B->done.push_back(S.id);
      B->maps_dirty = true;
      continue;
    }
    S.stop = std::max(0.5 * S.tol * S.bnorm, 2e-6 * S.beta);
  }
  // V_0 = r / |r|, g = [|r|, 0 ...] for the slots that go on
  st = push_flags(B, act, s);
  if(st != GFEM_OK) return st;
  hipLaunchKernelGGL(bk_scale, dim3((unsigned)B->n_vblk), dim3(256), 0, s, B->d_items, B->d_act, B->d_vblk, m, -1,
                     nullptr);
  hipLaunchKernelGGL(bk_cycle_start, dim3((unsigned)ns), dim3(64), 0, s, B->d_items, B->d_act, B->d_kcnt, m);
  GFEM_HIP(ctx, hipGetLastError());
  return GFEM_OK;
}

}  // namespace

namespace gfem {

// gfem_batch_submit, or (a64_src, rhs_src non-null) a request whose FP64
// matrix and right-hand side are copied from those device buffers instead of
// assembled (gfem_solve of a system with a multigrid hierarchy)
gfem_status submit_request(gfem_batch *B, gfem_system *sys, double tol, int max_iter, void *stream, int64_t *id,
                           const double2 *a64_src, const double2 *rhs_src)
{
  if(!B || !sys || !id) return fail(B ? B->ctx : nullptr, GFEM_ERR_INVALID_ARG, "batch_submit: null argument");
  gfem_ctx *ctx = B->ctx;
  if(sys->ctx != ctx) return fail(ctx, GFEM_ERR_INVALID_ARG, "batch_submit: system of another context");
  if(!sys->finalized) return fail(ctx, GFEM_ERR_STATE, "batch_submit: system not finalized");
  if(!sys->cplx) return fail(ctx, GFEM_ERR_UNSUPPORTED, "batch_submit: complex systems only");
  if(sys->symmetric) return fail(ctx, GFEM_ERR_UNSUPPORTED, "batch_submit: symmetric storage");
  if(sys->n >= (int64_t(1) << 31)) return fail(ctx, GFEM_ERR_TOO_LARGE, "batch_submit: system too large");
  if(!(tol > 0.) || max_iter < 1) return fail(ctx, GFEM_ERR_INVALID_ARG, "batch_submit: tol > 0, max_iter >= 1");
  int slot = -1;
  for(int i = 0; i < (int)B->slots.size(); ++i)
    if(!B->slots[i].used) {
      slot = i;
      break;
    }
  if(slot < 0) return fail(ctx, GFEM_ERR_BUSY, "batch_submit: every request slot is taken");
  const int m = B->m;
  const int64_t off = B->arena.alloc(request_bytes(sys, m));
  if(off < 0) return fail(ctx, GFEM_ERR_BUSY, "batch_submit: the HBM budget is taken");
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  auto &S = B->slots[slot];
  S = gfem_batch::Slot();
  S.used = true;
  S.live = true;
  S.id = B->next_id++;
  S.sys = sys;
  ++sys->batch_refs;
  S.off = off;
  S.tol = tol;
  S.max_iter = max_iter;
  const size_t n = (size_t)sys->n, nnz = (size_t)std::max<int64_t>(sys->nnz, 1);
  char *p = B->arena.at(off);
  auto take = [&](size_t bytes) {
    char *q = p;
    p += align256(bytes);
    return q;
  };
  BItem &I = S.item;
  I.n = sys->n;
  I.nnz = sys->nnz;
  I.ai = sys->d_ai;
  I.aj = sys->d_aj;
  double2 *a64 = reinterpret_cast<double2 *>(take(16 * nnz));
  float2 *a32 = reinterpret_cast<float2 *>(take(8 * nnz));
  double2 *rhs = reinterpret_cast<double2 *>(take(16 * n * (size_t)sys->nrhs));
  I.a64 = a64;
  I.a32 = a32;
  I.b = rhs;
  I.x = reinterpret_cast<double2 *>(take(16 * n));
  I.r = reinterpret_cast<double2 *>(take(16 * n));
  double2 *dinv = reinterpret_cast<double2 *>(take(16 * n));
  I.dinv = dinv;
  I.w = reinterpret_cast<float2 *>(take(8 * n));
  I.V = reinterpret_cast<float2 *>(take(8 * n * (size_t)(m + 1)));
  I.st = reinterpret_cast<double2 *>(take(16 * ((size_t)(m + 1) * m + 3 * (size_t)m + 2 + (size_t)(m + 1) + 2)));
  I.part = reinterpret_cast<double2 *>(take(16 * (size_t)(m + 2) * kPartMax));
  I.nparts = (int)((sys->n + dot_span(sys->n) - 1) / dot_span(sys->n));
  I.dc = reinterpret_cast<double2 *>(take(16 * krylov::dcgs_size(m)));
  I.nparts2 = (int)d2_parts(sys->n);
  I.part2 = reinterpret_cast<double2 *>(take(16 * 2 * (size_t)(m + 1) * (size_t)I.nparts2));
  I.mg = nullptr;
  I.z = I.u = nullptr;
  if(sys->mg) {
    I.z = reinterpret_cast<float2 *>(take(8 * n));
    I.u = reinterpret_cast<float2 *>(take(8 * n));
    char *mb = take(mg_request_bytes(sys));
    mg_bind(sys, mb, &B->mg_host[slot]);
    I.mg = reinterpret_cast<const MgReq *>(mb);  // the descriptor sits at the start of its block
  }
  B->where[S.id] = slot;
  *id = S.id;
  // assemble this request's system with the system's current parameters (or
  // take the given matrix and right-hand side)
  gfem_status st = GFEM_OK;
  if(a64_src) {
    const hipError_t e = hipMemcpyAsync(a64, a64_src, 16 * (size_t)sys->nnz, hipMemcpyDeviceToDevice, s);
    const hipError_t e2 =
      e == hipSuccess ? hipMemcpyAsync(rhs, rhs_src, 16 * n, hipMemcpyDeviceToDevice, s) : e;
    if(e2 != hipSuccess) st = hip_fail(ctx, e2, "batch_submit: copy of the assembled system");
  }
  else
    st = gfem_assemble_into(sys, GFEM_PREC_FP64, a64, rhs, stream);
  if(st != GFEM_OK) {
    B->where.erase(S.id);
    B->arena.release(off);
    --sys->batch_refs;
    S = gfem_batch::Slot();
    return st;
  }
  // right scaling by row norms, unless the multigrid cycle preconditions
  hipLaunchKernelGGL(bk_diag_inv, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int64_t)n, sys->d_ai, sys->d_aj,
                     a64, B->jacobi && !sys->mg ? 1 : 0, dinv);
  hipLaunchKernelGGL(bk_round_cols, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, s, (int64_t)sys->nnz, a64,
                     sys->d_aj, dinv, a32);
  hipLaunchKernelGGL(bk_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (int64_t)n, I.x);
  if(sys->mg) {
    // the cycle's level operators, then its descriptor (with the request's
    // coefficients), in stream order
    GFEM_HIP(ctx, mg_prepare(sys, B->mg_host[slot], a64, s));
    hipLaunchKernelGGL(k_put<MgReq>, dim3(1), dim3(64), 0, s, const_cast<MgReq *>(I.mg), B->mg_host[slot]);
  }
  // this slot's descriptor only, in stream order (no wait)
  hipLaunchKernelGGL(k_put<BItem>, dim3(1), dim3(64), 0, s, B->d_items + slot, S.item);
  GFEM_HIP(ctx, hipGetLastError());
  // block maps, initial residual (= -b) and |b| at the next step, for every
  // request submitted until then at once
  S.pending = true;
  B->maps_dirty = true;
  return GFEM_OK;
}

// gfem_solve of a system with a multigrid hierarchy: the last FP64 assembly
// as the only request of a batch kept by the system (one slot, its restart
// length), iterated to the end; x is copied to the system's solution
gfem_status solve_mg(gfem_system *sys, int rhs_idx, double tol, int restart, int max_iter, hipStream_t s,
                     gfem_solve_info *info)
{
  gfem_ctx *ctx = sys->ctx;
  const int64_t n = sys->n;
  const size_t bytes = request_bytes(sys, restart);
  if(sys->mg_batch && (sys->mg_batch_m != restart || sys->mg_batch_bytes < bytes)) {
    gfem_batch_destroy(sys->mg_batch);
    sys->mg_batch = nullptr;
  }
  if(!sys->mg_batch) {
    const gfem_status st = gfem_batch_create(ctx, restart, bytes, 1, &sys->mg_batch);
    if(st != GFEM_OK) return st;
    sys->mg_batch_m = restart;
    sys->mg_batch_bytes = bytes;
  }
  gfem_batch *B = sys->mg_batch;
  if(!sys->d_sol_x) GFEM_HIP(ctx, dalloc(reinterpret_cast<double2 **>(&sys->d_sol_x), (size_t)std::max<int64_t>(n, 1)));
  int64_t id = -1;
  gfem_status st = submit_request(B, sys, tol, max_iter, s, &id, reinterpret_cast<const double2 *>(sys->d_vals[0]),
                                  reinterpret_cast<const double2 *>(sys->d_rhs[0]) + (int64_t)rhs_idx * n);
  for(int live = 1; live > 0 && st == GFEM_OK;) st = gfem_batch_step(B, s, &live);
  const auto w = B->where.find(id);
  if(st == GFEM_OK && w != B->where.end()) {
    const hipError_t e = hipMemcpyAsync(sys->d_sol_x, B->slots[w->second].item.x, 16 * (size_t)n,
                                        hipMemcpyDeviceToDevice, s);
    const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(s) : e;
    if(e2 != hipSuccess) st = hip_fail(ctx, e2, "solve: solution copy");
  }
  if(st == GFEM_OK) st = gfem_batch_result(B, id, nullptr, 0, info);  // frees the slot
  if(st != GFEM_OK) {
    // a request may be left in the slot: the next solve starts from a new batch
    gfem_batch_destroy(B);
    sys->mg_batch = nullptr;
    return st;
  }
  sys->sol_valid = true;
  return GFEM_OK;
}

}  // namespace gfem

extern "C" {

gfem_status gfem_batch_request_bytes(const gfem_system *sys, int restart, size_t *bytes)
{
  if(!sys || !bytes || restart < 1) return fail(sys ? sys->ctx : nullptr, GFEM_ERR_INVALID_ARG, "request_bytes");
  *bytes = request_bytes(sys, restart);
  return GFEM_OK;
}

gfem_status gfem_batch_create(gfem_ctx *ctx, int restart, size_t hbm_bytes, int max_items, gfem_batch **out)
{
  if(!ctx || !out || restart < 1 || restart > 1000 || max_items < 1 || max_items > 65535 || hbm_bytes == 0)
    return fail(ctx, GFEM_ERR_INVALID_ARG, "batch_create: bad arguments");
  *out = nullptr;
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  gfem_batch *B = new(std::nothrow) gfem_batch();
  if(!B) return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "host allocation failed");
  B->ctx = ctx;
  B->m = restart;
  B->max_items = max_items;
  B->slots.resize((size_t)max_items);
  B->mg_host.resize((size_t)max_items);
  bool ok = B->arena.init(hbm_bytes) && dalloc(&B->d_items, (size_t)max_items) == hipSuccess &&
            dalloc(&B->d_act, (size_t)max_items) == hipSuccess && dalloc(&B->d_live, (size_t)max_items) == hipSuccess &&
            dalloc(&B->d_kcnt, (size_t)max_items) == hipSuccess && dalloc(&B->d_j, 1) == hipSuccess &&
            hipStreamCreate(&B->cap_stream) == hipSuccess &&
            dalloc(&B->d_scal, (size_t)(3 * max_items)) == hipSuccess &&
            hipHostMalloc(&B->h_flags, sizeof(int) * 2 * (size_t)max_items) == hipSuccess &&
            hipHostMalloc(&B->h_scal, sizeof(double) * 3 * (size_t)max_items) == hipSuccess;
  if(ok) ok = hipMemset(B->d_items, 0, sizeof(BItem) * (size_t)max_items) == hipSuccess;
  const char *ge = std::getenv("GFEM_BATCH_GRAPH");
  B->use_graph = !(ge && ge[0] == '0');
  const char *ce = std::getenv("GFEM_BATCH_CGS2");
  B->dcgs = !(ce && ce[0] == '1');
  const char *je = std::getenv("GFEM_KRYLOV_SCALE");
  B->jacobi = !(je && je[0] == '0');
  if(!ok) {
    gfem_batch_destroy(B);
    return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "batch_create: device memory");
  }
  *out = B;
  return GFEM_OK;
}

gfem_status gfem_batch_destroy(gfem_batch *B)
{
  if(!B) return GFEM_OK;
  (void)hipSetDevice(B->ctx->device);
  (void)hipDeviceSynchronize();
  for(auto &S : B->slots)  // requests never fetched release their system
    if(S.used && S.sys) --S.sys->batch_refs;
  dfree(B->d_items);
  dfree(B->d_act);
  dfree(B->d_live);
  dfree(B->d_kcnt);
  dfree(B->d_scal);
  dfree(B->d_rblk);
  dfree(B->d_vblk);
  dfree(B->d_dblk);
  dfree(B->d_d2blk);
  for(int l = 0; l < kMgMaxLevels; ++l) {
    dfree(B->d_mgblk[l]);
    dfree(B->d_mgcblk[l]);
  }
  drop_graphs(B);
  dfree(B->d_j);
  if(B->cap_stream) (void)hipStreamDestroy(B->cap_stream);
  if(B->h_flags) (void)hipHostFree(B->h_flags);
  if(B->h_scal) (void)hipHostFree(B->h_scal);
  delete B;
  return GFEM_OK;
}

gfem_status gfem_batch_submit(gfem_batch *B, gfem_system *sys, double tol, int max_iter, void *stream, int64_t *id)
{
  return gfem::submit_request(B, sys, tol, max_iter, stream, id, nullptr, nullptr);
}

gfem_status gfem_batch_step(gfem_batch *B, void *stream, int *n_live)
{
  if(!B) return fail(nullptr, GFEM_ERR_INVALID_ARG, "batch is null");
  gfem_ctx *ctx = B->ctx;
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const int ns = (int)B->slots.size(), m = B->m;
  gfem_status st = GFEM_OK;
  // requests submitted since the last step: initial residual and |b|, all at
  // once (those with r = 0 end here)
  std::vector<int> act(ns, 0);
  bool fresh = false;
  for(int i = 0; i < ns; ++i)
    if(B->slots[i].live && B->slots[i].pending) {
      B->slots[i].pending = false;
      act[i] = 1;
      fresh = true;
    }
  if(fresh) {
    if(B->maps_dirty) {
      st = build_maps(B, s);
      if(st != GFEM_OK) return st;
    }
    st = residual_and_prepare(B, act, s, true);
    if(st != GFEM_OK) return st;
    // coarsest-level inverses of the fresh requests with a hierarchy, one
    // block each, all in one launch (act is on the device)
    bool fresh_mg = false;
    for(int i = 0; i < ns; ++i) fresh_mg |= act[i] && B->slots[i].item.mg;
    if(fresh_mg) {
      hipLaunchKernelGGL(bk_mg_invert, dim3((unsigned)ns), dim3(kMgInvThreads), 0, s, B->d_items, B->d_act);
      GFEM_HIP(ctx, hipGetLastError());
    }
  }
  int live = 0;
  for(int i = 0; i < ns; ++i) {
    act[i] = B->slots[i].live ? 1 : 0;
    live += act[i];
  }
  if(live == 0) {
    if(n_live) *n_live = 0;
    return GFEM_OK;
  }
  if(B->maps_dirty) {
    st = build_maps(B, s);
    if(st != GFEM_OK) return st;
  }
  st = push_flags(B, act, s);
  if(st != GFEM_OK) return st;
  if(B->use_graph) hipLaunchKernelGGL(bk_set_j, dim3(1), dim3(64), 0, s, B->d_j, 0);
  for(int j = 0; j < m;) {
    int running = 0;
    for(int i = 0; i < ns; ++i) running += act[i];
    if(running == 0) break;
    // iterations up to the next convergence check (every kCheck iterations)
    const int seg = std::min(kCheck - j % kCheck, m - j);
    if(B->use_graph) {
      hipError_t e = hipSuccess;
      if(seg == kCheck && !B->g5) e = capture(B, kCheck, &B->g5);
      if(e == hipSuccess && seg != kCheck && !B->g1) e = capture(B, 1, &B->g1);
      if(e != hipSuccess) return hip_fail(ctx, e, "batch: graph capture");
      for(int r = 0; r < (seg == kCheck ? 1 : seg) && e == hipSuccess; ++r)
        e = hipGraphLaunch(seg == kCheck ? B->g5 : B->g1, s);
      if(e != hipSuccess) return hip_fail(ctx, e, "batch: graph launch");
    }
    else {
      for(int r = 0; r < seg; ++r) launch_iteration(B, s, j + r, nullptr);
      GFEM_HIP(ctx, hipGetLastError());
    }
    for(int i = 0; i < ns; ++i)
      if(act[i]) B->slots[i].iters += seg;
    j += seg;
    if(j % kCheck == 0 && j < m) {
      st = collect(B, s);
      if(st != GFEM_OK) return st;
      bool changed = false;
      for(int i = 0; i < ns; ++i) {
        auto &S = B->slots[i];
        if(act[i] && (!(B->h_scal[3 * i] > S.stop) || S.iters >= S.max_iter)) {
          act[i] = 0;  // stays live: its x is updated after the cycle
          changed = true;
        }
      }
      if(changed) {
        st = push_flags(B, act, s);
        if(st != GFEM_OK) return st;
      }
    }
  }
  if(B->dcgs) {
    // requests that ran the whole cycle: its last column
    int running = 0;
    for(int i = 0; i < ns; ++i) running += act[i];
    if(running > 0) {
      launch_dcgs(B, s, m, nullptr, true);
      GFEM_HIP(ctx, hipGetLastError());
    }
  }
  for(int i = 0; i < ns; ++i)
    if(B->slots[i].live) B->slots[i].cycles++;
  hipLaunchKernelGGL(bk_backsolve, dim3((unsigned)ns), dim3(64), 0, s, B->d_items, B->d_live, B->d_kcnt, m);
  hipLaunchKernelGGL(bk_update_x, dim3((unsigned)B->n_vblk), dim3(256), 0, s, B->d_items, B->d_live, B->d_kcnt,
                     B->d_vblk, m);
  if(B->n_mg) {  // x += M^-1 V y
    launch_mg(B, s, B->d_live, 1, 0, nullptr);
    hipLaunchKernelGGL(bk_mg_addz, dim3((unsigned)B->n_vblk), dim3(256), 0, s, B->d_items, B->d_live, B->d_vblk);
  }
  GFEM_HIP(ctx, hipGetLastError());
  for(int i = 0; i < ns; ++i) act[i] = B->slots[i].live ? 1 : 0;
  st = residual_and_prepare(B, act, s, false);
  if(st != GFEM_OK) return st;
  if(n_live) {
    int l = 0;
    for(const auto &S : B->slots) l += S.live ? 1 : 0;
    *n_live = l;
  }
  return GFEM_OK;
}

gfem_status gfem_batch_done(gfem_batch *B, int max, int64_t *ids, int *count)
{
  if(!B || !count || (max > 0 && !ids)) return fail(B ? B->ctx : nullptr, GFEM_ERR_INVALID_ARG, "batch_done");
  const int c = std::min<int>(max, (int)B->done.size());
  for(int i = 0; i < c; ++i) ids[i] = B->done[i];
  B->done.erase(B->done.begin(), B->done.begin() + c);
  *count = c;
  return GFEM_OK;
}

gfem_status gfem_batch_result(gfem_batch *B, int64_t id, void *x_host, size_t bytes, gfem_solve_info *info)
{
  if(!B) return fail(nullptr, GFEM_ERR_INVALID_ARG, "batch is null");
  gfem_ctx *ctx = B->ctx;
  auto w = B->where.find(id);
  if(w == B->where.end()) return fail(ctx, GFEM_ERR_INVALID_ARG, "batch_result: unknown request");
  auto &S = B->slots[w->second];
  if(S.live) return fail(ctx, GFEM_ERR_STATE, "batch_result: request still running");
  if(x_host) {
    const size_t need = 16 * (size_t)S.item.n;
    if(bytes < need) return fail(ctx, GFEM_ERR_INVALID_ARG, "batch_result: buffer too small");
    GFEM_HIP(ctx, hipSetDevice(ctx->device));
    GFEM_HIP(ctx, hipMemcpy(x_host, S.item.x, need, hipMemcpyDeviceToHost));
  }
  if(info) {
    info->iterations = S.iters;
    info->cycles = S.cycles;
    info->rel_residual = S.rel;
    info->converged = S.converged ? 1 : 0;
  }
  // an unused slot holds an empty descriptor (kernels over every slot skip it).
  // The clear completes before this call returns: the next submit writes the
  // slot on the caller's stream, which need not be ordered after the null
  // stream (hipStreamNonBlocking), so an unfinished clear could overwrite it.
  GFEM_HIP(ctx, hipSetDevice(ctx->device));
  hipLaunchKernelGGL(k_put<BItem>, dim3(1), dim3(64), 0, nullptr, B->d_items + w->second, BItem{});
  GFEM_HIP(ctx, hipGetLastError());
  GFEM_HIP(ctx, hipStreamSynchronize(nullptr));
  B->arena.release(S.off);
  B->where.erase(w);
  if(S.sys) --S.sys->batch_refs;
  S = gfem_batch::Slot();
  return GFEM_OK;
}

}  // extern "C"
