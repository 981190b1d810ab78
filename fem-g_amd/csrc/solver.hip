// Device solver for A x = -b (the reference's sign convention,
// fem-master/src/system/Solver.cpp:119-123), in place of the reference's
// PETSc/MUMPS LU (Solver.cpp:76-160).
//
// GMRES with iterative refinement: each outer step computes the true residual
// r = -b - A x in FP64 from the FP64 matrix, runs one restart cycle of
// GMRES(m) on A d = r in FP32 (matrix values rounded once to FP32, Krylov
// basis in FP32), and updates x += d in FP64.  Inner products are accumulated
// in FP64 with classical Gram-Schmidt applied twice (CGS2); every reduction
// sums block partials in a fixed order, so a solve is bitwise reproducible.
// The Hessenberg matrix and its Givens rotations stay on the device; the
// host reads the residual estimate every few iterations to stop a cycle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "internal.h"
#include "krylov.h"

namespace gfem {

namespace {

constexpr int kRowLanes = 8;       // lanes per CSR row in the SpMV kernels
constexpr int kRedBlocks = 512;    // partial sums per reduction
constexpr int kRedThreads = 256;
constexpr int kDotGroup = 8;       // basis vectors per pass of the dot kernel
constexpr int kSpmvXcds = 8;       // XCDs of an MI355X: the SpMV gives each a contiguous run of rows

template <bool C>
struct Sc;
template <>
struct Sc<true> {
  using f32 = float2;
  using f64 = double2;
};
template <>
struct Sc<false> {
  using f32 = float;
  using f64 = double;
};

__device__ __forceinline__ double2 to_d(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double to_d(float v) { return v; }
__device__ __forceinline__ double2 ld_d(const double2 *p, int64_t i) { return p[i]; }
__device__ __forceinline__ double2 ld_d(const double *p, int64_t i) { return make_double2(p[i], 0.); }
__device__ __forceinline__ double2 ld_d(const float2 *p, int64_t i) { return make_double2(p[i].x, p[i].y); }
__device__ __forceinline__ double2 ld_d(const float *p, int64_t i) { return make_double2(p[i], 0.); }
__device__ __forceinline__ void st(double2 *p, int64_t i, double2 v) { p[i] = v; }
__device__ __forceinline__ void st(double *p, int64_t i, double2 v) { p[i] = v.x; }
__device__ __forceinline__ void st(float2 *p, int64_t i, double2 v) { p[i] = make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ void st(float *p, int64_t i, double2 v) { p[i] = (float)v.x; }

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconjmul(double2 a, double2 b)  // conj(a) b
{
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

template <int L = kRowLanes>
__device__ __forceinline__ double2 lanes_sum(double2 v)
{
#pragma unroll
  for(int d = 1; d < L; d <<= 1) {
    v.x += __shfl_xor(v.x, d, L);
    v.y += __shfl_xor(v.y, d, L);
  }
  return v;
}

// y = A x (FP32 values and vectors, FP32 products summed per lane in FP32,
// L lanes per row combined in a fixed butterfly; L is autotuned)
template <bool C, int L>
__global__ void __launch_bounds__(256) k_spmv32(int64_t n, const int64_t *__restrict__ ai, const int32_t *__restrict__ aj,
                                                const typename Sc<C>::f32 *__restrict__ a,
                                                const typename Sc<C>::f32 *__restrict__ x,
                                                typename Sc<C>::f32 *__restrict__ y)
{
  // each XCD takes a contiguous run of blocks: its rows are one segment of
  // the (Hilbert-ordered) unknowns, so its L2 holds that segment of x and a
  // halo instead of all of x (the grid is a multiple of kSpmvXcds)
  const int64_t per = gridDim.x / kSpmvXcds;
  const int64_t blk = (int64_t)(blockIdx.x % kSpmvXcds) * per + blockIdx.x / kSpmvXcds;
  const int64_t row = (blk * blockDim.x + threadIdx.x) / L;
  const int l = (int)(threadIdx.x % L);
  if(row >= n) return;  // whole lane groups leave together
  float sr = 0.f, si = 0.f;
  // kU terms of the lane at a time: columns and values, then the x entries,
  // are in flight together; the products are still summed in term order
  constexpr int kU = 4;
  const int64_t k1 = ai[row + 1];
  for(int64_t k0 = ai[row] + l; k0 < k1; k0 += kU * L) {
    int32_t c[kU];
    double2 av[kU], xv[kU];
#pragma unroll
    for(int u = 0; u < kU; ++u) {
      const bool in = k0 + u * L < k1;
      c[u] = in ? aj[k0 + u * L] : 0;
      av[u] = in ? ld_d(a, k0 + u * L) : make_double2(0., 0.);
    }
#pragma unroll
    for(int u = 0; u < kU; ++u) xv[u] = k0 + u * L < k1 ? ld_d(x, c[u]) : make_double2(0., 0.);
#pragma unroll
    for(int u = 0; u < kU; ++u)
      if(k0 + u * L < k1) {
        sr += (float)av[u].x * (float)xv[u].x - (float)av[u].y * (float)xv[u].y;
        si += (float)av[u].x * (float)xv[u].y + (float)av[u].y * (float)xv[u].x;
      }
  }
  const double2 s = lanes_sum<L>(make_double2(sr, si));
  if(l == 0) st(y, row, s);
}

template <bool C>
hipError_t spmv32(int lanes, int64_t n, const int64_t *ai, const int32_t *aj, const typename Sc<C>::f32 *a,
                  const typename Sc<C>::f32 *x, typename Sc<C>::f32 *y, hipStream_t s)
{
  const unsigned nb = (unsigned)(((n * lanes + 255) / 256 + kSpmvXcds - 1) / kSpmvXcds * kSpmvXcds);
  if(lanes == 4)
    hipLaunchKernelGGL((k_spmv32<C, 4>), dim3(nb), dim3(256), 0, s, n, ai, aj, a, x, y);
  else if(lanes == 16)
    hipLaunchKernelGGL((k_spmv32<C, 16>), dim3(nb), dim3(256), 0, s, n, ai, aj, a, x, y);
  else
    hipLaunchKernelGGL((k_spmv32<C, 8>), dim3(nb), dim3(256), 0, s, n, ai, aj, a, x, y);
  return hipGetLastError();
}

// r = -b - A x in FP64
template <bool C>
__global__ void __launch_bounds__(256) k_residual64(int64_t n, const int64_t *__restrict__ ai,
                                                    const int32_t *__restrict__ aj,
                                                    const typename Sc<C>::f64 *__restrict__ a,
                                                    const typename Sc<C>::f64 *__restrict__ x,
                                                    const typename Sc<C>::f64 *__restrict__ b,
                                                    typename Sc<C>::f64 *__restrict__ r)
{
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes;
  const int l = (int)(threadIdx.x % kRowLanes);
  if(row >= n) return;
  double2 s = make_double2(0., 0.);
  for(int64_t k = ai[row] + l; k < ai[row + 1]; k += kRowLanes) {
    const double2 p = zmul(ld_d(a, k), ld_d(x, aj[k]));
    s.x += p.x;
    s.y += p.y;
  }
  s = lanes_sum(s);
  if(l == 0) {
    const double2 bv = ld_d(b, row);
    st(r, row, make_double2(-bv.x - s.x, -bv.y - s.y));
  }
}

template <typename T>
__device__ __forceinline__ void dots_group(int64_t n, int k0, int kg, const T *__restrict__ V, int64_t ld,
                                           const T *__restrict__ w, double2 *__restrict__ part,
                                           double2 (*s_red)[kDotGroup])
{
  double2 acc[kDotGroup];
#pragma unroll
  for(int q = 0; q < kDotGroup; ++q) acc[q] = make_double2(0., 0.);
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 wv = ld_d(w, i);
#pragma unroll
    for(int q = 0; q < kDotGroup; ++q)
      if(q < kg) {
        const double2 p = zconjmul(ld_d(V, (int64_t)(k0 + q) * ld + i), wv);
        acc[q].x += p.x;
        acc[q].y += p.y;
      }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for(int q = 0; q < kDotGroup; ++q) {
#pragma unroll
    for(int d = 32; d >= 1; d >>= 1) {
      acc[q].x += __shfl_xor(acc[q].x, d, 64);
      acc[q].y += __shfl_xor(acc[q].y, d, 64);
    }
    if(lane == 0) s_red[wv][q] = acc[q];
  }
  __syncthreads();
  if(threadIdx.x < kDotGroup && (int)threadIdx.x < kg) {
    double2 s = make_double2(0., 0.);
    for(int k = 0; k < kRedThreads / 64; ++k) {
      s.x += s_red[k][threadIdx.x].x;
      s.y += s_red[k][threadIdx.x].y;
    }
    part[(int64_t)(k0 + threadIdx.x) * kRedBlocks + blockIdx.x] = s;
  }
  __syncthreads();  // s_red is reused by the next group
}

// block partials of conj(V_k) . w for k in [0, nk), FP64 accumulation, the
// vectors kDotGroup at a time (w read once per group) inside one launch; V
// rows are `ld` apart.  part[(k) * kRedBlocks + block]
template <typename T>
__global__ void __launch_bounds__(kRedThreads) k_dots(int64_t n, int nk, const T *__restrict__ V, int64_t ld,
                                                      const T *__restrict__ w, double2 *__restrict__ part)
{
  __shared__ double2 s_red[kRedThreads / 64][kDotGroup];
  for(int k0 = 0; k0 < nk; k0 += kDotGroup) {
    const int kg = min(kDotGroup, nk - k0);
    dots_group<T>(n, k0, kg, V, ld, w, part, s_red);
  }
}

// out[k] (+)= sum over blocks of part[k][.], fixed order; one block per k
__global__ void __launch_bounds__(kRedThreads) k_dots_final(const double2 *__restrict__ part, double2 *__restrict__ out,
                                                            int accumulate)
{
  __shared__ double2 s_red[kRedThreads / 64];
  const int k = blockIdx.x;
  double2 s = make_double2(0., 0.);
  for(int b = threadIdx.x; b < kRedBlocks; b += kRedThreads) {
    s.x += part[(int64_t)k * kRedBlocks + b].x;
    s.y += part[(int64_t)k * kRedBlocks + b].y;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for(int d = 32; d >= 1; d >>= 1) {
    s.x += __shfl_xor(s.x, d, 64);
    s.y += __shfl_xor(s.y, d, 64);
  }
  if(lane == 0) s_red[wv] = s;
  __syncthreads();
  if(threadIdx.x == 0) {
    double2 t = make_double2(0., 0.);
    for(int q = 0; q < kRedThreads / 64; ++q) {
      t.x += s_red[q].x;
      t.y += s_red[q].y;
    }
    if(accumulate) {
      out[k].x += t.x;
      out[k].y += t.y;
    }
    else {
      out[k] = t;
    }
  }
}

// Delayed CGS2 (one reduction per Krylov step): block partials of
// conj(V_k) . u and conj(V_k) . w for k in [0, nk), both targets in one sweep.
// Block b owns rows [b kD2Span, (b+1) kD2Span); its waves take the basis
// vectors round robin over those rows (u and w, reused by the four waves,
// stay in L1), so the basis is read once from HBM and no barrier is needed.
// A lane sums its rows in order, the wave in a fixed butterfly, the blocks in
// a fixed order (k_dots2_final): bitwise reproducible.
// part[(2k + t) * gridDim.x + block], t = 0 for u, 1 for w.
constexpr int kD2Span = 256;  // rows per block
inline unsigned dots2_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kD2Span - 1) / kD2Span); }

template <typename T>
__global__ void __launch_bounds__(256) k_dots2(int64_t n, int nk, const T *__restrict__ V, int64_t ld,
                                               const T *__restrict__ u, const T *__restrict__ w,
                                               double2 *__restrict__ part)
{
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * kD2Span;
  const int64_t r1 = min(n, r0 + kD2Span);
  for(int k = wv; k < nk; k += 4) {
    double2 au = make_double2(0., 0.), aw = make_double2(0., 0.);
#pragma unroll 4
    for(int64_t i = r0 + lane; i < r1; i += 64) {
      const double2 v = ld_d(V, (int64_t)k * ld + i);
      const double2 pa = zconjmul(v, ld_d(u, i)), pb = zconjmul(v, ld_d(w, i));
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
      part[(int64_t)(2 * k) * gridDim.x + blockIdx.x] = au;
      part[(int64_t)(2 * k + 1) * gridDim.x + blockIdx.x] = aw;
    }
  }
}

// out[o] = sum over the nb block partials of part[o][.], fixed order; one block per output
__global__ void __launch_bounds__(kRedThreads) k_dots2_final(const double2 *__restrict__ part, int nb,
                                                             double2 *__restrict__ out)
{
  __shared__ double2 s_red[kRedThreads / 64];
  const int o = blockIdx.x;
  double2 s = make_double2(0., 0.);
  for(int b = threadIdx.x; b < nb; b += kRedThreads) {
    s.x += part[(int64_t)o * nb + b].x;
    s.y += part[(int64_t)o * nb + b].y;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for(int d = 32; d >= 1; d >>= 1) {
    s.x += __shfl_xor(s.x, d, 64);
    s.y += __shfl_xor(s.y, d, 64);
  }
  if(lane == 0) s_red[wv] = s;
  __syncthreads();
  if(threadIdx.x == 0) {
    double2 t = make_double2(0., 0.);
    for(int q = 0; q < kRedThreads / 64; ++q) {
      t.x += s_red[q].x;
      t.y += s_red[q].y;
    }
    out[o] = t;
  }
}

// Delayed CGS2 update of step j (u = V_j, the once-orthogonalised vector, v
// = A u in w):
//   q_j     = (u - sum_{k<j} V_k s_k) / r            into V_j (in place)
//   u_{j+1} = (v - sum_{k<j} V_k c_k - q_j c_j) / r  into V_{j+1}
// one read of V_0..V_j for both; s, c, 1/r from k_dcgs_coef
template <typename T>
__global__ void __launch_bounds__(256) k_dcgs_update(int64_t n, int j, T *__restrict__ V, int64_t ld,
                                                     const T *__restrict__ v, const double2 *__restrict__ sv,
                                                     const double2 *__restrict__ cv, const double2 *__restrict__ ir)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double2 q = ld_d(V, (int64_t)j * ld + i), un = ld_d(v, i);
  for(int k = 0; k < j; ++k) {
    const double2 vk = ld_d(V, (int64_t)k * ld + i);
    const double2 a = zmul(vk, sv[k]), b = zmul(vk, cv[k]);
    q.x -= a.x;
    q.y -= a.y;
    un.x -= b.x;
    un.y -= b.y;
  }
  const double inv = ir->x;
  q = make_double2(q.x * inv, q.y * inv);
  const double2 qc = zmul(q, cv[j]);
  un = make_double2((un.x - qc.x) * inv, (un.y - qc.y) * inv);
  st(V, (int64_t)j * ld + i, q);
  st(V, (int64_t)(j + 1) * ld + i, un);
}

// w -= sum_{k<=j} V_k h_k  (h FP64, arithmetic in FP64, w FP32)
template <typename T>
__global__ void __launch_bounds__(256) k_orth_update(int64_t n, int j1, const T *__restrict__ V, int64_t ld,
                                                     const double2 *__restrict__ h, T *__restrict__ w)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double2 s = ld_d(w, i);
  for(int k = 0; k < j1; ++k) {
    const double2 p = zmul(ld_d(V, (int64_t)k * ld + i), h[k]);
    s.x -= p.x;
    s.y -= p.y;
  }
  st(w, i, s);
}

// dst = src * scale (scale = 1/norm from a device scalar); FP64 or FP32 source
template <typename S, typename T>
__global__ void __launch_bounds__(256) k_scale(int64_t n, const S *__restrict__ src, const double2 *__restrict__ nrm2,
                                               T *__restrict__ dst)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double nv = sqrt(nrm2->x);
  const double inv = nv > 0. ? 1. / nv : 0.;
  const double2 v = ld_d(src, i);
  st(dst, i, make_double2(v.x * inv, v.y * inv));
}

// FP64 -> FP32 copy of the matrix values
template <typename S, typename T>
__global__ void __launch_bounds__(256) k_round(int64_t n, const S *__restrict__ src, T *__restrict__ dst)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) st(dst, i, ld_d(src, i));
}

// Right scaling of the Krylov iteration: it runs on A D^-1 with D the
// 2-norms of the rows of A (for the symmetric patterns assembled here, those
// of its columns), so the FP32 copy of the matrix has its columns scaled by
// D^-1 and a cycle's correction is D^-1 V y; the FP64 residual -b - A x is
// the unscaled one.  dinv = 1 / D (1 for an empty row, or without scaling).
template <typename S>
__global__ void __launch_bounds__(256) k_row_scale(int64_t n, const int64_t *__restrict__ ai, const S *__restrict__ a,
                                                   int scale, double *__restrict__ dinv)
{
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= n) return;
  double q = 0.;
  if(scale)
    for(int64_t k = ai[r]; k < ai[r + 1]; ++k) {
      const double2 v = ld_d(a, k);
      q += v.x * v.x + v.y * v.y;
    }
  dinv[r] = q > 0. ? 1. / sqrt(q) : 1.;
}

// FP32 copy of A D^-1
template <typename S, typename T>
__global__ void __launch_bounds__(256) k_round_cols(int64_t nnz, const S *__restrict__ a, const int32_t *__restrict__ aj,
                                                    const double *__restrict__ dinv, T *__restrict__ dst)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= nnz) return;
  const double2 v = ld_d(a, k);
  const double d = dinv[aj[k]];
  st(dst, k, make_double2(v.x * d, v.y * d));
}

// x (FP64) += D^-1 sum_k V_k y_k
template <typename T, typename X>
__global__ void __launch_bounds__(256) k_update_x(int64_t n, int j1, const T *__restrict__ V, int64_t ld,
                                                  const double2 *__restrict__ y, const double *__restrict__ dinv,
                                                  X *__restrict__ x)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double2 s = make_double2(0., 0.);
  for(int k = 0; k < j1; ++k) {
    const double2 p = zmul(ld_d(V, (int64_t)k * ld + i), y[k]);
    s.x += p.x;
    s.y += p.y;
  }
  const double2 xi = ld_d(x, i);
  st(x, i, make_double2(xi.x + dinv[i] * s.x, xi.y + dinv[i] * s.y));
}

// col[i] += y[i], i < k
__global__ void k_add_into(int k, const double2 *__restrict__ y, double2 *__restrict__ col)
{
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  if(i < k) {
    col[i].x += y[i].x;
    col[i].y += y[i].y;
  }
}

// Hessenberg bookkeeping, one thread.  st layout (double2):
//   H[(m+1) x m] column-major at 0, cs[m], sn[m], g[m+1], est at the end.
// Column j holds h_{0..j,j} (from the dots) and h_{j+1,j}^2 in slot j+1
// (squared norm from the norm pass).  Applies the previous rotations, makes
// rotation j, updates g and the residual estimate |g_{j+1}|.
__global__ void k_givens(int m, int j, double2 *__restrict__ stt)
{
  if(threadIdx.x != 0 || blockIdx.x != 0) return;
  double2 *col = stt + (int64_t)j * (m + 1);
  col[j + 1] = make_double2(sqrt(col[j + 1].x), 0.);
  krylov::givens_column(m, j, stt);
}

// Delayed CGS2 bookkeeping of step j (krylov.h), one wavefront
__global__ void __launch_bounds__(64) k_dcgs_coef(int m, int j, int tail, double2 *__restrict__ dc,
                                                  double2 *__restrict__ stt)
{
  krylov::dcgs_coef(m, j, tail, dc, stt);
}

// back substitution R y = g over the first k columns (one thread); y -> out
__global__ void k_backsolve(int m, int k, const double2 *__restrict__ stt, double2 *__restrict__ y)
{
  if(threadIdx.x != 0 || blockIdx.x != 0) return;
  const double2 *H = stt;
  const double2 *g = stt + (int64_t)(m + 1) * m + 2 * m;
  for(int i = k - 1; i >= 0; --i) {
    double2 s = g[i];
    for(int l = i + 1; l < k; ++l) {
      const double2 p = zmul(H[(int64_t)l * (m + 1) + i], y[l]);
      s.x -= p.x;
      s.y -= p.y;
    }
    const double2 d = H[(int64_t)i * (m + 1) + i];
    const double dd = d.x * d.x + d.y * d.y;
    y[i] = dd > 0. ? make_double2((s.x * d.x + s.y * d.y) / dd, (s.y * d.x - s.x * d.y) / dd) : make_double2(0., 0.);
  }
}

// start a cycle: g = [beta, 0, ...], beta = sqrt(nrm2)
__global__ void k_cycle_start(int m, const double2 *__restrict__ nrm2, double2 *__restrict__ stt)
{
  const int i = threadIdx.x + blockIdx.x * blockDim.x;
  double2 *g = stt + (int64_t)(m + 1) * m + 2 * m;
  if(i <= m) g[i] = make_double2(i == 0 ? sqrt(nrm2->x) : 0., 0.);
}

unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <bool C>
struct Gmres {
  using T32 = typename Sc<C>::f32;
  using T64 = typename Sc<C>::f64;
  gfem_system *sys;
  hipStream_t s;
  int64_t n;
  int m;

  // h[k0..k0+j1) (+)= V_k^H w  (FP64 sums in a fixed order)
  template <typename T>
  hipError_t dots(const T *V, int64_t ld, int j1, const T *w, double2 *out, int accumulate)
  {
    hipLaunchKernelGGL(k_dots<T>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, j1, V, ld, w, sys->d_sol_part);
    hipLaunchKernelGGL(k_dots_final, dim3((unsigned)j1), dim3(kRedThreads), 0, s, sys->d_sol_part, out, accumulate);
    return hipGetLastError();
  }
};

}  // namespace

template <bool C>
static gfem_status gmres_ir(gfem_system *sys, int rhs_idx, double tol, int m, int max_iter, hipStream_t s,
                            gfem_solve_info *info)
{
  using T32 = typename Sc<C>::f32;
  using T64 = typename Sc<C>::f64;
  gfem_ctx *ctx = sys->ctx;
  const int64_t n = sys->n, nnz = sys->nnz;
  auto chk = [&](hipError_t e) { return e == hipSuccess; };
  // workspace (kept between solves while the sizes fit)
  const size_t need_v = sizeof(T32) * (size_t)n * (size_t)(m + 1);
  if(sys->sol_m != m || !sys->d_sol_V) {
    // everything sized by the restart length m
    sys->sol_m = 0;
    dfree(sys->d_sol_V);
    dfree(sys->d_sol_state);
    dfree(sys->d_sol_dcgs);
    dfree(sys->d_sol_y);
    dfree(sys->d_sol_part);
    for(hipGraphExec_t g : sys->sol_graphs) (void)hipGraphExecDestroy(g);  // they hold the old workspace
    sys->sol_graphs.clear();
    if(!chk(hipMalloc(&sys->d_sol_V, need_v))) return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "solver basis allocation");
    if(!chk(hipMalloc(&sys->d_sol_part, sizeof(double2) * (size_t)std::max<int64_t>(kRedBlocks, dots2_blocks(n)) *
                                          (size_t)(2 * m + 4))))
      return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "solver reduction allocation");
    if(!chk(hipMalloc(&sys->d_sol_dcgs, sizeof(double2) * krylov::dcgs_size(m))))
      return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "solver state allocation");
    if(!chk(hipMalloc(&sys->d_sol_state, sizeof(double2) * ((size_t)(m + 1) * m + 3 * (size_t)m + 2))))
      return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "solver state allocation");
    if(!chk(hipMalloc(&sys->d_sol_y, sizeof(double2) * (size_t)(m + 1))))
      return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "solver state allocation");
    sys->sol_m = m;
  }
  if(!sys->d_sol_r) {  // (x alone may exist: a preconditioned solve allocates only it)
    if((!sys->d_sol_x && !chk(hipMalloc(&sys->d_sol_x, sizeof(T64) * (size_t)n))) ||
       !chk(hipMalloc(&sys->d_sol_r, sizeof(T64) * (size_t)n)) ||
       !chk(hipMalloc(&sys->d_sol_w, sizeof(T32) * (size_t)n)) ||
       !chk(hipMalloc(&sys->d_sol_a32, sizeof(T32) * (size_t)std::max<int64_t>(nnz, 1))) ||
       !chk(hipMalloc(&sys->d_sol_nrm, sizeof(double2) * 2)) ||
       !chk(hipMalloc(&sys->d_sol_dinv, sizeof(double) * (size_t)std::max<int64_t>(n, 1))))
      return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "solver workspace allocation");
    if(!chk(hipHostMalloc(&sys->h_sol_est, sizeof(double2) * 2)))
      return fail(ctx, GFEM_ERR_OUT_OF_MEMORY, "solver host scalar allocation");
  }
  T64 *x = reinterpret_cast<T64 *>(sys->d_sol_x);
  T64 *r = reinterpret_cast<T64 *>(sys->d_sol_r);
  T32 *w = reinterpret_cast<T32 *>(sys->d_sol_w);
  T32 *V = reinterpret_cast<T32 *>(sys->d_sol_V);
  T32 *a32 = reinterpret_cast<T32 *>(sys->d_sol_a32);
  const T64 *a64 = reinterpret_cast<const T64 *>(sys->d_vals[0]);
  const T64 *b = reinterpret_cast<const T64 *>(sys->d_rhs[0]) + (int64_t)rhs_idx * n;
  double2 *stt = sys->d_sol_state;
  double2 *H = stt;
  double2 *est = stt + (int64_t)(m + 1) * m + 3 * m + 1;
  double2 *nrm = sys->d_sol_nrm;
  double2 *hest = reinterpret_cast<double2 *>(sys->h_sol_est);
  Gmres<C> G{sys, s, n, m};
  const unsigned rb = blocks(n * kRowLanes, 256), vb = blocks(n, 256);

  GFEM_HIP(ctx, hipMemsetAsync(x, 0, sizeof(T64) * (size_t)n, s));
  {
    const char *ks = std::getenv("GFEM_KRYLOV_SCALE");
    const int scale = (ks && ks[0] == '0') ? 0 : 1;
    hipLaunchKernelGGL((k_row_scale<T64>), dim3(blocks(n, 256)), dim3(256), 0, s, n, sys->d_ai, a64, scale,
                       sys->d_sol_dinv);
    hipLaunchKernelGGL((k_round_cols<T64, T32>), dim3(blocks(nnz, 256)), dim3(256), 0, s, nnz, a64, sys->d_aj,
                       sys->d_sol_dinv, a32);
  }
  if(!sys->spmv_tuned) {
    // autotune the SpMV lanes per row on this matrix (one warm-up and three
    // timed products each; w is scratch here)
    const char *at = std::getenv("GFEM_AUTOTUNE");
    if(!(at && at[0] == '0')) {
      hipEvent_t e0, e1;
      GFEM_HIP(ctx, hipEventCreate(&e0));
      GFEM_HIP(ctx, hipEventCreate(&e1));
      // the lane count changes the rounding of the FP32 products: the default
      // (8) stays unless another is clearly faster (stable from run to run)
      float best = 1e30f, def = 1e30f;
      int pick = 8;
      hipLaunchKernelGGL((k_round<T64, T32>), dim3(blocks(n, 256)), dim3(256), 0, s, n, b, V);
      for(int lanes : {8, 4, 16}) {
        GFEM_HIP(ctx, spmv32<C>(lanes, n, sys->d_ai, sys->d_aj, a32, V, w, s));
        GFEM_HIP(ctx, hipEventRecord(e0, s));
        for(int r = 0; r < 3; ++r) GFEM_HIP(ctx, spmv32<C>(lanes, n, sys->d_ai, sys->d_aj, a32, V, w, s));
        GFEM_HIP(ctx, hipEventRecord(e1, s));
        GFEM_HIP(ctx, hipEventSynchronize(e1));
        float ms = 0.f;
        GFEM_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
        if(lanes == 8) def = ms;
        if(ms < best && (lanes == 8 || ms < 0.95f * def)) {
          best = ms;
          pick = lanes;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      sys->spmv_lanes = pick;
    }
    sys->spmv_tuned = true;
  }
  // Gram-Schmidt: delayed CGS2 (default; one reduction and two reads of the
  // basis per step) or, with GFEM_SOLVER_CGS2=1, classical CGS2 (two
  // reductions, four reads)
  const char *cg_env = std::getenv("GFEM_SOLVER_CGS2");
  const bool dcgs = !(cg_env && cg_env[0] == '1');
  double2 *dc = sys->d_sol_dcgs;
  double2 *dcD = dc + krylov::dcgs_dots_offset(m);
  double2 *dcS = dc + krylov::dcgs_s_offset(m), *dcC = dc + krylov::dcgs_c_offset(m), *dcR = dc + krylov::dcgs_r_offset(m);
  // Delayed CGS2 step j: V_j holds u_j (once orthogonalised; q_0 at j = 0).
  // w = A u_j; one sweep for Q_{j-1}^H [u_j, w] and u_j^H [u_j, w]; the
  // bookkeeping finishes column j-1 (rotated: the estimate lags one step)
  // and the coefficients; one sweep writes q_j into V_j and u_{j+1} into
  // V_{j+1}.  The tail (j = m) finishes column m-1 without a product.
  auto dcgs_step = [&](int j, bool tail, hipStream_t st) -> hipError_t {
    T32 *vj = V + (int64_t)j * n;
    if(!tail) {
      const hipError_t e = spmv32<C>(sys->spmv_lanes, n, sys->d_ai, sys->d_aj, a32, vj, w, st);
      if(e != hipSuccess) return e;
    }
    const unsigned nb2 = dots2_blocks(n);
    hipLaunchKernelGGL(k_dots2<T32>, dim3(nb2), dim3(256), 0, st, n, j + 1, V, n, vj, tail ? vj : w, sys->d_sol_part);
    hipLaunchKernelGGL(k_dots2_final, dim3((unsigned)(2 * (j + 1))), dim3(kRedThreads), 0, st, sys->d_sol_part,
                       (int)nb2, dcD);
    hipLaunchKernelGGL(k_dcgs_coef, dim3(1), dim3(64), 0, st, m, j, tail ? 1 : 0, dc, stt);
    if(!tail)
      hipLaunchKernelGGL(k_dcgs_update<T32>, dim3(vb), dim3(256), 0, st, n, j, V, n, w, dcS, dcC, dcR);
    return hipGetLastError();
  };
  // One Krylov iteration j on stream st: w = A v_j, two Gram-Schmidt passes,
  // |w|, v_{j+1}, Givens rotation.  Every argument is fixed for the system's
  // workspace, restart length and SpMV lanes, so a block of iterations can be
  // captured once and replayed.
  auto iteration = [&](int j, hipStream_t st) -> hipError_t {
    if(dcgs) return dcgs_step(j, false, st);
    Gmres<C> Gs{sys, st, n, m};
    T32 *vj = V + (int64_t)j * n;
    hipError_t e = spmv32<C>(sys->spmv_lanes, n, sys->d_ai, sys->d_aj, a32, vj, w, st);
    if(e != hipSuccess) return e;
    double2 *col = H + (int64_t)j * (m + 1);
    // CGS2: h = V^H w, w -= V h, twice
    if((e = Gs.dots(V, n, j + 1, w, col, 0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_orth_update<T32>, dim3(vb), dim3(256), 0, st, n, j + 1, V, n, col, w);
    if((e = Gs.dots(V, n, j + 1, w, sys->d_sol_y, 0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_orth_update<T32>, dim3(vb), dim3(256), 0, st, n, j + 1, V, n, sys->d_sol_y, w);
    hipLaunchKernelGGL(k_add_into, dim3(1), dim3(64 * ((j + 64) / 64)), 0, st, j + 1, sys->d_sol_y, col);
    // h_{j+1,j}^2 into the column, then v_{j+1} = w / h_{j+1,j}
    if((e = Gs.dots(w, n, 1, w, col + j + 1, 0)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_scale<T32, T32>), dim3(vb), dim3(256), 0, st, n, w, col + j + 1, V + (int64_t)(j + 1) * n);
    hipLaunchKernelGGL(k_givens, dim3(1), dim3(64), 0, st, m, j, stt);
    return hipGetLastError();
  };
  // HIP graphs of kSolGraphBlock iterations (launch overhead dominates small
  // systems); GFEM_SOLVER_GRAPHS=0 launches every kernel instead.  A failed
  // capture falls back to the direct launches.
  const char *gs_env = std::getenv("GFEM_SOLVER_GRAPHS");
  bool use_graphs = !(gs_env && gs_env[0] == '0');
  if(use_graphs && (sys->sol_graphs.empty() || sys->sol_graph_m != m || sys->sol_graph_lanes != sys->spmv_lanes ||
                    sys->sol_graph_dcgs != dcgs)) {
    for(hipGraphExec_t g : sys->sol_graphs) (void)hipGraphExecDestroy(g);
    sys->sol_graphs.clear();
    hipError_t e = sys->sol_cap ? hipSuccess : hipStreamCreateWithFlags(&sys->sol_cap, hipStreamNonBlocking);
    for(int j0 = 0; j0 < m && e == hipSuccess; j0 += kSolGraphBlock) {
      e = hipStreamBeginCapture(sys->sol_cap, hipStreamCaptureModeThreadLocal);
      if(e != hipSuccess) break;
      hipError_t le = hipSuccess;
      for(int j = j0; j < std::min(m, j0 + kSolGraphBlock) && le == hipSuccess; ++j) le = iteration(j, sys->sol_cap);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(sys->sol_cap, &g);
      if(le != hipSuccess) e = le;
      hipGraphExec_t ge = nullptr;
      if(e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      if(g) (void)hipGraphDestroy(g);
      if(e == hipSuccess) sys->sol_graphs.push_back(ge);
    }
    if(e != hipSuccess) {
      (void)hipGetLastError();
      for(hipGraphExec_t g : sys->sol_graphs) (void)hipGraphExecDestroy(g);
      sys->sol_graphs.clear();
      use_graphs = false;
    }
    else {
      sys->sol_graph_m = m;
      sys->sol_graph_lanes = sys->spmv_lanes;
      sys->sol_graph_dcgs = dcgs;
    }
  }
  // ||b||
  GFEM_HIP(ctx, G.dots(b, n, 1, b, nrm + 1, 0));
  GFEM_HIP(ctx, hipMemcpyAsync(hest, nrm, sizeof(double2) * 2, hipMemcpyDeviceToHost, s));
  GFEM_HIP(ctx, hipStreamSynchronize(s));
  const double bnorm = std::sqrt(hest[1].x);
  int iters = 0, cycles = 0;
  double rel = bnorm > 0. ? 1. : 0.;
  bool converged = bnorm == 0.;
  while(!converged && iters < max_iter) {
    // true residual in FP64
    hipLaunchKernelGGL(k_residual64<C>, dim3(rb), dim3(256), 0, s, n, sys->d_ai, sys->d_aj, a64, x, b, r);
    GFEM_HIP(ctx, G.dots(r, n, 1, r, nrm, 0));
    GFEM_HIP(ctx, hipMemcpyAsync(hest, nrm, sizeof(double2), hipMemcpyDeviceToHost, s));
    GFEM_HIP(ctx, hipStreamSynchronize(s));
    const double beta = std::sqrt(hest[0].x);
    rel = beta / bnorm;
    if(rel <= tol) {
      converged = true;
      break;
    }
    ++cycles;
    // v0 = r / beta (FP32), g = [beta, 0 ...]
    hipLaunchKernelGGL((k_scale<T64, T32>), dim3(vb), dim3(256), 0, s, n, r, nrm, V);
    hipLaunchKernelGGL(k_cycle_start, dim3(1), dim3(64 * ((m + 64) / 64)), 0, s, m, nrm, stt);
    // the cycle stops once the estimate reaches the outer target or the
    // FP32 floor relative to this cycle's residual
    const double stop = std::max(0.5 * tol * bnorm, 2e-6 * beta);
    // iterations in blocks of kSolGraphBlock, the estimate checked after each
    // block (and at the end of the cycle); a block that would pass max_iter
    // runs iteration by iteration
    int k = 0;
    for(int j = 0; j < m && iters < max_iter;) {
      const int jb = std::min(m, j + kSolGraphBlock);
      if(use_graphs && j % kSolGraphBlock == 0 && iters + (jb - j) <= max_iter) {
        GFEM_HIP(ctx, hipGraphLaunch(sys->sol_graphs[j / kSolGraphBlock], s));
        iters += jb - j;
        j = jb;
      }
      else {
        GFEM_HIP(ctx, iteration(j, s));
        ++iters;
        ++j;
      }
      if(dcgs && j == m) GFEM_HIP(ctx, dcgs_step(m, true, s));  // the cycle's last column
      // finished columns: delayed CGS2 finishes column j-1 at step j
      k = dcgs ? (j == m ? m : j - 1) : j;
      if(j % kSolGraphBlock == 0 || j == m) {
        GFEM_HIP(ctx, hipMemcpyAsync(hest, est, sizeof(double2), hipMemcpyDeviceToHost, s));
        GFEM_HIP(ctx, hipStreamSynchronize(s));
        if(!(hest[0].x > stop)) break;  // also stops on a NaN estimate
      }
    }
    // x += V y, y = R^-1 g
    hipLaunchKernelGGL(k_backsolve, dim3(1), dim3(64), 0, s, m, k, stt, sys->d_sol_y);
    hipLaunchKernelGGL((k_update_x<T32, T64>), dim3(vb), dim3(256), 0, s, n, k, V, n, sys->d_sol_y, sys->d_sol_dinv,
                       x);
    GFEM_HIP(ctx, hipGetLastError());
  }
  if(!converged) {
    // final true residual
    hipLaunchKernelGGL(k_residual64<C>, dim3(rb), dim3(256), 0, s, n, sys->d_ai, sys->d_aj, a64, x, b, r);
    GFEM_HIP(ctx, G.dots(r, n, 1, r, nrm, 0));
    GFEM_HIP(ctx, hipMemcpyAsync(hest, nrm, sizeof(double2), hipMemcpyDeviceToHost, s));
    GFEM_HIP(ctx, hipStreamSynchronize(s));
    rel = bnorm > 0. ? std::sqrt(hest[0].x) / bnorm : 0.;
    converged = rel <= tol;
  }
  sys->sol_valid = true;
  if(info) {
    info->iterations = iters;
    info->cycles = cycles;
    info->rel_residual = rel;
    info->converged = converged ? 1 : 0;
  }
  return GFEM_OK;
}

gfem_status solve_gmres_ir(gfem_system *sys, int rhs_idx, double tol, int restart, int max_iter, hipStream_t s,
                           gfem_solve_info *info)
{
  return sys->cplx ? gmres_ir<true>(sys, rhs_idx, tol, restart, max_iter, s, info)
                   : gmres_ir<false>(sys, rhs_idx, tol, restart, max_iter, s, info);
}

}  // namespace gfem
