// GMRES bookkeeping on the device, shared by the single-request solver
// (solver.hip) and the batched engine (batch.hip): Givens rotations of the
// Hessenberg matrix and the delayed-CGS2 step.
//
// Krylov state (double2): H[(m+1) x m] column-major (column j holds
// h_{0..j+1, j}) | cs[m] | sn[m] | g[m+1] | est (|g_{j+1}|, the residual
// estimate after column j).
//
// Delayed CGS2 (one global reduction per Krylov step; the second
// Gram-Schmidt pass of step j-1 is done at step j).  At step j the basis
// slot V_j holds u_j, the step j-1 vector after one orthogonalisation, and v
// = A u_j.  One sweep gives D[2k] = V_k^H u_j and D[2k+1] = V_k^H v for k <=
// j.  Then
//   s = Q_{j-1}^H u_j, r = sqrt(|u_j|^2 - |s|^2)   (u_j - Q s is orthogonal to Q)
//   column j-1 = h1 + s, h_{j,j-1} = r              (h1: its first pass)
//   q_j = (u_j - Q_{j-1} s) / r
//   A q_j = (v - Q_j Hu s) / r                       (A Q_{j-1} = Q_j Hu)
//   c = Q_j^H v = [z; (u_j^H v - s^H z) / r]
//   h1 (column j, first pass) = (c - Hu s) / r
//   u_{j+1} = A q_j - Q_j h1 = (v - Q_j c) / r
// Hu is the unrotated Hessenberg.  The tail (j = m) only finishes column
// m-1.  dc = Hu[(m+1) x m] | h1[m+1] | s[m] | c[m+1] | 1/r | D[2(m+1)].
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gfem {
namespace krylov {

__device__ __forceinline__ double2 mul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conjmul(double2 a, double2 b)  // conj(a) b
{
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__host__ __device__ inline size_t dcgs_s_offset(int m) { return (size_t)(m + 1) * m + (size_t)(m + 1); }
__host__ __device__ inline size_t dcgs_c_offset(int m) { return dcgs_s_offset(m) + m; }
__host__ __device__ inline size_t dcgs_r_offset(int m) { return dcgs_c_offset(m) + (size_t)(m + 1); }
__host__ __device__ inline size_t dcgs_dots_offset(int m) { return dcgs_r_offset(m) + 1; }
__host__ __device__ inline size_t dcgs_size(int m) { return dcgs_dots_offset(m) + 2 * (size_t)(m + 1); }

// Rotations of column j, whose entries h_{0..j+1, j} are in place with
// h_{j+1,j} real: the previous rotations, rotation j, g and the estimate.
__device__ inline void givens_column(int m, int j, double2 *__restrict__ stt)
{
  double2 *H = stt;
  double2 *cs = stt + (int64_t)(m + 1) * m;
  double2 *sn = cs + m;
  double2 *g = sn + m;
  double2 *est = g + m + 1;
  double2 *col = H + (int64_t)j * (m + 1);
  for(int i = 0; i < j; ++i) {
    // [c s; -conj(s) c] with real c
    const double c = cs[i].x;
    const double2 s = sn[i];
    const double2 a = col[i], b = col[i + 1];
    const double2 sb = mul(s, b);
    const double2 csa = conjmul(s, a);
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
    s = make_double2(b.x / nb, -b.y / nb);  // conj(b)/|b|: s b = |b|
  }
  else {
    const double t = hypot(na, nb);
    c = na / t;
    // s = (a/|a|) conj(b) / t
    const double2 ua = make_double2(a.x / na, a.y / na);
    s = make_double2((ua.x * b.x + ua.y * b.y) / t, (ua.y * b.x - ua.x * b.y) / t);
  }
  cs[j] = make_double2(c, 0.);
  sn[j] = s;
  const double2 sb = mul(s, b);
  col[j] = make_double2(c * a.x + sb.x, c * a.y + sb.y);
  col[j + 1] = make_double2(0., 0.);
  const double2 gj = g[j];
  const double2 csg = conjmul(s, gj);
  g[j] = make_double2(c * gj.x, c * gj.y);
  g[j + 1] = make_double2(-csg.x, -csg.y);
  est->x = hypot(g[j + 1].x, g[j + 1].y);
}

// The delayed-CGS2 bookkeeping of step j (see the top of this file), called
// by every lane of one 64-lane block.  Sums over k are lane-strided, then
// combined in a fixed butterfly (bitwise reproducible).
__device__ inline void dcgs_coef(int m, int j, int tail, double2 *__restrict__ dc, double2 *__restrict__ stt)
{
  const int t = threadIdx.x;
  double2 *Hu = dc;
  double2 *h1 = Hu + (int64_t)(m + 1) * m;
  double2 *sv = dc + dcgs_s_offset(m);
  double2 *cv = dc + dcgs_c_offset(m);
  double2 *ir = dc + dcgs_r_offset(m);
  const double2 *D = dc + dcgs_dots_offset(m);
  double ss = 0.;
  for(int k = t; k < j; k += 64) {
    const double2 d = D[2 * k];
    sv[k] = d;
    ss += d.x * d.x + d.y * d.y;
  }
#pragma unroll
  for(int d = 32; d >= 1; d >>= 1) ss += __shfl_xor(ss, d, 64);
  const double r2 = D[2 * j].x - ss;
  const double r = r2 > 0. ? sqrt(r2) : 0.;
  const double inv = r > 0. ? 1. / r : 0.;
  if(j >= 1) {
    double2 *hu = Hu + (int64_t)(j - 1) * (m + 1);
    double2 *col = stt + (int64_t)(j - 1) * (m + 1);
    for(int i = t; i < j; i += 64) {  // h1[i], sv[i] of this lane
      hu[i] = make_double2(h1[i].x + sv[i].x, h1[i].y + sv[i].y);
      col[i] = hu[i];
    }
    if(t == 0) {
      hu[j] = make_double2(r, 0.);
      col[j] = hu[j];
    }
    __syncthreads();
    if(t == 0) givens_column(m, j - 1, stt);
  }
  if(tail) return;
  double2 sz = make_double2(0., 0.);
  for(int k = t; k < j; k += 64) {
    const double2 z = D[2 * k + 1];
    cv[k] = z;
    const double2 p = conjmul(sv[k], z);
    sz.x += p.x;
    sz.y += p.y;
  }
#pragma unroll
  for(int d = 32; d >= 1; d >>= 1) {
    sz.x += __shfl_xor(sz.x, d, 64);
    sz.y += __shfl_xor(sz.y, d, 64);
  }
  if(t == 0) {
    const double2 beta = D[2 * j + 1];
    cv[j] = make_double2((beta.x - sz.x) * inv, (beta.y - sz.y) * inv);
    ir[0] = make_double2(inv, 0.);
  }
  __syncthreads();  // sv, cv and column j-1 of Hu seen by every lane
  for(int i = t; i <= j; i += 64) {
    double2 a = cv[i];
    for(int l = (i > 0 ? i - 1 : 0); l < j; ++l) {  // Hu column l has rows 0..l+1
      const double2 p = mul(Hu[(int64_t)l * (m + 1) + i], sv[l]);
      a.x -= p.x;
      a.y -= p.y;
    }
    h1[i] = make_double2(a.x * inv, a.y * inv);
  }
}

}  // namespace krylov
}  // namespace gfem
