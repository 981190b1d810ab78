// Element matrices of the per-point bilinear terms on the matrix cores.
//
//   dense[e*planes + p][pair] = sum_k A[e*planes + p][k] R[k][pair],  k = 7 q + s
//
// with, per Gauss point q, R(q,s; i*N+j) = w_q d_a phi_j d_b phi_i (the six
// components s of G = J^-1 J^-T, symmetrised) and w_q phi_j phi_i (s = 6),
// and A(q,s) = det G_s x the stiffness coefficient at x_q, det x the mass
// coefficient (s = 6): the reference's per-point B^T W B summed over the
// Gauss points (fem-master/src/term/BilinearTerm.cpp:303-343), as one product
// per group.  One kernel, three arithmetics:
//   FP64    v_mfma_f64_16x16x4f64: FP64 operands and sums;
//   FP32    v_mfma_f32_32x32x2f32: operands rounded to FP32, FP32 sums;
//   BF16X3  v_mfma_f32_32x32x16_bf16 on the bfloat16 split x = hi + lo of
//           both operands (from the FP64 values): hi*hi + lo*hi + hi*lo
//           summed in FP32 (lo*lo, ~2^-16 of a product, is left out).
// A block of 256 threads (4 waves, 2 x 2) computes a BM x BN tile of dense.
// K is walked in steps of KT in component-major order (k = s Qp + q, each
// component's points padded to Qp); the R tile of a step sits in LDS (padded
// rows: 16-lane groups of 16-byte reads land on distinct banks), double
// buffered: the next step's tile is loaded into registers while the MFMAs of
// this one run.  R is read from a transposed, zero-padded copy R_T[NNp][Kp]
// built once per group (k_ref_products_t).  A is either generated or read:
//   generated: every bilinear term of an affine group in the element
//     matrices, none with a coefficient that depends on x.  Then A(q,s) =
//     F_e[s] c_s does not depend on q, and a K step, inside one component s,
//     needs one value per row: the block keeps the 7 values of each of its
//     rows (FP64 element factors x the group's constant coefficients) and
//     the A panel never exists in memory;
//   read: rows written by k_coef_rows (expression coefficients, curved
//     elements with their own Jacobian at every point), staged per step.
#include <hip/hip_runtime.h>

#include "internal.h"

namespace gfem {
namespace {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr int kMmThreads = 256;

// Tile shapes: a block is 4 waves (2 x 2) over a BM x BN tile, each wave a
// WM x WN grid of MFMA micro-tiles; K steps of KT.  GEN blocks hold no A tile
// and take twice the rows.
template <int MODE, bool GEN, int KT_>
struct MmCfg;
template <bool GEN, int KT_>
struct MmCfg<kMmF64, GEN, KT_> {
  using E = double;                 // LDS operand element
  using Out = double;               // dense element
  static constexpr int MT = 16;     // micro-tile edge (16 x 16 x 4)
  static constexpr int BM = GEN ? 128 : 64, BN = 64, KT = KT_;
  static constexpr int WM = BM / 2, WN = BN / 2;
  static constexpr int LD = KT + 2;  // LDS row stride (elements): 16-byte multiple, odd in 16-byte units
};
template <bool GEN, int KT_>
struct MmCfg<kMmF32, GEN, KT_> {
  using E = float;
  using Out = float;
  static constexpr int MT = 32;     // 32 x 32 x 2
  static constexpr int BM = GEN ? 256 : 128, BN = 128, KT = KT_;
  static constexpr int WM = BM / 2, WN = BN / 2;
  static constexpr int LD = KT + 4;
};
template <bool GEN, int KT_>
struct MmCfg<kMmBf16x3, GEN, KT_> {
  using E = uint16_t;  // bfloat16 bits
  using Out = float;
  static constexpr int MT = 32;     // 32 x 32 x 16
  static constexpr int BM = GEN ? 256 : 128, BN = 128, KT = KT_;
  static constexpr int WM = BM / 2, WN = BN / 2;
  static constexpr int LD = KT + 8;
};

// K order: component-major, k = s Qp + q, with the Gauss points of each of
// the 7 components padded to Qp, a multiple of every K step (zero products
// in the padding).  A K step then lies inside one component s, where a
// constant-coefficient row of A is the single value F_e[s] c_s.  Rules of 48
// points and more are padded to a multiple of 64, smaller ones to 32.  K
// steps: 32 for read A; for generated A the step that measured fastest on
// the order-4 workload: 16 in FP64 (4 waves per SIMD), 32 in FP32, 64 (when
// the padding allows) for the bf16 split.
constexpr int kMmKT = 32;
inline int mm_qalign(int Q) { return Q >= 48 ? 64 : 32; }

// bfloat16 bits of x, rounded to nearest even (finite x)
__device__ __forceinline__ uint16_t bf16_of(float x)
{
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float float_of_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// x = hi + lo + O(2^-17 |x|)
__device__ __forceinline__ void split_bf16(double x, uint16_t &hi, uint16_t &lo)
{
  hi = bf16_of((float)x);
  lo = bf16_of((float)(x - (double)float_of_bf16(hi)));
}

// The transposed, zero-padded products R_T[NNp][Kp] of one group (FP64,
// FP32, or the bf16 hi / lo parts): R(q,s; i*N+j) as above, at k = s Qp + q.
template <int MODE>
__global__ void __launch_bounds__(256) k_ref_products_t(int N, int Q, int NNp, int Kp, const double *__restrict__ w,
                                                        const double *__restrict__ bf, const double *__restrict__ gbf,
                                                        void *__restrict__ Rt, uint16_t *__restrict__ Rt_lo)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= (int64_t)NNp * Kp) return;
  const int pair = (int)(t / Kp), k = (int)(t % Kp);
  const int NN = N * N, Qp = Kp / 7;
  const int s = k / Qp, q = k - s * Qp;
  double v = 0.;
  if(pair < NN && q < Q) {
    const int i = pair / N, j = pair % N;
    const double *gj = &gbf[(q * N + j) * 3], *gi = &gbf[(q * N + i) * 3];
    switch(s) {
    case 0: v = gj[0] * gi[0]; break;
    case 1: v = gj[1] * gi[1]; break;
    case 2: v = gj[2] * gi[2]; break;
    case 3: v = gj[0] * gi[1] + gj[1] * gi[0]; break;
    case 4: v = gj[0] * gi[2] + gj[2] * gi[0]; break;
    case 5: v = gj[1] * gi[2] + gj[2] * gi[1]; break;
    default: v = bf[q * N + j] * bf[q * N + i]; break;
    }
    v *= w[q];
  }
  if constexpr(MODE == kMmF64)
    static_cast<double *>(Rt)[t] = v;
  else if constexpr(MODE == kMmF32)
    static_cast<float *>(Rt)[t] = (float)v;
  else {
    uint16_t hi, lo;
    split_bf16(v, hi, lo);
    static_cast<uint16_t *>(Rt)[t] = hi;
    Rt_lo[t] = lo;
  }
}

// One K step of the block's A tile into LDS (hi / lo planes for the split),
// read from the rows A[M][7Q] (index 7 q + s) that k_coef_rows wrote.
template <int MODE>
__device__ __forceinline__ void stage_a(int t, int64_t row0, int64_t M, int Q, int Qp, int k0, const void *__restrict__ Ain,
                                        typename MmCfg<MODE, false, kMmKT>::E *sA,
                                        typename MmCfg<MODE, false, kMmKT>::E *sAlo)
{
  using C = MmCfg<MODE, false, kMmKT>;
  using E = typename C::E;
  const int s = k0 / Qp, q0 = k0 - s * Qp, K = 7 * Q;
  for(int idx = t; idx < C::BM * C::KT; idx += kMmThreads) {
    const int r = idx / C::KT, kk = idx % C::KT;
    const int q = q0 + kk;
    const int64_t row = row0 + r;
    E v = 0, vlo = 0;
    if(q < Q && row < M) {
      const int64_t at = row * K + 7 * q + s;
      if constexpr(MODE == kMmF64)
        v = static_cast<const double *>(Ain)[at];
      else if constexpr(MODE == kMmF32)
        v = static_cast<const float *>(Ain)[at];
      else {
        const float a = static_cast<const float *>(Ain)[at];
        v = bf16_of(a);
        vlo = bf16_of(a - float_of_bf16(v));
      }
    }
    sA[r * C::LD + kk] = v;
    if constexpr(MODE == kMmBf16x3) sAlo[r * C::LD + kk] = vlo;
  }
}

// The R tile of one K step (R_T rows col0 .. col0 + BN - 1, columns k0 ..
// k0 + KT - 1, always inside the padded copy) through registers: loads
// issued before the MFMAs of the current step, stored to LDS after them.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int MODE, bool GEN, int KT>
struct RTileShape {
  using C = MmCfg<MODE, GEN, KT>;
  static constexpr int per = 16 / sizeof(typename C::E);  // elements per 16-byte load
  static constexpr int chunks = C::KT / per;              // loads per row
  static constexpr int n = C::BN * chunks / kMmThreads;   // loads per thread
  static_assert(C::BN * chunks % kMmThreads == 0, "whole loads per thread");
};

template <int MODE, bool GEN, int KT, int N>
__device__ __forceinline__ void rtile_load(u32x4 (&v)[N], int t, int Kp, int64_t col0, int k0,
                                           const void *__restrict__ Rt)
{
  using S = RTileShape<MODE, GEN, KT>;
  using E = typename S::C::E;
  const E *src = static_cast<const E *>(Rt);
#pragma unroll
  for(int i = 0; i < N; ++i) {
    const int idx = t + i * kMmThreads, c = idx / S::chunks, part = idx % S::chunks;
    v[i] = *reinterpret_cast<const u32x4 *>(src + (col0 + c) * (int64_t)Kp + k0 + part * S::per);
  }
}

template <int MODE, bool GEN, int KT, int N>
__device__ __forceinline__ void rtile_store(const u32x4 (&v)[N], int t, typename MmCfg<MODE, GEN, KT>::E *sR)
{
  using S = RTileShape<MODE, GEN, KT>;
#pragma unroll
  for(int i = 0; i < N; ++i) {
    const int idx = t + i * kMmThreads, c = idx / S::chunks, part = idx % S::chunks;
    *reinterpret_cast<u32x4 *>(sR + c * S::C::LD + part * S::per) = v[i];
  }
}

template <int MODE, bool GEN, int KT>
__global__ void __launch_bounds__(kMmThreads) k_elem_mm(const GroupDev *__restrict__ gd, int64_t M, int planes,
                                                        int NN, int Q, int Kp, const double *__restrict__ F64,
                                                        const void *__restrict__ Ain, const void *__restrict__ Rt,
                                                        const void *__restrict__ Rt_lo, void *__restrict__ Cout)
{
  using C = MmCfg<MODE, GEN, KT>;
  using E = typename C::E;
  using Out = typename C::Out;
  static_assert(GEN || KT == kMmKT, "read A runs steps of kMmKT");
  constexpr bool SPLIT = MODE == kMmBf16x3;
  constexpr int MB = C::WM / C::MT, NB = C::WN / C::MT;  // micro-tiles per wave
  constexpr int ACC = C::MT * C::MT / 64;                 // accumulator entries per lane and micro-tile
  using Acc = typename std::conditional<MODE == kMmF64, v4d, v16f>::type;
  __shared__ __align__(16) E sR[2][C::BN * C::LD];
  __shared__ __align__(16) E sRlo[SPLIT ? 2 : 1][SPLIT ? C::BN * C::LD : 1];
  __shared__ __align__(16) E sA[GEN ? 1 : 2][GEN ? 1 : C::BM * C::LD];
  __shared__ __align__(16) E sAlo[!GEN && SPLIT ? 2 : 1][!GEN && SPLIT ? C::BM * C::LD : 1];
  __shared__ E s7[GEN ? C::BM * 8 : 1];
  __shared__ E s7lo[GEN && SPLIT ? C::BM * 8 : 1];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t row0 = (int64_t)blockIdx.x * C::BM;
  const int64_t col0 = (int64_t)blockIdx.y * C::BN;
  const int Qp = Kp / 7, nsteps = Kp / C::KT;
  if constexpr(GEN) {
    // the group's constant coefficients (stiffness, mass), as k_coef_rows
    // sums them, then each row's 7 values F_e[s] c_s (FP64, then rounded)
    double2 cs = make_double2(0., 0.), cm = make_double2(0., 0.);
    const int nterms = gd->nterms;
    for(int k = 0; k < nterms; ++k) {
      const DevTerm *dt = &gd->terms[k];
      if(dt->kind != 0) continue;
      const double2 c = make_double2(dt->c_re, dt->c_im), f = make_double2(dt->fac_re, dt->fac_im);
      const double2 v = make_double2(f.x * c.x - f.y * c.y, f.x * c.y + f.y * c.x);
      if(dt->form_dof == 1) {
        cs.x += v.x;
        cs.y += v.y;
      }
      else {
        cm.x += v.x;
        cm.y += v.y;
      }
    }
    for(int idx = t; idx < C::BM * 8; idx += kMmThreads) {
      const int r = idx >> 3, s = idx & 7;
      const int64_t row = row0 + r;
      double v = 0.;
      if(s < 7 && row < M) {
        const int64_t e = row / planes;
        const int p = (int)(row - e * planes);
        const double c = s < 6 ? (p == 0 ? cs.x : cs.y) : (p == 0 ? cm.x : cm.y);
        v = F64[e * 8 + s] * c;
      }
      if constexpr(MODE == kMmF64)
        s7[idx] = v;
      else if constexpr(MODE == kMmF32)
        s7[idx] = (float)v;
      else {
        uint16_t hi, lo;
        split_bf16(v, hi, lo);
        s7[idx] = hi;
        s7lo[idx] = lo;
      }
    }
  }
  Acc acc[MB][NB];
#pragma unroll
  for(int a = 0; a < MB; ++a)
#pragma unroll
    for(int b = 0; b < NB; ++b)
#pragma unroll
      for(int r = 0; r < ACC; ++r) acc[a][b][r] = 0;
  // prologue: step 0 into buffer 0
  constexpr int NL = RTileShape<MODE, GEN, KT>::n;
  u32x4 rt[NL], rtl[SPLIT ? NL : 1];
  rtile_load<MODE, GEN, KT>(rt, t, Kp, col0, 0, Rt);
  if constexpr(SPLIT) rtile_load<MODE, GEN, KT>(rtl, t, Kp, col0, 0, Rt_lo);
  rtile_store<MODE, GEN, KT>(rt, t, sR[0]);
  if constexpr(SPLIT) rtile_store<MODE, GEN, KT>(rtl, t, sRlo[0]);
  if constexpr(!GEN) stage_a<MODE>(t, row0, M, Q, Qp, 0, Ain, sA[0], sAlo[0]);
  __syncthreads();
  const int mrow = lane % C::MT;  // operand row / column of this lane inside a micro-tile
  for(int step = 0; step < nsteps; ++step) {
    const int cur = step & 1, k0 = step * C::KT;
    const bool more = step + 1 < nsteps;
    // the next step's tile (the last step reloads its own, stored into the
    // idle buffer and never read): loads without a branch around them
    const int kn = more ? k0 + C::KT : k0;
    rtile_load<MODE, GEN, KT>(rt, t, Kp, col0, kn, Rt);
    if constexpr(SPLIT) rtile_load<MODE, GEN, KT>(rtl, t, Kp, col0, kn, Rt_lo);
    const E *R = sR[cur];
    // a GEN row's A value of this step: component s = k0 / Qp
    const int s = k0 / Qp;
    if constexpr(MODE == kMmF64) {
      // lane l: row / column l % 16, K slot l / 16 takes k0 + 16 j4 + 4 (l / 16) + j
      double a7[MB];
      if constexpr(GEN) {
#pragma unroll
        for(int a = 0; a < MB; ++a) a7[a] = s7[(wm * C::WM + a * 16 + mrow) * 8 + s];
      }
#pragma unroll
      for(int j4 = 0; j4 < C::KT / 16; ++j4) {
        // the lane's 4 consecutive K entries, two 16-byte reads per operand
        const int kk = 16 * j4 + 4 * (lane >> 4);
        double bv[NB][4], av[MB][4];
#pragma unroll
        for(int b = 0; b < NB; ++b) {
          const double *p = &R[(wn * C::WN + b * 16 + mrow) * C::LD + kk];
          const double2 x = *reinterpret_cast<const double2 *>(p), y = *reinterpret_cast<const double2 *>(p + 2);
          bv[b][0] = x.x;
          bv[b][1] = x.y;
          bv[b][2] = y.x;
          bv[b][3] = y.y;
        }
#pragma unroll
        for(int a = 0; a < MB; ++a) {
          if constexpr(GEN) {
            av[a][0] = av[a][1] = av[a][2] = av[a][3] = a7[a];
          }
          else {
            const double *p = &sA[cur][(wm * C::WM + a * 16 + mrow) * C::LD + kk];
            const double2 x = *reinterpret_cast<const double2 *>(p), y = *reinterpret_cast<const double2 *>(p + 2);
            av[a][0] = x.x;
            av[a][1] = x.y;
            av[a][2] = y.x;
            av[a][3] = y.y;
          }
        }
#pragma unroll
        for(int j = 0; j < 4; ++j)
#pragma unroll
          for(int a = 0; a < MB; ++a)
#pragma unroll
            for(int b = 0; b < NB; ++b)
              acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a][j], bv[b][j], acc[a][b], 0, 0, 0);
      }
    }
    else if constexpr(MODE == kMmF32) {
      // lane l: row / column l % 32, K slot l / 32
      const int h = lane >> 5;
      float a7[MB];
      if constexpr(GEN) {
#pragma unroll
        for(int a = 0; a < MB; ++a) a7[a] = s7[(wm * C::WM + a * 32 + mrow) * 8 + s];
      }
      // K slot h of substep j: k0 + (KT / 2) h + j; consecutive MFMAs go to
      // different accumulators
#pragma unroll
      for(int j4 = 0; j4 < C::KT / 8; ++j4) {
        float4 bv[NB], av[MB];
#pragma unroll
        for(int b = 0; b < NB; ++b)
          bv[b] = *reinterpret_cast<const float4 *>(&R[(wn * C::WN + b * 32 + mrow) * C::LD + (C::KT / 2) * h + 4 * j4]);
#pragma unroll
        for(int a = 0; a < MB; ++a) {
          if constexpr(GEN)
            av[a] = make_float4(a7[a], a7[a], a7[a], a7[a]);
          else
            av[a] = *reinterpret_cast<const float4 *>(
              &sA[cur][(wm * C::WM + a * 32 + mrow) * C::LD + (C::KT / 2) * h + 4 * j4]);
        }
#pragma unroll
        for(int j = 0; j < 4; ++j)
#pragma unroll
          for(int a = 0; a < MB; ++a)
#pragma unroll
            for(int b = 0; b < NB; ++b)
              acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[a][j], bv[b][j], acc[a][b], 0, 0, 0);
      }
    }
    else {
      // lane l: row / column l % 32, elements i of k block kb at k0 + 16 kb + 8 (l / 32) + i
      const int h = lane >> 5;
      v8bf a7h[MB], a7l[MB];
      if constexpr(GEN) {
#pragma unroll
        for(int a = 0; a < MB; ++a) {
          const int o = (wm * C::WM + a * 32 + mrow) * 8 + s;
          const __bf16 vh = __builtin_bit_cast(__bf16, s7[o]), vl = __builtin_bit_cast(__bf16, s7lo[o]);
          a7h[a] = v8bf{vh, vh, vh, vh, vh, vh, vh, vh};
          a7l[a] = v8bf{vl, vl, vl, vl, vl, vl, vl, vl};
        }
      }
      const E *Rl = sRlo[cur];
#pragma unroll
      for(int kb = 0; kb < C::KT / 16; ++kb) {
        v8bf bh[NB], bl[NB];
#pragma unroll
        for(int b = 0; b < NB; ++b) {
          const int o = (wn * C::WN + b * 32 + mrow) * C::LD + 16 * kb + 8 * h;
          bh[b] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&R[o]));
          bl[b] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&Rl[o]));
        }
        v8bf ah[MB], al[MB];
#pragma unroll
        for(int a = 0; a < MB; ++a) {
          if constexpr(GEN) {
            ah[a] = a7h[a];
            al[a] = a7l[a];
          }
          else {
            const int o = (wm * C::WM + a * 32 + mrow) * C::LD + 16 * kb + 8 * h;
            ah[a] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&sA[cur][o]));
            al[a] = __builtin_bit_cast(v8bf, *reinterpret_cast<const uint4 *>(&sAlo[cur][o]));
          }
        }
        // hi*hi, lo*hi, hi*lo; consecutive MFMAs go to different accumulators
#pragma unroll
        for(int a = 0; a < MB; ++a)
#pragma unroll
          for(int b = 0; b < NB; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for(int a = 0; a < MB; ++a)
#pragma unroll
          for(int b = 0; b < NB; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[b], acc[a][b], 0, 0, 0);
#pragma unroll
        for(int a = 0; a < MB; ++a)
#pragma unroll
          for(int b = 0; b < NB; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[b], acc[a][b], 0, 0, 0);
      }
    }
    // the buffer written here was last read in the previous step, which
    // every wave finished before the barrier that ended it
    rtile_store<MODE, GEN, KT>(rt, t, sR[cur ^ 1]);
    if constexpr(SPLIT) rtile_store<MODE, GEN, KT>(rtl, t, sRlo[cur ^ 1]);
    if constexpr(!GEN)
      if(more) stage_a<MODE>(t, row0, M, Q, Qp, k0 + C::KT, Ain, sA[cur ^ 1], sAlo[cur ^ 1]);
    __syncthreads();
  }
  Out *Cd = static_cast<Out *>(Cout);
  if constexpr(MODE == kMmF64) {
    // D layout: register r of lane l is row 4 r + l / 16, column l % 16
#pragma unroll
    for(int a = 0; a < MB; ++a)
#pragma unroll
      for(int b = 0; b < NB; ++b) {
        const int64_t col = col0 + wn * C::WN + b * 16 + mrow;
#pragma unroll
        for(int r = 0; r < 4; ++r) {
          const int64_t row = row0 + wm * C::WM + a * 16 + 4 * r + (lane >> 4);
          if(row < M && col < NN) Cd[row * NN + col] = acc[a][b][r];
        }
      }
  }
  else {
    // D layout: register r of lane l is row 8 (r / 4) + 4 (l / 32) + r % 4,
    // column l % 32
    const int h = lane >> 5;
#pragma unroll
    for(int a = 0; a < MB; ++a)
#pragma unroll
      for(int b = 0; b < NB; ++b) {
        const int64_t col = col0 + wn * C::WN + b * 32 + mrow;
#pragma unroll
        for(int r = 0; r < 16; ++r) {
          const int64_t row = row0 + wm * C::WM + a * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
          if(row < M && col < NN) Cd[row * NN + col] = acc[a][b][r];
        }
      }
  }
}

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

}  // namespace

void elem_mm_dims(const Group &g, int64_t &NNp, int64_t &Kp)
{
  NNp = round_up((int64_t)g.N * g.N, 128);
  const int Q = g.rules[g.gemm_rule].Q;
  Kp = 7 * round_up((int64_t)Q, mm_qalign(Q));
}

hipError_t launch_ref_products_t(const Group &g, int mode, void *Rt, uint16_t *Rt_lo, hipStream_t s)
{
  const Rule &R = g.rules[g.gemm_rule];
  int64_t NNp, Kp;
  elem_mm_dims(g, NNp, Kp);
  const int64_t total = NNp * Kp;
  const dim3 grid((unsigned)((total + 255) / 256)), block(256);
  if(mode == kMmF64)
    hipLaunchKernelGGL(k_ref_products_t<kMmF64>, grid, block, 0, s, g.N, R.Q, (int)NNp, (int)Kp, R.d_w, R.d_bf,
                       R.d_gbf, Rt, nullptr);
  else if(mode == kMmF32)
    hipLaunchKernelGGL(k_ref_products_t<kMmF32>, grid, block, 0, s, g.N, R.Q, (int)NNp, (int)Kp, R.d_w, R.d_bf,
                       R.d_gbf, Rt, nullptr);
  else
    hipLaunchKernelGGL(k_ref_products_t<kMmBf16x3>, grid, block, 0, s, g.N, R.Q, (int)NNp, (int)Kp, R.d_w, R.d_bf,
                       R.d_gbf, Rt, Rt_lo);
  return hipGetLastError();
}

hipError_t launch_elem_mm(const Group &g, int mode, bool generated, int planes, const void *Ain, const void *Rt,
                          const void *Rt_lo, void *out, hipStream_t s)
{
  const int64_t M = g.ne * planes;
  if(M == 0) return hipSuccess;
  const int NN = g.N * g.N, Q = g.rules[g.gemm_rule].Q;
  int64_t NNp, Kp;
  elem_mm_dims(g, NNp, Kp);
#define GFEM_MM(MODE, GEN, KT)                                                                                       \
  hipLaunchKernelGGL((k_elem_mm<MODE, GEN, KT>),                                                                     \
                     dim3((unsigned)((M + MmCfg<MODE, GEN, KT>::BM - 1) / MmCfg<MODE, GEN, KT>::BM),                \
                          (unsigned)(NNp / MmCfg<MODE, GEN, KT>::BN)),                                               \
                     dim3(kMmThreads), 0, s, g.d_dev, M, planes, NN, Q, (int)Kp, g.d_F64, Ain, Rt, Rt_lo, out)
  if(mode == kMmF64) {
    if(generated)
      GFEM_MM(kMmF64, true, 16);
    else
      GFEM_MM(kMmF64, false, kMmKT);
  }
  else if(mode == kMmF32) {
    if(generated)
      GFEM_MM(kMmF32, true, kMmKT);
    else
      GFEM_MM(kMmF32, false, kMmKT);
  }
  else {
    if(!generated)
      GFEM_MM(kMmBf16x3, false, kMmKT);
    else if(mm_qalign(Q) == 64)
      GFEM_MM(kMmBf16x3, true, 64);
    else
      GFEM_MM(kMmBf16x3, true, kMmKT);
  }
#undef GFEM_MM
  return hipGetLastError();
}

}  // namespace gfem
