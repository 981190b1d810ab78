// Shifted-Laplacian multigrid preconditioner of the batched Helmholtz solver
// (csrc/mg.hip builds it, csrc/batch.hip applies it).
//
// The preconditioner of A = sum_t c_t X_t (constant-coefficient bilinear
// terms t with unit matrices X_t) is one V-cycle of the complex shifted
// operator  S = A + i beta c_s X_s  (s: the shifted term, the mass term of
// -lap - k^2 in a Helmholtz formulation, so S = K - (1 + i beta) k^2 M - i k R).
// The hierarchy is geometric and built once per system:
//   level 0  the system's DOFs, S assembled per request from A and X_s;
//   level 1  a uniform grid over the DOFs' bounding box with twice the mean DOF
//            spacing, P_0 = trilinear interpolation of grid values at the DOF
//            positions snapped to the lattice of half that spacing (weights
//            1, 1/2, 1/4, 1/8: short rows, small coarse stencils);
//   level l+1  every other grid line of level l (and the last), P_l trilinear;
//   coarsest  at most kMgCoarsest points per axis, solved exactly (dense
//            inverse per request).
// Coarse operators are Galerkin products P^T X_t P of every term class
// (terms of equal coefficient summed), computed once; a request combines them
// with its own coefficients.  Smoothing is one
// damped Jacobi sweep (weight kMgOmega) before and after the coarse
// correction.  Vectors and operators of the cycle are complex64.
#pragma once

#include "internal.h"

namespace gfem {

constexpr int kMgMaxLevels = 8;  // levels including the fine one
constexpr int kMgMaxTerms = 8;   // classes of constant bilinear terms combined per level
constexpr int kMgCoarsest = 5;   // grid points per axis of the coarsest level, at most
constexpr int kMgSmall = 512;    // levels of at most this many rows run inside one block per request
constexpr int kMgSmallThreads = 1024;
constexpr float kMgOmega = 0.5f;  // Jacobi damping

// One level of a request's cycle (device pointers; level 0: the system's
// pattern, its values and right-hand side come from the Krylov vector).
struct MgLevelReq {
  int n;
  const int64_t *ai64;  // level 0 pattern rows
  const int32_t *ai;    // coarse level pattern rows
  const int32_t *aj;
  const float2 *a;      // S on this level
  const float2 *dinv;   // omega / diag(S)
  float2 *b, *x, *r, *y;
  // transfer to the next coarser level: P (rows n, columns of level + 1) and
  // its transpose
  const int32_t *pi, *pj;
  const float *pw;
  const int32_t *ri, *rj;
  const float *rw;
};

// coefficient of each term class of a request, the shifted one times (1 + i beta)
struct Coefs {
  int nt;
  float2 c[kMgMaxTerms];
};

struct MgReq {
  int nlev;    // levels, the last one solved exactly
  int lsmall;  // first level run by the one-block kernel (levels lsmall .. nlev - 1)
  float2 *ainv;  // dense inverse of S on the coarsest level, column-major [n][n]
  // what the inversion at the request's first batch step reads: the
  // system's per-class dense coarsest matrices [nterms][n][n] and this
  // request's coefficients
  const float2 *gdense;
  Coefs cf;
  MgLevelReq lev[kMgMaxLevels];
};

constexpr int kMgInvMax = kMgCoarsest * kMgCoarsest * kMgCoarsest;  // coarsest rows, at most
constexpr int kMgInvThreads = 1024;

// Hierarchy of a system, device-resident (mg.hip).
struct MgSystem {
  int nlev = 0, lsmall = 0;
  int nterms = 0, shift = -1;           // term classes combined, index of the shifted one among them
  // (group, term) of each class's first member: the class's coefficient
  double beta = 0.5;
  int term_group[kMgMaxTerms], term_index[kMgMaxTerms];
  int n[kMgMaxLevels];
  int64_t nnz[kMgMaxLevels];
  // coarse patterns and per-term Galerkin values (levels 1 .. nlev - 2 sparse)
  int32_t *ai[kMgMaxLevels] = {}, *aj[kMgMaxLevels] = {};
  float2 *g[kMgMaxLevels] = {};  // [nterms][nnz]
  float2 *gdense = nullptr;      // coarsest level, [nterms][n][n]
  float2 *xs = nullptr;          // shifted term's unit matrix on level 0 (system pattern)
  int32_t *pi[kMgMaxLevels] = {}, *pj[kMgMaxLevels] = {}, *ri[kMgMaxLevels] = {}, *rj[kMgMaxLevels] = {};
  float *pw[kMgMaxLevels] = {}, *rw[kMgMaxLevels] = {};
};

void mg_free(MgSystem *m);
// bytes of one request's cycle state (vectors, level values, coarse inverse)
size_t mg_request_bytes(const gfem_system *sys);
// lay out a request's cycle state at `base` (mg_request_bytes bytes) and
// fill *req; then mg_prepare computes its level operators from the system's
// current coefficients and the request's assembled FP64 matrix a64 and
// stores the coefficients in *req (the coarsest inverse is formed at the
// request's first batch step, for all requests admitted since the last step
// in one launch)
void mg_bind(const gfem_system *sys, char *base, MgReq *req);
hipError_t mg_prepare(const gfem_system *sys, MgReq &req, const double2 *a64, hipStream_t s);

}  // namespace gfem
