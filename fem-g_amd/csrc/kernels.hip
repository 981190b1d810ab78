// CDNA4 (gfx950) kernels of the assembly hot path.
//
//   k_basis      Lagrange basis / reference gradients at Gauss points
//                (gmsh getBasisFunctions, FunctionSpaceLagrange.cpp:114-137)
//   k_tensor     coefficient-weighted reference tensors of the group's
//                constant-coefficient bilinear terms (the quadrature sum of
//                BilinearTerm::evaluate, BilinearTerm.cpp:315-325, with the
//                affine J^-1 J^-T pulled out of the sum)
//   k_geometry   per-element J, det, J^-1 J^-T in FP64 (gmsh getJacobians +
//                FieldEvaluator<Form1>, FieldEvaluator.h:69)
//   k_linear     per-element load vectors, per-Gauss-point source evaluation
//                (LinearTerm::evaluate, LinearTerm.cpp:176-217)
//   k_dense      per-Gauss-point element matrices (BilinearTerm::evaluate)
//   k_gather     atomic-free CSR assembly: one block per chunk of the
//                nonzero-sorted element map, segmented sum in element order
//                (replaces the binary search + omp atomic of
//                MatrixFactory::addValues, MatrixFactory.cpp:232-309)
//   k_rhs        RHS rows: linear terms + Dirichlet lifting
//                (VectorFactory::addValues / addValuesDC, VectorFactory.cpp:121-235)
//   k_ref_*      the plain reference-order GPU assembler (dense element
//                matrices + atomic scatter) used as the FP64 baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "internal.h"

namespace gfem {

namespace {

// ------------------------------------------------------------------ complex helpers
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 cexp_(double2 a)
{
  const double m = exp(a.x);
  double s, c;
  sincos(a.y, &s, &c);
  return make_double2(m * c, m * s);
}
__device__ __forceinline__ double2 clog_(double2 a) { return make_double2(log(hypot(a.x, a.y)), atan2(a.y, a.x)); }
__device__ __forceinline__ double2 csqrt_(double2 a)
{
  const double r = hypot(a.x, a.y);
  if(r == 0.) return make_double2(0., 0.);
  const double re = sqrt(0.5 * (r + fabs(a.x)));
  const double im = 0.5 * a.y / re;
  return (a.x >= 0.) ? make_double2(re, im) : make_double2(fabs(im), a.y >= 0. ? re : -re);
}

__device__ __forceinline__ double2 cunary(int op, double2 a)
{
  switch(op) {
  case GFEM_OP_NEG: return make_double2(-a.x, -a.y);
  case GFEM_OP_EXP:
    if(a.y == 0.) return make_double2(exp(a.x), 0.);
    return cexp_(a);
  case GFEM_OP_SQRT: return csqrt_(a);
  case GFEM_OP_SIN: {
    double s, c;
    sincos(a.x, &s, &c);
    return a.y == 0. ? make_double2(s, 0.) : make_double2(s * cosh(a.y), c * sinh(a.y));
  }
  case GFEM_OP_COS: {
    double s, c;
    sincos(a.x, &s, &c);
    return a.y == 0. ? make_double2(c, 0.) : make_double2(c * cosh(a.y), -s * sinh(a.y));
  }
  case GFEM_OP_ABS: return make_double2(hypot(a.x, a.y), 0.);
  default: return a;
  }
}

__device__ __forceinline__ double2 cbinary(int op, double2 a, double2 b)
{
  switch(op) {
  case GFEM_OP_ADD: return cadd(a, b);
  case GFEM_OP_SUB: return csub(a, b);
  case GFEM_OP_MUL: return cmul(a, b);
  case GFEM_OP_DIV: return cdiv(a, b);
  case GFEM_OP_POW:
    if(a.y == 0. && b.y == 0. && (a.x > 0. || b.x == floor(b.x))) return make_double2(pow(a.x, b.x), 0.);
    return (a.x == 0. && a.y == 0.) ? make_double2(0., 0.) : cexp_(cmul(b, clog_(a)));
  default: return a;
  }
}

// Coefficient program: complex stack machine (GFEM_OP_*) with the stack in
// eight registers (push shifts down, binary ops shift up).  The program is
// the same for every lane, so the switch is uniform and nothing spills to
// scratch memory.
__device__ __forceinline__ double runary(int op, double a)
{
  switch(op) {
  case GFEM_OP_NEG: return -a;
  case GFEM_OP_EXP: return exp(a);
  case GFEM_OP_SIN: return sin(a);
  case GFEM_OP_COS: return cos(a);
  case GFEM_OP_ABS: return fabs(a);
  default: return a;
  }
}

__device__ __forceinline__ double rbinary(int op, double a, double b)
{
  switch(op) {
  case GFEM_OP_ADD: return a + b;
  case GFEM_OP_SUB: return a - b;
  case GFEM_OP_MUL: return a * b;
  case GFEM_OP_DIV: return a / b;
  default: return a;
  }
}

// real-arithmetic evaluation of a real-only program (same stack discipline)
__device__ __forceinline__ double eval_expr_real(const DevExpr *__restrict__ ex, double x, double y, double z)
{
  double s0 = 0., s1 = 0., s2 = 0., s3 = 0., s4 = 0., s5 = 0., s6 = 0., s7 = 0.;
  const int n = __builtin_amdgcn_readfirstlane(ex->n_ops);
  for(int k = 0; k < n; ++k) {
    // the program is the same for every lane: keep op / arg in scalar registers
    const int op = __builtin_amdgcn_readfirstlane(ex->ops[2 * k]);
    const int arg = __builtin_amdgcn_readfirstlane(ex->ops[2 * k + 1]);
    if(op <= GFEM_OP_Z) {
      const double v = op == GFEM_OP_CONST ? ex->consts[2 * arg] : (op == GFEM_OP_X ? x : (op == GFEM_OP_Y ? y : z));
      s7 = s6; s6 = s5; s5 = s4; s4 = s3; s3 = s2; s2 = s1; s1 = s0; s0 = v;
    }
    else if(op >= GFEM_OP_ADD && op <= GFEM_OP_DIV) {
      s0 = rbinary(op, s1, s0);
      s1 = s2; s2 = s3; s3 = s4; s4 = s5; s5 = s6; s6 = s7;
    }
    else {
      s0 = runary(op, s0);
    }
  }
  return s0;
}

__device__ __forceinline__ double2 eval_expr(const DevExpr *__restrict__ ex, double x, double y, double z)
{
  if(__builtin_amdgcn_readfirstlane(ex->real_only)) return make_double2(eval_expr_real(ex, x, y, z), 0.);
  double2 s0 = make_double2(0., 0.), s1 = s0, s2 = s0, s3 = s0, s4 = s0, s5 = s0, s6 = s0, s7 = s0;
  const int n = __builtin_amdgcn_readfirstlane(ex->n_ops);
  for(int k = 0; k < n; ++k) {
    const int op = __builtin_amdgcn_readfirstlane(ex->ops[2 * k]);
    const int arg = __builtin_amdgcn_readfirstlane(ex->ops[2 * k + 1]);
    if(op <= GFEM_OP_Z) {
      double2 v;
      if(op == GFEM_OP_CONST)
        v = make_double2(ex->consts[2 * arg], ex->consts[2 * arg + 1]);
      else
        v = make_double2(op == GFEM_OP_X ? x : (op == GFEM_OP_Y ? y : z), 0.);
      s7 = s6; s6 = s5; s5 = s4; s4 = s3; s3 = s2; s2 = s1; s1 = s0; s0 = v;
    }
    else if((op >= GFEM_OP_ADD && op <= GFEM_OP_DIV) || op == GFEM_OP_POW) {
      s0 = cbinary(op, s1, s0);
      s1 = s2; s2 = s3; s3 = s4; s4 = s5; s5 = s6; s6 = s7;
    }
    else {
      s0 = cunary(op, s0);
    }
  }
  return s0;
}

// Affine element geometry from its vertices: J columns (col-major
// J(i,j) = dx_i/du_j on the gmsh reference element), det, Kinv = J^-T
// (row-major).  1-D/2-D elements get unit normals as missing columns and
// det = length/area element, as gmsh getJacobians does.
struct Geo {
  double c[3][3];  // c[j] = column j of J
  double det;
  double Ki[3][3];
};

__device__ __forceinline__ void cross(const double *a, const double *b, double *r)
{
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// det and Kinv from the first `dim` columns of J (c), completing the missing
// columns with unit normals
__device__ __forceinline__ void geometry_complete(int dim, Geo &g)
{
  if(dim == 3) {
    double n[3];
    cross(g.c[1], g.c[2], n);
    g.det = g.c[0][0] * n[0] + g.c[0][1] * n[1] + g.c[0][2] * n[2];
  }
  else if(dim == 2) {
    double n[3];
    cross(g.c[0], g.c[1], n);
    const double a = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for(int i = 0; i < 3; ++i) g.c[2][i] = n[i] / a;
    g.det = a;
  }
  else {
    const double a = sqrt(g.c[0][0] * g.c[0][0] + g.c[0][1] * g.c[0][1] + g.c[0][2] * g.c[0][2]);
    double t[3] = {g.c[0][0] / a, g.c[0][1] / a, g.c[0][2] / a};
    double z[3] = {0., 0., 1.};
    if(fabs(t[2]) > 0.9) {
      z[0] = 1.;
      z[2] = 0.;
    }
    double n1[3], n2[3];
    cross(z, t, n1);
    const double l1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
    for(int i = 0; i < 3; ++i) n1[i] /= l1;
    cross(t, n1, n2);
    for(int i = 0; i < 3; ++i) {
      g.c[1][i] = n1[i];
      g.c[2][i] = n2[i];
    }
    g.det = a;
  }
  // A = J^T: A[i][j] = c[i][j]; Ki = A^-1 by cofactors
  const double(*A)[3] = g.c;
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double d = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const double id = 1. / d;
  g.Ki[0][0] = c00 * id;
  g.Ki[1][0] = c01 * id;
  g.Ki[2][0] = c02 * id;
  g.Ki[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  g.Ki[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  g.Ki[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  g.Ki[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  g.Ki[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  g.Ki[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
}

// affine element: J's columns from the vertices (the gmsh line is [-1, 1])
__device__ __forceinline__ void element_geometry(int dim, const double X[4][3], Geo &g)
{
  // constant indices after unrolling (a runtime bound would put c in scratch)
#pragma unroll
  for(int j = 0; j < 3; ++j)
#pragma unroll
    for(int i = 0; i < 3; ++i)
      g.c[j][i] = j >= dim ? 0. : (dim == 1 ? 0.5 * (X[1][i] - X[0][i]) : X[j + 1][i] - X[0][i]);
  geometry_complete(dim, g);
}

// curved element at one Gauss point (isoparametric, the element's own
// Lagrange basis): c_j = sum_n X_n dpsi_n/du_j, x = sum_n psi_n X_n
// (gmsh getJacobians at the point), from all N nodes of the element
__device__ __forceinline__ void curved_geometry(int dim, int N, const int32_t *__restrict__ conn, int64_t e,
                                                const double *__restrict__ xyz, const double *__restrict__ bf_q,
                                                const double *__restrict__ gbf_q, Geo &g, double *x)
{
#pragma unroll
  for(int j = 0; j < 3; ++j)
#pragma unroll
    for(int i = 0; i < 3; ++i) g.c[j][i] = 0.;
  x[0] = x[1] = x[2] = 0.;
  for(int n = 0; n < N; ++n) {
    const int64_t v = conn[e * N + n];
    const double X[3] = {xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]};
    const double b = bf_q[n];
#pragma unroll
    for(int i = 0; i < 3; ++i) x[i] += b * X[i];
#pragma unroll
    for(int j = 0; j < 3; ++j) {
      const double d = j < dim ? gbf_q[3 * n + j] : 0.;  // gradients hold 3 components
#pragma unroll
      for(int i = 0; i < 3; ++i) g.c[j][i] += d * X[i];
    }
  }
  geometry_complete(dim, g);
}

// the 7 factors det G_00, G_11, G_22, G_01, G_02, G_12, det (G = J^-1 J^-T)
__device__ __forceinline__ void geometry_factors(const Geo &g, double f[7])
{
  double G[3][3];
  for(int a = 0; a < 3; ++a)
    for(int b = 0; b < 3; ++b) G[a][b] = g.Ki[0][a] * g.Ki[0][b] + g.Ki[1][a] * g.Ki[1][b] + g.Ki[2][a] * g.Ki[2][b];
  const double d = g.det;
  f[0] = d * G[0][0];
  f[1] = d * G[1][1];
  f[2] = d * G[2][2];
  f[3] = d * G[0][1];
  f[4] = d * G[0][2];
  f[5] = d * G[1][2];
  f[6] = d;
}

__device__ __forceinline__ void load_vertices(const int32_t *__restrict__ vtx, int nv, int64_t e,
                                              const double *__restrict__ xyz, double X[4][3])
{
  for(int k = 0; k < 4; ++k) {
    if(k < nv) {
      const int64_t n = vtx[e * nv + k];
      X[k][0] = xyz[3 * n];
      X[k][1] = xyz[3 * n + 1];
      X[k][2] = xyz[3 * n + 2];
    }
    else {
      X[k][0] = X[k][1] = X[k][2] = 0.;
    }
  }
}

// physical point of reference point u (affine map)
__device__ __forceinline__ void phys_point(int dim, const double X[4][3], const double *u, double *x)
{
  if(dim == 1) {
    const double t = 0.5 * (u[0] + 1.);
    for(int i = 0; i < 3; ++i) x[i] = X[0][i] + t * (X[1][i] - X[0][i]);
    return;
  }
  for(int i = 0; i < 3; ++i) {
    double s = X[0][i];
    for(int j = 0; j < dim; ++j) s += u[j] * (X[j + 1][i] - X[0][i]);
    x[i] = s;
  }
}

// Work-groups are dispatched round-robin over the XCDs, each with its own
// L2.  remap() sends block b to work item (b mod X) * per + b / X, so each XCD
// walks a contiguous range of items: neighbouring gather chunks and RHS rows
// share elements, which then hit in that XCD's L2.
constexpr int kXcds = 8;

__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t nb)
{
  const int64_t per = nb / kXcds;  // nb is a multiple of kXcds
  return (b % kXcds) * per + b / kXcds;
}

inline int64_t xcd_grid(int64_t n) { return (n + kXcds - 1) / kXcds * kXcds; }

// ------------------------------------------------------------------ k_basis
// Lagrange basis on the gmsh reference simplex, lattice node a with
// barycentric multi-index m: phi = prod_k prod_{l<m_k} (p lam_k - l)/(l+1).
__global__ void k_basis(int dim, int p, int N, int Q, const int *__restrict__ nodes, const double *__restrict__ pts,
                        double *__restrict__ bf, double *__restrict__ gbf)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= N * Q) return;
  const int q = t / N, a = t % N;
  const double *u = &pts[3 * q];
  double lam[4] = {0., 0., 0., 0.}, dlam[4][3] = {{0.}};
  if(dim == 1) {
    lam[0] = 0.5 * (1. - u[0]);
    lam[1] = 0.5 * (1. + u[0]);
    dlam[0][0] = -0.5;
    dlam[1][0] = 0.5;
  }
  else {
    double s = 0.;
    for(int j = 0; j < dim; ++j) {
      lam[j + 1] = u[j];
      dlam[j + 1][j] = 1.;
      dlam[0][j] = -1.;
      s += u[j];
    }
    lam[0] = 1. - s;
  }
  double f = 1., g[3] = {0., 0., 0.};
  for(int k = 0; k <= dim; ++k) {
    const int mk = nodes[4 * a + k];
    for(int l = 0; l < mk; ++l) {
      const double fac = (p * lam[k] - l) / (l + 1);
      const double dfac = (double)p / (l + 1);
      for(int j = 0; j < 3; ++j) g[j] = g[j] * fac + f * dfac * dlam[k][j];
      f *= fac;
    }
  }
  bf[q * N + a] = f;
  for(int j = 0; j < 3; ++j) gbf[(q * N + a) * 3 + j] = g[j];
}

// ------------------------------------------------------------------ k_tensor
// T[pair][c], pair = i_tf*N + j_dof:
//   c=0..2  sum_q w dphi_j/du_c dphi_i/du_c      c=3..5 symmetric cross terms
//   c=6     sum_q w phi_j phi_i
// weighted by each term's constant factor and summed over terms.
__device__ __forceinline__ void tensor_pair(int pair, const TensorArgs &a, double2 *__restrict__ T64,
                                            float2 *__restrict__ T32, double *__restrict__ T64r, float *__restrict__ T32r)
{
  const int i = pair / a.N, j = pair % a.N;
  double2 acc[8];
  for(int c = 0; c < 8; ++c) acc[c] = make_double2(0., 0.);
  for(int t = 0; t < a.nterms; ++t) {
    if(!a.active[t]) continue;
    const int r = a.rule[t];
    double s[7] = {0., 0., 0., 0., 0., 0., 0.};
    for(int q = 0; q < a.Q[r]; ++q) {
      const double w = a.w[r][q];
      if(a.form_dof[t] == 1) {
        const double *gj = &a.gbf[r][(q * a.N + j) * 3];
        const double *gi = &a.gbf[r][(q * a.N + i) * 3];
        s[0] += w * gj[0] * gi[0];
        s[1] += w * gj[1] * gi[1];
        s[2] += w * gj[2] * gi[2];
        s[3] += w * (gj[0] * gi[1] + gj[1] * gi[0]);
        s[4] += w * (gj[0] * gi[2] + gj[2] * gi[0]);
        s[5] += w * (gj[1] * gi[2] + gj[2] * gi[1]);
      }
      else {
        s[6] += w * a.bf[r][q * a.N + j] * a.bf[r][q * a.N + i];
      }
    }
    for(int c = 0; c < 7; ++c) acc[c] = cadd(acc[c], make_double2(a.fac[t].x * s[c], a.fac[t].y * s[c]));
  }
  for(int c = 0; c < 8; ++c) {
    T64[pair * 8 + c] = acc[c];
    T32[pair * 8 + c] = make_float2((float)acc[c].x, (float)acc[c].y);
    T64r[pair * 8 + c] = acc[c].x;
    T32r[pair * 8 + c] = (float)acc[c].x;
  }
}

__global__ void k_tensor(TensorArgs a, double2 *__restrict__ T64, float2 *__restrict__ T32, double *__restrict__ T64r,
                         float *__restrict__ T32r)
{
  const int pair = blockIdx.x * blockDim.x + threadIdx.x;
  if(pair >= a.N * a.N) return;
  tensor_pair(pair, a, T64, T32, T64r, T32r);
}

// the tensors of several groups in one launch: item blockIdx.y
__global__ void k_tensor_many(const TensorItem *__restrict__ items)
{
  const TensorItem &it = items[blockIdx.y];
  const int pair = blockIdx.x * blockDim.x + threadIdx.x;
  if(pair >= it.a.N * it.a.N) return;
  tensor_pair(pair, it.a, it.T64, it.T32, it.T64r, it.T32r);
}

// ------------------------------------------------------------------ k_geometry
__device__ __forceinline__ void geometry_one(int dim, int nv, int absdet, int64_t e, const int32_t *__restrict__ vtx,
                                             const double *__restrict__ xyz, double *__restrict__ F64,
                                             float *__restrict__ F32)
{
  double X[4][3];
  load_vertices(vtx, nv, e, xyz, X);
  Geo g;
  element_geometry(dim, X, g);
  if(absdet) g.det = fabs(g.det);
  double f[8];
  geometry_factors(g, f);
  f[7] = 0.;
  double2 *o2 = reinterpret_cast<double2 *>(&F64[e * 8]);
  for(int k = 0; k < 4; ++k) o2[k] = make_double2(f[2 * k], f[2 * k + 1]);
  if(F32) {  // the FP32 copy only for mixed-precision assemblies
    float4 *o4 = reinterpret_cast<float4 *>(&F32[e * 8]);
    o4[0] = make_float4((float)f[0], (float)f[1], (float)f[2], (float)f[3]);
    o4[1] = make_float4((float)f[4], (float)f[5], (float)f[6], 0.f);
  }
}

__global__ void __launch_bounds__(256) k_geometry(int dim, int nv, int absdet, int64_t ne,
                                                  const int32_t *__restrict__ vtx, const double *__restrict__ xyz,
                                                  double *__restrict__ F64, float *__restrict__ F32)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(e < ne) geometry_one(dim, nv, absdet, e, vtx, xyz, F64, F32);
}

// every group of a system in one launch: thread t takes element t - off[g]
// of the group g with off[g] <= t < off[g + 1] (groups are few and
// contiguous, so the scan is short and uniform within most waves)
struct GeomGroups {
  int ng;
  int dim[kMaxGroups], nv[kMaxGroups], absdet[kMaxGroups];
  int64_t off[kMaxGroups + 1];
  const int32_t *vtx[kMaxGroups];
  double *F64[kMaxGroups];
  float *F32[kMaxGroups];
};

// A block whose elements all belong to one group stages its factors in LDS
// and writes them as one contiguous run (each store instruction then covers
// whole cache lines, where a thread-per-element store of a 64-byte record
// touches 64 lines); a block that straddles two groups stores directly.
__global__ void __launch_bounds__(256) k_geometry_all(const GeomGroups gg, const double *__restrict__ xyz)
{
  constexpr int kB = 256, kPad = 5;  // double2 per element in LDS (4 + 1 padding)
  __shared__ double2 s_f[kB * kPad];
  const int64_t total = gg.off[gg.ng];
  const int64_t t0 = (int64_t)blockIdx.x * kB;
  const int64_t t = t0 + threadIdx.x;
  int g0 = 0;
  while(g0 + 1 < gg.ng && t0 >= gg.off[g0 + 1]) ++g0;
  const int64_t tend = min(t0 + kB, total);
  if(tend > gg.off[g0 + 1]) {  // uniform across the block
    if(t >= total) return;
    int g = g0;
    while(g + 1 < gg.ng && t >= gg.off[g + 1]) ++g;
    geometry_one(gg.dim[g], gg.nv[g], gg.absdet[g], t - gg.off[g], gg.vtx[g], xyz, gg.F64[g], gg.F32[g]);
    return;
  }
  if(t < tend) {
    const int64_t e = t - gg.off[g0];
    double X[4][3];
    load_vertices(gg.vtx[g0], gg.nv[g0], e, xyz, X);
    Geo g;
    element_geometry(gg.dim[g0], X, g);
    if(gg.absdet[g0]) g.det = fabs(g.det);
    double f[8];
    geometry_factors(g, f);
    f[7] = 0.;
#pragma unroll
    for(int k = 0; k < 4; ++k) s_f[threadIdx.x * kPad + k] = make_double2(f[2 * k], f[2 * k + 1]);
  }
  __syncthreads();
  const int cnt = (int)(tend - t0);
  const int64_t e0 = t0 - gg.off[g0];
  double2 *o2 = reinterpret_cast<double2 *>(gg.F64[g0] + e0 * 8);
  for(int p = threadIdx.x; p < 4 * cnt; p += kB) o2[p] = s_f[(p >> 2) * kPad + (p & 3)];
  if(gg.F32[g0]) {  // the FP32 copy only for mixed-precision assemblies
    float4 *o4 = reinterpret_cast<float4 *>(gg.F32[g0] + e0 * 8);
    for(int p = threadIdx.x; p < 2 * cnt; p += kB) {
      const double2 a = s_f[(p >> 1) * kPad + 2 * (p & 1)], b = s_f[(p >> 1) * kPad + 2 * (p & 1) + 1];
      o4[p] = make_float4((float)a.x, (float)a.y, (float)b.x, (float)b.y);
    }
  }
}

// ------------------------------------------------------------------ k_linear
// One linear term on one group, one thread per element:
//   b_e(i) (+)= sum_q w_q det conj(c_tf) phi_i(q) f(x_q)
// The coefficient program travels in the kernel arguments (scalar loads,
// uniform across the wave); the rule's weights, points and basis values are
// staged in LDS once per block.
struct LinTerm {
  DevExpr expr;
  double fac_re, fac_im, c_re, c_im;
  int per_point, Q, dim, nv, accumulate;
  int real_out;  // element loads stored as double (every linear term of the system real)
  const int32_t *conn;  // curved group: node indices [ne][N] (isoparametric points), else null
  const double *gbf;    // basis gradients [Q][N][3] for the curved map
};

// element load k after a group's base: complex (double2 units), or real
// (double units, imaginary part 0)
__device__ __forceinline__ double2 be_load(const void *be, int real_out, int64_t k)
{
  return real_out ? make_double2(reinterpret_cast<const double *>(be)[k], 0.)
                  : reinterpret_cast<const double2 *>(be)[k];
}
__device__ __forceinline__ void be_store(void *be, int real_out, int64_t k, double2 v)
{
  if(real_out)
    reinterpret_cast<double *>(be)[k] = v.x;
  else
    reinterpret_cast<double2 *>(be)[k] = v;
}

template <int N, bool STAGED>
__global__ void __launch_bounds__(128) k_linear(LinTerm a, const double *__restrict__ w, const double *__restrict__ pts,
                                                const double *__restrict__ bf, int64_t ne,
                                                const int32_t *__restrict__ vtx, const double *__restrict__ xyz,
                                                const double *__restrict__ F64, void *__restrict__ be)
{
  extern __shared__ double s_tab[];
  const int Q = a.Q;
  const double *tw = w, *tp = pts, *tb = bf;
  if(STAGED) {
    double *lw = s_tab, *lp = s_tab + Q, *lb = s_tab + 4 * Q;
    for(int k = threadIdx.x; k < Q; k += blockDim.x) lw[k] = w[k];
    for(int k = threadIdx.x; k < 3 * Q; k += blockDim.x) lp[k] = pts[k];
    for(int k = threadIdx.x; k < Q * N; k += blockDim.x) lb[k] = bf[k];
    __syncthreads();
    tw = lw;
    tp = lp;
    tb = lb;
  }
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= ne) return;
  double X[4][3];
  load_vertices(vtx, a.nv, e, xyz, X);
  const double det = F64[e * 8 + 6];
  double2 acc[N];
#pragma unroll
  for(int i = 0; i < N; ++i) acc[i] = a.accumulate ? be_load(be, a.real_out, e * N + i) : make_double2(0., 0.);
  const double2 fac = make_double2(a.fac_re, a.fac_im), cst = make_double2(a.c_re, a.c_im);
  for(int q = 0; q < Q; ++q) {
    double2 f = cst;
    double dq = det;
    if(a.conn) {  // curved group: this point's position and det
      Geo gq;
      double x[3];
      curved_geometry(a.dim, N, a.conn, e, xyz, &tb[q * N], &a.gbf[3 * q * N], gq, x);
      dq = gq.det;
      if(a.per_point) f = eval_expr(&a.expr, x[0], x[1], x[2]);
    }
    else if(a.per_point) {
      double x[3];
      phys_point(a.dim, X, &tp[3 * q], x);
      f = eval_expr(&a.expr, x[0], x[1], x[2]);
    }
    const double wd = tw[q] * dq;
    const double2 cf = cmul(make_double2(fac.x * wd, fac.y * wd), f);
#pragma unroll
    for(int i = 0; i < N; ++i) {
      const double b = tb[q * N + i];
      acc[i].x += cf.x * b;
      acc[i].y += cf.y * b;
    }
  }
#pragma unroll
  for(int i = 0; i < N; ++i) be_store(be, a.real_out, e * N + i, acc[i]);
}

// Point-parallel form of the same term.  A block takes EB consecutive
// elements: vertex coordinates and determinants are staged in LDS once, one
// thread per (element, Gauss point) evaluates c_q = w_q det conj(c_tf) f(x_q)
// into LDS, then one thread per (element, local dof) forms
// b_e(i) (+)= sum_q phi_i(q) c_q.  Same layout as the compiled kernel.
__global__ void __launch_bounds__(256) k_linear_pp(LinTerm a, int N, int EB, const double *__restrict__ w,
                                                   const double *__restrict__ pts, const double *__restrict__ bf,
                                                   int64_t ne, const int32_t *__restrict__ vtx,
                                                   const double *__restrict__ xyz, const double *__restrict__ F64,
                                                   void *__restrict__ be)
{
  extern __shared__ double s_dyn[];
  const int Q = a.Q, nv = a.nv;
  double *s_bf = s_dyn;           // [Q][N]
  double *s_u = s_bf + Q * N;     // [Q][3]
  double *s_w = s_u + 3 * Q;      // [Q]
  double *s_X = s_w + Q;          // [EB][4][3]
  double *s_det = s_X + 12 * EB;  // [EB]
  double2 *s_c = reinterpret_cast<double2 *>(s_det + EB + ((Q * N + 4 * Q + 13 * EB) & 1));  // [EB][Q]
  const int t = threadIdx.x;
  const int64_t e0 = (int64_t)blockIdx.x * EB;
  const int nb = (int)min((int64_t)EB, ne - e0);
  for(int k = t; k < Q * N; k += 256) s_bf[k] = bf[k];
  for(int k = t; k < 3 * Q; k += 256) s_u[k] = pts[k];
  for(int k = t; k < Q; k += 256) s_w[k] = w[k];
  for(int k = t; k < nb * nv * 3; k += 256) {
    const int le = k / (3 * nv), r = k - le * 3 * nv, v = r / 3, c = r - 3 * v;
    s_X[12 * le + r] = xyz[3 * (int64_t)vtx[(e0 + le) * nv + v] + c];
  }
  for(int k = t; k < nb; k += 256) s_det[k] = F64[(e0 + k) * 8 + 6];
  __syncthreads();
  for(int p = t; p < nb * Q; p += 256) {
    const int le = p / Q, q = p - le * Q;
    double2 f = make_double2(a.c_re, a.c_im);
    if(a.per_point) {
      double X[4][3];
      for(int v = 0; v < 4; ++v)
        for(int i = 0; i < 3; ++i) X[v][i] = v < nv ? s_X[12 * le + 3 * v + i] : 0.;
      double x[3];
      phys_point(a.dim, X, &s_u[3 * q], x);
      f = eval_expr(&a.expr, x[0], x[1], x[2]);
    }
    const double wd = s_w[q] * s_det[le];
    s_c[p] = cmul(make_double2(a.fac_re * wd, a.fac_im * wd), f);
  }
  __syncthreads();
  for(int k = t; k < nb * N; k += 256) {
    const int le = k / N, i = k - le * N;
    const int64_t o = (e0 + le) * N + i;
    double2 acc = a.accumulate ? be_load(be, a.real_out, o) : make_double2(0., 0.);
    const double2 *c = &s_c[le * Q];
    for(int q = 0; q < Q; ++q) {
      const double b = s_bf[q * N + i];
      acc.x += c[q].x * b;
      acc.y += c[q].y * b;
    }
    be_store(be, a.real_out, o, acc);
  }
}

// ------------------------------------------------------------------ k_dense
// one thread per (element, pair): the reference's per-Gauss-point loop
// A_e(j,i) += w det (c E_j) . conj(c_tf E_i)
__global__ void __launch_bounds__(256) k_dense(const GroupDev *__restrict__ gd, int all_terms, int64_t ne,
                                               const int32_t *__restrict__ vtx, const double *__restrict__ xyz,
                                               int planes, double *__restrict__ Ae)
{
  const int N = gd->N;
  const int NN = N * N;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= ne * NN) return;
  const int64_t e = t / NN;
  const int pair = (int)(t % NN);
  const int i = pair / N, j = pair % N;
  const int dim = gd->dim;
  const int32_t *conn = gd->conn;  // curved group: geometry at every point
  double X[4][3];
  load_vertices(vtx, gd->nv, e, xyz, X);
  Geo g;
  if(!conn) {
    element_geometry(dim, X, g);
    if(gd->absdet) g.det = fabs(g.det);
  }
  double2 acc = make_double2(0., 0.);
  const int nterms = __builtin_amdgcn_readfirstlane(gd->nterms);
  for(int k = 0; k < nterms; ++k) {
    const DevTerm *dt = &gd->terms[k];
    if(__builtin_amdgcn_readfirstlane(dt->kind) != 0) continue;
    const bool pp = __builtin_amdgcn_readfirstlane(dt->per_point) != 0;
    if(!all_terms && !pp) continue;
    const RuleDev &R = gd->rules[__builtin_amdgcn_readfirstlane(dt->rule)];
    const int Q = __builtin_amdgcn_readfirstlane(R.Q);
    const int fd = __builtin_amdgcn_readfirstlane(dt->form_dof);
    double2 sum = make_double2(0., 0.);
    for(int q = 0; q < Q; ++q) {
      Geo G;
      double xq[3];
      if(conn)
        curved_geometry(dim, N, conn, e, xyz, &R.bf[q * N], &R.gbf[3 * q * N], G, xq);
      else
        G = g;
      double s;
      if(fd == 1) {
        const double *gj = &R.gbf[(q * N + j) * 3];
        const double *gi = &R.gbf[(q * N + i) * 3];
        s = 0.;
        for(int c = 0; c < 3; ++c) {
          const double ej = G.Ki[c][0] * gj[0] + G.Ki[c][1] * gj[1] + G.Ki[c][2] * gj[2];
          const double ei = G.Ki[c][0] * gi[0] + G.Ki[c][1] * gi[1] + G.Ki[c][2] * gi[2];
          s += ej * ei;
        }
      }
      else {
        s = R.bf[q * N + j] * R.bf[q * N + i];
      }
      double2 c = make_double2(dt->c_re, dt->c_im);
      if(pp) {
        double x[3];
        if(conn) {
          x[0] = xq[0];
          x[1] = xq[1];
          x[2] = xq[2];
        }
        else
          phys_point(dim, X, &R.pts[3 * q], x);
        c = eval_expr(&dt->expr, x[0], x[1], x[2]);
      }
      const double wd = R.w[q] * G.det * s;
      sum.x += wd * c.x;
      sum.y += wd * c.y;
    }
    acc = cadd(acc, cmul(make_double2(dt->fac_re, dt->fac_im), sum));
  }
  // planes layout: [e][real, imaginary][N*N]
  Ae[e * planes * NN + pair] = acc.x;
  if(planes == 2) Ae[e * planes * NN + NN + pair] = acc.y;
}

__global__ void __launch_bounds__(256) k_dense_to32(int64_t total, const double *__restrict__ a, float *__restrict__ b)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < total) b[i] = (float)a[i];
}

// ------------------------------------------------------------------ GEMM form of the per-point terms
// For affine elements, A_e(j,i) = sum_q sum_s A_e(q,s) R(q,s; i*N+j) with
//   R(q,s; pair) = w_q * (d_a phi_j d_b phi_i [+ d_b phi_j d_a phi_i]) for the
//                  6 components s = (00,11,22,01,02,12) of G = J^-1 J^-T,
//                  w_q * phi_j phi_i for s = 6 (mass),
//   A_e(q,s)     = det G_s * sum over stiffness terms of fac c(x_q), and
//                  det * sum over mass terms of fac c(x_q) for s = 6,
// the same quantity as the reference's per-point B^T W B
// (fem-master/src/term/BilinearTerm.cpp:303-343): the element matrices are
// one GEMM  dense[e*planes + p][pair] = sum_k A[e*planes + p][k] R[k][pair],
// computed on the matrix cores by csrc/elemmm.hip.  k_coef_rows writes the
// rows of A where that kernel cannot generate them itself.
//
// One thread per (element, Gauss point): the 7 entries of A for that point,
// per plane (real part, then imaginary part for complex systems).  A block's
// rows are staged in LDS and written back with consecutive lanes on
// consecutive addresses (a block's points are consecutive entries of each
// row of A).  PP: some coefficient is an expression of x (without it the
// kernel leaves out the point coordinates and the interpreter).
constexpr int kCoefBlock = 256;

template <typename T, bool PP>
__global__ void __launch_bounds__(kCoefBlock) k_coef_rows(const GroupDev *__restrict__ gd, int all_terms, int rule,
                                                          int64_t ne, const int32_t *__restrict__ vtx,
                                                          const double *__restrict__ xyz,
                                                          const double *__restrict__ F64, int planes,
                                                          T *__restrict__ A)
{
  __shared__ T st[2][kCoefBlock * 7];
  const RuleDev &R = gd->rules[rule];
  const int Q = R.Q;
  const int64_t total = ne * Q, t0 = (int64_t)blockIdx.x * kCoefBlock;
  // point t0 + i is (e0 + (q0 + i) / Q, (q0 + i) % Q): one wide division per block
  const int64_t e0 = t0 / Q;
  const uint32_t q0 = (uint32_t)(t0 - e0 * Q);
  if(t0 + threadIdx.x < total) {
    const uint32_t r = q0 + threadIdx.x, de = r / (uint32_t)Q;
    const int64_t e = e0 + de;
    const int q = (int)(r - de * (uint32_t)Q);
    double x[3] = {0., 0., 0.};
    double F[7];  // det G00, G11, G22, G01, G02, G12, det
    if(gd->conn) {  // curved group: this point's own Jacobian and position
      const int N = gd->N;
      Geo gq;
      curved_geometry(gd->dim, N, gd->conn, e, xyz, &R.bf[q * N], &R.gbf[3 * q * N], gq, x);
      geometry_factors(gq, F);
    }
    else {
#pragma unroll
      for(int s = 0; s < 7; ++s) F[s] = F64[e * 8 + s];
    }
    if constexpr(PP) {
      if(!gd->conn) {
        double X[4][3];
        load_vertices(vtx, gd->nv, e, xyz, X);
        phys_point(gd->dim, X, &R.pts[3 * q], x);
      }
    }
    double2 stiff = make_double2(0., 0.), mass = make_double2(0., 0.);
    const int nterms = __builtin_amdgcn_readfirstlane(gd->nterms);
    for(int k = 0; k < nterms; ++k) {
      const DevTerm *dt = &gd->terms[k];
      if(__builtin_amdgcn_readfirstlane(dt->kind) != 0) continue;
      const bool pp = PP && __builtin_amdgcn_readfirstlane(dt->per_point) != 0;
      if(!all_terms && !pp) continue;
      double2 c = make_double2(dt->c_re, dt->c_im);
      if constexpr(PP)
        if(pp) c = eval_expr(&dt->expr, x[0], x[1], x[2]);
      const double2 v = cmul(make_double2(dt->fac_re, dt->fac_im), c);
      if(__builtin_amdgcn_readfirstlane(dt->form_dof) == 1)
        stiff = cadd(stiff, v);
      else
        mass = cadd(mass, v);
    }
    for(int p = 0; p < planes; ++p) {
      const double cs = p == 0 ? stiff.x : stiff.y, cm = p == 0 ? mass.x : mass.y;
#pragma unroll
      for(int s = 0; s < 7; ++s) {
        st[p][threadIdx.x * 7 + s] = (T)(F[s] * (s < 6 ? cs : cm));
      }
    }
  }
  __syncthreads();
  const int npts = (int)min((int64_t)kCoefBlock, total - t0);
  const int64_t K = (int64_t)7 * Q;
  for(int p = 0; p < planes; ++p)
    for(int j = threadIdx.x; j < 7 * npts; j += kCoefBlock) {
      const uint32_t tl = (uint32_t)j / 7u, r = q0 + tl, de = r / (uint32_t)Q;
      const int64_t e = e0 + de;
      const int64_t col = (int64_t)(r - de * (uint32_t)Q) * 7 + (j - 7 * (int)tl);
      A[(e * planes + p) * K + col] = st[p][j];
    }
}

// ------------------------------------------------------------------ element value
__device__ __forceinline__ int find_group(const GroupTable &gt, int64_t eg)
{
  int g = 0;
  while(g + 1 < gt.ng && eg >= gt.e_off[g + 1]) ++g;
  return g;
}

// value of local entry `pair` of global element slot `eg`
template <bool CPLX>
__device__ __forceinline__ double2 elem_value64(const GroupTable &gt, int tensor, int64_t eg, int pair)
{
  const int g = find_group(gt, eg);
  const int64_t el = eg - gt.e_off[g];
  double2 v = make_double2(0., 0.);
  if(tensor && gt.T64[g]) {  // curved groups have no reference tensor
    const double2 *F = reinterpret_cast<const double2 *>(gt.F64[g] + el * 8);
    const double2 *T = gt.T64[g] + (int64_t)pair * 8;
#pragma unroll
    for(int k = 0; k < 4; ++k) {
      const double2 f = F[k];
      const double2 t0 = T[2 * k], t1 = T[2 * k + 1];
      v.x += f.x * t0.x + f.y * t1.x;
      if(CPLX) v.y += f.x * t0.y + f.y * t1.y;
    }
  }
  if(gt.dense64[g]) {
    const int NN = gt.N[g] * gt.N[g];
    const double *d = gt.dense64[g] + el * (CPLX ? 2 : 1) * NN + pair;
    v.x += d[0];
    if(CPLX) v.y += d[NN];
  }
  return v;
}

template <bool CPLX>
__device__ __forceinline__ float2 elem_value32(const GroupTable &gt, int tensor, int64_t eg, int pair)
{
  const int g = find_group(gt, eg);
  const int64_t el = eg - gt.e_off[g];
  float2 v = make_float2(0.f, 0.f);
  if(tensor && gt.T32[g]) {  // curved groups have no reference tensor
    const float4 *F = reinterpret_cast<const float4 *>(gt.F32[g] + el * 8);
    const float4 *T = reinterpret_cast<const float4 *>(gt.T32[g] + (int64_t)pair * 8);
#pragma unroll
    for(int k = 0; k < 2; ++k) {
      const float4 f = F[k];
      const float4 ta = T[2 * k], tb = T[2 * k + 1];
      v.x += f.x * ta.x + f.y * ta.z + f.z * tb.x + f.w * tb.z;
      if(CPLX) v.y += f.x * ta.y + f.y * ta.w + f.z * tb.y + f.w * tb.w;
    }
  }
  if(gt.dense32[g]) {
    const int NN = gt.N[g] * gt.N[g];
    const float *d = gt.dense32[g] + el * (CPLX ? 2 : 1) * NN + pair;
    v.x += d[0];
    if(CPLX) v.y += d[NN];
  }
  return v;
}

template <typename V>
struct ValTraits;
template <>
struct ValTraits<double> {
  static constexpr bool cplx = false;
  __device__ static double make(double2 v) { return v.x; }
};
template <>
struct ValTraits<double2> {
  static constexpr bool cplx = true;
  __device__ static double2 make(double2 v) { return v; }
};

// ------------------------------------------------------------------ k_gather
// Atomic-free CSR assembly, one 256-thread block per chunk of whole nonzeros
// (<= kChunk contributions, <= kChunkElems distinct elements).  The chunk is
// one fixed-stride record (header, words in nonzero order, element list), so
// the block issues every load of its map at once.  Then:
//   1. the chunk's element factors F (FP64 geometry) are staged in LDS;
//   2. thread t evaluates the words of slots 2t, 2t+1 as F . T[pair] into
//      registers (the words are already in nonzero order);
//   3. each nonzero's run of slots is summed by a segmented scan and written
//      once.  No atomics; the order of additions is fixed, so results are
//      bitwise reproducible.

// F . T[pair] for one tensor row `row` (real rows: 8 components; complex
// rows: 8 (re, im) pairs), added to (vr, vi).  The row may sit in memory or
// in LDS: the arithmetic is this one function either way, so both give the
// same bits.
template <bool CPLX, bool MIXED>
__device__ __forceinline__ void tensor_row(const typename std::conditional<MIXED, float, double>::type *F, bool treal,
                                           const unsigned char *row,
                                           typename std::conditional<MIXED, float, double>::type &vr,
                                           typename std::conditional<MIXED, float, double>::type &vi)
{
  if(!CPLX || treal) {
    if(MIXED) {
      const float4 *T = reinterpret_cast<const float4 *>(row);
      const float4 a = T[0], b = T[1];
      vr += F[0] * a.x + F[1] * a.y + F[2] * a.z + F[3] * a.w + F[4] * b.x + F[5] * b.y + F[6] * b.z;
    }
    else {
      const double2 *T = reinterpret_cast<const double2 *>(row);
      const double2 a = T[0], b = T[1], c = T[2], d = T[3];
      vr += F[0] * a.x + F[1] * a.y + F[2] * b.x + F[3] * b.y + F[4] * c.x + F[5] * c.y + F[6] * d.x;
    }
  }
  else {
    if(MIXED) {
      const float4 *T = reinterpret_cast<const float4 *>(row);
#pragma unroll
      for(int q = 0; q < 4; ++q) {
        const float4 tq = T[q];  // components 2q (x,y) and 2q+1 (z,w)
        vr += F[2 * q] * tq.x + F[2 * q + 1] * tq.z;
        vi += F[2 * q] * tq.y + F[2 * q + 1] * tq.w;
      }
    }
    else {
      const double2 *T = reinterpret_cast<const double2 *>(row);
#pragma unroll
      for(int q = 0; q < 7; ++q) {
        const double2 tq = T[q];
        vr += F[q] * tq.x;
        vi += F[q] * tq.y;
      }
    }
  }
}

// bytes of one tensor row in memory
template <bool CPLX, bool MIXED>
__host__ __device__ constexpr int tensor_row_bytes(bool treal)
{
  return 8 * (MIXED ? 4 : 8) * ((!CPLX || treal) ? 1 : 2);
}

// Value of one contribution word: pair (i, j), chunk element le.  The
// element's factors F (7 per element, structure of arrays) sit in LDS with
// its group and index; the group's reference tensor row [pair][8] is read
// from memory (it stays in the caches), or, with LDS_T and a staged group
// (s_toff[g] >= 0), from the block's LDS copy at s_tl + s_toff[g] + pair *
// s_trs[g].  Per-point element matrices (planes [real, imaginary][N*N]) are
// added first.
template <bool CPLX, bool MIXED, bool LDS_T = false>
__device__ __forceinline__ void gather_eval(uint32_t w, int logP, uint32_t pmask, int tensor,
                                            const typename std::conditional<MIXED, float, double>::type (*s_F)[kChunkElems],
                                            const uint8_t *s_eg, const uint32_t *s_el, const uint8_t *s_gR,
                                            const int *s_gNN, const void *const *s_gT,
                                            const typename std::conditional<MIXED, float, double>::type *const *s_gD,
                                            typename std::conditional<MIXED, float, double>::type &vr,
                                            typename std::conditional<MIXED, float, double>::type &vi,
                                            const unsigned char *s_tl = nullptr, const int *s_toff = nullptr,
                                            const int *s_trs = nullptr)
{
  using R = typename std::conditional<MIXED, float, double>::type;
  const int pair = (int)(w & pmask);
  const int le = (int)((w >> logP) & 0x7Fu);
  const int g = s_eg[le];
  vr = 0;
  vi = 0;
  if(s_gD[g]) {  // per-point element matrix, planes [real, imaginary][N*N]
    const int NN = s_gNN[g];
    const R *d = s_gD[g] + (int64_t)s_el[le] * (CPLX ? 2 : 1) * NN + pair;
    vr = d[0];
    if(CPLX) vi = d[NN];
  }
  if(!tensor || !s_gT[g]) return;  // curved groups have no reference tensor
  R F[8];
#pragma unroll
  for(int q = 0; q < 7; ++q) F[q] = s_F[q][le];  // structure-of-arrays: no bank conflicts across elements
  F[7] = 0;
  // real and complex rows at separate call sites (a run-time flag inside the
  // helper keeps the slot arrays out of registers)
  const bool treal = !CPLX || s_gR[g];
  if constexpr(LDS_T) {
    const int off = s_toff[g];
    if(off >= 0) {
      const unsigned char *row = s_tl + off + pair * s_trs[g];
      if(treal)
        tensor_row<CPLX, MIXED>(F, true, row, vr, vi);
      else
        tensor_row<CPLX, MIXED>(F, false, row, vr, vi);
      return;
    }
  }
  const unsigned char *T = reinterpret_cast<const unsigned char *>(s_gT[g]);
  if(treal)
    tensor_row<CPLX, MIXED>(F, true, T + (int64_t)pair * tensor_row_bytes<CPLX, MIXED>(true), vr, vi);
  else
    tensor_row<CPLX, MIXED>(F, false, T + (int64_t)pair * tensor_row_bytes<CPLX, MIXED>(false), vr, vi);
}

// Sums of the nonzeros' runs of slots and their output positions.  Thread t
// holds S consecutive slots (values v, first-of-nonzero flags h; slots past
// the chunk end are flagged as heads and not written).  A run's sum is its
// slots in this thread plus a carry from the threads to its right: the carry
// is a segmented suffix scan over threads (shuffles within a wave, one LDS
// exchange across waves), so long runs cost log steps instead of a serial
// walk.  The order of additions is fixed: results are bitwise reproducible.
// Output position of a run = first nonzero of the chunk + rank of its head.
//
// runs_scan: the wave's part.  (xr, xi) = the sum of the slots that continue
// the run from the left, taken over this lane and the lanes to its right up
// to the first head (f = a head was met); incl = inclusive prefix count of
// the heads over the lanes.
template <bool CPLX, typename R, int S>
__device__ __forceinline__ void runs_scan(int lane, const int *h, const bool *in, const R *vr, const R *vi, R &xr,
                                          R &xi, int &f, int &incl)
{
  xr = 0;
  xi = 0;
  f = 0;
#pragma unroll
  for(int k = 0; k < S; ++k) {
    if(h[k]) f = 1;
    if(!f) {
      xr += vr[k];
      xi += vi[k];
    }
  }
#pragma unroll
  for(int d = 1; d < 64; d <<= 1) {
    const R yr = __shfl_down(xr, d, 64);
    const R yi = CPLX ? __shfl_down(xi, d, 64) : (R)0;
    const int g = __shfl_down(f, d, 64);
    if(lane + d < 64 && !f) {
      xr += yr;
      if(CPLX) xi += yi;
      f = g;
    }
  }
  // ranks: inclusive scan of the (in-range) head counts
  int cnt = 0;
#pragma unroll
  for(int k = 0; k < S; ++k) cnt += (in[k] && h[k]) ? 1 : 0;
  incl = cnt;
#pragma unroll
  for(int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d, 64);
    if(lane >= d) incl += v;
  }
}

// runs_store: with the carry (cr, ci) into the wave from its right and
// rank = one past this lane's last head, sum each run and write it once
template <bool CPLX, typename R, int S>
__device__ __forceinline__ void runs_store(int lane, const int *h, const bool *in, const R *vr, const R *vi, R xr,
                                           R xi, int f, R cr, R ci, int rank, uint32_t ns, void *__restrict__ out)
{
  if(!f) {
    xr += cr;
    xi += ci;
  }
  // carry into this thread = suffix value of the next thread
  R ar = __shfl_down(xr, 1, 64), ai = CPLX ? __shfl_down(xi, 1, 64) : (R)0;
  if(lane == 63) {
    ar = cr;
    ai = ci;
  }
  // runs inside the thread, right to left
#pragma unroll
  for(int k = S - 1; k >= 0; --k) {
    ar = vr[k] + ar;
    ai = vi[k] + ai;
    if(h[k]) {
      if(in[k]) {
        const int64_t pos = (int64_t)ns + (--rank);
        if(CPLX) {
          if(std::is_same<R, float>::value)
            reinterpret_cast<float2 *>(out)[pos] = make_float2((float)ar, (float)ai);
          else
            reinterpret_cast<double2 *>(out)[pos] = make_double2((double)ar, (double)ai);
        }
        else {
          reinterpret_cast<R *>(out)[pos] = ar;
        }
      }
      ar = 0;
      ai = 0;
    }
  }
}

// the block's runs: each wave scans, the waves exchange their counts and
// carries through LDS
template <bool CPLX, typename R, int S, int B>
__device__ __forceinline__ void gather_runs(int t, const int *h, const bool *in, const R *vr, const R *vi,
                                            uint32_t ns, int *s_wsum, R *s_wx, R *s_wy, int *s_wf,
                                            void *__restrict__ out)
{
  constexpr int kW = B / 64;
  const int lane = t & 63, wv = t >> 6;
  R xr, xi;
  int f, incl;
  runs_scan<CPLX, R, S>(lane, h, in, vr, vi, xr, xi, f, incl);
  if(lane == 63) s_wsum[wv] = incl;
  if(lane == 0) {
    s_wx[wv] = xr;
    s_wy[wv] = xi;
    s_wf[wv] = f;
  }
  __syncthreads();
  int rank = incl;  // one past this thread's last head
  for(int k = 0; k < wv; ++k) rank += s_wsum[k];
  // carry into this wave from the waves to its right
  R cr = 0, ci = 0;
  for(int k = kW - 1; k > wv; --k) {
    if(s_wf[k]) {
      cr = s_wx[k];
      ci = s_wy[k];
    }
    else {
      cr += s_wx[k];
      ci += s_wy[k];
    }
  }
  runs_store<CPLX, R, S>(lane, h, in, vr, vi, xr, xi, f, cr, ci, rank, ns, out);
}

// S slots per thread, kChunk / S threads per block
template <bool CPLX, bool MIXED, bool WIDE, int S>
__device__ __forceinline__ void gather_chunk(const GroupTable *__restrict__ gtp, const uint32_t *__restrict__ recs,
                                             int64_t c, int tensor, void *__restrict__ out)
{
  using R = typename std::conditional<MIXED, float, double>::type;
  constexpr int B = kChunk / S;
  static_assert(S == 2 || S == 4 || S == 8, "2, 4 or 8 slots per thread");
  constexpr int kStride = WIDE ? kRecWide : kRecNarrow;
  constexpr int kList = kStride - kChunkElems;  // element-list region of a record
  using R2 = typename std::conditional<MIXED, float2, double2>::type;
  __shared__ const void *s_gF[kMaxGroups];
  __shared__ const void *s_gT[kMaxGroups];
  __shared__ const R *s_gD[kMaxGroups];
  __shared__ int s_gNN[kMaxGroups];
  __shared__ uint8_t s_gR[kMaxGroups];
  __shared__ R s_F[7][kChunkElems];
  __shared__ uint32_t s_el[kChunkElems];
  __shared__ uint8_t s_eg[kChunkElems];
  __shared__ int s_wsum[B / 64];
  __shared__ R s_wx[B / 64], s_wy[B / 64];
  __shared__ int s_wf[B / 64];
  const GroupTable &gt = *gtp;
  const int t = threadIdx.x;
  // the whole record in one round of loads: words of slots S t .. S t + S - 1
  const uint32_t *rec = recs + c * kStride;
  const uint32_t ns = rec[0], clen = rec[1], nel = rec[2];
  uint32_t w[S];
  if constexpr(WIDE) {
    if constexpr(S == 8) {
      const uint4 p = reinterpret_cast<const uint4 *>(rec + kRecHead)[2 * t];
      const uint4 q = reinterpret_cast<const uint4 *>(rec + kRecHead)[2 * t + 1];
      w[0] = p.x; w[1] = p.y; w[2] = p.z; w[3] = p.w;
      w[4] = q.x; w[5] = q.y; w[6] = q.z; w[7] = q.w;
    }
    else if constexpr(S == 4) {
      const uint4 p = reinterpret_cast<const uint4 *>(rec + kRecHead)[t];
      w[0] = p.x; w[1] = p.y; w[2] = p.z; w[3] = p.w;
    }
    else {
      const uint2 p = reinterpret_cast<const uint2 *>(rec + kRecHead)[t];
      w[0] = p.x; w[1] = p.y;
    }
  }
  else {
    if constexpr(S == 8) {
      const uint4 p = reinterpret_cast<const uint4 *>(rec + kRecHead)[t];
      w[0] = p.x & 0xFFFFu; w[1] = p.x >> 16; w[2] = p.y & 0xFFFFu; w[3] = p.y >> 16;
      w[4] = p.z & 0xFFFFu; w[5] = p.z >> 16; w[6] = p.w & 0xFFFFu; w[7] = p.w >> 16;
    }
    else if constexpr(S == 4) {
      const uint2 p = reinterpret_cast<const uint2 *>(rec + kRecHead)[t];
      w[0] = p.x & 0xFFFFu; w[1] = p.x >> 16; w[2] = p.y & 0xFFFFu; w[3] = p.y >> 16;
    }
    else {
      const uint32_t p = rec[kRecHead + t];
      w[0] = p & 0xFFFFu;
      w[1] = p >> 16;
    }
  }
  // the element-list region (all kChunkElems slots, fixed offset) in the same
  // round of loads as the words: the list's length (header) is not waited for
  constexpr int kLP = (kChunkElems + B - 1) / B;  // list slots per thread
  uint32_t lv[kLP];
#pragma unroll
  for(int k = 0; k < kLP; ++k) lv[k] = (k * B + t < kChunkElems) ? rec[kList + k * B + t] : 0u;
  const int logP = gt.logP;
  const uint32_t pmask = (1u << logP) - 1u;
  const int ng = gt.ng;
  if(t < ng) {
    const bool treal = !CPLX || gt.treal[t];
    s_gR[t] = treal ? 1 : 0;
    s_gF[t] = MIXED ? (const void *)gt.F32[t] : (const void *)gt.F64[t];
    s_gT[t] = MIXED ? (treal ? (const void *)gt.T32r[t] : (const void *)gt.T32[t])
                    : (treal ? (const void *)gt.T64r[t] : (const void *)gt.T64[t]);
    s_gD[t] = MIXED ? (const R *)gt.dense32[t] : (const R *)gt.dense64[t];
    s_gNN[t] = gt.N[t] * gt.N[t];
  }
  // chunk elements: entry j of the list sits in record slot kStride - nel + j,
  // list slot i = kChunkElems - nel + j
#pragma unroll
  for(int k = 0; k < kLP; ++k) {
    const int j = k * B + t - (kChunkElems - (int)nel);
    if(j >= 0 && k * B + t < kChunkElems) {
      s_eg[j] = (uint8_t)(lv[k] >> 27);
      s_el[j] = lv[k] & 0x7FFFFFFu;
    }
  }
  __syncthreads();
  // element factors: 4 lanes per element, 2 factors each (one 64-byte row)
  for(int idx = t; idx < (int)nel * 4; idx += B) {
    const int jj = idx >> 2, q = idx & 3;
    const R2 f = reinterpret_cast<const R2 *>(reinterpret_cast<const R *>(s_gF[s_eg[jj]]) + (int64_t)s_el[jj] * 8)[q];
    s_F[2 * q][jj] = f.x;
    if(q < 3) s_F[2 * q + 1][jj] = f.y;
  }
  __syncthreads();
  const uint32_t hb = WIDE ? 31u : 15u;
  R vr[S], vi[S];
  int h[S];
  bool in[S];
#pragma unroll
  for(int k = 0; k < S; ++k) {
    in[k] = S * t + k < (int)clen;
    h[k] = in[k] ? (int)((w[k] >> hb) & 1u) : 1;  // past the end: closed, not written
    vr[k] = 0;
    vi[k] = 0;
    if(in[k])
      gather_eval<CPLX, MIXED>(w[k], logP, pmask, tensor, s_F, s_eg, s_el, s_gR, s_gNN, s_gT, s_gD, vr[k], vi[k]);
  }
  gather_runs<CPLX, R, S, B>(t, h, in, vr, vi, ns, s_wsum, s_wx, s_wy, s_wf, out);
}

template <bool CPLX, bool MIXED, bool WIDE, int S>
__global__ void __launch_bounds__(kChunk / S, S == 8 ? 5 : 8)
  k_gather(const GroupTable *__restrict__ gtp, const uint32_t *__restrict__ recs, int64_t nchunks, int tensor,
           void *__restrict__ out)
{
  const int64_t c = xcd_remap(blockIdx.x, gridDim.x);
  if(c >= nchunks) return;  // uniform per block
  gather_chunk<CPLX, MIXED, WIDE, S>(gtp, recs, c, tensor, out);
}

// item of a block range of a launch over several systems: the global block
// index of its first block, and what the standalone launch would take
__device__ __forceinline__ int find_item(const int64_t *__restrict__ first, int nitems, int64_t b)
{
  int lo = 0, hi = nitems - 1;
  while(lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if(first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// the chunks of several systems (tensor form) in one launch: items
// [first chunk, table, records, output], grid over all their chunks
template <bool CPLX, bool MIXED, bool WIDE, int S>
__global__ void __launch_bounds__(kChunk / S, S == 8 ? 5 : 8)
  k_gather_many(const GatherItem *__restrict__ items, const int64_t *__restrict__ first, int nitems, int64_t total)
{
  const int64_t cg = xcd_remap(blockIdx.x, gridDim.x);
  if(cg >= total) return;  // uniform per block
  const int i = find_item(first, nitems, cg);
  const GatherItem it = items[i];
  gather_chunk<CPLX, MIXED, WIDE, S>(it.table, it.recs, cg - first[i], 1, it.out);
}

// ------------------------------------------------------------------ k_gather_wave
// The 4-slot gather with one wavefront per chunk, persistent.  The 4-slot
// block above gives a chunk's left half (slots 0..255) to one wave and its
// right half to another, which meet at a barrier; here one wave does the
// right half, then the left half with the right half's carry, in registers:
// the same operations in the same order, so the values are bitwise those of
// the 4-slot block.  With no barrier per chunk the waves of a block run
// independently:
//   - each block stages the reference tensors of the groups that carry most
//     contributions in LDS once (TensorLds, sized by the launch), so a contribution's
//     tensor row is four LDS reads instead of four vector-memory loads (the
//     tensor-row loads were ~40 % of the block kernel's time);
//   - each wave loads its next chunk's record while it evaluates the current
//     one, so the record's HBM latency is off the chain.
// Chunks: each XCD takes a contiguous eighth of them (elements shared by
// neighbouring chunks stay in that XCD's L2), its waves round robin.
constexpr int kGwMaxWaves = 16;  // waves per block (at most)

// LDS bytes of one wave: factors [7][kChunkElems], element list, groups
template <bool MIXED>
__host__ __device__ constexpr int gather_wave_bytes()
{
  return 7 * kChunkElems * (MIXED ? 4 : 8) + kChunkElems * 4 + kChunkElems;
}

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool CPLX, bool MIXED, bool WIDE>
__global__ void __launch_bounds__(64 * kGwMaxWaves)
  k_gather_wave(const GroupTable *__restrict__ gtp, const uint32_t *__restrict__ recs, int64_t nchunks,
                void *__restrict__ out, const TensorLds tl)
{
  using R = typename std::conditional<MIXED, float, double>::type;
  using R2 = typename std::conditional<MIXED, float2, double2>::type;
  using WV = typename std::conditional<WIDE, uint4, uint2>::type;  // 4 words of one lane and half
  constexpr int S = 4;
  constexpr int kStride = WIDE ? kRecWide : kRecNarrow;
  constexpr int kList = kStride - kChunkElems;
  extern __shared__ uint4 s_gw[];
  __shared__ const void *s_gF[kMaxGroups];
  __shared__ const void *s_gT[kMaxGroups];
  __shared__ const R *s_gD[kMaxGroups];
  __shared__ int s_gNN[kMaxGroups];
  __shared__ int s_toff[kMaxGroups], s_trs[kMaxGroups];
  __shared__ uint8_t s_gR[kMaxGroups];
  const GroupTable &gt = *gtp;
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6), nw = (int)(blockDim.x >> 6);
  const int ng = gt.ng;
  unsigned char *lds = reinterpret_cast<unsigned char *>(s_gw);
  if(t < ng) {
    const bool treal = !CPLX || gt.treal[t];
    s_gR[t] = treal ? 1 : 0;
    s_gF[t] = MIXED ? (const void *)gt.F32[t] : (const void *)gt.F64[t];
    s_gT[t] = MIXED ? (treal ? (const void *)gt.T32r[t] : (const void *)gt.T32[t])
                    : (treal ? (const void *)gt.T64r[t] : (const void *)gt.T64[t]);
    s_gD[t] = MIXED ? (const R *)gt.dense32[t] : (const R *)gt.dense64[t];
    s_gNN[t] = gt.N[t] * gt.N[t];
    s_toff[t] = tl.off[t];
    s_trs[t] = tl.rs[t];
  }
  // the staged groups' tensor rows, 16 bytes per thread and step, in the
  // layout the launch sized the LDS for
  for(int g = 0; g < ng; ++g) {
    const int off = tl.off[g];
    if(off < 0) continue;
    const uint4 *T = reinterpret_cast<const uint4 *>(
      MIXED ? (tl.real[g] ? (const void *)gt.T32r[g] : (const void *)gt.T32[g])
            : (tl.real[g] ? (const void *)gt.T64r[g] : (const void *)gt.T64[g]));
    const int q16 = tl.q16[g], rs = tl.rs[g];
    const int n16 = gt.N[g] * gt.N[g] * q16;
    for(int idx = t; idx < n16; idx += (int)blockDim.x) {
      const int r = idx / q16, p = idx - r * q16;
      *reinterpret_cast<uint4 *>(lds + off + r * rs + p * 16) = T[idx];
    }
  }
  // this wave's region: factors (structure of arrays), element list
  unsigned char *wbase = lds + tl.bytes + wv * gather_wave_bytes<MIXED>();
  R(*s_F)[kChunkElems] = reinterpret_cast<R(*)[kChunkElems]>(wbase);
  uint32_t *s_el = reinterpret_cast<uint32_t *>(wbase + 7 * kChunkElems * sizeof(R));
  uint8_t *s_eg = reinterpret_cast<uint8_t *>(s_el + kChunkElems);
  __syncthreads();
  const int logP = gt.logP;
  const uint32_t pmask = (1u << logP) - 1u;
  const uint32_t hb = WIDE ? 31u : 15u;
  const int64_t nbx = gridDim.x / kXcds;  // the grid is a multiple of kXcds
  const int64_t per = (nchunks + kXcds - 1) / kXcds;
  const int64_t c0 = (int64_t)(blockIdx.x % kXcds) * per, c1 = min(nchunks, c0 + per);
  const int64_t step = nbx * nw;
  int64_t c = c0 + (int64_t)(blockIdx.x / kXcds) * nw + wv;
  WV wq[2], wn[2];  // words of the halves: this chunk, next chunk
  uint32_t lv[2], ln[2];
  uint32_t ns = 0, clen = 0, nel = 0, nsn = 0, clenn = 0, neln = 0;
  auto load = [&](int64_t cc, WV *wd, uint32_t *ld, uint32_t &a, uint32_t &b, uint32_t &e) {
    const uint32_t *rec = recs + cc * kStride;
    a = rec[0];
    b = rec[1];
    e = rec[2];
    wd[0] = reinterpret_cast<const WV *>(rec + kRecHead)[lane];
    wd[1] = reinterpret_cast<const WV *>(rec + kRecHead)[64 + lane];
    ld[0] = rec[kList + lane];
    ld[1] = rec[kList + 64 + lane];
  };
  auto words = [&](const WV &p, uint32_t *w) {
    if constexpr(WIDE) {
      w[0] = p.x; w[1] = p.y; w[2] = p.z; w[3] = p.w;
    }
    else {
      w[0] = p.x & 0xFFFFu; w[1] = p.x >> 16; w[2] = p.y & 0xFFFFu; w[3] = p.y >> 16;
    }
  };
  if(c < c1) load(c, wq, lv, ns, clen, nel);
  while(c < c1) {
    const int64_t cn = c + step;
    // chunk elements: entry j of the list sits in list slot kChunkElems - nel + j
#pragma unroll
    for(int k = 0; k < 2; ++k) {
      const int j = k * 64 + lane - (kChunkElems - (int)nel);
      if(j >= 0) {
        s_eg[j] = (uint8_t)(lv[k] >> 27);
        s_el[j] = lv[k] & 0x7FFFFFFu;
      }
    }
    wave_sync();
    // element factors: 4 lanes per element, 2 factors each (one 64-byte
    // row), in two batches of loads (registers); the next record's loads go
    // out after the first batch and stay in flight through this chunk
    constexpr int kFB = 2 * kChunkElems / 64;  // loads per lane and batch
#pragma unroll
    for(int bt = 0; bt < 2; ++bt) {
      R2 fv[kFB];
#pragma unroll
      for(int k = 0; k < kFB; ++k) {
        const int idx = (bt * kFB + k) * 64 + lane;
        if(idx < (int)nel * 4) {
          const int jj = idx >> 2, q = idx & 3;
          fv[k] = reinterpret_cast<const R2 *>(reinterpret_cast<const R *>(s_gF[s_eg[jj]]) + (int64_t)s_el[jj] * 8)[q];
        }
      }
      if(bt == 0 && cn < c1) load(cn, wn, ln, nsn, clenn, neln);
#pragma unroll
      for(int k = 0; k < kFB; ++k) {
        const int idx = (bt * kFB + k) * 64 + lane;
        if(idx < (int)nel * 4) {
          const int jj = idx >> 2, q = idx & 3;
          s_F[2 * q][jj] = fv[k].x;
          if(q < 3) s_F[2 * q + 1][jj] = fv[k].y;
        }
      }
    }
    wave_sync();
    // heads of the left half: the right half's ranks start after them
    uint32_t w[S];
    int h[S];
    bool in[S];
    int cnt0 = 0;
    words(wq[0], w);
#pragma unroll
    for(int k = 0; k < S; ++k)
      if(S * lane + k < (int)clen && ((w[k] >> hb) & 1u)) ++cnt0;
#pragma unroll
    for(int d = 32; d >= 1; d >>= 1) cnt0 += __shfl_xor(cnt0, d, 64);
    R vr[S], vi[S];
    R xr, xi, x1r, x1i;
    int f, incl, f1;
    // right half (slots 256..511): no carry from its right
    words(wq[1], w);
#pragma unroll
    for(int k = 0; k < S; ++k) {
      in[k] = S * (64 + lane) + k < (int)clen;
      h[k] = in[k] ? (int)((w[k] >> hb) & 1u) : 1;  // past the end: closed, not written
      vr[k] = 0;
      vi[k] = 0;
      if(in[k])
        gather_eval<CPLX, MIXED, true>(w[k], logP, pmask, 1, s_F, s_eg, s_el, s_gR, s_gNN, s_gT, s_gD, vr[k], vi[k],
                                       lds, s_toff, s_trs);
    }
    runs_scan<CPLX, R, S>(lane, h, in, vr, vi, xr, xi, f, incl);
    x1r = __shfl(xr, 0, 64);
    x1i = __shfl(xi, 0, 64);
    f1 = __shfl(f, 0, 64);
    runs_store<CPLX, R, S>(lane, h, in, vr, vi, xr, xi, f, (R)0, (R)0, incl + cnt0, ns, out);
    // left half (slots 0..255), carry from the right half
    words(wq[0], w);
#pragma unroll
    for(int k = 0; k < S; ++k) {
      in[k] = S * lane + k < (int)clen;
      h[k] = in[k] ? (int)((w[k] >> hb) & 1u) : 1;
      vr[k] = 0;
      vi[k] = 0;
      if(in[k])
        gather_eval<CPLX, MIXED, true>(w[k], logP, pmask, 1, s_F, s_eg, s_el, s_gR, s_gNN, s_gT, s_gD, vr[k], vi[k],
                                       lds, s_toff, s_trs);
    }
    runs_scan<CPLX, R, S>(lane, h, in, vr, vi, xr, xi, f, incl);
    R cr = 0, ci = 0;
    if(f1) {
      cr = x1r;
      ci = x1i;
    }
    else {
      cr += x1r;
      ci += x1i;
    }
    runs_store<CPLX, R, S>(lane, h, in, vr, vi, xr, xi, f, cr, ci, incl, ns, out);
    wave_sync();  // this chunk's list and factors are read before the next one's are written
    c = cn;
    wq[0] = wn[0];
    wq[1] = wn[1];
    lv[0] = ln[0];
    lv[1] = ln[1];
    ns = nsn;
    clen = clenn;
    nel = neln;
  }
}

// ------------------------------------------------------------------ k_rhs
// L lanes per row (2, 4 or 8, autotuned): Dirichlet lifting
// A_e(i_dof, j_tf) u_D for fixed columns (transposed local index,
// VectorFactory.cpp:184), then the element load vectors.  Lane l takes the
// row's terms l, l + L, ... in element order, and the lanes combine in a
// fixed butterfly, so sums are deterministic (FP64) for a given L.  Few lanes
// give each lane several independent loads in flight (rows have ~7 terms on
// P2 tetrahedra); more lanes suit long rows (many fixed columns).
template <int L>
__device__ __forceinline__ double2 lane_sum(double2 v)
{
#pragma unroll
  for(int d = 1; d < L; d <<= 1) {
    v.x += __shfl_xor(v.x, d, L);
    v.y += __shfl_xor(v.y, d, L);
  }
  return v;
}

template <bool CPLX, bool MIXED, int L>
__device__ __forceinline__ void rhs_rows(int64_t blk, const GroupTable *__restrict__ gtp, int tensor, int64_t n, int nrhs,
                                         const int64_t *__restrict__ dc_ptr, const uint32_t *__restrict__ dc_words,
                                         const int32_t *__restrict__ dc_fix, const double2 *__restrict__ uD,
                                         const int64_t *__restrict__ lin_ptr, const uint32_t *__restrict__ lin_words,
                                         const double2 *__restrict__ be_all, int64_t be_total, int be_real,
                                         void *__restrict__ out)
{
  const GroupTable &gt = *gtp;
  const int64_t row = (blk * blockDim.x + threadIdx.x) / L;
  const int l = (int)(threadIdx.x % L);
  if(row >= n) return;  // whole lane groups leave together
  double2 lift = make_double2(0., 0.);
  for(int64_t k = dc_ptr[row] + l; k < dc_ptr[row + 1]; k += L) {
    const uint32_t w = dc_words[k];
    const int64_t eg = w >> gt.logP;
    const int pair = (int)(w & ((1u << gt.logP) - 1u));
    double2 a;
    if(MIXED) {
      const float2 v = elem_value32<CPLX>(gt, tensor, eg, pair);
      a = make_double2(v.x, v.y);
    }
    else {
      a = elem_value64<CPLX>(gt, tensor, eg, pair);
    }
    lift = cadd(lift, cmul(a, uD[dc_fix[k]]));
  }
  lift = lane_sum<L>(lift);
  const int64_t l0 = lin_ptr[row], l1 = lin_ptr[row + 1];
  // kU terms of the lane at a time: their words, then their element loads,
  // are all in flight together; the sum still runs in term order
  constexpr int kU = 4;
  for(int r = 0; r < nrhs; ++r) {
    double2 acc = make_double2(0., 0.);
    for(int64_t k0 = l0 + l; k0 < l1; k0 += kU * L) {
      uint32_t wv[kU];
#pragma unroll
      for(int u = 0; u < kU; ++u) wv[u] = k0 + u * L < l1 ? lin_words[k0 + u * L] : 0u;
      if(be_real) {
        const double *ber = reinterpret_cast<const double *>(be_all) + (int64_t)r * be_total;
        double v[kU];
#pragma unroll
        for(int u = 0; u < kU; ++u) v[u] = k0 + u * L < l1 ? ber[wv[u]] : 0.;
#pragma unroll
        for(int u = 0; u < kU; ++u)
          if(k0 + u * L < l1) acc.x += v[u];
      }
      else {
        double2 v[kU];
#pragma unroll
        for(int u = 0; u < kU; ++u)
          v[u] = k0 + u * L < l1 ? be_all[(int64_t)r * be_total + wv[u]] : make_double2(0., 0.);
#pragma unroll
        for(int u = 0; u < kU; ++u)
          if(k0 + u * L < l1) {
            acc.x += v[u].x;
            acc.y += v[u].y;
          }
      }
    }
    acc = cadd(lift, lane_sum<L>(acc));
    if(l != 0) continue;
    const int64_t pos = (int64_t)r * n + row;
    if(MIXED) {
      if(CPLX)
        reinterpret_cast<float2 *>(out)[pos] = make_float2((float)acc.x, (float)acc.y);
      else
        reinterpret_cast<float *>(out)[pos] = (float)acc.x;
    }
    else {
      if(CPLX)
        reinterpret_cast<double2 *>(out)[pos] = acc;
      else
        reinterpret_cast<double *>(out)[pos] = acc.x;
    }
  }
}

template <bool CPLX, bool MIXED, int L>
__global__ void __launch_bounds__(256) k_rhs(const GroupTable *__restrict__ gtp, int tensor, int64_t n, int nrhs, const int64_t *__restrict__ dc_ptr,
                                             const uint32_t *__restrict__ dc_words, const int32_t *__restrict__ dc_fix,
                                             const double2 *__restrict__ uD, const int64_t *__restrict__ lin_ptr,
                                             const uint32_t *__restrict__ lin_words, const double2 *__restrict__ be_all,
                                             int64_t be_total, int be_real, void *__restrict__ out)
{
  rhs_rows<CPLX, MIXED, L>(xcd_remap(blockIdx.x, gridDim.x), gtp, tensor, n, nrhs, dc_ptr, dc_words, dc_fix, uD,
                           lin_ptr, lin_words, be_all, be_total, be_real, out);
}

// the right-hand sides of several systems (tensor form) in one launch
template <bool CPLX, bool MIXED, int L>
__global__ void __launch_bounds__(256) k_rhs_many(const RhsItem *__restrict__ items, const int64_t *__restrict__ first,
                                                  int nitems, int64_t total)
{
  const int64_t bg = xcd_remap(blockIdx.x, gridDim.x);
  if(bg >= total) return;
  const int i = find_item(first, nitems, bg);
  const RhsItem &a = items[i];
  rhs_rows<CPLX, MIXED, L>(bg - first[i], a.table, 1, a.n, a.nrhs, a.dc_ptr, a.dc_words, a.dc_fix, a.uD, a.lin_ptr,
                           a.lin_words, a.be_all, a.be_total, a.be_real, a.out);
}

// ------------------------------------------------------------------ reference strategy
template <bool CPLX>
__global__ void __launch_bounds__(256) k_ref_scatter(int64_t total, int NN, const double *__restrict__ Ae,
                                                     const int32_t *__restrict__ e2n, void *__restrict__ out)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= total) return;
  const int32_t nz = e2n[t];
  if(nz < 0) return;
  const int64_t e = t / NN, pair = t % NN;
  const double *a = Ae + e * (CPLX ? 2 : 1) * NN + pair;  // planes [real, imaginary]
  const double2 v = make_double2(a[0], CPLX ? a[NN] : 0.);
  if(CPLX) {
    double *o = reinterpret_cast<double *>(out) + 2 * (int64_t)nz;
    unsafeAtomicAdd(o, v.x);
    unsafeAtomicAdd(o + 1, v.y);
  }
  else {
    unsafeAtomicAdd(reinterpret_cast<double *>(out) + nz, v.x);
  }
}

template <bool CPLX>
__global__ void __launch_bounds__(256) k_ref_rhs(int64_t ne, int N, int nrhs, int64_t n,
                                                 const int32_t *__restrict__ num, const int32_t *__restrict__ type,
                                                 const double *__restrict__ Ae, const void *__restrict__ be,
                                                 int64_t be_stride, int be_real, const double2 *__restrict__ uD,
                                                 void *__restrict__ out)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= ne * N) return;
  const int64_t e = t / N;
  const int i = (int)(t % N);
  const int ti = type[t];
  if(ti != GFEM_AT_UNKNOWN && ti != GFEM_AT_UNKNOWN_BUBBLE) return;
  const int64_t row = num[t];
  double2 lift = make_double2(0., 0.);
  if(Ae) {
    for(int j = 0; j < N; ++j) {
      if(type[e * N + j] != GFEM_AT_FIXED) continue;
      const int64_t NN = (int64_t)N * N;
      const double *ap = Ae + e * (CPLX ? 2 : 1) * NN + (int64_t)j * N + i;
      const double2 a = make_double2(ap[0], CPLX ? ap[NN] : 0.);
      lift = cadd(lift, cmul(a, uD[num[e * N + j]]));
    }
  }
  for(int r = 0; r < nrhs; ++r) {
    double2 v = lift;
    if(be) {
      const double2 b = be_load(be, be_real, (int64_t)r * be_stride + e * N + i);
      v.x += b.x;
      v.y += b.y;
    }
    double *o = reinterpret_cast<double *>(out);
    if(CPLX) {
      unsafeAtomicAdd(o + 2 * ((int64_t)r * n + row), v.x);
      unsafeAtomicAdd(o + 2 * ((int64_t)r * n + row) + 1, v.y);
    }
    else {
      unsafeAtomicAdd(o + (int64_t)r * n + row, v.x);
    }
  }
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// a group's element loads of right-hand side r inside gfem_system::d_be_all
// ([nrhs][be_total], flat index be_off + e N + i): in double2 units, or in
// double units when every linear term is real (the index k_rhs reads with)
inline void *group_be(const Group &g, int r, bool be_real)
{
  double2 *all = g.d_be - g.be_off;
  if(be_real) return reinterpret_cast<double *>(all) + (int64_t)r * g.be_stride + g.be_off;
  return g.d_be + (int64_t)r * g.be_stride;
}

}  // namespace

// ====================================================================== launchers

hipError_t launch_basis(const Group &g, Rule &r, hipStream_t s)
{
  const int n = g.N * r.Q;
  hipLaunchKernelGGL(k_basis, dim3(blocks_for(n, 128)), dim3(128), 0, s, g.dim, g.order, g.N, r.Q, g.d_nodes,
                     r.d_pts, r.d_bf, r.d_gbf);
  return hipGetLastError();
}

hipError_t launch_tensor(const Group &g, hipStream_t s)
{
  const TensorArgs a = tensor_args(g);
  hipLaunchKernelGGL(k_tensor, dim3(blocks_for(g.N * g.N, 64)), dim3(64), 0, s, a, g.d_T64, g.d_T32, g.d_T64r,
                     g.d_T32r);
  return hipGetLastError();
}

hipError_t launch_tensor_many(const TensorItem *items, int nitems, int maxN, hipStream_t s)
{
  if(nitems == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tensor_many, dim3(blocks_for(maxN * maxN, 64), (unsigned)nitems), dim3(64), 0, s, items);
  return hipGetLastError();
}

TensorArgs tensor_args(const Group &g)
{
  TensorArgs a{};
  a.N = g.N;
  a.nterms = 0;
  for(size_t k = 0; k < g.dterms.size() && a.nterms < 16; ++k) {
    const DevTerm &t = g.dterms[k];
    const int i = a.nterms++;
    a.form_dof[i] = t.form_dof;
    a.form_tf[i] = t.form_tf;
    a.rule[i] = t.rule;
    a.active[i] = (t.kind == 0 && !t.per_point) ? 1 : 0;
    // constant coefficient * tf factor (conjugated tf coefficient and dt factor)
    const double cr = t.c_re * t.fac_re - t.c_im * t.fac_im;
    const double ci = t.c_re * t.fac_im + t.c_im * t.fac_re;
    a.fac[i] = make_double2(cr, ci);
  }
  for(size_t r = 0; r < g.rules.size() && r < 8; ++r) {
    a.Q[r] = g.rules[r].Q;
    a.w[r] = g.rules[r].d_w;
    a.bf[r] = g.rules[r].d_bf;
    a.gbf[r] = g.rules[r].d_gbf;
  }
  return a;
}

hipError_t launch_geometry_all(const gfem_system &sys, const double *xyz, bool fp32, hipStream_t s)
{
  GeomGroups gg{};
  gg.ng = (int)sys.groups.size();
  if(gg.ng == 0) return hipSuccess;
  if(gg.ng > kMaxGroups) return hipErrorInvalidValue;
  for(int g = 0; g < gg.ng; ++g) {
    const Group &G = sys.groups[g];
    gg.dim[g] = G.dim;
    gg.nv[g] = G.nv;
    gg.absdet[g] = G.absdet ? 1 : 0;
    gg.off[g + 1] = gg.off[g] + G.ne;
    gg.vtx[g] = G.d_vtx;
    gg.F64[g] = G.d_F64;
    gg.F32[g] = fp32 ? G.d_F32 : nullptr;
  }
  const int64_t total = gg.off[gg.ng];
  if(total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_geometry_all, dim3(blocks_for(total, 256)), dim3(256), 0, s, gg, xyz);
  return hipGetLastError();
}

hipError_t launch_geometry(const Group &g, const double *xyz, hipStream_t s)
{
  if(g.ne == 0) return hipSuccess;
  hipLaunchKernelGGL(k_geometry, dim3(blocks_for(g.ne, 256)), dim3(256), 0, s, g.dim, g.nv, g.absdet ? 1 : 0, g.ne, g.d_vtx, xyz,
                     g.d_F64, nullptr);  // FP64 only
  return hipGetLastError();
}

template <int N>
static void linear_n(const LinTerm &a, const Rule &R, size_t lds, const Group &g, const double *xyz, void *be,
                     hipStream_t s)
{
  if(lds)
    hipLaunchKernelGGL((k_linear<N, true>), dim3(blocks_for(g.ne, 128)), dim3(128), lds, s, a, R.d_w, R.d_pts, R.d_bf,
                       g.ne, g.d_vtx, xyz, g.d_F64, be);
  else
    hipLaunchKernelGGL((k_linear<N, false>), dim3(blocks_for(g.ne, 128)), dim3(128), 0, s, a, R.d_w, R.d_pts, R.d_bf,
                       g.ne, g.d_vtx, xyz, g.d_F64, be);
}

hipError_t launch_linear(const Group &g, const double *xyz, int nrhs, int eb_shift, bool be_real, hipStream_t s)
{
  if(g.ne == 0 || !g.has_linear) return hipSuccess;
  for(int r = 0; r < nrhs; ++r) {
    void *be = group_be(g, r, be_real);
    bool first = true;
    for(const DevTerm &t : g.dterms) {
      if(t.kind != 1 || t.rhs_idx != r) continue;
      const Rule &R = g.rules[t.rule];
      LinTerm a{};
      a.expr = t.expr;
      a.fac_re = t.fac_re;
      a.fac_im = t.fac_im;
      a.c_re = t.c_re;
      a.c_im = t.c_im;
      a.per_point = t.per_point;
      a.Q = R.Q;
      a.dim = g.dim;
      a.nv = g.nv;
      a.accumulate = first ? 0 : 1;
      a.real_out = be_real ? 1 : 0;
      a.conn = g.curved ? g.d_conn : nullptr;  // curved: per-point geometry, interpreted kernel
      a.gbf = R.d_gbf;
      // point-parallel kernels: EB elements per block, about 4 points per thread
      const int EB = std::max(1, std::min(128, (256 << eb_shift) / std::max(R.Q, 1)));
      const size_t lds_pp = sizeof(double) * (size_t)(R.Q * g.N + 4 * R.Q + 13 * EB + 1) +
                            sizeof(double2) * (size_t)EB * R.Q;
      const unsigned nblk = (unsigned)((g.ne + EB - 1) / EB);
      const size_t ti = (size_t)(&t - g.dterms.data());
      hipFunction_t jf = (ti < g.jit.size() && !g.curved) ? g.jit[ti] : nullptr;
      if(jf) {
        // compiled coefficient program (jit.cpp), one thread per element in
        // blocks of 64 << eb_shift threads; argument layout of JitLinArgs
        JitLinArgs ja{};
        for(int k = 0; k < 2 * GFEM_MAX_CONSTS; ++k) ja.consts[k] = t.expr.consts[k];
        ja.fac_re = t.fac_re;
        ja.fac_im = t.fac_im;
        ja.Q = R.Q;
        ja.dim = g.dim;
        ja.nv = g.nv;
        ja.N = g.N;
        ja.accumulate = first ? 0 : 1;
        ja.EB = EB;
        ja.real_out = be_real ? 1 : 0;
        int64_t ne = g.ne;
        const double *pw = R.d_w, *pp = R.d_pts, *pb = R.d_bf, *pF = g.d_F64;
        const int32_t *pv = g.d_vtx;
        void *pbe = be;
        void *params[] = {&ja, &pw, &pp, &pb, &ne, &pv, &xyz, &pF, &pbe};
        const unsigned bs = 64u << std::min(std::max(eb_shift, 1), 3);
        const unsigned jblk = (unsigned)((g.ne + bs - 1) / bs);
        const hipError_t e = hipModuleLaunchKernel(jf, jblk, 1, 1, bs, 1, 1, 0, s, params, nullptr);
        if(e != hipSuccess) return e;
        first = false;
        continue;
      }
      if(lds_pp <= 64 * 1024 && !g.curved) {
        hipLaunchKernelGGL(k_linear_pp, dim3(nblk), dim3(256), lds_pp, s, a, g.N, EB, R.d_w, R.d_pts, R.d_bf, g.ne,
                           g.d_vtx, xyz, g.d_F64, be);
        first = false;
        continue;
      }
      const size_t lds = sizeof(double) * (size_t)R.Q * (4 + g.N);
      const size_t dyn = lds <= 48 * 1024 ? lds : 0;  // stage the rule in LDS when it fits
      switch(g.N) {
      case 1: linear_n<1>(a, R, dyn, g, xyz, be, s); break;
      case 2: linear_n<2>(a, R, dyn, g, xyz, be, s); break;
      case 3: linear_n<3>(a, R, dyn, g, xyz, be, s); break;
      case 4: linear_n<4>(a, R, dyn, g, xyz, be, s); break;
      case 5: linear_n<5>(a, R, dyn, g, xyz, be, s); break;
      case 6: linear_n<6>(a, R, dyn, g, xyz, be, s); break;
      case 10: linear_n<10>(a, R, dyn, g, xyz, be, s); break;
      case 15: linear_n<15>(a, R, dyn, g, xyz, be, s); break;
      case 20: linear_n<20>(a, R, dyn, g, xyz, be, s); break;
      case 35: linear_n<35>(a, R, dyn, g, xyz, be, s); break;
      default: return hipErrorInvalidValue;
      }
      first = false;
    }
  }
  return hipGetLastError();
}

hipError_t launch_dense(const Group &g, const double *xyz, bool all_terms, int planes, hipStream_t s)
{
  if(g.ne == 0) return hipSuccess;
  const int64_t total = g.ne * (int64_t)g.N * g.N;
  hipLaunchKernelGGL(k_dense, dim3(blocks_for(total, 256)), dim3(256), 0, s, g.d_dev, all_terms ? 1 : 0, g.ne,
                     g.d_vtx, xyz, planes, g.d_dense64);
  return hipGetLastError();
}

hipError_t launch_dense_to32(const Group &g, int planes, hipStream_t s)
{
  const int64_t total = g.ne * planes * (int64_t)g.N * g.N;
  if(total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_dense_to32, dim3(blocks_for(total, 256)), dim3(256), 0, s, total, g.d_dense64, g.d_dense32);
  return hipGetLastError();
}

// ------------------------------------------------------------------ descriptor updates
hipError_t launch_put_term(DevTerm *dst, const DevTerm &v, hipStream_t s)
{
  hipLaunchKernelGGL(k_put<DevTerm>, dim3(1), dim3(64), 0, s, dst, v);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_zero(uint64_t *__restrict__ p, int64_t words)
{
  for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

hipError_t launch_zero(void *p, size_t bytes, hipStream_t s)
{
  const int64_t words = (int64_t)(bytes / 8);
  if(words == 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int64_t>((words + 255) / 256, 65536);
  hipLaunchKernelGGL(k_zero, dim3(grid), dim3(256), 0, s, reinterpret_cast<uint64_t *>(p), words);
  return hipGetLastError();
}

hipError_t launch_put_table(GroupTable *dst, const GroupTable &v, hipStream_t s)
{
  hipLaunchKernelGGL(k_put<GroupTable>, dim3(1), dim3(64), 0, s, dst, v);
  return hipGetLastError();
}

template <typename T>
static void coef_rows_pp(const Group &g, const double *xyz, bool all_terms, int planes, T *A, hipStream_t s)
{
  const int64_t total = g.ne * (int64_t)g.rules[g.gemm_rule].Q;
  const dim3 grid(blocks_for(total, kCoefBlock)), block(kCoefBlock);
  if(g.has_dense)  // a bilinear term with a coefficient that depends on x
    hipLaunchKernelGGL((k_coef_rows<T, true>), grid, block, 0, s, g.d_dev, all_terms ? 1 : 0, g.gemm_rule, g.ne,
                       g.d_vtx, xyz, g.d_F64, planes, A);
  else
    hipLaunchKernelGGL((k_coef_rows<T, false>), grid, block, 0, s, g.d_dev, all_terms ? 1 : 0, g.gemm_rule, g.ne,
                       g.d_vtx, xyz, g.d_F64, planes, A);
}

hipError_t launch_coef_rows(const Group &g, const double *xyz, bool all_terms, int planes, bool fp32, hipStream_t s)
{
  if(g.ne == 0) return hipSuccess;
  if(fp32)
    coef_rows_pp<float>(g, xyz, all_terms, planes, g.d_A32, s);
  else
    coef_rows_pp<double>(g, xyz, all_terms, planes, g.d_A64, s);
  return hipGetLastError();
}

constexpr int kTPad = 16;  // bytes after each tensor row in LDS: rows start on different banks

TensorLds gather_lds_layout(const gfem_system &sys, int prec)
{
  TensorLds tl{};
  const int ng = (int)sys.groups.size();
  const int budget = prec ? 8192 : 16384;  // bytes of staged rows
  int64_t total = 0;
  std::vector<std::pair<int64_t, int>> cand;
  for(int g = 0; g < ng; ++g) {
    tl.off[g] = -1;
    const Group &G = sys.groups[g];
    const bool treal = !sys.cplx || G.treal;
    const void *T = prec ? (treal ? (const void *)G.d_T32r : (const void *)G.d_T32)
                         : (treal ? (const void *)G.d_T64r : (const void *)G.d_T64);
    if(G.curved || !T || G.ne == 0) continue;
    const int64_t w = G.ne * G.N * G.N;  // contributions of the group
    total += w;
    cand.push_back({w, g});
  }
  std::stable_sort(cand.begin(), cand.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
  int bytes = 0;
  for(const auto &c : cand) {
    if(c.first * 32 < total) break;  // groups with < 1/32 of the contributions stay in memory
    const Group &G = sys.groups[c.second];
    const bool treal = !sys.cplx || G.treal;
    const int rb = 8 * (prec ? 4 : 8) * (treal ? 1 : 2);
    const int need = G.N * G.N * (rb + kTPad);
    if(bytes + need > budget) continue;
    tl.off[c.second] = bytes;
    tl.rs[c.second] = rb + kTPad;
    tl.q16[c.second] = rb / 16;
    tl.real[c.second] = treal ? 1 : 0;
    bytes += need;
  }
  tl.bytes = bytes;
  return tl;
}

template <bool C, bool M, bool W>
const void *gather_wave_fn()
{
  return reinterpret_cast<const void *>(&k_gather_wave<C, M, W>);
}

// waves per block and blocks of the one-wave gather: the most resident waves
// per CU (LDS and registers, from the runtime's occupancy), cached per
// precision for the staged bytes
static hipError_t gather_wave_plan(const gfem_system &sys, int prec, const void *fn, int tbytes)
{
  if(sys.gw_tbytes[prec] == tbytes) return hipSuccess;
  const int wb = prec ? gather_wave_bytes<true>() : gather_wave_bytes<false>();
  int dev = 0, ncu = 0;
  hipError_t e = hipGetDevice(&dev);
  if(e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  // one limit for every launch of the function: the LDS of a CU less the
  // kernel's static arrays
  constexpr int kMaxDyn = 160 * 1024 - 1024;
  if(e == hipSuccess) e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxDyn);
  int best_nw = 0, best_occ = 0;
  for(int nw : {4, 6, 8, 10, 12, 16}) {
    if(e != hipSuccess) break;
    const size_t lds = (size_t)tbytes + (size_t)nw * wb;
    if(lds > (size_t)kMaxDyn) continue;
    int occ = 0;
    if(e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 64 * nw, lds);
    if(e == hipSuccess && occ * nw > best_occ * best_nw) {
      best_nw = nw;
      best_occ = occ;
    }
  }
  if(e != hipSuccess) return e;
  if(best_nw == 0) return hipErrorInvalidConfiguration;
  int64_t nb = std::min<int64_t>((int64_t)ncu * best_occ, (sys.n_chunks + best_nw - 1) / best_nw);
  sys.gw_blocks[prec] = xcd_grid(std::max<int64_t>(nb, 1));
  sys.gw_nw[prec] = best_nw;
  sys.gw_tbytes[prec] = tbytes;
  return hipSuccess;
}

template <bool C, bool M, bool W>
static hipError_t gather_wave_launch(const gfem_system &sys, int prec, void *out, hipStream_t s)
{
  const TensorLds tl = gather_lds_layout(sys, prec);
  const void *fn = gather_wave_fn<C, M, W>();
  hipError_t e = gather_wave_plan(sys, prec, fn, tl.bytes);
  if(e != hipSuccess) return e;
  const int nw = sys.gw_nw[prec];
  const size_t lds = (size_t)tl.bytes + (size_t)nw * (M ? gather_wave_bytes<true>() : gather_wave_bytes<false>());
  hipLaunchKernelGGL((k_gather_wave<C, M, W>), dim3((unsigned)sys.gw_blocks[prec]), dim3(64 * nw), lds, s,
                     sys.d_table, sys.d_rec, sys.n_chunks, out, tl);
  return hipGetLastError();
}

hipError_t launch_gather(const gfem_system &sys, int prec, int tensor, void *out, hipStream_t s)
{
  if(sys.n_chunks == 0) return hipSuccess;
  if(tensor && sys.gather_wave[prec] && sys.gather_slots[prec] == 4) {
    if(sys.cplx)
      return prec ? (sys.rec_wide ? gather_wave_launch<true, true, true>(sys, prec, out, s)
                                  : gather_wave_launch<true, true, false>(sys, prec, out, s))
                  : (sys.rec_wide ? gather_wave_launch<true, false, true>(sys, prec, out, s)
                                  : gather_wave_launch<true, false, false>(sys, prec, out, s));
    return prec ? (sys.rec_wide ? gather_wave_launch<false, true, true>(sys, prec, out, s)
                                : gather_wave_launch<false, true, false>(sys, prec, out, s))
                : (sys.rec_wide ? gather_wave_launch<false, false, true>(sys, prec, out, s)
                                : gather_wave_launch<false, false, false>(sys, prec, out, s));
  }
  const GroupTable *gt = tensor ? sys.d_table : sys.d_table_dense;
  const dim3 grid((unsigned)xcd_grid(sys.n_chunks));
#define GFEM_GATHER_S(C, M, W, S)                                                                            \
  hipLaunchKernelGGL((k_gather<C, M, W, S>), grid, dim3(kChunk / S), 0, s, gt, sys.d_rec, sys.n_chunks, tensor, out)
#define GFEM_GATHER(C, M)                                                                                    \
  if(sys.rec_wide) {                                                                                         \
    if(sys.gather_slots[prec] == 8)                                                                          \
      GFEM_GATHER_S(C, M, true, 8);                                                                          \
    else if(sys.gather_slots[prec] == 4)                                                                     \
      GFEM_GATHER_S(C, M, true, 4);                                                                          \
    else                                                                                                     \
      GFEM_GATHER_S(C, M, true, 2);                                                                          \
  }                                                                                                          \
  else {                                                                                                     \
    if(sys.gather_slots[prec] == 8)                                                                          \
      GFEM_GATHER_S(C, M, false, 8);                                                                         \
    else if(sys.gather_slots[prec] == 4)                                                                     \
      GFEM_GATHER_S(C, M, false, 4);                                                                         \
    else                                                                                                     \
      GFEM_GATHER_S(C, M, false, 2);                                                                         \
  }
  if(sys.cplx) {
    if(prec) {
      GFEM_GATHER(true, true)
    }
    else {
      GFEM_GATHER(true, false)
    }
  }
  else {
    if(prec) {
      GFEM_GATHER(false, true)
    }
    else {
      GFEM_GATHER(false, false)
    }
  }
#undef GFEM_GATHER
#undef GFEM_GATHER_S
  return hipGetLastError();
}

template <int L>
static void rhs_lanes(const gfem_system &sys, int prec, int tensor, void *out, hipStream_t s)
{
  const GroupTable *gt = tensor ? sys.d_table : sys.d_table_dense;
  const dim3 grid((unsigned)xcd_grid(blocks_for(sys.n * L, 256))), block(256);
#define GFEM_RHS_ARGS gt, tensor, sys.n, sys.nrhs, sys.d_dc_ptr, sys.d_dc_words, sys.d_dc_fix, sys.d_dc, sys.d_lin_ptr, sys.d_lin_words, sys.d_be_all, sys.be_total, sys.be_real ? 1 : 0, out
  if(sys.cplx) {
    if(prec)
      hipLaunchKernelGGL((k_rhs<true, true, L>), grid, block, 0, s, GFEM_RHS_ARGS);
    else
      hipLaunchKernelGGL((k_rhs<true, false, L>), grid, block, 0, s, GFEM_RHS_ARGS);
  }
  else {
    if(prec)
      hipLaunchKernelGGL((k_rhs<false, true, L>), grid, block, 0, s, GFEM_RHS_ARGS);
    else
      hipLaunchKernelGGL((k_rhs<false, false, L>), grid, block, 0, s, GFEM_RHS_ARGS);
  }
#undef GFEM_RHS_ARGS
}

hipError_t launch_rhs(const gfem_system &sys, int prec, int tensor, void *out, hipStream_t s)
{
  if(sys.n == 0) return hipSuccess;
  switch(sys.rhs_lanes[prec ? 1 : 0]) {
  case 2: rhs_lanes<2>(sys, prec, tensor, out, s); break;
  case 4: rhs_lanes<4>(sys, prec, tensor, out, s); break;
  case 8: rhs_lanes<8>(sys, prec, tensor, out, s); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

GatherItem gather_item(const gfem_system &sys, void *out) { return GatherItem{sys.d_table, sys.d_rec, out}; }

int64_t gather_item_blocks(const gfem_system &sys) { return sys.n_chunks; }

hipError_t launch_gather_many(bool cplx, bool mixed, bool wide, int slots, const GatherItem *items,
                              const int64_t *first, int nitems, int64_t total, hipStream_t s)
{
  if(total == 0) return hipSuccess;
  const dim3 grid((unsigned)xcd_grid(total));
#define GFEM_GM(C, M, W, S)                                                                                        \
  hipLaunchKernelGGL((k_gather_many<C, M, W, S>), grid, dim3(kChunk / S), 0, s, items, first, nitems, total)
#define GFEM_GM_S(C, M, W)                                                                                         \
  if(slots == 8)                                                                                                   \
    GFEM_GM(C, M, W, 8);                                                                                           \
  else if(slots == 4)                                                                                              \
    GFEM_GM(C, M, W, 4);                                                                                           \
  else                                                                                                             \
    GFEM_GM(C, M, W, 2);
#define GFEM_GM_W(C, M)                                                                                            \
  if(wide) {                                                                                                       \
    GFEM_GM_S(C, M, true)                                                                                          \
  }                                                                                                                \
  else {                                                                                                           \
    GFEM_GM_S(C, M, false)                                                                                         \
  }
  if(cplx) {
    if(mixed) {
      GFEM_GM_W(true, true)
    }
    else {
      GFEM_GM_W(true, false)
    }
  }
  else {
    if(mixed) {
      GFEM_GM_W(false, true)
    }
    else {
      GFEM_GM_W(false, false)
    }
  }
#undef GFEM_GM_W
#undef GFEM_GM_S
#undef GFEM_GM
  return hipGetLastError();
}

RhsItem rhs_item(const gfem_system &sys, void *out)
{
  return RhsItem{sys.d_table, sys.n, sys.nrhs, sys.be_real ? 1 : 0, sys.d_dc_ptr, sys.d_dc_words, sys.d_dc_fix, sys.d_dc,
                 sys.d_lin_ptr, sys.d_lin_words, sys.d_be_all, sys.be_total, out};
}

int64_t rhs_item_blocks(const gfem_system &sys, int prec) { return blocks_for(sys.n * sys.rhs_lanes[prec ? 1 : 0], 256); }

hipError_t launch_rhs_many(bool cplx, bool mixed, int lanes, const RhsItem *items, const int64_t *first, int nitems,
                           int64_t total, hipStream_t s)
{
  if(total == 0) return hipSuccess;
  const dim3 grid((unsigned)xcd_grid(total)), block(256);
#define GFEM_RM(L)                                                                                                 \
  if(cplx) {                                                                                                       \
    if(mixed)                                                                                                      \
      hipLaunchKernelGGL((k_rhs_many<true, true, L>), grid, block, 0, s, items, first, nitems, total);             \
    else                                                                                                           \
      hipLaunchKernelGGL((k_rhs_many<true, false, L>), grid, block, 0, s, items, first, nitems, total);            \
  }                                                                                                                \
  else {                                                                                                           \
    if(mixed)                                                                                                      \
      hipLaunchKernelGGL((k_rhs_many<false, true, L>), grid, block, 0, s, items, first, nitems, total);            \
    else                                                                                                           \
      hipLaunchKernelGGL((k_rhs_many<false, false, L>), grid, block, 0, s, items, first, nitems, total);           \
  }
  switch(lanes) {
  case 2: GFEM_RM(2) break;
  case 4: GFEM_RM(4) break;
  case 8: GFEM_RM(8) break;
  default: return hipErrorInvalidValue;
  }
#undef GFEM_RM
  return hipGetLastError();
}

hipError_t launch_ref_scatter(const gfem_system &sys, const Group &g, hipStream_t s)
{
  if(g.ne == 0 || !g.has_bilinear) return hipSuccess;
  const int64_t total = g.ne * (int64_t)g.N * g.N;
  const int NN = g.N * g.N;
  if(sys.cplx)
    hipLaunchKernelGGL(k_ref_scatter<true>, dim3(blocks_for(total, 256)), dim3(256), 0, s, total, NN, g.d_dense64,
                       g.d_e2n, sys.d_vals[0]);
  else
    hipLaunchKernelGGL(k_ref_scatter<false>, dim3(blocks_for(total, 256)), dim3(256), 0, s, total, NN, g.d_dense64,
                       g.d_e2n, sys.d_vals[0]);
  return hipGetLastError();
}

hipError_t launch_ref_rhs(const gfem_system &sys, const Group &g, hipStream_t s)
{
  if(g.ne == 0) return hipSuccess;
  const int64_t total = g.ne * (int64_t)g.N;
  const double *Ae = g.has_bilinear ? g.d_dense64 : nullptr;
  const void *be = g.has_linear ? group_be(g, 0, sys.be_real) : nullptr;
  if(sys.cplx)
    hipLaunchKernelGGL(k_ref_rhs<true>, dim3(blocks_for(total, 256)), dim3(256), 0, s, g.ne, g.N, sys.nrhs, sys.n,
                       g.d_num, g.d_type, Ae, be, g.be_stride, sys.be_real ? 1 : 0, sys.d_dc, sys.d_rhs[0]);
  else
    hipLaunchKernelGGL(k_ref_rhs<false>, dim3(blocks_for(total, 256)), dim3(256), 0, s, g.ne, g.N, sys.nrhs, sys.n,
                       g.d_num, g.d_type, Ae, be, g.be_stride, sys.be_real ? 1 : 0, sys.d_dc, sys.d_rhs[0]);
  return hipGetLastError();
}

}  // namespace gfem
