/*
 * CPU oracle for the GmshFEM pre()/assemble() hot path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library, and only as the checker.  Single-threaded, FP64 /
 * complex128, element order, dense A_e, binary-search scatter: a plain
 * restatement of the reference algorithm (file:line cited per function).
 *
 * Pinning (tests/test_oracle.py, tests/test_golden.py):
 *   hilbert index / key     visiting orders of 2x2, 4x4, 2x2x2 and 4x4x4 grids
 *                           derived by hand from the reference tables
 *                           (tests/golden/hilbert_orders.json), raw index and
 *                           quantised key; each fixture rejects the transposed
 *                           and the reflected curve; face-adjacent path property
 *   gauss rules             exactness on monomials up to the stated degree
 *   lagrange basis          Kronecker property at nodes, partition of unity,
 *                           polynomial reproduction
 *   geometry                det = 2*area / 6*vol; J^-T grad reproduces linears
 *   element terms           P1 cotangent stiffness, P1 mass closed forms,
 *                           load = area/3, K*1 = 0, brute-force dense assembly
 *   AFrequency folding      lone dt^2 / dt mass terms on a P1 tet equal
 *                           -w^2 / -i w times the closed-form mass matrix
 *   scatter / pattern       structured-grid nnz closed form, sorted rows,
 *                           dense brute force on tiny meshes, A x = -b patch test
 *   DOF numbering           hand-derived numbering of 4x4 and 4x4x4 Dirichlet
 *                           P1 grids (tests/golden), grid-neighbour walk
 * parity unpinned (gmsh-side, no gmsh offline): exact gmsh Gauss tables,
 * gmsh's getBoundingBox tolerance, robin_hood iteration order (only breaks
 * Hilbert-key ties), gmsh node ordering of order>=3 elements.
 */
#include "oracle.h"

#include <complex.h>
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define OR_DOF_FIELD_OFFSET 1024ULL /* GMSHFEM_DOF_FIELD_OFFSET, fem-master/CMakeLists.txt:173-199 */
#define OR_FIELD_TAG 1ULL

/* AssembleType, fem-master/src/dofs/Dof.h:27-35 */
enum { AT_UNDEFINED = 0, AT_FIXED = 1, AT_UNKNOWN = 2, AT_UNKNOWN_BUBBLE = 3 };

typedef struct {
  int dim, tag, type, order, npe;
  int64_t ne;
  int64_t *conn; /* node tags */
  /* element -> dof map, filled by pre (FieldInterface::fillIndices) */
  int64_t *num;
  int32_t *atype;
} or_block;

typedef struct {
  int n;
  int *dim, *tag;
} or_domain;

typedef struct {
  int nops, nconst;
  int *ops;
  double *consts;
  cplx c;
} or_coef;

typedef struct {
  int bilinear;
  or_domain dom;
  int gauss;
  int form_dof, form_tf, hermitian, dt_order, rhs_idx;
  or_coef cl;
  cplx cr;
  int insertion;
} or_term;

typedef struct {
  uint64_t num_type, entity;
  int64_t num_dof; /* 1-based */
  int type;        /* 1 unknown, 2 fixed (dofs::Type) */
  int bubble;
  float xyz[3];
  cplx value;
} or_dof;

struct or_problem {
  int64_t n_nodes;
  double *xyz;
  int nblocks;
  or_block *blocks;
  int field_order;
  or_domain field_dom;
  int nconstr;
  or_domain *constr_dom;
  cplx *constr_val;
  int nterms;
  or_term *terms;
  int symmetric, dofs_sort;
  double frequency;
  int threads; /* element-loop threads of or_assemble (1: sequential, the checker) */
  /* dofs: open-addressing hash (numType, entity) -> index into dofs[] */
  int64_t ndofs, cap;
  or_dof *dofs;
  int64_t hcap;
  int64_t *htab;
  int64_t n_unknown, n_fixed, n_bubble;
  cplx *values_dc;
  /* csr */
  int64_t *ai, *aj;
  cplx *val;
  int nrhs;
  cplx **rhs;
};

/* ------------------------------------------------------------------ */
/* small utilities                                                    */
/* ------------------------------------------------------------------ */

static void *xcalloc(size_t n, size_t s)
{
  void *p = calloc(n ? n : 1, s);
  if(!p) abort();
  return p;
}

static void dom_copy(or_domain *d, int n, const int *dims, const int *tags)
{
  d->n = n;
  d->dim = xcalloc(n, sizeof(int));
  d->tag = xcalloc(n, sizeof(int));
  for(int i = 0; i < n; ++i) {
    d->dim[i] = dims[i];
    d->tag[i] = tags[i];
  }
}

static int dom_has(const or_domain *d, int dim, int tag)
{
  for(int i = 0; i < d->n; ++i)
    if(d->dim[i] == dim && d->tag[i] == tag) return 1;
  return 0;
}

static int dom_maxdim(const or_domain *d)
{
  int m = 0;
  for(int i = 0; i < d->n; ++i)
    if(d->dim[i] > m) m = d->dim[i];
  return m;
}

/* ------------------------------------------------------------------ */
/* element types (gmsh numbering, gmsh reference manual "MSH" section) */
/* ------------------------------------------------------------------ */

int or_elem_nodes(int t, int *dim, int *order)
{
  static const int tab[][3] = {
    /* type, dim, order */
    {15, 0, 0}, {1, 1, 1}, {8, 1, 2}, {26, 1, 3}, {27, 1, 4},
    {2, 2, 1},  {9, 2, 2}, {21, 2, 3}, {23, 2, 4},
    {4, 3, 1},  {11, 3, 2}, {29, 3, 3}, {30, 3, 4}};
  for(unsigned i = 0; i < sizeof(tab) / sizeof(tab[0]); ++i) {
    if(tab[i][0] == t) {
      *dim = tab[i][1];
      *order = tab[i][2];
      if(*dim == 0) return 1;
      if(*dim == 1) return *order + 1;
      if(*dim == 2) return (*order + 1) * (*order + 2) / 2;
      return (*order + 1) * (*order + 2) * (*order + 3) / 6;
    }
  }
  return -1;
}

/* Reference-node lattice of an order-p simplex, as integer barycentric
 * multi-indices (vertex k weight).  Ordering convention: vertices, edge
 * interiors, face interiors, volume interior, recursively; the gmsh
 * documented ordering for lines / triangles / tetrahedra.  parity unpinned
 * for order >= 3 (gmsh's exact face-node order cannot be checked offline). */
static int nodes_tri(int q, const int P[3][4], int out[][4], int n)
{
  if(q == 0) {
    for(int k = 0; k < 4; ++k) out[n][k] = P[0][k];
    return n + 1;
  }
  for(int v = 0; v < 3; ++v, ++n)
    for(int k = 0; k < 4; ++k) out[n][k] = P[v][k];
  static const int ed[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  for(int e = 0; e < 3; ++e)
    for(int i = 1; i < q; ++i, ++n)
      for(int k = 0; k < 4; ++k) out[n][k] = P[ed[e][0]][k] + i * (P[ed[e][1]][k] - P[ed[e][0]][k]) / q;
  if(q >= 3) {
    int Q[3][4];
    for(int v = 0; v < 3; ++v)
      for(int k = 0; k < 4; ++k) {
        int s = P[v][k];
        for(int w = 0; w < 3; ++w)
          if(w != v) s += (P[w][k] - P[v][k]) / q;
        Q[v][k] = s;
      }
    n = nodes_tri(q - 3, (const int(*)[4])Q, out, n);
  }
  return n;
}

static int nodes_tet(int q, const int P[4][4], int out[][4], int n)
{
  if(q == 0) {
    for(int k = 0; k < 4; ++k) out[n][k] = P[0][k];
    return n + 1;
  }
  for(int v = 0; v < 4; ++v, ++n)
    for(int k = 0; k < 4; ++k) out[n][k] = P[v][k];
  static const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
  for(int e = 0; e < 6; ++e)
    for(int i = 1; i < q; ++i, ++n)
      for(int k = 0; k < 4; ++k) out[n][k] = P[ed[e][0]][k] + i * (P[ed[e][1]][k] - P[ed[e][0]][k]) / q;
  if(q >= 3) {
    static const int fc[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
    for(int f = 0; f < 4; ++f) {
      int Q[3][4];
      for(int v = 0; v < 3; ++v)
        for(int k = 0; k < 4; ++k) {
          int a = fc[f][v];
          int s = P[a][k];
          for(int w = 0; w < 3; ++w)
            if(w != v) s += (P[fc[f][w]][k] - P[a][k]) / q;
          Q[v][k] = s;
        }
      n = nodes_tri(q - 3, (const int(*)[4])Q, out, n);
    }
  }
  if(q >= 4) {
    int Q[4][4];
    for(int v = 0; v < 4; ++v)
      for(int k = 0; k < 4; ++k) {
        int s = P[v][k];
        for(int w = 0; w < 4; ++w)
          if(w != v) s += (P[w][k] - P[v][k]) / q;
        Q[v][k] = s;
      }
    n = nodes_tet(q - 4, (const int(*)[4])Q, out, n);
  }
  return n;
}

static int ref_nodes(int dim, int p, int out[][4])
{
  int P[4][4] = {{0}};
  for(int v = 0; v <= dim; ++v) P[v][v] = p;
  if(dim == 1) {
    int n = 0;
    for(int v = 0; v < 2; ++v, ++n)
      for(int k = 0; k < 4; ++k) out[n][k] = P[v][k];
    for(int i = 1; i < p; ++i, ++n) {
      out[n][0] = p - i;
      out[n][1] = i;
      out[n][2] = out[n][3] = 0;
    }
    return n;
  }
  if(dim == 2) return nodes_tri(p, (const int(*)[4])P, out, 0);
  return nodes_tet(p, (const int(*)[4])P, out, 0);
}

int or_ref_nodes(int dim, int p, int *out)
{
  int nodes[64][4];
  const int n = ref_nodes(dim, p, nodes);
  for(int a = 0; a < n; ++a)
    for(int k = 0; k < 4; ++k) out[4 * a + k] = nodes[a][k];
  return n;
}

/* barycentric coordinates of a point of the gmsh reference element */
static void bary(int dim, const double *u, double *lam, double dlam[4][3])
{
  memset(dlam, 0, sizeof(double) * 12);
  if(dim == 1) { /* line on [-1, 1] */
    lam[0] = 0.5 * (1. - u[0]);
    lam[1] = 0.5 * (1. + u[0]);
    dlam[0][0] = -0.5;
    dlam[1][0] = 0.5;
    return;
  }
  double s = 0.;
  for(int j = 0; j < dim; ++j) {
    lam[j + 1] = u[j];
    dlam[j + 1][j] = 1.;
    dlam[0][j] = -1.;
    s += u[j];
  }
  lam[0] = 1. - s;
}

/* Lagrange basis of order p on the gmsh reference simplex (gmsh
 * "Lagrange"/"GradLagrange", fem-master/src/field/FunctionSpaceLagrange.cpp:114-137)
 * in closed form: phi_a = prod_k prod_{m<a_k} (p*lam_k - m)/(m+1).
 * grad has 3 components per function (3rd = 0 in 2-D, 2nd/3rd = 0 in 1-D). */
int or_lagrange(int dim, int p, const double *uvw, double *val, double *grad)
{
  int nodes[64][4];
  const int N = ref_nodes(dim, p, nodes);
  double lam[4] = {0}, dlam[4][3];
  bary(dim, uvw, lam, dlam);
  for(int a = 0; a < N; ++a) {
    double f = 1., g[3] = {0., 0., 0.};
    for(int k = 0; k <= dim; ++k) {
      for(int m = 0; m < nodes[a][k]; ++m) {
        const double fac = (p * lam[k] - m) / (m + 1);
        const double dfac = (double)p / (m + 1);
        for(int j = 0; j < 3; ++j) g[j] = g[j] * fac + f * dfac * dlam[k][j];
        f *= fac;
      }
    }
    if(val) val[a] = f;
    if(grad)
      for(int j = 0; j < 3; ++j) grad[3 * a + j] = g[j];
  }
  return N;
}

/* ------------------------------------------------------------------ */
/* Gauss rules ("GaussN" = exact for total degree N).  gmsh's own tables
 * (gmsh::model::mesh::getIntegrationPoints, called at
 * fem-master/src/field/FunctionSpaceInterface.cpp:470-481) are not
 * available offline: collapsed Gauss-Jacobi product rules are used
 * instead (exact to the same degree).  parity unpinned vs gmsh's tables;
 * pinned by exactness tests.                                            */
/* ------------------------------------------------------------------ */

/* Golub-Welsch for weight (1-x)^alpha on [-1,1], cyclic Jacobi eigensolver */
static void gauss_jacobi01(int n, double alpha, double *x, double *w)
{
  double A[16][16], V[16][16];
  memset(A, 0, sizeof(A));
  memset(V, 0, sizeof(V));
  for(int k = 0; k < n; ++k) {
    V[k][k] = 1.;
    const double s = 2. * k + alpha;
    A[k][k] = (k == 0) ? -alpha / (alpha + 2.) : -alpha * alpha / (s * (s + 2.));
    if(k >= 1) {
      const double b = 2. * k * (k + alpha) / s / sqrt((s + 1.) * (s - 1.));
      A[k][k - 1] = A[k - 1][k] = b;
    }
  }
  for(int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.;
    for(int p = 0; p < n; ++p)
      for(int q = p + 1; q < n; ++q) off += A[p][q] * A[p][q];
    if(off < 1e-32) break;
    for(int p = 0; p < n; ++p)
      for(int q = p + 1; q < n; ++q) {
        if(fabs(A[p][q]) < 1e-300) continue;
        const double th = (A[q][q] - A[p][p]) / (2. * A[p][q]);
        const double t = (th >= 0 ? 1. : -1.) / (fabs(th) + sqrt(th * th + 1.));
        const double c = 1. / sqrt(t * t + 1.), s = t * c;
        for(int k = 0; k < n; ++k) {
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for(int k = 0; k < n; ++k) {
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for(int k = 0; k < n; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  const double mu0 = pow(2., alpha + 1.) / (alpha + 1.);
  int idx[16];
  for(int i = 0; i < n; ++i) idx[i] = i;
  for(int i = 0; i < n; ++i)
    for(int j = i + 1; j < n; ++j)
      if(A[idx[j]][idx[j]] < A[idx[i]][idx[i]]) {
        int t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
      }
  for(int i = 0; i < n; ++i) {
    x[i] = A[idx[i]][idx[i]];
    w[i] = mu0 * V[0][idx[i]] * V[0][idx[i]];
  }
}

/* Symmetric simplex rules (positive weights, interior points): rows       */
/* (u, v, w, weight) on the reference triangle (area 1/2) / tetrahedron    */
/* (volume 1/6), the same table as the engine's refelem.h; pinned by the   */
/* monomial-exactness test.  Beyond the table: collapsed Gauss-Jacobi.      */
// 1 point, exact to degree 1
static const double kTri1[1][4] = {
  {0.33333333333333331, 0.33333333333333331, 0, 0.5},
};
// 3 points, exact to degree 2
static const double kTri2[3][4] = {
  {0.16666666666666666, 0.66666666666666674, 0, 0.16666666666666666},
  {0.66666666666666674, 0.16666666666666666, 0, 0.16666666666666666},
  {0.16666666666666666, 0.16666666666666666, 0, 0.16666666666666666},
};
// 6 points, exact to degree 4
static const double kTri4[6][4] = {
  {0.44594849091596472, 0.44594849091596472, 0, 0.11169079483900558},
  {0.10810301816807055, 0.44594849091596472, 0, 0.11169079483900558},
  {0.44594849091596472, 0.10810301816807055, 0, 0.11169079483900558},
  {0.091576213509770979, 0.81684757298045807, 0, 0.054975871827661095},
  {0.81684757298045807, 0.091576213509770979, 0, 0.054975871827661095},
  {0.091576213509770979, 0.091576213509770979, 0, 0.054975871827661095},
};
// 7 points, exact to degree 5
static const double kTri5[7][4] = {
  {0.33333333333333331, 0.33333333333333331, 0, 0.11250000000000335},
  {0.47014206410511605, 0.47014206410511605, 0, 0.066197076394251611},
  {0.059715871789767894, 0.47014206410511605, 0, 0.066197076394251611},
  {0.47014206410511605, 0.059715871789767894, 0, 0.066197076394251611},
  {0.10128650732345675, 0.79742698535308654, 0, 0.062969590272413958},
  {0.79742698535308654, 0.10128650732345675, 0, 0.062969590272413958},
  {0.10128650732345675, 0.10128650732345675, 0, 0.062969590272413958},
};
// 12 points, exact to degree 6
static const double kTri6[12][4] = {
  {0.24928674517092644, 0.50142650965814717, 0, 0.05839313786317591},
  {0.50142650965814717, 0.24928674517092644, 0, 0.05839313786317591},
  {0.24928674517092644, 0.24928674517092644, 0, 0.05839313786317591},
  {0.31035245103377096, 0.63650249912139989, 0, 0.041425537809195001},
  {0.63650249912139989, 0.31035245103377096, 0, 0.041425537809195001},
  {0.053145049844829151, 0.63650249912139989, 0, 0.041425537809195001},
  {0.63650249912139989, 0.053145049844829151, 0, 0.041425537809195001},
  {0.053145049844829151, 0.31035245103377096, 0, 0.041425537809195001},
  {0.31035245103377096, 0.053145049844829151, 0, 0.041425537809195001},
  {0.063089014491498452, 0.8738219710170031, 0, 0.025422453185100776},
  {0.8738219710170031, 0.063089014491498452, 0, 0.025422453185100776},
  {0.063089014491498452, 0.063089014491498452, 0, 0.025422453185100776},
};
// 16 points, exact to degree 8
static const double kTri8[16][4] = {
  {0.33333333333333331, 0.33333333333333331, 0, 0.07215780383892359},
  {0.17056930775180987, 0.65886138449638021, 0, 0.051608685267353016},
  {0.65886138449638021, 0.17056930775180987, 0, 0.051608685267353016},
  {0.17056930775180987, 0.17056930775180987, 0, 0.051608685267353016},
  {0.45929258829276531, 0.45929258829276531, 0, 0.047545817133622131},
  {0.081414823414469373, 0.45929258829276531, 0, 0.047545817133622131},
  {0.45929258829276531, 0.081414823414469373, 0, 0.047545817133622131},
  {0.050547228317028847, 0.89890554336594231, 0, 0.016229248811595872},
  {0.89890554336594231, 0.050547228317028847, 0, 0.016229248811595872},
  {0.050547228317028847, 0.050547228317028847, 0, 0.016229248811595872},
  {0.26311282963448862, 0.72849239295548485, 0, 0.013615157087227228},
  {0.72849239295548485, 0.26311282963448862, 0, 0.013615157087227228},
  {0.0083947774100264407, 0.72849239295548485, 0, 0.013615157087227228},
  {0.72849239295548485, 0.0083947774100264407, 0, 0.013615157087227228},
  {0.0083947774100264407, 0.26311282963448862, 0, 0.013615157087227228},
  {0.26311282963448862, 0.0083947774100264407, 0, 0.013615157087227228},
};
// 1 point, exact to degree 1
static const double kTet1[1][4] = {
  {0.25, 0.25, 0.25, 0.16666666666666666},
};
// 4 points, exact to degree 2
static const double kTet2[4][4] = {
  {0.1381966011250105, 0.1381966011250105, 0.58541019662496852, 0.041666666666666664},
  {0.1381966011250105, 0.58541019662496852, 0.1381966011250105, 0.041666666666666664},
  {0.58541019662496852, 0.1381966011250105, 0.1381966011250105, 0.041666666666666664},
  {0.1381966011250105, 0.1381966011250105, 0.1381966011250105, 0.041666666666666664},
};
// 14 points, exact to degree 5
static const double kTet5[14][4] = {
  {0.31088591926330045, 0.31088591926330045, 0.31088591926330045, 0.018781320953002337},
  {0.067342242210098657, 0.31088591926330045, 0.31088591926330045, 0.018781320953002337},
  {0.31088591926330045, 0.067342242210098657, 0.31088591926330045, 0.018781320953002337},
  {0.31088591926330045, 0.31088591926330045, 0.067342242210098657, 0.018781320953002337},
  {0.092735250310890707, 0.092735250310890707, 0.72179424906732792, 0.012248840519393511},
  {0.092735250310890707, 0.72179424906732792, 0.092735250310890707, 0.012248840519393511},
  {0.72179424906732792, 0.092735250310890707, 0.092735250310890707, 0.012248840519393511},
  {0.092735250310890707, 0.092735250310890707, 0.092735250310890707, 0.012248840519393511},
  {0.045503704125651745, 0.45449629587434826, 0.45449629587434826, 0.007091003462847207},
  {0.45449629587434826, 0.045503704125651745, 0.45449629587434826, 0.007091003462847207},
  {0.45449629587434826, 0.45449629587434826, 0.045503704125651745, 0.007091003462847207},
  {0.045503704125651745, 0.045503704125651745, 0.45449629587434826, 0.007091003462847207},
  {0.045503704125651745, 0.45449629587434826, 0.045503704125651745, 0.007091003462847207},
  {0.45449629587434826, 0.045503704125651745, 0.045503704125651745, 0.007091003462847207},
};
// 24 points, exact to degree 6
static const double kTet6[24][4] = {
  {0.32233789014227748, 0.32233789014227748, 0.32233789014227748, 0.0092261969239417659},
  {0.032986329573167605, 0.32233789014227748, 0.32233789014227748, 0.0092261969239417659},
  {0.32233789014227748, 0.032986329573167605, 0.32233789014227748, 0.0092261969239417659},
  {0.32233789014227748, 0.32233789014227748, 0.032986329573167605, 0.0092261969239417659},
  {0.063661001875016499, 0.26967233145831676, 0.60300566479165019, 0.0080357142857140525},
  {0.063661001875016499, 0.60300566479165019, 0.26967233145831676, 0.0080357142857140525},
  {0.26967233145831676, 0.063661001875016499, 0.60300566479165019, 0.0080357142857140525},
  {0.26967233145831676, 0.60300566479165019, 0.063661001875016499, 0.0080357142857140525},
  {0.60300566479165019, 0.063661001875016499, 0.26967233145831676, 0.0080357142857140525},
  {0.60300566479165019, 0.26967233145831676, 0.063661001875016499, 0.0080357142857140525},
  {0.063661001875016499, 0.063661001875016499, 0.60300566479165019, 0.0080357142857140525},
  {0.063661001875016499, 0.60300566479165019, 0.063661001875016499, 0.0080357142857140525},
  {0.60300566479165019, 0.063661001875016499, 0.063661001875016499, 0.0080357142857140525},
  {0.063661001875016499, 0.063661001875016499, 0.26967233145831676, 0.0080357142857140525},
  {0.063661001875016499, 0.26967233145831676, 0.063661001875016499, 0.0080357142857140525},
  {0.26967233145831676, 0.063661001875016499, 0.063661001875016499, 0.0080357142857140525},
  {0.21460287125914035, 0.21460287125914035, 0.35619138622257895, 0.0066537917096959063},
  {0.21460287125914035, 0.35619138622257895, 0.21460287125914035, 0.0066537917096959063},
  {0.35619138622257895, 0.21460287125914035, 0.21460287125914035, 0.0066537917096959063},
  {0.21460287125914035, 0.21460287125914035, 0.21460287125914035, 0.0066537917096959063},
  {0.040673958534612614, 0.040673958534612614, 0.87797812439616219, 0.0016795351758868431},
  {0.040673958534612614, 0.87797812439616219, 0.040673958534612614, 0.0016795351758868431},
  {0.87797812439616219, 0.040673958534612614, 0.040673958534612614, 0.0016795351758868431},
  {0.040673958534612614, 0.040673958534612614, 0.040673958534612614, 0.0016795351758868431},
};

static const double (*simplex_table(int dim, int deg, int *q))[4]
{
  *q = 0;
  if(dim == 2) {
    if(deg <= 1) { *q = 1; return kTri1; }
    if(deg == 2) { *q = 3; return kTri2; }
    if(deg <= 4) { *q = 6; return kTri4; }
    if(deg == 5) { *q = 7; return kTri5; }
    if(deg == 6) { *q = 12; return kTri6; }
    if(deg <= 8) { *q = 16; return kTri8; }
  }
  if(dim == 3) {
    if(deg <= 1) { *q = 1; return kTet1; }
    if(deg == 2) { *q = 4; return kTet2; }
    if(deg <= 5) { *q = 14; return kTet5; }
    if(deg == 6) { *q = 24; return kTet6; }
  }
  return NULL;
}

int or_gauss_rule(int dim, int deg, double *pts, double *w, int maxq)
{
  if(deg < 0) deg = 0;
  {
    int nq = 0;
    const double(*t)[4] = simplex_table(dim, deg, &nq);
    if(t) {
      if(nq > maxq) return -1;
      for(int k = 0; k < nq; ++k) {
        pts[3 * k] = t[k][0];
        pts[3 * k + 1] = t[k][1];
        pts[3 * k + 2] = t[k][2];
        w[k] = t[k][3];
      }
      return nq;
    }
  }
  const int n = (deg + 2) / 2; /* ceil((deg+1)/2) */
  if(n > 16) return -1;
  double x[3][16], wx[3][16];
  for(int d = 0; d < dim; ++d) {
    const double alpha = (dim == 1) ? 0. : (double)(dim - 1 - d);
    gauss_jacobi01(n, alpha, x[d], wx[d]);
  }
  int q = 0;
  if(dim == 1) {
    if(n > maxq) return -1;
    for(int i = 0; i < n; ++i, ++q) {
      pts[3 * q] = x[0][i];
      pts[3 * q + 1] = pts[3 * q + 2] = 0.;
      w[q] = wx[0][i];
    }
    return q;
  }
  /* map [-1,1] with weight (1-x)^a to [0,1] with weight (1-s)^a */
  double s[3][16], ws[3][16];
  for(int d = 0; d < dim; ++d) {
    const double alpha = (double)(dim - 1 - d);
    for(int i = 0; i < n; ++i) {
      s[d][i] = 0.5 * (x[d][i] + 1.);
      ws[d][i] = wx[d][i] * pow(2., -alpha - 1.);
    }
  }
  if(dim == 2) {
    if(n * n > maxq) return -1;
    for(int i = 0; i < n; ++i)
      for(int j = 0; j < n; ++j, ++q) {
        pts[3 * q] = s[0][i];
        pts[3 * q + 1] = (1. - s[0][i]) * s[1][j];
        pts[3 * q + 2] = 0.;
        w[q] = ws[0][i] * ws[1][j];
      }
    return q;
  }
  if(n * n * n > maxq) return -1;
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < n; ++j)
      for(int k = 0; k < n; ++k, ++q) {
        pts[3 * q] = s[0][i];
        pts[3 * q + 1] = (1. - s[0][i]) * s[1][j];
        pts[3 * q + 2] = (1. - s[0][i]) * (1. - s[1][j]) * s[2][k];
        w[q] = ws[0][i] * ws[1][j] * ws[2][k];
      }
  return q;
}

/* ------------------------------------------------------------------ */
/* Hilbert key, restated from fem-master/src/reorder/Hilbert.cpp:17-198 */
/* ------------------------------------------------------------------ */

typedef struct {
  int dim, N;
  uint8_t ti[8][3][8];   /* transinvgraycode[e][d][k] */
  uint8_t nextED[32][8];
  uint8_t tiED[32][8];
} hilbert_tab;

static void hilbert_init(hilbert_tab *H, int dim)
{
  const int N = 1 << dim;
  const unsigned mask = (~0u) ^ ((~0u) << dim);
  memset(H, 0, sizeof(*H));
  H->dim = dim;
  H->N = N;
  uint8_t gray[8], invgray[8], e[8], d[8];
  uint8_t tg[8][3][8];
  for(int i = 0; i < N; ++i) gray[i] = (uint8_t)(i ^ (i >> 1));
  for(int s = 0; s < N; ++s)
    for(int r = 0; r < dim; ++r)
      for(int i = 0; i < N; ++i) {
        const unsigned k = (unsigned)(i ^ s);
        tg[s][r][i] = (uint8_t)((k >> (r + 1)) | ((k << (dim - r - 1)) & mask));
      }
  invgray[0] = 0;
  for(int i = 1; i < N; ++i) {
    unsigned v = (unsigned)i;
    for(int j = 1; j < dim; ++j) v ^= (unsigned)i >> j;
    invgray[i] = (uint8_t)v;
  }
  for(int s = 0; s < N; ++s)
    for(int r = 0; r < dim; ++r)
      for(int k = 0; k < N; ++k) H->ti[s][r][k] = invgray[tg[s][r][k]];
  e[0] = 0;
  for(int i = 1; i < N; ++i) e[i] = gray[2 * ((i - 1) / 2)];
  d[0] = 0;
  for(int i = 1; i < N; ++i) {
    if(i < N / 2)
      d[i] = (uint8_t)((i % 2 == 0) ? gray[i - 1] % dim : gray[i] % dim);
    else
      d[i] = d[N - i - 1];
  }
  for(int ed = 0; ed < N * dim; ++ed) {
    const int ee = ed & (int)mask, dd = ed >> dim;
    if(dd >= dim) continue;
    for(int k = 0; k < N; ++k) {
      const unsigned l = e[H->ti[ee][dd][k]];
      const unsigned nE = (unsigned)ee ^ (((l << (dd + 1)) & mask) | (l >> (dim - dd - 1)));
      const unsigned nD = (unsigned)(dd + d[H->ti[ee][dd][k]] + 1) % (unsigned)dim;
      H->nextED[ed][k] = (uint8_t)((nD << dim) | nE);
      H->tiED[ed][k] = H->ti[ee][dd][k];
    }
  }
}

static uint64_t hilbert_index(const hilbert_tab *H, const uint32_t *u, int begbit)
{
  uint64_t h = 0;
  unsigned ed = 0;
  for(int m = begbit; m >= 0; --m) {
    unsigned l = 0;
    for(int n = H->dim - 1; n >= 0; --n) l ^= ((u[n] >> m) & 1u) << n;
    h <<= H->dim;
    h ^= H->tiED[ed][l];
    ed = H->nextED[ed][l];
  }
  return h;
}

/* quantisation of Hilbert<dim>::apply, fem-master/src/reorder/Hilbert.cpp:225-271
 * (float arithmetic, no FMA: this file is compiled with -ffp-contract=off) */
static void hilbert_scale(int dim, const float *min, const float *max, float *scale, float *smin, int *nbits)
{
  *nbits = (int)floor(64.0 / dim);
  const unsigned two = 1u << (*nbits - 1);
  for(int i = 0; i < dim; ++i) {
    scale[i] = (float)two / (max[i] - min[i]);
    smin[i] = -min[i] * scale[i];
  }
}

uint64_t or_hilbert_index(int dim, const uint32_t *u, int begbit)
{
  hilbert_tab H;
  hilbert_init(&H, dim);
  return hilbert_index(&H, u, begbit);
}

uint64_t or_hilbert_key(int dim, const float *f, const float *min, const float *max)
{
  hilbert_tab H;
  hilbert_init(&H, dim);
  float scale[3], smin[3];
  int nbits;
  hilbert_scale(dim, min, max, scale, smin, &nbits);
  uint32_t u[3];
  for(int j = 0; j < dim; ++j) {
    volatile float t = f[j] * scale[j];
    u[j] = (uint32_t)(t + smin[j]);
  }
  return hilbert_index(&H, u, nbits - 1);
}

/* ------------------------------------------------------------------ */
/* DOF dictionary (FieldInterface::_values, robin_hood map keyed by     */
/* (numType, entity)) -- a linear-probe table here                      */
/* ------------------------------------------------------------------ */

static uint64_t mix64(uint64_t a, uint64_t b)
{
  uint64_t h = a * 0x9E3779B97F4A7C15ULL ^ (b + 0x632BE59BD9B4E019ULL);
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ULL;
  h ^= h >> 32;
  return h;
}

static int64_t dof_find(const or_problem *p, uint64_t nt, uint64_t ent)
{
  if(!p->htab) return -1;
  uint64_t h = mix64(nt, ent) & (uint64_t)(p->hcap - 1);
  while(p->htab[h] >= 0) {
    const or_dof *d = &p->dofs[p->htab[h]];
    if(d->num_type == nt && d->entity == ent) return p->htab[h];
    h = (h + 1) & (uint64_t)(p->hcap - 1);
  }
  return -1;
}

static void dof_rehash(or_problem *p, int64_t newcap)
{
  free(p->htab);
  p->hcap = newcap;
  p->htab = xcalloc(newcap, sizeof(int64_t));
  for(int64_t i = 0; i < newcap; ++i) p->htab[i] = -1;
  for(int64_t i = 0; i < p->ndofs; ++i) {
    uint64_t h = mix64(p->dofs[i].num_type, p->dofs[i].entity) & (uint64_t)(newcap - 1);
    while(p->htab[h] >= 0) h = (h + 1) & (uint64_t)(newcap - 1);
    p->htab[h] = i;
  }
}

/* insert if absent; returns index of the new dof or -1 if it already existed
 * (FieldInterface::setDof / setValue, FieldInterface.cpp:457-464) */
static int64_t dof_insert(or_problem *p, uint64_t nt, uint64_t ent)
{
  if(dof_find(p, nt, ent) >= 0) return -1;
  if(p->ndofs + 1 > p->cap) {
    p->cap = p->cap ? 2 * p->cap : 1024;
    p->dofs = realloc(p->dofs, (size_t)p->cap * sizeof(or_dof));
    if(!p->dofs) abort();
  }
  if(2 * (p->ndofs + 1) > p->hcap) dof_rehash(p, p->hcap ? 2 * p->hcap : 2048);
  or_dof *d = &p->dofs[p->ndofs];
  memset(d, 0, sizeof(*d));
  d->num_type = nt;
  d->entity = ent;
  uint64_t h = mix64(nt, ent) & (uint64_t)(p->hcap - 1);
  while(p->htab[h] >= 0) h = (h + 1) & (uint64_t)(p->hcap - 1);
  p->htab[h] = p->ndofs;
  return p->ndofs++;
}

/* ------------------------------------------------------------------ */
/* problem construction                                               */
/* ------------------------------------------------------------------ */

or_problem *or_new(void)
{
  or_problem *p = xcalloc(1, sizeof(or_problem));
  p->nrhs = 1;
  p->threads = 1;
  return p;
}

static void coef_free(or_coef *c)
{
  free(c->ops);
  free(c->consts);
}

void or_free(or_problem *p)
{
  if(!p) return;
  free(p->xyz);
  for(int b = 0; b < p->nblocks; ++b) {
    free(p->blocks[b].conn);
    free(p->blocks[b].num);
    free(p->blocks[b].atype);
  }
  free(p->blocks);
  free(p->field_dom.dim);
  free(p->field_dom.tag);
  for(int c = 0; c < p->nconstr; ++c) {
    free(p->constr_dom[c].dim);
    free(p->constr_dom[c].tag);
  }
  free(p->constr_dom);
  free(p->constr_val);
  for(int t = 0; t < p->nterms; ++t) {
    free(p->terms[t].dom.dim);
    free(p->terms[t].dom.tag);
    coef_free(&p->terms[t].cl);
  }
  free(p->terms);
  free(p->dofs);
  free(p->htab);
  free(p->values_dc);
  free(p->ai);
  free(p->aj);
  free(p->val);
  if(p->rhs)
    for(int i = 0; i < p->nrhs; ++i) free(p->rhs[i]);
  free(p->rhs);
  free(p);
}

int or_set_nodes(or_problem *p, int64_t n, const double *xyz)
{
  free(p->xyz);
  p->n_nodes = n;
  p->xyz = xcalloc((size_t)(3 * n), sizeof(double));
  memcpy(p->xyz, xyz, sizeof(double) * 3 * (size_t)n);
  return 0;
}

int or_add_block(or_problem *p, int dim, int tag, int type, int64_t ne, const int64_t *conn)
{
  int d, o;
  const int npe = or_elem_nodes(type, &d, &o);
  if(npe < 0 || d != dim) return -1;
  p->blocks = realloc(p->blocks, sizeof(or_block) * (size_t)(p->nblocks + 1));
  or_block *b = &p->blocks[p->nblocks++];
  memset(b, 0, sizeof(*b));
  b->dim = dim;
  b->tag = tag;
  b->type = type;
  b->order = o;
  b->npe = npe;
  b->ne = ne;
  b->conn = xcalloc((size_t)(ne * npe), sizeof(int64_t));
  memcpy(b->conn, conn, sizeof(int64_t) * (size_t)(ne * npe));
  return p->nblocks - 1;
}

int or_set_field(or_problem *p, int order, int nent, const int *dims, const int *tags)
{
  p->field_order = order;
  dom_copy(&p->field_dom, nent, dims, tags);
  return 0;
}

int or_add_dirichlet(or_problem *p, int nent, const int *dims, const int *tags, double re, double im)
{
  p->constr_dom = realloc(p->constr_dom, sizeof(or_domain) * (size_t)(p->nconstr + 1));
  p->constr_val = realloc(p->constr_val, sizeof(cplx) * (size_t)(p->nconstr + 1));
  dom_copy(&p->constr_dom[p->nconstr], nent, dims, tags);
  p->constr_val[p->nconstr] = re + I * im;
  return p->nconstr++;
}

static void coef_set(or_coef *c, double re, double im, int nops, const int *ops, int nconst, const double *consts)
{
  memset(c, 0, sizeof(*c));
  c->c = re + I * im;
  c->nops = nops;
  c->nconst = nconst;
  if(nops > 0) {
    c->ops = xcalloc((size_t)(2 * nops), sizeof(int));
    memcpy(c->ops, ops, sizeof(int) * 2 * (size_t)nops);
  }
  if(nconst > 0) {
    c->consts = xcalloc((size_t)(2 * nconst), sizeof(double));
    memcpy(c->consts, consts, sizeof(double) * 2 * (size_t)nconst);
  }
}

static or_term *term_new(or_problem *p, int nent, const int *dims, const int *tags, int gauss)
{
  p->terms = realloc(p->terms, sizeof(or_term) * (size_t)(p->nterms + 1));
  or_term *t = &p->terms[p->nterms];
  memset(t, 0, sizeof(*t));
  dom_copy(&t->dom, nent, dims, tags);
  t->gauss = gauss;
  t->insertion = p->nterms++;
  return t;
}

int or_add_bilinear(or_problem *p, int nent, const int *dims, const int *tags, int gauss, int form_dof, int form_tf,
                    int hermitian, int dt_order, double cl_re, double cl_im, int cl_nops, const int *cl_ops,
                    int cl_nconst, const double *cl_consts, double cr_re, double cr_im)
{
  if(form_dof < 0 || form_dof > 1 || form_tf < 0 || form_tf > 1) return -1;
  or_term *t = term_new(p, nent, dims, tags, gauss);
  t->bilinear = 1;
  t->form_dof = form_dof;
  t->form_tf = form_tf;
  t->hermitian = hermitian;
  t->dt_order = dt_order;
  coef_set(&t->cl, cl_re, cl_im, cl_nops, cl_ops, cl_nconst, cl_consts);
  t->cr = cr_re + I * cr_im;
  return t->insertion;
}

int or_add_linear(or_problem *p, int nent, const int *dims, const int *tags, int gauss, int form_tf, int hermitian,
                  double f_re, double f_im, int f_nops, const int *f_ops, int f_nconst, const double *f_consts,
                  double cr_re, double cr_im, int rhs_idx)
{
  if(form_tf != 0 || rhs_idx < 0) return -1;
  or_term *t = term_new(p, nent, dims, tags, gauss);
  t->bilinear = 0;
  t->form_tf = form_tf;
  t->hermitian = hermitian;
  coef_set(&t->cl, f_re, f_im, f_nops, f_ops, f_nconst, f_consts);
  t->cr = cr_re + I * cr_im;
  t->rhs_idx = rhs_idx;
  if(rhs_idx + 1 > p->nrhs) p->nrhs = rhs_idx + 1;
  return t->insertion;
}

int or_set_threads(or_problem *p, int threads)
{
  p->threads = threads > 1 ? threads : 1;
  return 0;
}

int or_set_options(or_problem *p, int symmetric, double frequency, int dofs_sort)
{
  p->symmetric = symmetric;
  p->frequency = frequency;
  p->dofs_sort = dofs_sort;
  return 0;
}

/* ------------------------------------------------------------------ */
/* pre(): dofs, numbering, maps, pattern                               */
/* ------------------------------------------------------------------ */

/* entity iteration order: std::set<pair<int,int>> = (dim, tag) ascending */
static int ent_cmp(const void *a, const void *b)
{
  const int *x = a, *y = b;
  if(x[0] != y[0]) return x[0] - y[0];
  return x[1] - y[1];
}

static int sorted_entities(const or_domain *d, int (*out)[2])
{
  int n = 0;
  for(int i = 0; i < d->n; ++i) {
    int dup = 0;
    for(int j = 0; j < n; ++j)
      if(out[j][0] == d->dim[i] && out[j][1] == d->tag[i]) dup = 1;
    if(!dup) {
      out[n][0] = d->dim[i];
      out[n][1] = d->tag[i];
      ++n;
    }
  }
  qsort(out, (size_t)n, sizeof(int[2]), ent_cmp);
  return n;
}

/* element types of an entity, ascending type code (gmsh getElementTypes) */
static int entity_blocks(const or_problem *p, int dim, int tag, int *out)
{
  int n = 0;
  for(int b = 0; b < p->nblocks; ++b)
    if(p->blocks[b].dim == dim && p->blocks[b].tag == tag) out[n++] = b;
  for(int i = 0; i < n; ++i)
    for(int j = i + 1; j < n; ++j)
      if(p->blocks[out[j]].type < p->blocks[out[i]].type) {
        int t = out[i];
        out[i] = out[j];
        out[j] = t;
      }
  return n;
}

/* Lagrange keys of a block (gmsh getKeys with "Lagrange": typeKey 0,
 * entityKey = node tag, one key per element node).  gmsh-side convention,
 * parity unpinned. */
static uint64_t key_num_type(void) { return 0ULL + OR_DOF_FIELD_OFFSET * OR_FIELD_TAG; }

/* DofsFactory::_generateFixedDofs for constant Lagrange constraints,
 * fem-master/src/dofs/DofsFactory.cpp:39-74 */
static void generate_fixed(or_problem *p)
{
  for(int c = 0; c < p->nconstr; ++c) {
    int ents[256][2];
    const int ne = sorted_entities(&p->constr_dom[c], ents);
    for(int e = 0; e < ne; ++e) {
      int bl[16];
      const int nb = entity_blocks(p, ents[e][0], ents[e][1], bl);
      for(int k = 0; k < nb; ++k) {
        const or_block *b = &p->blocks[bl[k]];
        for(int64_t i = 0; i < b->ne * b->npe; ++i) {
          const int64_t id = dof_insert(p, key_num_type(), (uint64_t)b->conn[i]);
          if(id >= 0) {
            or_dof *d = &p->dofs[id];
            d->type = 2;
            const double *x = &p->xyz[3 * (b->conn[i] - 1)];
            d->xyz[0] = (float)x[0];
            d->xyz[1] = (float)x[1];
            d->xyz[2] = (float)x[2];
            d->value = p->constr_val[c];
            d->num_dof = ++p->n_fixed;
          }
        }
      }
    }
  }
}

/* DofsFactory::_generateUnknownDofs, fem-master/src/dofs/DofsFactory.cpp:406-460.
 * Lagrange keys are vertex-type keys (gmsh getKeysInformation gives dim 0
 * for node keys), so no key is a bubble.  parity unpinned (gmsh-side). */
static void generate_unknown(or_problem *p)
{
  int ents[256][2];
  const int ne = sorted_entities(&p->field_dom, ents);
  for(int e = 0; e < ne; ++e) {
    int bl[16];
    const int nb = entity_blocks(p, ents[e][0], ents[e][1], bl);
    for(int k = 0; k < nb; ++k) {
      const or_block *b = &p->blocks[bl[k]];
      for(int64_t i = 0; i < b->ne * b->npe; ++i) {
        const int64_t id = dof_insert(p, key_num_type(), (uint64_t)b->conn[i]);
        if(id >= 0) {
          or_dof *d = &p->dofs[id];
          d->type = 1;
          const double *x = &p->xyz[3 * (b->conn[i] - 1)];
          d->xyz[0] = (float)x[0];
          d->xyz[1] = (float)x[1];
          d->xyz[2] = (float)x[2];
          d->num_dof = ++p->n_unknown;
        }
      }
    }
  }
}

typedef struct {
  uint64_t h;
  int64_t idx;
} keyed;

/* stable LSD radix sort on the 64-bit key: same permutation as the
 * 2048-bucket + 7 byte-pass sort of Hilbert.cpp:241-336 (both are stable
 * sorts by the full key) */
static void stable_sort_keys(keyed *a, int64_t n)
{
  keyed *tmp = xcalloc((size_t)n, sizeof(keyed));
  for(int pass = 0; pass < 8; ++pass) {
    int64_t cnt[257] = {0};
    for(int64_t i = 0; i < n; ++i) cnt[((a[i].h >> (8 * pass)) & 0xFF) + 1]++;
    for(int k = 0; k < 256; ++k) cnt[k + 1] += cnt[k];
    for(int64_t i = 0; i < n; ++i) tmp[cnt[(a[i].h >> (8 * pass)) & 0xFF]++] = a[i];
    memcpy(a, tmp, sizeof(keyed) * (size_t)n);
  }
  free(tmp);
}

/* DofsManager::reorderWithHilbert(bubble = true), DofsManager.cpp:118-239.
 * Bounding box: gmsh::model::getBoundingBox(-1,-1) is taken as the node
 * bounding box (parity unpinned: gmsh may pad it by a geometric tolerance). */
static void reorder_hilbert(or_problem *p)
{
  int dim = 0;
  for(int b = 0; b < p->nblocks; ++b)
    if(p->blocks[b].dim > dim) dim = p->blocks[b].dim;
  if(dim < 2 || p->n_unknown == 0) return;
  double bmin[3], bmax[3];
  for(int j = 0; j < 3; ++j) {
    bmin[j] = 1e300;
    bmax[j] = -1e300;
  }
  for(int64_t i = 0; i < p->n_nodes; ++i)
    for(int j = 0; j < 3; ++j) {
      if(p->xyz[3 * i + j] < bmin[j]) bmin[j] = p->xyz[3 * i + j];
      if(p->xyz[3 * i + j] > bmax[j]) bmax[j] = p->xyz[3 * i + j];
    }
  float fmin[3], fmax[3];
  for(int j = 0; j < 3; ++j) {
    fmin[j] = (float)(bmin[j] - 0.000001 * fabs(bmin[j]));
    fmax[j] = (float)(bmax[j] + 0.000001 * fabs(bmax[j]));
  }
  hilbert_tab H;
  hilbert_init(&H, dim);
  float scale[3], smin[3];
  int nbits;
  hilbert_scale(dim, fmin, fmax, scale, smin, &nbits);
  keyed *nb = xcalloc((size_t)p->n_unknown, sizeof(keyed));
  keyed *bb = xcalloc((size_t)p->n_bubble + 1, sizeof(keyed));
  int64_t n1 = 0, n2 = 0;
  /* collection order: dictionary order (creation order here; the reference
   * uses robin_hood iteration order, which only matters for equal keys) */
  for(int64_t i = 0; i < p->ndofs; ++i) {
    const or_dof *d = &p->dofs[i];
    if(d->type != 1) continue;
    uint32_t u[3];
    for(int j = 0; j < dim; ++j) {
      volatile float t = d->xyz[j] * scale[j];
      u[j] = (uint32_t)(t + smin[j]);
    }
    keyed k = {hilbert_index(&H, u, nbits - 1), i};
    if(d->bubble)
      bb[n2++] = k;
    else
      nb[n1++] = k;
  }
  stable_sort_keys(nb, n1);
  stable_sort_keys(bb, n2);
  for(int64_t i = 0; i < n1; ++i) p->dofs[nb[i].idx].num_dof = i + 1;
  for(int64_t i = 0; i < n2; ++i) p->dofs[bb[i].idx].num_dof = p->n_unknown - p->n_bubble + i + 1;
  free(nb);
  free(bb);
}

/* FieldInterface::getIndices / fillIndices, FieldInterface.cpp:959-1076 */
static void fill_indices(or_problem *p, or_block *b)
{
  if(b->num) return;
  b->num = xcalloc((size_t)(b->ne * b->npe), sizeof(int64_t));
  b->atype = xcalloc((size_t)(b->ne * b->npe), sizeof(int32_t));
  for(int64_t i = 0; i < b->ne * b->npe; ++i) {
    const int64_t id = dof_find(p, key_num_type(), (uint64_t)b->conn[i]);
    if(id >= 0) {
      const or_dof *d = &p->dofs[id];
      b->num[i] = d->num_dof - 1;
      b->atype[i] = (d->type == 2) ? AT_FIXED : (d->bubble ? AT_UNKNOWN_BUBBLE : AT_UNKNOWN);
    }
    else {
      b->num[i] = 0;
      b->atype[i] = AT_UNDEFINED;
    }
  }
}

static int term_on_entity(const or_term *t, int dim, int tag) { return dom_has(&t->dom, dim, tag); }

/* dynamic sorted row used by addPattern (MatrixFactory.cpp:477-503) */
typedef struct {
  int64_t n, cap;
  int64_t *c;
} prow;

static void prow_insert(prow *r, int64_t j)
{
  int64_t lo = 0, hi = r->n;
  while(lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if(r->c[mid] < j)
      lo = mid + 1;
    else
      hi = mid;
  }
  if(lo < r->n && r->c[lo] == j) return;
  if(r->n + 1 > r->cap) {
    r->cap = r->cap ? 2 * r->cap : 16;
    r->c = realloc(r->c, sizeof(int64_t) * (size_t)r->cap);
  }
  memmove(&r->c[lo + 1], &r->c[lo], sizeof(int64_t) * (size_t)(r->n - lo));
  r->c[lo] = j;
  r->n++;
}

/* Formulation::pre, fem-master/src/problem/Formulation.cpp:897-1039 */
int or_pre(or_problem *p)
{
  if(!p->field_dom.n || !p->xyz) return -1;
  generate_fixed(p);
  generate_unknown(p);
  free(p->values_dc);
  p->values_dc = xcalloc((size_t)p->n_fixed + 1, sizeof(cplx));
  for(int64_t i = 0; i < p->ndofs; ++i)
    if(p->dofs[i].type == 2) p->values_dc[p->dofs[i].num_dof - 1] = p->dofs[i].value;
  if(p->dofs_sort == 0) reorder_hilbert(p);

  /* _buildPattern, Formulation.cpp:1370-1415: entities in gmsh order
   * (dim ascending, tag ascending); only entities carrying a bilinear term */
  const int64_t n = p->n_unknown;
  prow *rows = xcalloc((size_t)n + 1, sizeof(prow));
  int ents[1024][2];
  int nent = 0;
  for(int b = 0; b < p->nblocks; ++b) {
    int dup = 0;
    for(int k = 0; k < nent; ++k)
      if(ents[k][0] == p->blocks[b].dim && ents[k][1] == p->blocks[b].tag) dup = 1;
    if(!dup && nent < 1024) {
      ents[nent][0] = p->blocks[b].dim;
      ents[nent][1] = p->blocks[b].tag;
      ++nent;
    }
  }
  qsort(ents, (size_t)nent, sizeof(int[2]), ent_cmp);
  for(int e = 0; e < nent; ++e) {
    int hasbil = 0;
    for(int t = 0; t < p->nterms; ++t)
      if(p->terms[t].bilinear && term_on_entity(&p->terms[t], ents[e][0], ents[e][1])) hasbil = 1;
    if(!hasbil) continue;
    int bl[16];
    const int nb = entity_blocks(p, ents[e][0], ents[e][1], bl);
    for(int k = 0; k < nb; ++k) {
      or_block *b = &p->blocks[bl[k]];
      fill_indices(p, b);
      /* MatrixFactory::addPattern(m = #tf, n = #dof, tf, dof), MatrixFactory.cpp:458-539 */
      for(int64_t el = 0; el < b->ne; ++el) {
        const int64_t *num = &b->num[el * b->npe];
        const int32_t *at = &b->atype[el * b->npe];
        for(int i = 0; i < b->npe; ++i) {
          if(at[i] != AT_UNKNOWN) continue;
          for(int j = 0; j < b->npe; ++j) {
            if(at[j] != AT_UNKNOWN) continue;
            if(p->symmetric && num[j] < num[i]) continue;
            prow_insert(&rows[num[i]], num[j]);
          }
        }
      }
    }
  }
  /* finalizePattern, MatrixFactory.cpp:187-213 */
  free(p->ai);
  free(p->aj);
  p->ai = xcalloc((size_t)n + 1, sizeof(int64_t));
  for(int64_t i = 0; i < n; ++i) p->ai[i + 1] = p->ai[i] + rows[i].n;
  p->aj = xcalloc((size_t)p->ai[n] + 1, sizeof(int64_t));
  for(int64_t i = 0; i < n; ++i) {
    memcpy(&p->aj[p->ai[i]], rows[i].c, sizeof(int64_t) * (size_t)rows[i].n);
    free(rows[i].c);
  }
  free(rows);
  /* element maps of every block (also blocks carrying only linear terms) */
  for(int b = 0; b < p->nblocks; ++b) fill_indices(p, &p->blocks[b]);
  return 0;
}

int64_t or_n_unknown(const or_problem *p) { return p->n_unknown; }
int64_t or_n_fixed(const or_problem *p) { return p->n_fixed; }
int64_t or_nnz(const or_problem *p) { return p->ai ? p->ai[p->n_unknown] : 0; }

int or_get_csr(const or_problem *p, int64_t *ai, int64_t *aj)
{
  if(!p->ai) return -1;
  memcpy(ai, p->ai, sizeof(int64_t) * (size_t)(p->n_unknown + 1));
  memcpy(aj, p->aj, sizeof(int64_t) * (size_t)p->ai[p->n_unknown]);
  return 0;
}

int or_get_block_dofs(const or_problem *p, int block, int64_t *num, int32_t *type)
{
  if(block < 0 || block >= p->nblocks || !p->blocks[block].num) return -1;
  const or_block *b = &p->blocks[block];
  memcpy(num, b->num, sizeof(int64_t) * (size_t)(b->ne * b->npe));
  memcpy(type, b->atype, sizeof(int32_t) * (size_t)(b->ne * b->npe));
  return 0;
}

int or_get_node_dof(const or_problem *p, int64_t tag, int64_t *num, int32_t *type)
{
  const int64_t id = dof_find(p, key_num_type(), (uint64_t)tag);
  if(id < 0) {
    *num = 0;
    *type = 0;
    return 0;
  }
  *num = p->dofs[id].num_dof;
  *type = p->dofs[id].type == 2 ? AT_FIXED : AT_UNKNOWN;
  return 0;
}

/* ------------------------------------------------------------------ */
/* expression evaluation (ExecutionTree walk, fem-master/src/function/ */
/* executionTree/ExecutionTree.h:58-153; X node nullaryOperations.h:   */
/* 1412-1418; Exp unaryOperations.h:968-973)                           */
/* ------------------------------------------------------------------ */

static cplx coef_eval(const or_coef *c, const double *x)
{
  if(c->nops == 0) return c->c;
  cplx st[64];
  int sp = 0;
  for(int i = 0; i < c->nops; ++i) {
    const int op = c->ops[2 * i], arg = c->ops[2 * i + 1];
    switch(op) {
    case OR_OP_CONST: st[sp++] = c->consts[2 * arg] + I * c->consts[2 * arg + 1]; break;
    case OR_OP_X: st[sp++] = x[0]; break;
    case OR_OP_Y: st[sp++] = x[1]; break;
    case OR_OP_Z: st[sp++] = x[2]; break;
    case OR_OP_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
    case OR_OP_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
    case OR_OP_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
    case OR_OP_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
    case OR_OP_NEG: st[sp - 1] = -st[sp - 1]; break;
    case OR_OP_EXP: st[sp - 1] = cexp(st[sp - 1]); break;
    case OR_OP_SQRT: st[sp - 1] = csqrt(st[sp - 1]); break;
    case OR_OP_SIN: st[sp - 1] = csin(st[sp - 1]); break;
    case OR_OP_COS: st[sp - 1] = ccos(st[sp - 1]); break;
    case OR_OP_POW: --sp; st[sp - 1] = cpow(st[sp - 1], st[sp]); break;
    case OR_OP_ABS: st[sp - 1] = cabs(st[sp - 1]); break;
    default: return NAN;
    }
  }
  return sp > 0 ? st[sp - 1] : 0.;
}

/* ------------------------------------------------------------------ */
/* geometry: gmsh getJacobians restated (Assembler.cpp:130-145): J at   */
/* each Gauss point from the element's full geometric Lagrange basis,   */
/* column-major J(i,j) = dx_i/du_j; for 1-D/2-D elements the missing     */
/* columns are unit normals and det is the length / area element.       */
/* ------------------------------------------------------------------ */

typedef struct {
  int Q, N, dim;
  double w[512];
  double pts[3 * 512];
  double *bf;  /* [Q][N] field basis values     */
  double *gbf; /* [Q][N][3] field basis gradients */
  double *gg;  /* [Q][npe][3] geometric basis gradients */
  double *gv;  /* [Q][npe] geometric basis values */
} ruletab;

static int rule_build(ruletab *r, int dim, int order_field, int order_geo, int gauss)
{
  memset(r, 0, sizeof(*r));
  r->dim = dim;
  r->Q = or_gauss_rule(dim, gauss, r->pts, r->w, 512);
  if(r->Q <= 0) return -1;
  int nodes[64][4];
  r->N = ref_nodes(dim, order_field, nodes);
  const int G = ref_nodes(dim, order_geo, nodes);
  r->bf = xcalloc((size_t)(r->Q * r->N), sizeof(double));
  r->gbf = xcalloc((size_t)(r->Q * r->N * 3), sizeof(double));
  r->gv = xcalloc((size_t)(r->Q * G), sizeof(double));
  r->gg = xcalloc((size_t)(r->Q * G * 3), sizeof(double));
  for(int q = 0; q < r->Q; ++q) {
    or_lagrange(dim, order_field, &r->pts[3 * q], &r->bf[q * r->N], &r->gbf[q * r->N * 3]);
    or_lagrange(dim, order_geo, &r->pts[3 * q], &r->gv[q * G], &r->gg[q * G * 3]);
  }
  return 0;
}

static void rule_free(ruletab *r)
{
  free(r->bf);
  free(r->gbf);
  free(r->gv);
  free(r->gg);
}

static void cross3(const double *a, const double *b, double *c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* J (col-major), det, physical point of element el at Gauss point q */
static void geometry(const or_problem *p, const or_block *b, int64_t el, const ruletab *r, int q, double *J,
                     double *det, double *x)
{
  double c[3][3] = {{0}};
  x[0] = x[1] = x[2] = 0.;
  for(int k = 0; k < b->npe; ++k) {
    const double *X = &p->xyz[3 * (b->conn[el * b->npe + k] - 1)];
    const double *g = &r->gg[(q * b->npe + k) * 3];
    const double v = r->gv[q * b->npe + k];
    for(int i = 0; i < 3; ++i) {
      x[i] += v * X[i];
      for(int j = 0; j < b->dim; ++j) c[j][i] += X[i] * g[j];
    }
  }
  if(b->dim == 3) {
    double n[3];
    cross3(c[1], c[2], n);
    *det = c[0][0] * n[0] + c[0][1] * n[1] + c[0][2] * n[2];
  }
  else if(b->dim == 2) {
    double n[3];
    cross3(c[0], c[1], n);
    const double a = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for(int i = 0; i < 3; ++i) c[2][i] = n[i] / a;
    *det = a;
  }
  else {
    const double a = sqrt(c[0][0] * c[0][0] + c[0][1] * c[0][1] + c[0][2] * c[0][2]);
    double t[3] = {c[0][0] / a, c[0][1] / a, c[0][2] / a};
    double z[3] = {0., 0., 1.};
    if(fabs(t[2]) > 0.9) {
      z[0] = 1.;
      z[2] = 0.;
    }
    double n1[3], n2[3];
    cross3(z, t, n1);
    const double l1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
    for(int i = 0; i < 3; ++i) n1[i] /= l1;
    cross3(t, n1, n2);
    for(int i = 0; i < 3; ++i) {
      c[1][i] = n1[i];
      c[2][i] = n2[i];
    }
    *det = a;
  }
  for(int j = 0; j < 3; ++j)
    for(int i = 0; i < 3; ++i) J[i + 3 * j] = c[j][i];
}

/* inverse of J^T (FieldEvaluator<Form1>, FieldEvaluator.h:69) */
static void inv_transpose(const double *J, double *Ki)
{
  /* A = J^T: A(i,j) = J(j,i) = J[j + 3 i] */
  double A[3][3];
  for(int i = 0; i < 3; ++i)
    for(int j = 0; j < 3; ++j) A[i][j] = J[j + 3 * i];
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double d = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  double inv[3][3];
  inv[0][0] = c00 / d;
  inv[1][0] = c01 / d;
  inv[2][0] = c02 / d;
  inv[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / d;
  inv[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / d;
  inv[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / d;
  inv[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / d;
  inv[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / d;
  inv[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / d;
  for(int i = 0; i < 3; ++i)
    for(int j = 0; j < 3; ++j) Ki[3 * i + j] = inv[i][j]; /* row-major */
}

/* fieldExpression (C x N) of FieldEvaluator<Form0/Form1> at point q */
static int field_expr(int form, const ruletab *r, int q, const double *J, double *out)
{
  if(form == 0) {
    for(int a = 0; a < r->N; ++a) out[a] = r->bf[q * r->N + a];
    return 1;
  }
  double Ki[9];
  inv_transpose(J, Ki);
  for(int a = 0; a < r->N; ++a) {
    const double *g = &r->gbf[(q * r->N + a) * 3];
    for(int i = 0; i < 3; ++i) out[i * r->N + a] = Ki[3 * i] * g[0] + Ki[3 * i + 1] * g[1] + Ki[3 * i + 2] * g[2];
  }
  return 3;
}

/* ------------------------------------------------------------------ */
/* scatter: MatrixFactory::addValues (MatrixFactory.cpp:232-309, Unknown */
/* rows / Unknown + Fixed cols) and VectorFactory::addValues[DC]          */
/* (VectorFactory.cpp:121-235)                                            */
/* ------------------------------------------------------------------ */

static int64_t csr_pos(const or_problem *p, int64_t row, int64_t col)
{
  int64_t lo = p->ai[row], hi = p->ai[row + 1];
  while(lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if(p->aj[mid] < col)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

/* dst += v; with several element-loop threads the reference adds under
 * `omp atomic`, re and im separately (MatrixFactory.cpp:267-283,
 * VectorFactory.cpp:121-260) */
static void accum(cplx *dst, cplx v, int atomic)
{
  if(!atomic) {
    *dst += v;
    return;
  }
  double *d = (double *)dst;
  const double re = creal(v), im = cimag(v);
#pragma omp atomic
  d[0] += re;
#pragma omp atomic
  d[1] += im;
}

/* m = #tf (rows), n = #dof (cols), values = A_e col-major (n x m): values[i*n + j] */
static int add_values(or_problem *p, int m, int n, const int64_t *ni, const int32_t *ti, const int64_t *nj,
                      const int32_t *tj, const cplx *values, int atomic)
{
  int haveDC = 0;
  for(int i = 0; i < m; ++i) {
    if(ti[i] == AT_FIXED) {
      haveDC = 1;
      continue;
    }
    if(ti[i] != AT_UNKNOWN && ti[i] != AT_UNKNOWN_BUBBLE) continue;
    for(int j = 0; j < n; ++j) {
      if(tj[j] == AT_FIXED) {
        haveDC = 1;
        continue;
      }
      if(tj[j] != AT_UNKNOWN && tj[j] != AT_UNKNOWN_BUBBLE) continue;
      if(p->symmetric && nj[j] < ni[i]) continue;
      const int64_t pos = csr_pos(p, ni[i], nj[j]);
      accum(&p->val[pos], values[i * n + j], atomic);
    }
  }
  return haveDC;
}

static void add_values_dc(or_problem *p, cplx *b, int m, int n, const int64_t *ni, const int32_t *ti,
                          const int64_t *nj, const int32_t *tj, const cplx *values, int atomic)
{
  for(int i = 0; i < m; ++i) {
    if(ti[i] != AT_UNKNOWN && ti[i] != AT_UNKNOWN_BUBBLE) continue;
    for(int j = 0; j < n; ++j)
      if(tj[j] == AT_FIXED) accum(&b[ni[i]], values[j * n + i] * p->values_dc[nj[j]], atomic);
  }
}

static void add_rhs(cplx *b, int n, const int64_t *ni, const int32_t *ti, const cplx *values, int atomic)
{
  for(int i = 0; i < n; ++i)
    if(ti[i] == AT_UNKNOWN || ti[i] == AT_UNKNOWN_BUBBLE) accum(&b[ni[i]], values[i], atomic);
}

/* ------------------------------------------------------------------ */
/* assemble()                                                          */
/* ------------------------------------------------------------------ */

static int cmp_int(const void *a, const void *b) { return *(const int *)a - *(const int *)b; }

/* terms of an entity ordered as Assembler::sort does (map keyed by the
 * integration-type string "GaussN", then insertion order),
 * fem-master/src/term/Assembler.cpp:52-66 */
static int sorted_terms(const or_problem *p, int dim, int tag, int bilinear, int *out)
{
  int n = 0;
  for(int t = 0; t < p->nterms; ++t)
    if(p->terms[t].bilinear == bilinear && term_on_entity(&p->terms[t], dim, tag)) out[n++] = t;
  for(int i = 0; i < n; ++i)
    for(int j = i + 1; j < n; ++j) {
      char si[32], sj[32];
      snprintf(si, sizeof(si), "Gauss%d", p->terms[out[i]].gauss);
      snprintf(sj, sizeof(sj), "Gauss%d", p->terms[out[j]].gauss);
      const int c = strcmp(si, sj);
      if(c > 0 || (c == 0 && p->terms[out[j]].insertion < p->terms[out[i]].insertion)) {
        int t = out[i];
        out[i] = out[j];
        out[j] = t;
      }
    }
  return n;
}

/* BilinearTerm::evaluate, fem-master/src/term/BilinearTerm.cpp:303-343:
 * A_e(j_dof, i_tf) += w det (cl * E_dof)^T conj(cr * E_tf)  (Hermitian) */
static void bilinear_eval(const or_problem *p, const or_block *b, int64_t el, const or_term *t, const ruletab *r,
                          cplx *Ae)
{
  const int N = r->N;
  double ed[3 * 64], et[3 * 64];
  for(int q = 0; q < r->Q; ++q) {
    double J[9], det, x[3];
    geometry(p, b, el, r, q, J, &det, x);
    const int cd = field_expr(t->form_dof, r, q, J, ed);
    field_expr(t->form_tf, r, q, J, et);
    const cplx cl = coef_eval(&t->cl, x);
    const cplx cr = t->hermitian ? conj(t->cr) : t->cr;
    const double wd = r->w[q] * det;
    for(int i = 0; i < N; ++i)
      for(int j = 0; j < N; ++j) {
        double s = 0.;
        for(int c = 0; c < cd; ++c) s += ed[c * N + j] * et[c * N + i];
        Ae[j + i * N] += wd * (cl * cr * s);
      }
  }
}

/* LinearTerm::evaluate, fem-master/src/term/LinearTerm.cpp:176-217:
 * b_e(i) += w det conj(cr * E_tf(i)) f   (Hermitian) */
static void linear_eval(const or_problem *p, const or_block *b, int64_t el, const or_term *t, const ruletab *r,
                        cplx *be)
{
  const int N = r->N;
  for(int q = 0; q < r->Q; ++q) {
    double J[9], det, x[3];
    geometry(p, b, el, r, q, J, &det, x);
    const cplx f = coef_eval(&t->cl, x);
    const cplx cr = t->hermitian ? conj(t->cr) : t->cr;
    const double wd = r->w[q] * det;
    for(int i = 0; i < N; ++i) be[i] += wd * (cr * r->bf[q * N + i]) * f;
  }
}

/* Formulation::assemble / _assembleDim, Formulation.cpp:1151-1215, 1417-1493;
 * Assembler::_assembleBilinear / _assembleLinear, Assembler.cpp:157-479 */
int or_assemble(or_problem *p)
{
  if(!p->ai) return -1;
  const int64_t nnz = p->ai[p->n_unknown];
  free(p->val);
  p->val = xcalloc((size_t)nnz + 1, sizeof(cplx));
  if(p->rhs)
    for(int i = 0; i < p->nrhs; ++i) free(p->rhs[i]);
  free(p->rhs);
  p->rhs = xcalloc((size_t)p->nrhs, sizeof(cplx *));
  for(int i = 0; i < p->nrhs; ++i) p->rhs[i] = xcalloc((size_t)p->n_unknown + 1, sizeof(cplx));

  /* module: AFrequency when any bilinear term has a time derivative
   * (Formulation.cpp:1095-1149) */
  int freq_module = 0;
  for(int t = 0; t < p->nterms; ++t)
    if(p->terms[t].bilinear && p->terms[t].dt_order != 0) freq_module = 1;
  const double w = p->frequency;

  for(int dim = 3; dim >= 0; --dim) {
    int types[32], ntypes = 0;
    for(int b = 0; b < p->nblocks; ++b) {
      if(p->blocks[b].dim != dim) continue;
      int dup = 0;
      for(int k = 0; k < ntypes; ++k)
        if(types[k] == p->blocks[b].type) dup = 1;
      if(!dup) types[ntypes++] = p->blocks[b].type;
    }
    qsort(types, (size_t)ntypes, sizeof(int), cmp_int);
    for(int ti = 0; ti < ntypes; ++ti) {
      int tags[1024], ntags = 0;
      for(int b = 0; b < p->nblocks; ++b)
        if(p->blocks[b].dim == dim && p->blocks[b].type == types[ti] && ntags < 1024) tags[ntags++] = p->blocks[b].tag;
      qsort(tags, (size_t)ntags, sizeof(int), cmp_int);
      for(int ei = 0; ei < ntags; ++ei) {
        int bi = -1;
        for(int b = 0; b < p->nblocks; ++b)
          if(p->blocks[b].dim == dim && p->blocks[b].tag == tags[ei] && p->blocks[b].type == types[ti]) bi = b;
        or_block *b = &p->blocks[bi];
        int bil[64], lin[64];
        const int nbil = sorted_terms(p, dim, tags[ei], 1, bil);
        const int nlin = sorted_terms(p, dim, tags[ei], 0, lin);
        if(nbil + nlin == 0) continue;
        fill_indices(p, b);
        const int N = b->npe;
        if(p->field_order != b->order) return -2; /* Lagrange space must match the mesh order */
        if(nbil) {
          ruletab *rt = xcalloc((size_t)nbil, sizeof(ruletab));
          for(int k = 0; k < nbil; ++k)
            if(rule_build(&rt[k], dim, b->order, b->order, p->terms[bil[k]].gauss)) return -3;
          const int nth = p->threads > 1 ? p->threads : 1, atomic = nth > 1;
#pragma omp parallel num_threads(nth) if(nth > 1)
          {
            cplx *Ae = xcalloc((size_t)(N * N), sizeof(cplx));
            cplx *tmp = xcalloc((size_t)(N * N), sizeof(cplx));
#pragma omp for schedule(static)
            for(int64_t el = 0; el < b->ne; ++el) {
              memset(Ae, 0, sizeof(cplx) * (size_t)(N * N));
              for(int k = 0; k < nbil; ++k) {
                const or_term *t = &p->terms[bil[k]];
                if(!freq_module || t->dt_order == 0) {
                  bilinear_eval(p, b, el, t, &rt[k], Ae);
                }
                else {
                  memset(tmp, 0, sizeof(cplx) * (size_t)(N * N));
                  bilinear_eval(p, b, el, t, &rt[k], tmp);
                  const cplx f = (t->dt_order == 2) ? (cplx)(-w * w) : -I * w;
                  for(int k2 = 0; k2 < N * N; ++k2) Ae[k2] += f * tmp[k2];
                }
              }
              const int64_t *num = &b->num[el * N];
              const int32_t *at = &b->atype[el * N];
              if(add_values(p, N, N, num, at, num, at, Ae, atomic))
                for(int r = 0; r < p->nrhs; ++r) add_values_dc(p, p->rhs[r], N, N, num, at, num, at, Ae, atomic);
            }
            free(Ae);
            free(tmp);
          }
          for(int k = 0; k < nbil; ++k) rule_free(&rt[k]);
          free(rt);
        }
        if(nlin) {
          ruletab *rt = xcalloc((size_t)nlin, sizeof(ruletab));
          for(int k = 0; k < nlin; ++k)
            if(rule_build(&rt[k], dim, b->order, b->order, p->terms[lin[k]].gauss)) return -3;
          const int nth = p->threads > 1 ? p->threads : 1, atomic = nth > 1;
#pragma omp parallel num_threads(nth) if(nth > 1)
          {
            cplx *be = xcalloc((size_t)(N * p->nrhs), sizeof(cplx));
#pragma omp for schedule(static)
            for(int64_t el = 0; el < b->ne; ++el) {
              memset(be, 0, sizeof(cplx) * (size_t)(N * p->nrhs));
              for(int k = 0; k < nlin; ++k) {
                const or_term *t = &p->terms[lin[k]];
                linear_eval(p, b, el, t, &rt[k], &be[N * t->rhs_idx]);
              }
              for(int r = 0; r < p->nrhs; ++r)
                add_rhs(p->rhs[r], N, &b->num[el * N], &b->atype[el * N], &be[N * r], atomic);
            }
            free(be);
          }
          for(int k = 0; k < nlin; ++k) rule_free(&rt[k]);
          free(rt);
        }
      }
    }
  }
  return 0;
}

int or_get_values(const or_problem *p, double *re, double *im)
{
  if(!p->val) return -1;
  for(int64_t i = 0; i < p->ai[p->n_unknown]; ++i) {
    re[i] = creal(p->val[i]);
    im[i] = cimag(p->val[i]);
  }
  return 0;
}

int or_n_rhs(const or_problem *p) { return p->nrhs; }

int or_get_rhs(const or_problem *p, int idx, double *re, double *im)
{
  if(!p->rhs || idx < 0 || idx >= p->nrhs) return -1;
  for(int64_t i = 0; i < p->n_unknown; ++i) {
    re[i] = creal(p->rhs[idx][i]);
    im[i] = cimag(p->rhs[idx][i]);
  }
  return 0;
}
