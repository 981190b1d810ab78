// Hierarchical H1 basis on simplices (lines, triangles, tetrahedra) of order
// p >= 1: the space the reference builds with gmsh's "H1Legendre<p>" basis and
// "HierarchicalSolin" orientations (fem-master/src/field/
// FunctionSpaceHierarchicalH1.cpp:32-109, FunctionSpaceInterface.cpp:319-331,
// 495-519).  gmsh's own definitions are not part of the reference; this is
// the published hierarchic family of P. Solin, K. Segeth, I. Dolezel,
// "Higher-Order Finite Element Methods" (2004), sections 1.2 and 2.2-2.3:
//
//   Legendre L_n; Lobatto l_k = (L_k - L_{k-2}) / sqrt(2(2k-1)), k >= 2;
//   kernel functions phi_{k-2} = 4 l_k / (1 - x^2)
//                              = -4 / (k (k-1)) sqrt((2k-1)/2) L'_{k-1}
//   vertex a:            lambda_a
//   edge (a, b), k>=2:   lambda_a lambda_b phi_{k-2}(lambda_b - lambda_a)
//   face (a, b, c):      lambda_a lambda_b lambda_c phi_{n1-1}(lambda_b - lambda_a)
//                        phi_{n2-1}(lambda_c - lambda_b),  n1, n2 >= 1, n1 + n2 <= p - 1
//   interior (tets):     lambda_0..3 phi_{n1-1}(l1 - l0) phi_{n2-1}(l2 - l1)
//                        phi_{n3-1}(l3 - l2),  n1 + n2 + n3 <= p - 1
//
// Local vertices are the element's vertices in ascending global order, so an
// edge or face shared by two elements is traversed the same way in both and
// the traces agree (conforming without per-element orientation tables).
// Order of the local functions: vertices, edges (a < b lexicographic), faces
// (a < b < c lexicographic), interior; the modes of one entity by total
// degree, then by n1, then n2.
//
// Reference elements are gmsh's: line [-1, 1] (lambda_0 = (1 - xi)/2),
// triangle and tetrahedron with vertex 0 at the origin (lambda_0 = 1 - sum xi).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

namespace gfem {
namespace hier {

// functions of order p on a simplex of dimension dim (= dim P_p)
inline int count(int dim, int p)
{
  if(dim == 1) return p + 1;
  if(dim == 2) return (p + 1) * (p + 2) / 2;
  return (p + 1) * (p + 2) * (p + 3) / 6;
}

// Legendre L_n and its first two derivatives at x, n = 0..nmax (nmax < 16)
inline void legendre(int nmax, double x, double *L, double *dL, double *d2L)
{
  L[0] = 1.;
  dL[0] = 0.;
  d2L[0] = 0.;
  if(nmax < 1) return;
  L[1] = x;
  dL[1] = 1.;
  d2L[1] = 0.;
  for(int n = 1; n < nmax; ++n) {
    L[n + 1] = ((2. * n + 1.) * x * L[n] - n * L[n - 1]) / (n + 1.);
    dL[n + 1] = dL[n - 1] + (2. * n + 1.) * L[n];
    d2L[n + 1] = d2L[n - 1] + (2. * n + 1.) * dL[n];
  }
}

// kernel function phi_j and its derivative at x
inline void kernel(int j, double x, double &f, double &df)
{
  double L[16], dL[16], d2L[16];
  legendre(j + 1, x, L, dL, d2L);
  const double c = -4. / ((j + 1.) * (j + 2.)) * std::sqrt((2. * j + 3.) / 2.);
  f = c * dL[j + 1];
  df = c * d2L[j + 1];
}

// local entities of a simplex: edges and faces as local vertex tuples
inline std::vector<std::array<int, 2>> edges(int dim)
{
  if(dim == 1) return {{0, 1}};
  if(dim == 2) return {{0, 1}, {0, 2}, {1, 2}};
  return {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
}
inline std::vector<std::array<int, 3>> faces(int dim)
{
  if(dim == 2) return {{0, 1, 2}};
  if(dim == 3) return {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  return {};
}

// bubble modes of a face ((n1, n2)) or a tetrahedron ((n1, n2, n3)) of order p
inline std::vector<std::array<int, 3>> modes(int nvars, int p)
{
  std::vector<std::array<int, 3>> m;
  for(int s = nvars; s <= p - 1; ++s) {
    if(nvars == 2) {
      for(int a = 1; a <= s - 1; ++a) m.push_back({a, s - a, 0});
    }
    else {
      for(int a = 1; a <= s - 2; ++a)
        for(int b = 1; a + b <= s - 1; ++b) m.push_back({a, b, s - a - b});
    }
  }
  return m;
}

// One basis function's key: the entity (its vertices' global tags, in
// ascending order) and the mode on it.  n = number of vertices of the
// entity: 1 vertex, 2 edge, 3 face, 4 tetrahedron.
struct Key {
  int n;
  int64_t v[4];
  int mode;
};

// keys of the local functions of an element whose vertex tags, ascending,
// are vt[0..dim]
inline std::vector<Key> keys(int dim, int p, const int64_t *vt)
{
  std::vector<Key> k;
  const int nv = dim + 1;
  for(int a = 0; a < nv; ++a) k.push_back({1, {vt[a], 0, 0, 0}, 0});
  for(const auto &e : edges(dim))
    for(int m = 2; m <= p; ++m) k.push_back({2, {vt[e[0]], vt[e[1]], 0, 0}, m});
  const auto fm = modes(2, p);
  for(const auto &f : faces(dim))
    for(size_t m = 0; m < fm.size(); ++m) k.push_back({3, {vt[f[0]], vt[f[1]], vt[f[2]], 0}, (int)m});
  if(dim == 3) {
    const auto im = modes(3, p);
    for(size_t m = 0; m < im.size(); ++m) k.push_back({4, {vt[0], vt[1], vt[2], vt[3]}, (int)m});
  }
  return k;
}

// Values val[q*N + i] and reference gradients grad[(q*N + i)*3 + c] of the N
// functions at npts reference points pts[3*q ..] (unused coordinates 0).
inline void eval(int dim, int p, int npts, const double *pts, double *val, double *grad)
{
  const int N = count(dim, p), nv = dim + 1;
  double dl[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if(dim == 1) {
    dl[0][0] = -0.5;
    dl[1][0] = 0.5;
  }
  else {
    for(int j = 0; j < dim; ++j) {
      dl[0][j] = -1.;
      dl[j + 1][j] = 1.;
    }
  }
  const auto E = edges(dim);
  const auto Fc = faces(dim);
  const auto fm = modes(2, p), im = modes(3, p);
  for(int q = 0; q < npts; ++q) {
    const double *x = &pts[3 * q];
    double l[4] = {0, 0, 0, 0};
    if(dim == 1) {
      l[0] = 0.5 * (1. - x[0]);
      l[1] = 0.5 * (1. + x[0]);
    }
    else {
      l[0] = 1.;
      for(int j = 0; j < dim; ++j) {
        l[j + 1] = x[j];
        l[0] -= x[j];
      }
    }
    double *v = &val[(int64_t)q * N];
    double *g = &grad[(int64_t)q * N * 3];
    int i = 0;
    auto put = [&](double f, const double *df) {
      v[i] = f;
      for(int c = 0; c < 3; ++c) g[3 * i + c] = df[c];
      ++i;
    };
    for(int a = 0; a < nv; ++a) put(l[a], dl[a]);
    for(const auto &e : E) {
      const int a = e[0], b = e[1];
      const double s = l[b] - l[a], w = l[a] * l[b];
      for(int m = 2; m <= p; ++m) {
        double ph, dph;
        kernel(m - 2, s, ph, dph);
        double df[3];
        for(int c = 0; c < 3; ++c)
          df[c] = (dl[a][c] * l[b] + l[a] * dl[b][c]) * ph + w * dph * (dl[b][c] - dl[a][c]);
        put(w * ph, df);
      }
    }
    for(const auto &f : Fc) {
      const int a = f[0], b = f[1], cc = f[2];
      const double s1 = l[b] - l[a], s2 = l[cc] - l[b], w = l[a] * l[b] * l[cc];
      double dw[3];
      for(int c = 0; c < 3; ++c)
        dw[c] = dl[a][c] * l[b] * l[cc] + l[a] * dl[b][c] * l[cc] + l[a] * l[b] * dl[cc][c];
      for(const auto &md : fm) {
        double p1, dp1, p2, dp2;
        kernel(md[0] - 1, s1, p1, dp1);
        kernel(md[1] - 1, s2, p2, dp2);
        double df[3];
        for(int c = 0; c < 3; ++c)
          df[c] = dw[c] * p1 * p2 + w * (dp1 * (dl[b][c] - dl[a][c]) * p2 + p1 * dp2 * (dl[cc][c] - dl[b][c]));
        put(w * p1 * p2, df);
      }
    }
    if(dim == 3) {
      const double s1 = l[1] - l[0], s2 = l[2] - l[1], s3 = l[3] - l[2];
      const double w = l[0] * l[1] * l[2] * l[3];
      double dw[3];
      for(int c = 0; c < 3; ++c)
        dw[c] = dl[0][c] * l[1] * l[2] * l[3] + l[0] * dl[1][c] * l[2] * l[3] + l[0] * l[1] * dl[2][c] * l[3] +
                l[0] * l[1] * l[2] * dl[3][c];
      for(const auto &md : im) {
        double p1, dp1, p2, dp2, p3, dp3;
        kernel(md[0] - 1, s1, p1, dp1);
        kernel(md[1] - 1, s2, p2, dp2);
        kernel(md[2] - 1, s3, p3, dp3);
        double df[3];
        for(int c = 0; c < 3; ++c)
          df[c] = dw[c] * p1 * p2 * p3 + w * (dp1 * (dl[1][c] - dl[0][c]) * p2 * p3 +
                                              p1 * dp2 * (dl[2][c] - dl[1][c]) * p3 +
                                              p1 * p2 * dp3 * (dl[3][c] - dl[2][c]));
        put(w * p1 * p2 * p3, df);
      }
    }
  }
}

}  // namespace hier
}  // namespace gfem
