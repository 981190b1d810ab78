// Reference elements, node lattices and Gauss rules (host-side helpers).
//
// These are the gmsh-side services the reference obtains from libgmsh:
//   element properties     gmsh::model::mesh::getElementProperties
//                          (FunctionSpaceLagrange.cpp:26-40)
//   integration points     gmsh::model::mesh::getIntegrationPoints
//                          (FunctionSpaceInterface.cpp:470-481)
// gmsh reference elements: line u in [-1,1]; triangle (0,0),(1,0),(0,1);
// tetrahedron (0,0,0),(1,0,0),(0,1,0),(0,0,1).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

namespace gfem {

struct ElemType {
  int type, dim, order, npe;
};

// gmsh element type codes of the Lagrange simplices handled here
inline bool elem_type_info(int type, ElemType &out)
{
  static const int tab[][3] = {{15, 0, 0}, {1, 1, 1},  {8, 1, 2},  {26, 1, 3}, {27, 1, 4}, {2, 2, 1},  {9, 2, 2},
                               {21, 2, 3}, {23, 2, 4}, {4, 3, 1},  {11, 3, 2}, {29, 3, 3}, {30, 3, 4}};
  for(const auto &r : tab) {
    if(r[0] != type) continue;
    const int p = r[2];
    int n = 1;
    if(r[1] == 1) n = p + 1;
    if(r[1] == 2) n = (p + 1) * (p + 2) / 2;
    if(r[1] == 3) n = (p + 1) * (p + 2) * (p + 3) / 6;
    out = {type, r[1], p, n};
    return true;
  }
  return false;
}

inline int elem_type_code(int dim, int order)
{
  static const int codes[4][5] = {{15, 15, 15, 15, 15}, {0, 1, 8, 26, 27}, {0, 2, 9, 21, 23}, {0, 4, 11, 29, 30}};
  if(dim < 0 || dim > 3 || order < 1 || order > 4) return -1;
  return codes[dim][order];
}

using Lat = std::array<int, 4>;  // integer barycentric multi-index, sums to the order

// Node order of an order-q sub-simplex with corner lattice points P: corners,
// then edge interiors (edges in gmsh order, each walked from its first
// vertex), then face interiors, then the volume interior, recursively.
namespace detail {
inline Lat lerp(const Lat &a, const Lat &b, int i, int q)
{
  Lat r;
  for(int k = 0; k < 4; ++k) r[k] = a[k] + i * (b[k] - a[k]) / q;
  return r;
}

inline void tri_nodes(int q, const std::array<Lat, 3> &P, std::vector<Lat> &out)
{
  if(q == 0) {
    out.push_back(P[0]);
    return;
  }
  for(int v = 0; v < 3; ++v) out.push_back(P[v]);
  static const int ed[3][2] = {{0, 1}, {1, 2}, {2, 0}};
  for(const auto &e : ed)
    for(int i = 1; i < q; ++i) out.push_back(lerp(P[e[0]], P[e[1]], i, q));
  if(q >= 3) {
    std::array<Lat, 3> Q;
    for(int v = 0; v < 3; ++v) {
      Lat s = P[v];
      for(int w = 0; w < 3; ++w)
        if(w != v)
          for(int k = 0; k < 4; ++k) s[k] += (P[w][k] - P[v][k]) / q;
      Q[v] = s;
    }
    tri_nodes(q - 3, Q, out);
  }
}

inline void tet_nodes(int q, const std::array<Lat, 4> &P, std::vector<Lat> &out)
{
  if(q == 0) {
    out.push_back(P[0]);
    return;
  }
  for(int v = 0; v < 4; ++v) out.push_back(P[v]);
  static const int ed[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
  for(const auto &e : ed)
    for(int i = 1; i < q; ++i) out.push_back(lerp(P[e[0]], P[e[1]], i, q));
  if(q >= 3) {
    static const int fc[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
    for(const auto &f : fc) {
      std::array<Lat, 3> Q;
      for(int v = 0; v < 3; ++v) {
        Lat s = P[f[v]];
        for(int w = 0; w < 3; ++w)
          if(w != v)
            for(int k = 0; k < 4; ++k) s[k] += (P[f[w]][k] - P[f[v]][k]) / q;
        Q[v] = s;
      }
      tri_nodes(q - 3, Q, out);
    }
  }
  if(q >= 4) {
    std::array<Lat, 4> Q;
    for(int v = 0; v < 4; ++v) {
      Lat s = P[v];
      for(int w = 0; w < 4; ++w)
        if(w != v)
          for(int k = 0; k < 4; ++k) s[k] += (P[w][k] - P[v][k]) / q;
      Q[v] = s;
    }
    tet_nodes(q - 4, Q, out);
  }
}
}  // namespace detail

// Lattice multi-indices of the nodes of the order-p element of dimension dim.
inline std::vector<Lat> lattice_nodes(int dim, int p)
{
  std::vector<Lat> out;
  if(dim == 0) {
    out.push_back({0, 0, 0, 0});
    return out;
  }
  if(dim == 1) {
    out.push_back({p, 0, 0, 0});
    out.push_back({0, p, 0, 0});
    for(int i = 1; i < p; ++i) out.push_back({p - i, i, 0, 0});
    return out;
  }
  if(dim == 2) {
    detail::tri_nodes(p, {Lat{p, 0, 0, 0}, Lat{0, p, 0, 0}, Lat{0, 0, p, 0}}, out);
    return out;
  }
  detail::tet_nodes(p, {Lat{p, 0, 0, 0}, Lat{0, p, 0, 0}, Lat{0, 0, p, 0}, Lat{0, 0, 0, p}}, out);
  return out;
}

// Gauss-Jacobi nodes/weights for weight (1-x)^alpha on [-1,1]: eigenvalues of
// the Jacobi matrix by symmetric QR sweeps (Golub & Welsch, 1969).
inline void gauss_jacobi(int n, double alpha, std::vector<double> &x, std::vector<double> &w)
{
  std::vector<double> d(n), e(n, 0.), z0(n, 0.);
  for(int k = 0; k < n; ++k) {
    const double s = 2. * k + alpha;
    d[k] = (k == 0) ? -alpha / (alpha + 2.) : -alpha * alpha / (s * (s + 2.));
    if(k + 1 < n) {
      const double kk = k + 1, t = 2. * kk + alpha;
      e[k] = 2. * kk * (kk + alpha) / t / std::sqrt((t + 1.) * (t - 1.));
    }
  }
  // eigen-decomposition of the symmetric tridiagonal (d, e), tracking the first
  // component of each eigenvector (implicit QL with Wilkinson shifts)
  std::vector<std::vector<double>> Z(n, std::vector<double>(n, 0.));
  for(int i = 0; i < n; ++i) Z[i][i] = 1.;
  for(int l = 0; l < n; ++l) {
    for(int iter = 0; iter < 60; ++iter) {
      int m = l;
      for(; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if(std::fabs(e[m]) <= 1e-17 * dd) break;
      }
      if(m == l) break;
      double g = (d[l + 1] - d[l]) / (2. * e[l]);
      double r = std::hypot(g, 1.);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
      double s = 1., c = 1., p = 0.;
      int i = m - 1;
      for(; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if(r == 0.) {
          d[i + 1] -= p;
          e[m] = 0.;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2. * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        for(int k = 0; k < n; ++k) {
          f = Z[k][i + 1];
          Z[k][i + 1] = s * Z[k][i] + c * f;
          Z[k][i] = c * Z[k][i] - s * f;
        }
      }
      if(r == 0. && i >= l) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.;
    }
  }
  const double mu0 = std::pow(2., alpha + 1.) / (alpha + 1.);
  std::vector<int> idx(n);
  for(int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  x.resize(n);
  w.resize(n);
  for(int i = 0; i < n; ++i) {
    x[i] = d[idx[i]];
    w[i] = mu0 * Z[0][idx[i]] * Z[0][idx[i]];
  }
}

// "GaussN" on the gmsh reference element: collapsed (Duffy) Gauss-Jacobi
// product rule, exact for polynomials of total degree N.  pts are (u,v,w).
inline int gauss_rule(int dim, int degree, std::vector<double> &pts, std::vector<double> &w)
{
  const int n = std::max(degree, 0) / 2 + 1;
  pts.clear();
  w.clear();
  if(dim == 0) {
    pts = {0., 0., 0.};
    w = {1.};
    return 1;
  }
  std::vector<double> x[3], wx[3];
  for(int d = 0; d < dim; ++d) gauss_jacobi(n, dim == 1 ? 0. : double(dim - 1 - d), x[d], wx[d]);
  if(dim == 1) {
    for(int i = 0; i < n; ++i) {
      pts.insert(pts.end(), {x[0][i], 0., 0.});
      w.push_back(wx[0][i]);
    }
    return n;
  }
  std::vector<double> s[3], ws[3];
  for(int d = 0; d < dim; ++d) {
    const double a = double(dim - 1 - d);
    for(int i = 0; i < n; ++i) {
      s[d].push_back(0.5 * (x[d][i] + 1.));
      ws[d].push_back(wx[d][i] * std::pow(2., -a - 1.));
    }
  }
  if(dim == 2) {
    for(int i = 0; i < n; ++i)
      for(int j = 0; j < n; ++j) {
        pts.insert(pts.end(), {s[0][i], (1. - s[0][i]) * s[1][j], 0.});
        w.push_back(ws[0][i] * ws[1][j]);
      }
    return n * n;
  }
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < n; ++j)
      for(int k = 0; k < n; ++k) {
        pts.insert(pts.end(), {s[0][i], (1. - s[0][i]) * s[1][j], (1. - s[0][i]) * (1. - s[1][j]) * s[2][k]});
        w.push_back(ws[0][i] * ws[1][j] * ws[2][k]);
      }
  return n * n * n;
}

}  // namespace gfem
