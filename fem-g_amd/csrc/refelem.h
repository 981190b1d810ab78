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

// Fully symmetric simplex rules with positive weights and interior points,
// rows (u, v, w, weight) on the gmsh reference triangle (area 1/2) and
// tetrahedron (volume 1/6).  Orbit structures of the published rules
// (Strang & Fix / Dunavant 1985 for triangles, Keast 1986 / Walkington
// 2000 for tetrahedra), the values solved to machine precision from the
// moment equations (csrc/simplex_rules.py), exact for every monomial of the
// stated degree.
namespace rules {
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
}  // namespace rules

// table rule exact to `degree`, or null (then the collapsed rule)
inline const double (*simplex_table(int dim, int degree, int &q))[4]
{
  using namespace rules;
  q = 0;
  if(dim == 2) {
    if(degree <= 1) return q = 1, kTri1;
    if(degree == 2) return q = 3, kTri2;
    if(degree <= 4) return q = 6, kTri4;
    if(degree == 5) return q = 7, kTri5;
    if(degree == 6) return q = 12, kTri6;
    if(degree <= 8) return q = 16, kTri8;
  }
  if(dim == 3) {
    if(degree <= 1) return q = 1, kTet1;
    if(degree == 2) return q = 4, kTet2;
    if(degree <= 5) return q = 14, kTet5;
    if(degree == 6) return q = 24, kTet6;
  }
  return nullptr;
}

// "GaussN" on the gmsh reference element, exact for polynomials of total
// degree N: the symmetric table rule above (triangles to degree 8,
// tetrahedra to degree 6), else the collapsed (Duffy) Gauss-Jacobi product
// rule; Gauss-Legendre on lines.  pts are (u,v,w).
inline int gauss_rule(int dim, int degree, std::vector<double> &pts, std::vector<double> &w)
{
  const int n = std::max(degree, 0) / 2 + 1;
  pts.clear();
  w.clear();
  {
    int q = 0;
    const double(*t)[4] = simplex_table(dim, std::max(degree, 0), q);
    if(t) {
      for(int k = 0; k < q; ++k) {
        pts.insert(pts.end(), {t[k][0], t[k][1], t[k][2]});
        w.push_back(t[k][3]);
      }
      return q;
    }
  }
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
