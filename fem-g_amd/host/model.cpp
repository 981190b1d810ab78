#include "model.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <random>
#include <stdexcept>

#include "../csrc/refelem.h"

namespace gmshfem::model {

namespace {
std::shared_ptr<Model> g_current;
std::atomic<uint64_t> g_uid{1};
}  // namespace

void setCurrent(std::shared_ptr<Model> m) { g_current = std::move(m); }
std::shared_ptr<Model> current() { return g_current; }

std::vector<std::pair<int, int>> Model::entities(int d) const
{
  std::vector<std::pair<int, int>> out;
  for(const auto &b : blocks)
    if(d < 0 || b.dim == d) out.emplace_back(b.dim, b.tag);
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

std::vector<int> Model::elementTypes(int d, int tag) const
{
  std::vector<int> out;
  for(const auto &b : blocks)
    if(b.dim == d && b.tag == tag) out.push_back(b.type);
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

std::vector<int> Model::elementTypes(int d) const
{
  std::vector<int> out;
  for(const auto &b : blocks)
    if(b.dim == d) out.push_back(b.type);
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

const Block *Model::block(int d, int tag, int type) const
{
  for(const auto &b : blocks)
    if(b.dim == d && b.tag == tag && b.type == type) return &b;
  return nullptr;
}

void Model::boundingBox(double mn[3], double mx[3]) const
{
  for(int j = 0; j < 3; ++j) {
    mn[j] = 1e300;
    mx[j] = -1e300;
  }
  for(int64_t i = 0; i < numNodes(); ++i)
    for(int j = 0; j < 3; ++j) {
      mn[j] = std::min(mn[j], xyz[3 * i + j]);
      mx[j] = std::max(mx[j], xyz[3 * i + j]);
    }
}

namespace {

// Lagrange nodes of every element live on the refined lattice of spacing
// h/p; a node is identified by its refined lattice coordinates, so elements
// sharing an edge/face share its nodes.
struct Lattice {
  int n, p, dim;
  int64_t side;
  std::vector<int64_t> tag;  // refined lattice index -> node tag (0 = none yet)
  std::vector<double> *xyz;

  int64_t index(const int64_t *c) const
  {
    int64_t id = 0;
    for(int d = dim - 1; d >= 0; --d) id = id * side + c[d];
    return id;
  }

  // node of lattice multi-index `a` (sums to p) in the element with
  // vertices V (refined coords) and physical vertex positions X
  int64_t node(const gfem::Lat &a, const std::vector<std::array<int64_t, 3>> &V,
               const std::vector<std::array<double, 3>> &X)
  {
    int64_t c[3] = {0, 0, 0};
    for(int d = 0; d < dim; ++d) {
      int64_t s = 0;
      for(size_t k = 0; k < V.size(); ++k) s += a[k] * V[k][d];
      c[d] = s / p;
    }
    const int64_t id = index(c);
    if(tag[id] == 0) {
      tag[id] = (int64_t)xyz->size() / 3 + 1;
      for(int d = 0; d < 3; ++d) {
        double s = 0.;
        for(size_t k = 0; k < X.size(); ++k) s += a[k] * X[k][d];
        xyz->push_back(s / p);
      }
    }
    return tag[id];
  }
};

void addElements(Block &b, int dim, int p, Lattice &L, const std::vector<std::vector<std::array<int64_t, 3>>> &cells,
                 const std::vector<double> &vx, int64_t vside)
{
  const std::vector<gfem::Lat> lat = gfem::lattice_nodes(dim, p);
  std::vector<std::array<double, 3>> X(dim + 1);
  std::vector<std::array<int64_t, 3>> V(dim + 1);
  for(const auto &cv : cells) {
    for(int k = 0; k <= dim; ++k) {
      int64_t vid = 0;
      for(int d = L.dim - 1; d >= 0; --d) vid = vid * vside + cv[k][d];
      for(int d = 0; d < 3; ++d) X[k][d] = vx[3 * vid + d];
      for(int d = 0; d < 3; ++d) V[k][d] = cv[k][d] * p;
    }
    for(const auto &a : lat) b.conn.push_back(L.node(a, V, X));
  }
}

}  // namespace

std::shared_ptr<Model> unitSquare(int n, int order, double jitter, uint64_t seed)
{
  if(n < 1 || order < 1 || order > 4) throw std::invalid_argument("unitSquare: bad n/order");
  auto m = std::make_shared<Model>();
  m->name = "unit_square";
  m->dim = 2;
  m->uid = g_uid++;
  const double h = 1.0 / n;
  const int64_t vs = n + 1;
  // coarse vertices, jittered in the interior
  std::vector<double> vx(3 * vs * vs, 0.);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-jitter * h, jitter * h);
  for(int64_t j = 0; j <= n; ++j)
    for(int64_t i = 0; i <= n; ++i) {
      const int64_t id = j * vs + i;
      vx[3 * id] = i * h;
      vx[3 * id + 1] = j * h;
      if(i > 0 && i < n && j > 0 && j < n && jitter != 0.) {
        vx[3 * id] += U(rng);
        vx[3 * id + 1] += U(rng);
      }
    }
  Lattice L{n, order, 2, (int64_t)n * order + 1, {}, &m->xyz};
  L.tag.assign(L.side * L.side, 0);
  using C = std::array<int64_t, 3>;
  // surface triangles
  {
    std::vector<std::vector<C>> cells;
    for(int64_t j = 0; j < n; ++j)
      for(int64_t i = 0; i < n; ++i) {
        const C v00{i, j, 0}, v10{i + 1, j, 0}, v01{i, j + 1, 0}, v11{i + 1, j + 1, 0};
        cells.push_back({v00, v10, v11});
        cells.push_back({v00, v11, v01});
      }
    Block b;
    b.dim = 2;
    b.tag = 1;
    b.order = order;
    b.type = gfem::elem_type_code(2, order);
    b.npe = (int)gfem::lattice_nodes(2, order).size();
    addElements(b, 2, order, L, cells, vx, vs);
    m->blocks.push_back(std::move(b));
  }
  // boundary curves, counter-clockwise
  for(int c = 1; c <= 4; ++c) {
    std::vector<std::vector<C>> cells;
    for(int64_t k = 0; k < n; ++k) {
      C a, b;
      if(c == 1) { a = {k, 0, 0}; b = {k + 1, 0, 0}; }
      if(c == 2) { a = {n, k, 0}; b = {n, k + 1, 0}; }
      if(c == 3) { a = {n - k, n, 0}; b = {n - k - 1, n, 0}; }
      if(c == 4) { a = {0, n - k, 0}; b = {0, n - k - 1, 0}; }
      cells.push_back({a, b});
    }
    Block b;
    b.dim = 1;
    b.tag = c;
    b.order = order;
    b.type = gfem::elem_type_code(1, order);
    b.npe = order + 1;
    addElements(b, 1, order, L, cells, vx, vs);
    m->blocks.push_back(std::move(b));
  }
  // corner points
  const C corners[4] = {{0, 0, 0}, {n, 0, 0}, {n, n, 0}, {0, n, 0}};
  for(int c = 0; c < 4; ++c) {
    Block b;
    b.dim = 0;
    b.tag = c + 1;
    b.type = 15;
    b.npe = 1;
    const int64_t c2[3] = {corners[c][0] * order, corners[c][1] * order, 0};
    b.conn.push_back(L.tag[L.index(c2)]);
    m->blocks.push_back(std::move(b));
  }
  return m;
}

std::shared_ptr<Model> unitCube(int n, int order, double jitter, uint64_t seed)
{
  if(n < 1 || order < 1 || order > 4) throw std::invalid_argument("unitCube: bad n/order");
  auto m = std::make_shared<Model>();
  m->name = "unit_cube";
  m->dim = 3;
  m->uid = g_uid++;
  const double h = 1.0 / n;
  const int64_t vs = n + 1;
  std::vector<double> vx(3 * vs * vs * vs, 0.);
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(-jitter * h, jitter * h);
  for(int64_t k = 0; k <= n; ++k)
    for(int64_t j = 0; j <= n; ++j)
      for(int64_t i = 0; i <= n; ++i) {
        const int64_t id = (k * vs + j) * vs + i;
        vx[3 * id] = i * h;
        vx[3 * id + 1] = j * h;
        vx[3 * id + 2] = k * h;
        if(i > 0 && i < n && j > 0 && j < n && k > 0 && k < n && jitter != 0.) {
          vx[3 * id] += U(rng);
          vx[3 * id + 1] += U(rng);
          vx[3 * id + 2] += U(rng);
        }
      }
  Lattice L{n, order, 3, (int64_t)n * order + 1, {}, &m->xyz};
  L.tag.assign(L.side * L.side * L.side, 0);
  using C = std::array<int64_t, 3>;
  {
    static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    static const int sign[6] = {1, -1, -1, 1, 1, -1};
    std::vector<std::vector<C>> cells;
    cells.reserve((size_t)n * n * n * 6);
    for(int64_t k = 0; k < n; ++k)
      for(int64_t j = 0; j < n; ++j)
        for(int64_t i = 0; i < n; ++i)
          for(int t = 0; t < 6; ++t) {
            C v[4];
            v[0] = {i, j, k};
            v[1] = v[0];
            v[1][perm[t][0]]++;
            v[2] = v[1];
            v[2][perm[t][1]]++;
            v[3] = {i + 1, j + 1, k + 1};
            if(sign[t] < 0) std::swap(v[2], v[3]);
            cells.push_back({v[0], v[1], v[2], v[3]});
          }
    Block b;
    b.dim = 3;
    b.tag = 1;
    b.order = order;
    b.type = gfem::elem_type_code(3, order);
    b.npe = (int)gfem::lattice_nodes(3, order).size();
    b.conn.reserve(cells.size() * b.npe);
    addElements(b, 3, order, L, cells, vx, vs);
    m->blocks.push_back(std::move(b));
  }
  // boundary faces: each face cell split along its min-max corner diagonal,
  // which is an edge of the adjacent tetrahedra
  for(int s = 1; s <= 6; ++s) {
    const int ax = (s - 1) / 2;            // normal axis
    const int64_t fix = ((s - 1) % 2) ? n : 0;
    const int a1 = (ax + 1) % 3, a2 = (ax + 2) % 3;
    std::vector<std::vector<C>> cells;
    for(int64_t q = 0; q < n; ++q)
      for(int64_t r = 0; r < n; ++r) {
        C c00, c10, c01, c11;
        c00[ax] = c10[ax] = c01[ax] = c11[ax] = fix;
        c00[a1] = r;     c00[a2] = q;
        c10[a1] = r + 1; c10[a2] = q;
        c01[a1] = r;     c01[a2] = q + 1;
        c11[a1] = r + 1; c11[a2] = q + 1;
        cells.push_back({c00, c10, c11});
        cells.push_back({c00, c11, c01});
      }
    Block b;
    b.dim = 2;
    b.tag = s;
    b.order = order;
    b.type = gfem::elem_type_code(2, order);
    b.npe = (int)gfem::lattice_nodes(2, order).size();
    addElements(b, 2, order, L, cells, vx, vs);
    m->blocks.push_back(std::move(b));
  }
  return m;
}

std::shared_ptr<Model> fromArrays(const std::string &name, int dim, const std::vector<double> &xyz,
                                  const std::vector<Block> &blocks)
{
  auto m = std::make_shared<Model>();
  m->name = name;
  m->dim = dim;
  m->xyz = xyz;
  m->blocks = blocks;
  m->uid = g_uid++;
  for(auto &b : m->blocks) {
    gfem::ElemType et;
    if(!gfem::elem_type_info(b.type, et)) throw std::invalid_argument("unknown element type");
    b.dim = et.dim;
    b.order = et.order;
    b.npe = et.npe;
    if(b.conn.size() % (size_t)b.npe) throw std::invalid_argument("connectivity size");
  }
  return m;
}

}  // namespace gmshfem::model
