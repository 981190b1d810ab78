#include "gmshfem.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "../csrc/hilbert.h"
#include "../csrc/refelem.h"

namespace gmshfem {

// ====================================================================== common

namespace common {

static double now()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
void Timer::tick() { _start = now(); }
void Timer::tock() { _t = now() - _start; }

namespace {
std::mutex g_ctx_mutex;
gfem_ctx *g_ctx = nullptr;
int g_device = 0;
uint64_t g_nodes_uid = 0;
}  // namespace

void setDevice(int device)
{
  std::lock_guard<std::mutex> lk(g_ctx_mutex);
  if(g_ctx && device != g_device) {
    gfem_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    g_nodes_uid = 0;
  }
  g_device = device;
}

gfem_ctx *deviceContext()
{
  std::lock_guard<std::mutex> lk(g_ctx_mutex);
  if(!g_ctx) {
    const gfem_status s = gfem_ctx_create(g_device, &g_ctx);
    if(s != GFEM_OK) throw Exception(std::string("gfem_ctx_create: ") + gfem_ctx_last_error(nullptr));
  }
  return g_ctx;
}

static void uploadNodes(const model::Model &m)
{
  gfem_ctx *ctx = deviceContext();
  std::lock_guard<std::mutex> lk(g_ctx_mutex);
  if(g_nodes_uid == m.uid) return;
  const gfem_status s = gfem_nodes_upload(ctx, m.numNodes(), m.xyz.data());
  if(s != GFEM_OK) throw Exception(std::string("gfem_nodes_upload: ") + gfem_ctx_last_error(ctx));
  g_nodes_uid = m.uid;
}

}  // namespace common

// ====================================================================== domain

namespace domain {
Domain Domain::operator|(const Domain &o) const
{
  Domain d = *this;
  d._ents.insert(o._ents.begin(), o._ents.end());
  return d;
}
int Domain::maxDim() const
{
  int m = 0;
  for(auto &e : _ents) m = std::max(m, e.first);
  return m;
}
}  // namespace domain

// ====================================================================== function

namespace function {

ScalarFunction::ScalarFunction(std::complex<double> v) : _constant(true), _value(v)
{
  _consts = {v.real(), v.imag()};
  _ops = {GFEM_OP_CONST, 0};
}

ScalarFunction ScalarFunction::coordinate(int axis)
{
  ScalarFunction f;
  f._constant = false;
  f._consts.clear();
  f._ops = {GFEM_OP_X + axis, 0};
  return f;
}

// append the program of b to a, shifting b's constant indices
static void append(std::vector<int32_t> &ops, std::vector<double> &consts, const ScalarFunction &b)
{
  const int shift = (int)consts.size() / 2;
  for(size_t k = 0; k < b.ops().size(); k += 2) {
    ops.push_back(b.ops()[k]);
    ops.push_back(b.ops()[k] == GFEM_OP_CONST ? b.ops()[k + 1] + shift : b.ops()[k + 1]);
  }
  consts.insert(consts.end(), b.consts().begin(), b.consts().end());
}

static std::complex<double> apply1(int op, std::complex<double> a)
{
  switch(op) {
  case GFEM_OP_NEG: return -a;
  case GFEM_OP_EXP: return std::exp(a);
  case GFEM_OP_SQRT: return std::sqrt(a);
  case GFEM_OP_SIN: return std::sin(a);
  case GFEM_OP_COS: return std::cos(a);
  case GFEM_OP_ABS: return std::abs(a);
  }
  throw common::Exception("unknown unary operation");
}

static std::complex<double> apply2(int op, std::complex<double> a, std::complex<double> b)
{
  switch(op) {
  case GFEM_OP_ADD: return a + b;
  case GFEM_OP_SUB: return a - b;
  case GFEM_OP_MUL: return a * b;
  case GFEM_OP_DIV: return a / b;
  case GFEM_OP_POW:
    if(a.imag() == 0. && b.imag() == 0. && (a.real() > 0. || b.real() == std::floor(b.real())))
      return std::pow(a.real(), b.real());
    return std::pow(a, b);
  }
  throw common::Exception("unknown binary operation");
}

ScalarFunction ScalarFunction::unary(int op, const ScalarFunction &a)
{
  if(a._constant) return ScalarFunction(apply1(op, a._value));
  ScalarFunction f;
  f._constant = false;
  f._ops.clear();
  f._consts.clear();
  append(f._ops, f._consts, a);
  f._ops.push_back(op);
  f._ops.push_back(0);
  return f;
}

ScalarFunction ScalarFunction::binary(int op, const ScalarFunction &a, const ScalarFunction &b)
{
  if(a._constant && b._constant) return ScalarFunction(apply2(op, a._value, b._value));
  ScalarFunction f;
  f._constant = false;
  f._ops.clear();
  f._consts.clear();
  append(f._ops, f._consts, a);
  append(f._ops, f._consts, b);
  f._ops.push_back(op);
  f._ops.push_back(0);
  return f;
}

std::complex<double> ScalarFunction::evaluate(double x, double y, double z) const
{
  if(_constant) return _value;
  std::vector<std::complex<double>> st;
  for(size_t k = 0; k < _ops.size(); k += 2) {
    const int op = _ops[k], arg = _ops[k + 1];
    if(op == GFEM_OP_CONST)
      st.emplace_back(_consts[2 * arg], _consts[2 * arg + 1]);
    else if(op == GFEM_OP_X)
      st.emplace_back(x);
    else if(op == GFEM_OP_Y)
      st.emplace_back(y);
    else if(op == GFEM_OP_Z)
      st.emplace_back(z);
    else if((op >= GFEM_OP_ADD && op <= GFEM_OP_DIV) || op == GFEM_OP_POW) {
      const auto b = st.back();
      st.pop_back();
      st.back() = apply2(op, st.back(), b);
    }
    else
      st.back() = apply1(op, st.back());
  }
  return st.back();
}

ScalarFunction operator+(const ScalarFunction &a, const ScalarFunction &b) { return ScalarFunction::binary(GFEM_OP_ADD, a, b); }
ScalarFunction operator-(const ScalarFunction &a, const ScalarFunction &b) { return ScalarFunction::binary(GFEM_OP_SUB, a, b); }
ScalarFunction operator*(const ScalarFunction &a, const ScalarFunction &b) { return ScalarFunction::binary(GFEM_OP_MUL, a, b); }
ScalarFunction operator/(const ScalarFunction &a, const ScalarFunction &b) { return ScalarFunction::binary(GFEM_OP_DIV, a, b); }
ScalarFunction operator-(const ScalarFunction &a) { return ScalarFunction::unary(GFEM_OP_NEG, a); }
ScalarFunction x() { return ScalarFunction::coordinate(0); }
ScalarFunction y() { return ScalarFunction::coordinate(1); }
ScalarFunction z() { return ScalarFunction::coordinate(2); }
ScalarFunction exp(const ScalarFunction &a) { return ScalarFunction::unary(GFEM_OP_EXP, a); }
ScalarFunction sqrt(const ScalarFunction &a) { return ScalarFunction::unary(GFEM_OP_SQRT, a); }
ScalarFunction sin(const ScalarFunction &a) { return ScalarFunction::unary(GFEM_OP_SIN, a); }
ScalarFunction cos(const ScalarFunction &a) { return ScalarFunction::unary(GFEM_OP_COS, a); }
ScalarFunction abs(const ScalarFunction &a) { return ScalarFunction::unary(GFEM_OP_ABS, a); }
ScalarFunction pow(const ScalarFunction &a, const ScalarFunction &b) { return ScalarFunction::binary(GFEM_OP_POW, a, b); }

}  // namespace function

// ====================================================================== field / equation

namespace field {
static int g_field_tag = 0;
Field::Field(const std::string &name, const domain::Domain &dom, int order)
  : _name(name), _domain(dom), _order(order), _tag(g_field_tag++)
{
  if(order < 1 || order > 4) throw common::Exception("Lagrange order must be in 1..4");
}
Field::Field(const std::string &name, const domain::Domain &dom, FunctionSpaceTypeForm0 type, int order)
  : _name(name), _domain(dom), _type(type), _order(order), _tag(g_field_tag++)
{
  if(type == FunctionSpaceTypeForm0::Lagrange && (order < 1 || order > 4))
    throw common::Exception("Lagrange order must be in 1..4");
  if(type == FunctionSpaceTypeForm0::HierarchicalH1 && (order < 1 || order > GFEM_HIER_MAX_ORDER))
    throw common::Exception("HierarchicalH1 order must be in 1.." + std::to_string(GFEM_HIER_MAX_ORDER));
}
void Field::addConstraint(const domain::Domain &dom, const function::ScalarFunction &value)
{
  _constraints.emplace_back(dom, value);
}
}  // namespace field

namespace equation {
Unknown dof(const field::Field &u)
{
  Unknown r;
  r.field = &u;
  return r;
}
Unknown tf(const field::Field &u)
{
  Unknown r = dof(u);
  r.tf = true;
  return r;
}
Unknown dt_dof(const field::Field &u)
{
  Unknown r = dof(u);
  r.dt = 1;
  return r;
}
Unknown dt2_dof(const field::Field &u)
{
  Unknown r = dof(u);
  r.dt = 2;
  return r;
}
Unknown grad(const Unknown &u)
{
  if(u.form != 0) throw common::Exception("grad of a gradient");
  Unknown r = u;
  r.form = 1;
  return r;
}
Unknown operator*(const function::ScalarFunction &c, const Unknown &u)
{
  Unknown r = u;
  r.coef = c * u.coef;
  return r;
}
}  // namespace equation

// ====================================================================== reorder

namespace reorder {

// curve tables and quantisation: csrc/hilbert.h (Hilbert.cpp:17-271)
using Curve = gfem::HilbertCurve;
using Quant = gfem::HilbertQuant;

uint64_t hilbertIndex(int dim, const uint32_t *u, int top) { return Curve(dim).index(u, top); }

uint64_t hilbertKey(int dim, const float *f, const float *min, const float *max)
{
  const Curve c(dim);
  const Quant q(dim, min, max);
  uint32_t u[3] = {0, 0, 0};
  for(int i = 0; i < dim; ++i) u[i] = q(i, f[i]);
  return c.index(u, q.bits - 1);
}

std::vector<int64_t> hilbertOrder(int dim, const std::vector<float> &coords, const float *min, const float *max)
{
  const Curve c(dim);
  const Quant q(dim, min, max);
  const int64_t n = (int64_t)coords.size() / 3;
  std::vector<std::pair<uint64_t, int64_t>> keys(n);
  for(int64_t i = 0; i < n; ++i) {
    uint32_t u[3] = {0, 0, 0};
    for(int k = 0; k < dim; ++k) u[k] = q(k, coords[3 * i + k]);
    keys[i] = {c.index(u, q.bits - 1), i};
  }
  std::stable_sort(keys.begin(), keys.end(),
                   [](const std::pair<uint64_t, int64_t> &a, const std::pair<uint64_t, int64_t> &b) {
                     return a.first < b.first;
                   });
  std::vector<int64_t> order(n);
  for(int64_t i = 0; i < n; ++i) order[i] = keys[i].second;
  return order;
}

}  // namespace reorder

// ====================================================================== problem

namespace problem {

template <class T>
struct ScalarKind;
template <>
struct ScalarKind<double> {
  static constexpr gfem_scalar value = GFEM_SCALAR_F64;
  static constexpr bool complex = false;
};
template <>
struct ScalarKind<float> {
  static constexpr gfem_scalar value = GFEM_SCALAR_F32;
  static constexpr bool complex = false;
};
template <>
struct ScalarKind<std::complex<double>> {
  static constexpr gfem_scalar value = GFEM_SCALAR_C128;
  static constexpr bool complex = true;
};
template <>
struct ScalarKind<std::complex<float>> {
  static constexpr gfem_scalar value = GFEM_SCALAR_C64;
  static constexpr bool complex = true;
};

template <class T>
T scalarFrom(const std::complex<double> &v)
{
  if constexpr(ScalarKind<T>::complex)
    return T((typename T::value_type)v.real(), (typename T::value_type)v.imag());
  else
    return T(v.real());
}

static int parseGauss(const std::string &s)
{
  if(s.rfind("Gauss", 0) != 0) throw common::Exception("unknown integration type '" + s + "'");
  try {
    return std::stoi(s.substr(5));
  }
  catch(...) {
    throw common::Exception("unknown integration type '" + s + "'");
  }
}

template <class T>
Formulation<T>::Formulation(const std::string &name) : _name(name)
{
}

template <class T>
Formulation<T>::~Formulation()
{
  releaseSystem();
  releasePre();
}

template <class T>
void Formulation<T>::releaseSystem()
{
  if(_sys) gfem_system_destroy(_sys);
  _sys = nullptr;
}

template <class T>
unsigned int Formulation<T>::integral(const equation::Unknown &lhs, const equation::Unknown &rhs,
                                      const domain::Domain &dom, const std::string &integrationType,
                                      ProductType product)
{
  if(lhs.tf || !rhs.tf) throw common::Exception("integral(lhs, rhs): lhs must be a dof, rhs a test function");
  if(lhs.field != rhs.field) throw common::Exception("one unknown field per formulation");
  if(lhs.form != rhs.form) throw common::Exception("bilinear forms must be (dof, tf) or (grad dof, grad tf)");
  if(rhs.dt) throw common::Exception("time derivative on a test function");
  if(_field && _field != lhs.field) throw common::Exception("one unknown field per formulation");
  _field = lhs.field;
  Term t;
  t.bilinear = true;
  t.lhs = lhs;
  t.rhs = rhs;
  t.dom = dom;
  t.gauss = parseGauss(integrationType);
  t.product = product;
  _terms.push_back(t);
  return (unsigned)_terms.size() - 1;
}

template <class T>
unsigned int Formulation<T>::integral(const function::ScalarFunction &f, const equation::Unknown &rhs,
                                      const domain::Domain &dom, const std::string &integrationType,
                                      ProductType product, unsigned int rhsIdx)
{
  if(!rhs.tf) throw common::Exception("integral(f, rhs): rhs must be a test function");
  if(rhs.form != 0) throw common::Exception("linear terms take tf(u)");
  if(_field && _field != rhs.field) throw common::Exception("one unknown field per formulation");
  _field = rhs.field;
  Term t;
  t.bilinear = false;
  t.rhs = rhs;
  t.source = f;
  t.dom = dom;
  t.gauss = parseGauss(integrationType);
  t.product = product;
  t.rhsIdx = rhsIdx;
  _numRHS = std::max(_numRHS, rhsIdx + 1);
  _terms.push_back(t);
  return (unsigned)_terms.size() - 1;
}

template <class T>
void Formulation<T>::setTermCoefficient(unsigned int tag, const function::ScalarFunction &f)
{
  if(tag >= _terms.size()) throw common::Exception("setTermCoefficient: unknown term " + std::to_string(tag));
  Term &t = _terms[tag];
  function::ScalarFunction &c = t.bilinear ? t.lhs.coef : t.source;
  if(c.isConstant() != f.isConstant() || (!f.isConstant() && f.ops() != c.ops()))
    throw common::Exception("setTermCoefficient: the coefficient must keep its form (constant, or the same program)");
  c = f;
  if(!_sys) return;
  gfem_ctx *ctx = common::deviceContext();
  for(const auto &gk : _termSlots[tag]) {
    gfem_status st;
    if(f.isConstant())
      st = gfem_term_update(_sys, gk.first, gk.second, f.value().real(), f.value().imag(), 0, nullptr, nullptr);
    else
      st = gfem_term_update(_sys, gk.first, gk.second, 0., 0., (int)f.consts().size() / 2, f.consts().data(), nullptr);
    if(st != GFEM_OK) throw common::Exception(std::string("gfem_term_update: ") + gfem_ctx_last_error(ctx));
  }
}

template <class T>
std::vector<int> Formulation<T>::setShiftedLaplacian(unsigned int tag, double beta)
{
  if(!_sys) throw common::Exception("setShiftedLaplacian: the system exists after the first assemble()");
  if(tag >= _terms.size() || !_terms[tag].bilinear)
    throw common::Exception("setShiftedLaplacian: unknown bilinear term " + std::to_string(tag));
  if(_termSlots[tag].empty()) throw common::Exception("setShiftedLaplacian: the term has no elements");
  std::vector<int> g, t;  // every (group, term) slot of the tag: one class of equal coefficients
  for(const auto &s : _termSlots[tag]) {
    g.push_back(s.first);
    t.push_back(s.second);
  }
  gfem_ctx *ctx = common::deviceContext();
  if(gfem_mg_setup(_sys, (int)g.size(), g.data(), t.data(), beta) != GFEM_OK)
    throw common::Exception(std::string("gfem_mg_setup: ") + gfem_ctx_last_error(ctx));
  int levels = 0, rows[8] = {0};
  gfem_mg_info(_sys, &levels, rows, nullptr);
  return std::vector<int>(rows, rows + levels);
}

template <class T>
void Formulation<T>::instanceParams(const std::vector<std::pair<unsigned int, function::ScalarFunction>> &params,
                                   std::vector<gfem_term_param> &out, std::vector<std::vector<double>> &store)
{
  if(!_sys) throw common::Exception("instanceParams: the system exists after the first assemble()");
  for(const auto &tf : params) {
    const unsigned int tag = tf.first;
    const function::ScalarFunction &f = tf.second;
    if(tag >= _terms.size()) throw common::Exception("instanceParams: unknown term " + std::to_string(tag));
    Term &t = _terms[tag];
    function::ScalarFunction &c = t.bilinear ? t.lhs.coef : t.source;
    if(c.isConstant() != f.isConstant() || (!f.isConstant() && f.ops() != c.ops()))
      throw common::Exception("instanceParams: the coefficient must keep its form (constant, or the same program)");
    c = f;
    const double *consts = nullptr;
    if(!f.isConstant()) {
      store.push_back(f.consts());
      consts = store.back().data();
    }
    for(const auto &gk : _termSlots[tag]) {
      gfem_term_param q{};
      q.group = gk.first;
      q.term = gk.second;
      if(f.isConstant()) {
        q.coef_re = f.value().real();
        q.coef_im = f.value().imag();
      }
      else {
        q.n_consts = (int)f.consts().size() / 2;
        q.consts = consts;
      }
      out.push_back(q);
    }
  }
}

template <class T>
std::pair<int64_t, int> Formulation<T>::nodeDof(int64_t tag) const
{
  if(tag < 1 || tag >= (int64_t)_dofNum.size()) return {0, 0};
  return {_dofNum[tag], _dofType[tag]};
}

template <class T>
const typename Formulation<T>::Pass *Formulation<T>::findPass(int dim, int tag, int type) const
{
  for(const auto &p : _passes)
    if(p.dim == dim && p.tag == tag && p.type == type) return &p;
  return nullptr;
}

// the pass with its element maps on the host (copied from the device after
// a device pre())
template <class T>
const typename Formulation<T>::Pass *Formulation<T>::hostPass(int dim, int tag, int type) const
{
  const Pass *cp = findPass(dim, tag, type);
  if(!cp || cp->onHost || !_pre) return cp;
  Pass &p = _passes[(size_t)(cp - _passes.data())];
  const model::Block *b = _model->block(dim, tag, type);
  p.num.resize(b->conn.size());
  p.atype.resize(b->conn.size());
  p.e2n.assign(p.bilinear ? b->conn.size() * (size_t)b->npe : 0, 0);
  const gfem_status st = gfem_pre_pass_maps(_pre, (int)(cp - _passes.data()), p.num.data(), p.atype.data(),
                                            p.bilinear ? p.e2n.data() : nullptr);
  if(st != GFEM_OK)
    throw common::Exception(std::string("gfem_pre_pass_maps: ") + gfem_ctx_last_error(common::deviceContext()));
  p.onHost = true;
  return &p;
}

template <class T>
const std::vector<int64_t> *Formulation<T>::blockDofs(int dim, int tag, int type) const
{
  const Pass *p = hostPass(dim, tag, type);
  return p ? &p->num : nullptr;
}

template <class T>
const std::vector<int32_t> *Formulation<T>::blockTypes(int dim, int tag, int type) const
{
  const Pass *p = hostPass(dim, tag, type);
  return p ? &p->atype : nullptr;
}

template <class T>
const std::vector<int64_t> *Formulation<T>::blockNonzeros(int dim, int tag, int type) const
{
  const Pass *p = hostPass(dim, tag, type);
  return p ? &p->e2n : nullptr;
}

// Formulation::pre, fem-master/src/problem/Formulation.cpp:897-1039
template <class T>
common::Timer Formulation<T>::pre(DofsSort sort)
{
  common::Timer time;
  time.tick();
  releaseSystem();
  releasePre();
  _ai.clear();
  _aj.clear();
  _passes.clear();
  _values.clear();
  _rhs.clear();
  _lastError.clear();
  _model = model::current();
  if(!_model) throw common::Exception("no current model");
  if(!_field) throw common::Exception("formulation '" + _name + "' has no term");
  const model::Model &M = *_model;
  const int64_t nn = M.numNodes();
  _dofNum.assign(nn + 1, 0);
  _dofType.assign(nn + 1, 0);
  _fixedValues.clear();
  _keyDof.clear();
  _unknownKeys.clear();
  _nFixed = _nUnknown = 0;
  if(!_field->hierarchical())
    for(const auto &ent : _field->domain().entities())
      for(int type : M.elementTypes(ent.first, ent.second)) {
        const model::Block *b = M.block(ent.first, ent.second, type);
        if(b->order != _field->order() && b->dim > 0)
          throw common::Exception("Lagrange order differs from the mesh order");
      }
  // assembly passes: every (entity, element type) carrying a term
  for(const auto &ent : M.entities()) {
    std::vector<unsigned> terms;
    bool bilinear = false;
    for(unsigned t = 0; t < _terms.size(); ++t)
      if(_terms[t].activated && _terms[t].dom.have(ent)) {
        terms.push_back(t);
        bilinear = bilinear || _terms[t].bilinear;
      }
    if(terms.empty()) continue;
    for(int type : M.elementTypes(ent.first, ent.second)) {
      if(M.block(ent.first, ent.second, type)->dim == 0) continue;
      Pass p;
      p.dim = ent.first;
      p.tag = ent.second;
      p.type = type;
      p.terms = terms;
      p.bilinear = bilinear;
      _passes.push_back(std::move(p));
    }
  }
  if(_field->hierarchical())
    preHier(sort);  // entity keys: on the host
  else if(_preDevice)
    preDevice(sort);
  else
    preHost(sort);
  time.tock();
  _preSeconds = time.time();
  return time;
}

// pre() of a hierarchical H1 field (FunctionSpaceHierarchicalH1 keys,
// FunctionSpaceInterface.cpp:319-331): a key is (entity, mode), the entity
// named by its vertex tags in ascending order (csrc/hbasis.h).  Numbering is
// first touch over the same block sequences as preHost (fixed dofs
// DofsFactory.cpp:39-74, unknown dofs :406-460), the unknowns renumbered
// along the Hilbert curve of the key coordinates (an entity's barycentre,
// DofsManager.cpp:118-239).  Fixed values: at order 1 the constraint at the
// vertex; above, the L2 projection of the constraint on its domain, solved
// as a formulation of its own as the reference does (DofsFactory.cpp:104-131).
template <class T>
void Formulation<T>::preHier(DofsSort sort)
{
  using gfem::hier::Key;
  const model::Model &M = *_model;
  const int p = _field->order();
  auto checkBlock = [](const model::Block *b) {
    if(b->dim > 0 && b->order != 1)
      throw common::Exception("hierarchical fields take a first-order (straight-sided) mesh");
  };
  auto elemKeys = [&](const model::Block *b, int64_t e, int64_t *vt) {
    const int nv = b->dim + 1;
    for(int k = 0; k < nv; ++k) vt[k] = b->conn[e * b->npe + k];
    std::sort(vt, vt + nv);
    if(b->dim == 0) return std::vector<Key>{Key{1, {vt[0], 0, 0, 0}, 0}};
    return gfem::hier::keys(b->dim, p, vt);
  };
  auto keyCoord = [&](const Key &k, double *x) {
    for(int j = 0; j < 3; ++j) {
      double c = 0.;
      for(int i = 0; i < k.n; ++i) c += M.xyz[3 * (k.v[i] - 1) + j];
      x[j] = c / k.n;
    }
  };
  int64_t vt[4];
  // fixed dofs, per constraint in order; a key keeps its first constraint
  const auto &cons = _field->constraints();
  std::vector<std::vector<Key>> fixedKeys(cons.size());
  for(size_t c = 0; c < cons.size(); ++c)
    for(const auto &ent : cons[c].first.entities())
      for(int type : M.elementTypes(ent.first, ent.second)) {
        const model::Block *b = M.block(ent.first, ent.second, type);
        checkBlock(b);
        for(int64_t e = 0; e < b->size(); ++e)
          for(const Key &k : elemKeys(b, e, vt))
            if(_keyDof.emplace(k, std::make_pair((int64_t)(_nFixed + 1), (int8_t)2)).second) {
              ++_nFixed;
              fixedKeys[c].push_back(k);
            }
      }
  // unknown dofs over the field's domain
  for(const auto &ent : _field->domain().entities())
    for(int type : M.elementTypes(ent.first, ent.second)) {
      const model::Block *b = M.block(ent.first, ent.second, type);
      checkBlock(b);
      for(int64_t e = 0; e < b->size(); ++e)
        for(const Key &k : elemKeys(b, e, vt))
          if(_keyDof.emplace(k, std::make_pair((int64_t)(_nUnknown + 1), (int8_t)1)).second) {
            ++_nUnknown;
            _unknownKeys.push_back(k);
          }
    }
  if(_nUnknown == 0) {
    _lastError = "no unknown dofs";
    return;
  }
  if(sort != DofsSort::None && M.dim >= 2) {
    double mn[3], mx[3];
    M.boundingBox(mn, mx);
    float fmn[3], fmx[3];
    for(int j = 0; j < 3; ++j) {
      fmn[j] = (float)(mn[j] - 0.000001 * std::fabs(mn[j]));
      fmx[j] = (float)(mx[j] + 0.000001 * std::fabs(mx[j]));
    }
    std::vector<float> coords(3 * _unknownKeys.size());
    for(size_t i = 0; i < _unknownKeys.size(); ++i) {
      double x[3];
      keyCoord(_unknownKeys[i], x);
      for(int j = 0; j < 3; ++j) coords[3 * i + j] = (float)x[j];
    }
    const std::vector<int64_t> order = reorder::hilbertOrder(M.dim, coords, fmn, fmx);
    std::vector<Key> sorted(_unknownKeys.size());
    for(size_t i = 0; i < order.size(); ++i) {
      sorted[i] = _unknownKeys[order[i]];
      _keyDof[sorted[i]].first = (int64_t)i + 1;
    }
    _unknownKeys.swap(sorted);
  }
  for(const auto &kv : _keyDof)
    if(kv.first.n == 1) {
      _dofNum[kv.first.v[0]] = kv.second.first;
      _dofType[kv.first.v[0]] = kv.second.second;
    }
  // fixed values
  _fixedValues.assign(_nFixed, std::complex<double>(0.));
  for(size_t c = 0; c < cons.size(); ++c) {
    if(fixedKeys[c].empty()) continue;
    if(p == 1) {
      for(const Key &k : fixedKeys[c]) {
        double x[3];
        keyCoord(k, x);
        _fixedValues[(size_t)_keyDof[k].first - 1] = cons[c].second.evaluate(x[0], x[1], x[2]);
      }
      continue;
    }
    Formulation<T> proj("boundary");
    field::Field fcp("fieldCp", cons[c].first, field::FunctionSpaceTypeForm0::HierarchicalH1, p);
    const std::string gauss = "Gauss" + std::to_string(2 * (p + 1));
    proj.integral(equation::dof(fcp), equation::tf(fcp), cons[c].first, gauss);
    proj.integral(-cons[c].second, equation::tf(fcp), cons[c].first, gauss);
    proj.pre(DofsSort::None);
    if(!proj.lastError().empty()) throw common::Exception("constraint projection: " + proj.lastError());
    proj.assemble();
    if(!proj.lastError().empty()) throw common::Exception("constraint projection: " + proj.lastError());
    SolveOptions so;
    so.tol = 1e-13;
    so.restart = 100;
    so.maxIter = 20000;
    proj.solve(so);
    if(!proj.lastError().empty()) throw common::Exception("constraint projection: " + proj.lastError());
    const std::vector<T> &x = proj.getSolution();
    for(const Key &k : fixedKeys[c]) {
      const auto d = proj.keyDof(k);
      if(d.second == 1 && d.first >= 1 && (size_t)d.first <= x.size())
        _fixedValues[(size_t)_keyDof[k].first - 1] = std::complex<double>(x[(size_t)d.first - 1]);
    }
  }
  // element maps of every pass, keys in the basis' local order
  for(auto &pp : _passes) {
    const model::Block *b = M.block(pp.dim, pp.tag, pp.type);
    checkBlock(b);
    const int nv = pp.dim + 1;
    const int N = gfem::hier::count(pp.dim, p);
    const int64_t ne = b->size();
    pp.N = N;
    pp.verts.resize((size_t)(ne * nv));
    pp.num.assign((size_t)(ne * N), 0);
    pp.atype.assign((size_t)(ne * N), GFEM_AT_UNDEFINED);
    for(int64_t e = 0; e < ne; ++e) {
      const std::vector<Key> keys = elemKeys(b, e, vt);
      for(int k = 0; k < nv; ++k) pp.verts[(size_t)(e * nv + k)] = vt[k];
      for(int i = 0; i < N; ++i) {
        const auto it = _keyDof.find(keys[i]);
        if(it == _keyDof.end()) continue;
        pp.num[(size_t)(e * N + i)] = it->second.first - 1;
        pp.atype[(size_t)(e * N + i)] = it->second.second == 1 ? GFEM_AT_UNKNOWN : GFEM_AT_FIXED;
      }
    }
  }
  buildPattern();
}

template <class T>
std::pair<int64_t, int> Formulation<T>::keyDof(const gfem::hier::Key &k) const
{
  const auto it = _keyDof.find(k);
  if(it == _keyDof.end()) return {0, 0};
  return {it->second.first, it->second.second};
}

template <class T>
void Formulation<T>::preHost(DofsSort sort)
{
  const model::Model &M = *_model;
  // fixed dofs: constraint order, entity order, element type order, key order
  // (DofsFactory.cpp:39-74); Lagrange keys are the element node tags
  for(const auto &c : _field->constraints()) {
    for(const auto &ent : c.first.entities()) {
      for(int type : M.elementTypes(ent.first, ent.second)) {
        const model::Block *b = M.block(ent.first, ent.second, type);
        for(int64_t tag : b->conn) {
          if(_dofType[tag]) continue;
          _dofType[tag] = 2;
          _dofNum[tag] = (int64_t)++_nFixed;
          const double *X = &M.xyz[3 * (tag - 1)];
          _fixedValues.push_back(c.second.evaluate(X[0], X[1], X[2]));
        }
      }
    }
  }
  // unknown dofs (DofsFactory.cpp:406-460); Lagrange keys are never bubbles
  std::vector<int64_t> unknownTags;
  for(const auto &ent : _field->domain().entities()) {
    for(int type : M.elementTypes(ent.first, ent.second)) {
      const model::Block *b = M.block(ent.first, ent.second, type);
      if(b->order != _field->order() && b->dim > 0)
        throw common::Exception("Lagrange order differs from the mesh order");
      for(int64_t tag : b->conn) {
        if(_dofType[tag]) continue;
        _dofType[tag] = 1;
        _dofNum[tag] = (int64_t)++_nUnknown;
        unknownTags.push_back(tag);
      }
    }
  }
  if(_nUnknown == 0) {
    _lastError = "no unknown dofs";
    return;
  }
  // Hilbert renumbering (DofsManager.cpp:118-239)
  if(sort != DofsSort::None && M.dim >= 2) {
    double mn[3], mx[3];
    M.boundingBox(mn, mx);
    float fmn[3], fmx[3];
    for(int j = 0; j < 3; ++j) {
      fmn[j] = (float)(mn[j] - 0.000001 * std::fabs(mn[j]));
      fmx[j] = (float)(mx[j] + 0.000001 * std::fabs(mx[j]));
    }
    std::vector<float> coords(3 * unknownTags.size());
    for(size_t i = 0; i < unknownTags.size(); ++i)
      for(int j = 0; j < 3; ++j) coords[3 * i + j] = (float)M.xyz[3 * (unknownTags[i] - 1) + j];
    const std::vector<int64_t> order = reorder::hilbertOrder(M.dim, coords, fmn, fmx);
    for(size_t i = 0; i < order.size(); ++i) _dofNum[unknownTags[order[i]]] = (int64_t)i + 1;
  }

  // element maps of every pass (FieldInterface::fillIndices)
  for(auto &p : _passes) {
    const model::Block *b = M.block(p.dim, p.tag, p.type);
    p.N = b->npe;
    p.num.resize(b->conn.size());
    p.atype.resize(b->conn.size());
    for(size_t k = 0; k < b->conn.size(); ++k) {
      const int64_t tag = b->conn[k];
      const int t = _dofType[tag];
      p.num[k] = t ? _dofNum[tag] - 1 : 0;
      p.atype[k] = t == 1 ? GFEM_AT_UNKNOWN : (t == 2 ? GFEM_AT_FIXED : GFEM_AT_UNDEFINED);
    }
  }

  buildPattern();
}

// sparsity pattern and element -> nonzero maps of the passes' element maps
// (p.num / p.atype, p.N keys per element)
template <class T>
void Formulation<T>::buildPattern()
{
  // sparsity pattern (Pattern::build + MatrixFactory::addPattern/finalizePattern):
  // row = tf dof, col = dof; unknown x unknown; upper triangle if symmetric
  const int64_t n = (int64_t)_nUnknown;
  std::vector<int64_t> cnt(n + 1, 0);
  for(const auto &p : _passes) {
    if(!p.bilinear) continue;
    const int N = p.N;
    const int64_t ne = (int64_t)p.num.size() / N;
    for(int64_t e = 0; e < ne; ++e) {
      int nu = 0;
      for(int j = 0; j < N; ++j) nu += p.atype[e * N + j] == GFEM_AT_UNKNOWN;
      for(int i = 0; i < N; ++i)
        if(p.atype[e * N + i] == GFEM_AT_UNKNOWN) cnt[p.num[e * N + i] + 1] += nu;
    }
  }
  for(int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  std::vector<int64_t> cols((size_t)cnt[n]);
  {
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for(const auto &p : _passes) {
      if(!p.bilinear) continue;
      const int N = p.N;
      const int64_t ne = (int64_t)p.num.size() / N;
      for(int64_t e = 0; e < ne; ++e)
        for(int i = 0; i < N; ++i) {
          if(p.atype[e * N + i] != GFEM_AT_UNKNOWN) continue;
          const int64_t r = p.num[e * N + i];
          for(int j = 0; j < N; ++j)
            if(p.atype[e * N + j] == GFEM_AT_UNKNOWN) cols[fill[r]++] = p.num[e * N + j];
        }
    }
  }
  _ai.assign(n + 1, 0);
  _aj.clear();
  _aj.reserve(cols.size() / 2);
  for(int64_t r = 0; r < n; ++r) {
    auto b = cols.begin() + cnt[r], e = cols.begin() + cnt[r + 1];
    std::sort(b, e);
    e = std::unique(b, e);
    for(auto it = b; it != e; ++it)
      if(!_symmetric || *it >= r) _aj.push_back(*it);
    _ai[r + 1] = (int64_t)_aj.size();
  }
  std::vector<int64_t>().swap(cols);
  // element -> nonzero map
  for(auto &p : _passes) {
    if(!p.bilinear) continue;
    const int N = p.N;
    const int64_t ne = (int64_t)p.num.size() / N;
    p.e2n.assign((size_t)(ne * N * N), -1);
    for(int64_t e = 0; e < ne; ++e)
      for(int i = 0; i < N; ++i) {
        if(p.atype[e * N + i] != GFEM_AT_UNKNOWN) continue;
        const int64_t r = p.num[e * N + i];
        const int64_t *rb = _aj.data() + _ai[r], *re = _aj.data() + _ai[r + 1];
        for(int j = 0; j < N; ++j) {
          if(p.atype[e * N + j] != GFEM_AT_UNKNOWN) continue;
          const int64_t c = p.num[e * N + j];
          if(_symmetric && c < r) continue;
          p.e2n[(e * N + i) * N + j] = std::lower_bound(rb, re, c) - _aj.data();
        }
      }
  }
}

// pre() on the device (include/gfem.h, gfem_pre_*): the same block
// sequences as preHost, numbered, renumbered and patterned on the GPU.  The
// host keeps the DOF numbering per node, the Dirichlet values and ai; aj and
// the element maps stay on the device until asked for.
template <class T>
void Formulation<T>::preDevice(DofsSort sort)
{
  const model::Model &M = *_model;
  common::uploadNodes(M);
  gfem_ctx *ctx = common::deviceContext();
  auto check = [ctx](gfem_status st, const char *what) {
    if(st != GFEM_OK) throw common::Exception(std::string(what) + ": " + gfem_ctx_last_error(ctx));
  };
  check(gfem_pre_create(ctx, &_pre), "gfem_pre_create");
  std::map<std::tuple<int, int, int>, int> ids;
  auto blockOf = [&](int dim, int tag, int type) {
    const auto key = std::make_tuple(dim, tag, type);
    const auto it = ids.find(key);
    if(it != ids.end()) return it->second;
    const model::Block *b = M.block(dim, tag, type);
    int id = -1;
    check(gfem_pre_add_block(_pre, b->size(), b->npe, b->conn.data(), &id), "gfem_pre_add_block");
    ids[key] = id;
    return id;
  };
  // fixed dofs: constraint order, entity order, element type order
  // (DofsFactory.cpp:39-74); unknown dofs over the field's domain
  // (DofsFactory.cpp:406-460)
  std::vector<int> fseq, fcon, useq;
  const auto &cons = _field->constraints();
  for(size_t c = 0; c < cons.size(); ++c)
    for(const auto &ent : cons[c].first.entities())
      for(int type : M.elementTypes(ent.first, ent.second)) {
        fseq.push_back(blockOf(ent.first, ent.second, type));
        fcon.push_back((int)c);
      }
  for(const auto &ent : _field->domain().entities())
    for(int type : M.elementTypes(ent.first, ent.second)) useq.push_back(blockOf(ent.first, ent.second, type));
  // Hilbert renumbering box (DofsManager.cpp:118-239)
  int hdim = 0;
  float fmn[3] = {0.f, 0.f, 0.f}, fmx[3] = {0.f, 0.f, 0.f};
  if(sort != DofsSort::None && M.dim >= 2) {
    hdim = M.dim;
    double mn[3], mx[3];
    M.boundingBox(mn, mx);
    for(int j = 0; j < 3; ++j) {
      fmn[j] = (float)(mn[j] - 0.000001 * std::fabs(mn[j]));
      fmx[j] = (float)(mx[j] + 0.000001 * std::fabs(mx[j]));
    }
  }
  int64_t nf = 0, nu = 0;
  check(gfem_pre_number(_pre, (int)fseq.size(), fseq.data(), (int)useq.size(), useq.data(), hdim, fmn, fmx, &nf, &nu),
        "gfem_pre_number");
  _nFixed = (uint64_t)nf;
  _nUnknown = (uint64_t)nu;
  {
    const int64_t nn = M.numNodes();
    std::vector<int64_t> num((size_t)nn);
    std::vector<int8_t> type((size_t)nn);
    std::vector<int64_t> ftag((size_t)nf);
    std::vector<int32_t> fseqOf((size_t)nf);
    check(gfem_pre_dofs(_pre, num.data(), type.data(), ftag.data(), fseqOf.data()), "gfem_pre_dofs");
    for(int64_t t = 1; t <= nn; ++t) {
      _dofNum[t] = num[t - 1];
      _dofType[t] = type[t - 1];
    }
    _fixedValues.resize((size_t)nf);
    for(int64_t f = 0; f < nf; ++f) {
      const double *X = &M.xyz[3 * (ftag[f] - 1)];
      _fixedValues[f] = cons[fcon[fseqOf[f]]].second.evaluate(X[0], X[1], X[2]);
    }
  }
  if(_nUnknown == 0) {
    _lastError = "no unknown dofs";
    releasePre();
    return;
  }
  std::vector<int> pblocks, pbil;
  for(auto &p : _passes) {
    pblocks.push_back(blockOf(p.dim, p.tag, p.type));
    pbil.push_back(p.bilinear ? 1 : 0);
    p.onHost = false;
  }
  int64_t nnz = 0;
  check(gfem_pre_pattern(_pre, (int)pblocks.size(), pblocks.data(), pbil.data(), _symmetric ? 1 : 0, &nnz),
        "gfem_pre_pattern");
  _ai.assign((size_t)_nUnknown + 1, 0);
  check(gfem_pre_csr(_pre, _ai.data(), nullptr), "gfem_pre_csr");
}

template <class T>
void Formulation<T>::releasePre()
{
  if(_pre) gfem_pre_destroy(_pre);
  _pre = nullptr;
}

template <class T>
const std::vector<int64_t> &Formulation<T>::aj() const
{
  if(_pre && _aj.empty() && !_ai.empty() && _ai.back() > 0) {
    std::vector<int64_t> aj((size_t)_ai.back());
    if(gfem_pre_csr(_pre, nullptr, aj.data()) != GFEM_OK)
      throw common::Exception(std::string("gfem_pre_csr: ") + gfem_ctx_last_error(common::deviceContext()));
    _aj.swap(aj);
  }
  return _aj;
}

// Hand the pattern, the Dirichlet values and one group per assembly pass to
// the device, in the traversal order of Formulation::assemble (dimension
// 3..0, element type ascending, entity tag ascending; Formulation.cpp:1417-1493).
// Runs once, at the first assemble() after pre().
template <class T>
void Formulation<T>::upload()
{
  const model::Model &M = *_model;
  const int64_t n = (int64_t)_nUnknown;
  gfem_ctx *ctx = common::deviceContext();
  std::vector<double> dc(2 * std::max<uint64_t>(_nFixed, 1), 0.);
  for(uint64_t i = 0; i < _nFixed; ++i) {
    dc[2 * i] = _fixedValues[i].real();
    dc[2 * i + 1] = _fixedValues[i].imag();
  }
  gfem_status st;
  if(_pre) {
    // pattern and maps already on the device (pre() on the device)
    st = gfem_system_create_pre(ctx, ScalarKind<T>::value, _pre, dc.data(), (int)_numRHS, _frequency, &_sys);
  }
  else {
    common::uploadNodes(M);
    st = gfem_system_create(ctx, ScalarKind<T>::value, n, _ai.data(), _aj.data(), _symmetric ? 1 : 0,
                            (int64_t)_nFixed, dc.data(), (int)_numRHS, _frequency, &_sys);
  }
  if(st != GFEM_OK) throw common::Exception(std::string("gfem_system_create: ") + gfem_ctx_last_error(ctx));
  std::vector<const Pass *> order;
  for(int dim = 3; dim >= 0; --dim)
    for(int type : M.elementTypes(dim))
      for(const auto &ent : M.entities(dim))
        if(const Pass *p = findPass(dim, ent.second, type)) order.push_back(p);
  _termSlots.assign(_terms.size(), {});
  for(const Pass *p : order) {
    const model::Block *b = M.block(p->dim, p->tag, p->type);
    int g = -1;
    int k = 0;
    if(_pre)
      st = gfem_group_add_pre(_sys, _pre, (int)(p - _passes.data()), p->dim, p->type, &g);
    else if(_field->hierarchical())
      st = gfem_group_add_hierarchical(_sys, p->dim, _field->order(), (int64_t)(p->verts.size() / (p->dim + 1)),
                                       p->verts.data(), p->num.data(), p->atype.data(),
                                       p->e2n.empty() ? nullptr : p->e2n.data(), &g);
    else
      st = gfem_group_add(_sys, p->dim, p->type, b->size(), b->conn.data(), p->num.data(), p->atype.data(),
                          p->e2n.empty() ? nullptr : p->e2n.data(), &g);
    if(st != GFEM_OK) throw common::Exception(std::string("gfem_group_add: ") + gfem_ctx_last_error(ctx));
    for(unsigned ti : p->terms) {
      const Term &t = _terms[ti];
      gfem_term_desc d{};
      d.kind = t.bilinear ? 0 : 1;
      d.gauss = t.gauss;
      d.form_dof = t.bilinear ? t.lhs.form : 0;
      d.form_tf = t.rhs.form;
      d.hermitian = t.product == ProductType::Hermitian ? 1 : 0;
      d.dt_order = t.bilinear ? t.lhs.dt : 0;
      d.rhs_idx = t.bilinear ? 0 : (int)t.rhsIdx;
      const function::ScalarFunction &c = t.bilinear ? t.lhs.coef : t.source;
      if(!t.rhs.coef.isConstant()) throw common::Exception("test-function coefficients must be constant");
      d.tf_coef_re = t.rhs.coef.value().real();
      d.tf_coef_im = t.rhs.coef.value().imag();
      if(c.isConstant()) {
        d.coef_re = c.value().real();
        d.coef_im = c.value().imag();
      }
      else {
        d.n_ops = (int)c.ops().size() / 2;
        d.ops = c.ops().data();
        d.n_consts = (int)c.consts().size() / 2;
        d.consts = c.consts().data();
      }
      st = gfem_group_add_term(_sys, g, &d);
      if(st != GFEM_OK) throw common::Exception(std::string("gfem_group_add_term: ") + gfem_ctx_last_error(ctx));
      _termSlots[ti].emplace_back(g, k++);
    }
  }
  st = gfem_system_finalize(_sys);
  if(st != GFEM_OK) {
    const std::string msg = gfem_ctx_last_error(ctx);
    releaseSystem();
    throw common::Exception("gfem_system_finalize: " + msg);
  }
}

template <class T>
static void convert(const std::vector<char> &raw, bool mixed, bool cplx, std::vector<T> &out, int64_t count)
{
  out.resize(count);
  for(int64_t i = 0; i < count; ++i) {
    double re, im = 0.;
    if(mixed) {
      const float *f = reinterpret_cast<const float *>(raw.data());
      re = f[cplx ? 2 * i : i];
      if(cplx) im = f[2 * i + 1];
    }
    else {
      const double *f = reinterpret_cast<const double *>(raw.data());
      re = f[cplx ? 2 * i : i];
      if(cplx) im = f[2 * i + 1];
    }
    if constexpr(ScalarKind<T>::complex)
      out[i] = T(re, im);
    else
      out[i] = (T)re;
  }
}

// Formulation::assemble, fem-master/src/problem/Formulation.cpp:1151-1215:
// on failure the matrix and every right-hand side are zeroed and the error
// is logged; a Timer is still returned.
template <class T>
common::Timer Formulation<T>::assemble(const AssembleOptions &opt)
{
  common::Timer time;
  time.tick();
  _lastError.clear();
  if(_nUnknown == 0) {
    // Formulation.cpp:1179-1184: a warning, nothing assembled (A and every b
    // stay empty), no error
    std::fprintf(stderr, "Warning : There are no unknown dofs to assemble. Did you forget to pre-process the formulation?\n");
    _values.clear();
    _rhs.assign(_numRHS, std::vector<T>());
    time.tock();
    return time;
  }
  try {
    if(!_sys) {
      try {
        upload();
      }
      catch(...) {
        releaseSystem();
        throw;
      }
    }
    gfem_ctx *ctx = common::deviceContext();
    gfem_status st = gfem_assemble(_sys, opt.precision, opt.strategy, opt.stream);
    if(st != GFEM_OK) throw common::Exception(std::string("gfem_assemble: ") + gfem_ctx_last_error(ctx));
    if(opt.download) {
      int64_t nnz = 0, n = 0;
      int vb = 0;
      gfem_system_info(_sys, opt.precision, &nnz, &n, &vb);
      std::vector<char> raw((size_t)(vb * std::max(nnz, n)));
      st = gfem_values_download(_sys, opt.precision, raw.data(), raw.size());
      if(st != GFEM_OK) throw common::Exception(std::string("gfem_values_download: ") + gfem_ctx_last_error(ctx));
      const bool mixed = opt.precision == GFEM_PREC_MIXED;
      convert(raw, mixed, ScalarKind<T>::complex, _values, nnz);
      _rhs.resize(_numRHS);
      for(unsigned r = 0; r < _numRHS; ++r) {
        st = gfem_rhs_download(_sys, opt.precision, (int)r, raw.data(), raw.size());
        if(st != GFEM_OK) throw common::Exception(std::string("gfem_rhs_download: ") + gfem_ctx_last_error(ctx));
        convert(raw, mixed, ScalarKind<T>::complex, _rhs[r], n);
      }
    }
  }
  catch(const std::exception &e) {
    _lastError = e.what();
    std::fprintf(stderr, "Error   : Unable to assemble the system\nError   : %s\n", e.what());
    std::fill(_values.begin(), _values.end(), T(0));
    for(auto &b : _rhs) std::fill(b.begin(), b.end(), T(0));
  }
  time.tock();
  return time;
}

template <class T>
common::Timer Formulation<T>::solve(const SolveOptions &opt)
{
  common::Timer time;
  time.tick();
  _lastError.clear();
  _solveInfo = SolveInfo();
  try {
    if(!_sys) throw common::Exception("solve before assemble");
    gfem_ctx *ctx = common::deviceContext();
    gfem_solve_info info{};
    gfem_status st = gfem_solve(_sys, (int)opt.rhsIdx, opt.tol, opt.restart, opt.maxIter, opt.stream, &info);
    if(st != GFEM_OK) throw common::Exception(std::string("gfem_solve: ") + gfem_ctx_last_error(ctx));
    _solveInfo.iterations = info.iterations;
    _solveInfo.cycles = info.cycles;
    _solveInfo.relResidual = info.rel_residual;
    _solveInfo.converged = info.converged != 0;
    if(opt.download) {
      int64_t nnz = 0, n = 0;
      int vb = 0;
      gfem_system_info(_sys, GFEM_PREC_FP64, &nnz, &n, &vb);
      std::vector<char> raw((size_t)(vb * n));
      st = gfem_solution_download(_sys, raw.data(), raw.size());
      if(st != GFEM_OK) throw common::Exception(std::string("gfem_solution_download: ") + gfem_ctx_last_error(ctx));
      convert(raw, false, ScalarKind<T>::complex, _solution, n);
    }
    if(!_solveInfo.converged)
      std::fprintf(stderr, "Warning : solver stopped at relative residual %g after %d iterations\n",
                   _solveInfo.relResidual, _solveInfo.iterations);
  }
  catch(const std::exception &e) {
    _lastError = e.what();
    std::fprintf(stderr, "Error   : Unable to solve the system\nError   : %s\n", e.what());
    std::fill(_solution.begin(), _solution.end(), T(0));
  }
  time.tock();
  return time;
}

template <class T>
T Formulation<T>::nodeValue(int64_t tag) const
{
  const auto d = nodeDof(tag);
  if(d.second == 1 && d.first >= 1 && (size_t)d.first <= _solution.size()) return _solution[(size_t)d.first - 1];
  if(d.second == 2 && d.first >= 1 && (size_t)d.first <= _fixedValues.size())
    return scalarFrom<T>(_fixedValues[(size_t)d.first - 1]);
  return T(0);
}

template class Formulation<double>;
template class Formulation<float>;
template class Formulation<std::complex<double>>;
template class Formulation<std::complex<float>>;

}  // namespace problem

}  // namespace gmshfem
