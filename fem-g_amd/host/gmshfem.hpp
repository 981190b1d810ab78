// Host side of the engine: the gmshfem::problem::Formulation<T> API
// (integral / pre / assemble / getLHS / getRHS) over the gfem C ABI.
//
// Mirrors fem-master/src/problem/Formulation.h:50-232 for the terms of the
// assembly hot path: scalar H1 Lagrange fields (Form0), dof / tf / grad /
// dt_dof / dt2_dof, constant or expression coefficients, Hermitian or Scalar
// products, several right-hand sides, constant Dirichlet constraints.
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "gfem.h"
#include "hbasis.h"
#include "model.hpp"

namespace gmshfem {

namespace common {
// common::Exception (fem-master/src/common/Exception.h)
class Exception : public std::runtime_error {
 public:
  explicit Exception(const std::string &w) : std::runtime_error(w) {}
};

// common::Timer (fem-master/src/common/Timer.h:23-63): wall seconds
class Timer {
 public:
  void tick();
  void tock();
  double time() const { return _t; }
  operator double() const { return _t; }

 private:
  double _start = 0., _t = 0.;
};

// the device context shared by every formulation (GmshFem initialisation)
gfem_ctx *deviceContext();
void setDevice(int device);
}  // namespace common

namespace domain {
// Domain (fem-master/src/domain/Domain.h:18-54): a set of (dim, tag) entities
class Domain {
 public:
  Domain() = default;
  Domain(int dim, int tag) { _ents.insert({dim, tag}); }
  Domain operator|(const Domain &o) const;
  bool have(const std::pair<int, int> &e) const { return _ents.count(e) != 0; }
  int maxDim() const;
  const std::set<std::pair<int, int>> &entities() const { return _ents; }

 private:
  std::set<std::pair<int, int>> _ents;
};
}  // namespace domain

namespace function {
// ScalarFunction: an expression of x, y, z compiled to the gfem bytecode
// (fem-master/src/function/ScalarFunction.cpp:203-222, executionTree/*)
class ScalarFunction {
 public:
  ScalarFunction() : ScalarFunction(0.) {}
  ScalarFunction(double v) : ScalarFunction(std::complex<double>(v, 0.)) {}
  ScalarFunction(std::complex<double> v);
  bool isConstant() const { return _constant; }
  std::complex<double> value() const { return _value; }
  const std::vector<int32_t> &ops() const { return _ops; }
  const std::vector<double> &consts() const { return _consts; }
  std::complex<double> evaluate(double x, double y, double z) const;

  static ScalarFunction coordinate(int axis);
  static ScalarFunction unary(int op, const ScalarFunction &a);
  static ScalarFunction binary(int op, const ScalarFunction &a, const ScalarFunction &b);

 private:
  bool _constant = true;
  std::complex<double> _value;
  std::vector<int32_t> _ops;
  std::vector<double> _consts;
};

ScalarFunction operator+(const ScalarFunction &a, const ScalarFunction &b);
ScalarFunction operator-(const ScalarFunction &a, const ScalarFunction &b);
ScalarFunction operator*(const ScalarFunction &a, const ScalarFunction &b);
ScalarFunction operator/(const ScalarFunction &a, const ScalarFunction &b);
ScalarFunction operator-(const ScalarFunction &a);
ScalarFunction x();
ScalarFunction y();
ScalarFunction z();
ScalarFunction exp(const ScalarFunction &a);
ScalarFunction sqrt(const ScalarFunction &a);
ScalarFunction sin(const ScalarFunction &a);
ScalarFunction cos(const ScalarFunction &a);
ScalarFunction abs(const ScalarFunction &a);
ScalarFunction pow(const ScalarFunction &a, const ScalarFunction &b);
}  // namespace function

namespace field {
// FunctionSpaceTypeForm0 (fem-master/src/field/FunctionSpace.h): the scalar
// H1 spaces of the hot path
enum class FunctionSpaceTypeForm0 { Lagrange, HierarchicalH1 };

// Field<T, Form0> with a Lagrange or hierarchical H1 function space
// (fem-master/src/field/Field.h:30-44, FunctionSpaceLagrange.cpp:21-50,
// FunctionSpaceHierarchicalH1.cpp:20-24)
class Field {
 public:
  Field(const std::string &name, const domain::Domain &dom, int order);
  Field(const std::string &name, const domain::Domain &dom, FunctionSpaceTypeForm0 type, int order);
  void addConstraint(const domain::Domain &dom, const function::ScalarFunction &value);
  const std::string &name() const { return _name; }
  const domain::Domain &domain() const { return _domain; }
  int order() const { return _order; }
  FunctionSpaceTypeForm0 type() const { return _type; }
  bool hierarchical() const { return _type == FunctionSpaceTypeForm0::HierarchicalH1; }
  int tag() const { return _tag; }
  const std::vector<std::pair<domain::Domain, function::ScalarFunction>> &constraints() const { return _constraints; }

 private:
  std::string _name;
  domain::Domain _domain;
  FunctionSpaceTypeForm0 _type = FunctionSpaceTypeForm0::Lagrange;
  int _order;
  int _tag;
  std::vector<std::pair<domain::Domain, function::ScalarFunction>> _constraints;
};
}  // namespace field

namespace equation {
// dof / tf / grad / dt_dof / dt2_dof (fem-master/src/equation/UnknownField.h:102-139)
struct Unknown {
  const field::Field *field = nullptr;
  bool tf = false;
  int form = 0;  // 0 value, 1 gradient
  int dt = 0;    // UnknownFieldType NotDt / Dt / DtDt
  function::ScalarFunction coef = function::ScalarFunction(1.);
};
Unknown dof(const field::Field &u);
Unknown tf(const field::Field &u);
Unknown dt_dof(const field::Field &u);
Unknown dt2_dof(const field::Field &u);
Unknown grad(const Unknown &u);
Unknown operator*(const function::ScalarFunction &c, const Unknown &u);
}  // namespace equation

namespace problem {

enum class DofsSort { Default, Hilbert, None };
enum class ProductType { Hermitian, Scalar };

struct Term {
  bool bilinear = true;
  equation::Unknown lhs, rhs;           // bilinear: dof side, tf side
  function::ScalarFunction source;      // linear: source
  domain::Domain dom;
  int gauss = 0;
  ProductType product = ProductType::Hermitian;
  int rhsIdx = 0;
  bool activated = true;
};

struct AssembleOptions {
  gfem_precision precision = GFEM_PREC_FP64;
  gfem_strategy strategy = GFEM_STRATEGY_FUSED;
  void *stream = nullptr;   // hipStream_t
  bool download = true;     // copy A and b back to the host
};

// Solver options (device GMRES with FP64 iterative refinement, A x = -b)
struct SolveOptions {
  double tol = 1e-8;      // ||-b - A x|| / ||b||
  int restart = 50;       // GMRES(restart)
  int maxIter = 20000;    // inner iterations
  unsigned int rhsIdx = 0;
  void *stream = nullptr; // hipStream_t
  bool download = true;   // copy x back to the host
};

struct SolveInfo {
  int iterations = 0, cycles = 0;
  double relResidual = 0.;
  bool converged = false;
};

template <class T>
class Formulation {
 public:
  explicit Formulation(const std::string &name);
  ~Formulation();
  Formulation(const Formulation &) = delete;
  Formulation &operator=(const Formulation &) = delete;

  // integral(dof-expression, tf-expression, domain, "GaussN", product)
  unsigned int integral(const equation::Unknown &lhs, const equation::Unknown &rhs, const domain::Domain &dom,
                        const std::string &integrationType, ProductType product = ProductType::Hermitian);
  // integral(source, tf-expression, domain, "GaussN", product, rhsIdx)
  unsigned int integral(const function::ScalarFunction &f, const equation::Unknown &rhs, const domain::Domain &dom,
                        const std::string &integrationType, ProductType product = ProductType::Hermitian,
                        unsigned int rhsIdx = 0);

  // New coefficient of term `tag`: the dof-side coefficient of a bilinear
  // term or the source of a linear term.  A constant stays a constant, an
  // expression keeps its program and changes its constants; applied on the
  // device from the next assembly.  (The reference binds coefficients when a
  // term is created; this serves repeated solves on one mesh.)
  void setTermCoefficient(unsigned int tag, const function::ScalarFunction &f);
  // Shifted-Laplacian multigrid preconditioner of this system's batched
  // solves (gfem_mg_setup): the term `tag` (the mass term of a Helmholtz
  // formulation) has its coefficient multiplied by (1 + i beta) in the
  // preconditioned operator.  Requests submitted afterwards use it, and so
  // does solve().  Needs the system (after the first assemble()).  Returns
  // the rows per level.
  std::vector<int> setShiftedLaplacian(unsigned int tag, double beta = 1.0);
  // The parameters of one instance of gfem_assemble_many: the (group, term)
  // slots of each term tag with its new coefficient, after the checks of
  // setTermCoefficient.  The formulation's own coefficients take the new
  // values too (the system keeps its last instance's parameters); expression
  // constants are held in `store`, which must outlive the call.
  void instanceParams(const std::vector<std::pair<unsigned int, function::ScalarFunction>> &params,
                      std::vector<gfem_term_param> &out, std::vector<std::vector<double>> &store);

  void setSymmetric(bool s) { _symmetric = s; }
  void setAngularFrequency(double w) { _frequency = w; }
  unsigned int numRHS() const { return _numRHS; }

  // pre(): DOF numbering, pattern and element maps.  On the device by
  // default (gfem_pre_*: bit-identical results, kept on the GPU for the
  // system); setPreOnDevice(false) runs the host implementation.
  common::Timer pre(DofsSort sort = DofsSort::Default);
  void setPreOnDevice(bool d) { _preDevice = d; }
  bool preOnDevice() const { return _preDevice; }
  common::Timer assemble(const AssembleOptions &opt = AssembleOptions());
  // reference: Timer solve(bool reusePreconditioner); needs an FP64 assembly
  common::Timer solve(const SolveOptions &opt = SolveOptions());

  // results of pre()
  uint64_t getNumberOfUnknownDof() const { return _nUnknown; }
  uint64_t getNumberOfFixedDof() const { return _nFixed; }
  uint64_t getNumberOfNonZeros() const { return _ai.empty() ? 0 : (uint64_t)_ai.back(); }
  const std::vector<int64_t> &ai() const { return _ai; }
  const std::vector<int64_t> &aj() const;
  // numDof (1-based) and type (1 unknown, 2 fixed, 0 none) of the dof of a node tag
  // (hierarchical fields: of the vertex function of that node)
  std::pair<int64_t, int> nodeDof(int64_t tag) const;
  // hierarchical fields: numDof and type of a key (entity vertex tags
  // ascending, mode), and the key of every unknown dof in numDof order
  std::pair<int64_t, int> keyDof(const gfem::hier::Key &k) const;
  const std::vector<gfem::hier::Key> &unknownKeys() const { return _unknownKeys; }
  // element maps of the block (dim, tag, type): numDof-1 / AssembleType per key
  const std::vector<int64_t> *blockDofs(int dim, int tag, int type) const;
  const std::vector<int32_t> *blockTypes(int dim, int tag, int type) const;
  const std::vector<int64_t> *blockNonzeros(int dim, int tag, int type) const;

  // results of assemble() (host copies, when downloaded)
  const std::vector<T> &getLHS() const { return _values; }
  const std::vector<T> &getRHS(unsigned int i = 0) const { return _rhs.at(i); }
  const std::string &lastError() const { return _lastError; }
  // results of solve(): unknown dofs in numDof order, and the value at a node
  // (the unknown's value, or the Dirichlet value of a fixed dof)
  const std::vector<T> &getSolution() const { return _solution; }
  const SolveInfo &solveInfo() const { return _solveInfo; }
  T nodeValue(int64_t tag) const;
  gfem_system *system() const { return _sys; }
  double preHostSeconds() const { return _preSeconds; }

 private:
  struct Pass {
    int dim, tag, type;
    int N = 0;                   // keys per element
    std::vector<int64_t> verts;  // hierarchical fields: vertex tags per element, ascending
    std::vector<int64_t> num;
    std::vector<int32_t> atype;
    std::vector<int64_t> e2n;
    std::vector<unsigned> terms;
    bool bilinear = false;
    bool onHost = true;  // maps above filled (device pre(): copied on request)
  };

  std::string _name;
  std::shared_ptr<model::Model> _model;
  std::vector<Term> _terms;
  const field::Field *_field = nullptr;
  unsigned int _numRHS = 1;
  bool _symmetric = false;
  double _frequency = 0.;
  uint64_t _nUnknown = 0, _nFixed = 0;
  std::vector<int64_t> _dofNum;     // node tag -> numDof (1-based)
  std::vector<int8_t> _dofType;     // node tag -> 1 unknown, 2 fixed
  std::vector<std::complex<double>> _fixedValues;
  struct KeyHash {
    size_t operator()(const gfem::hier::Key &k) const
    {
      uint64_t h = (uint64_t)k.n * 0x9E3779B97F4A7C15ull ^ (uint64_t)k.mode;
      for(int i = 0; i < k.n; ++i) h = (h ^ (uint64_t)k.v[i]) * 0x100000001B3ull;
      return (size_t)h;
    }
  };
  struct KeyEq {
    bool operator()(const gfem::hier::Key &a, const gfem::hier::Key &b) const
    {
      if(a.n != b.n || a.mode != b.mode) return false;
      for(int i = 0; i < a.n; ++i)
        if(a.v[i] != b.v[i]) return false;
      return true;
    }
  };
  std::unordered_map<gfem::hier::Key, std::pair<int64_t, int8_t>, KeyHash, KeyEq> _keyDof;
  std::vector<gfem::hier::Key> _unknownKeys;
  std::vector<int64_t> _ai;
  mutable std::vector<int64_t> _aj;  // device pre(): copied on request
  mutable std::vector<Pass> _passes;
  bool _preDevice = true;
  gfem_pre *_pre = nullptr;  // device pre() results
  std::vector<std::vector<std::pair<int, int>>> _termSlots;  // term -> (group, index in group) on the device
  gfem_system *_sys = nullptr;
  std::vector<T> _values;
  std::vector<std::vector<T>> _rhs;
  std::vector<T> _solution;
  SolveInfo _solveInfo;
  std::string _lastError;
  double _preSeconds = 0.;

  const Pass *findPass(int dim, int tag, int type) const;
  const Pass *hostPass(int dim, int tag, int type) const;  // with its maps on the host
  void releaseSystem();
  void releasePre();
  void preHost(DofsSort sort);
  void preHier(DofsSort sort);
  void preDevice(DofsSort sort);
  void buildPattern();
  void upload();
};

}  // namespace problem

namespace reorder {
// Hilbert key of a point of a 2-D / 3-D box (fem-master/src/reorder/Hilbert.cpp:17-271)
uint64_t hilbertKey(int dim, const float *f, const float *min, const float *max);
// Hilbert index of integer cell coordinates u, levels top..0
uint64_t hilbertIndex(int dim, const uint32_t *u, int top);
// stable ordering of points along the Hilbert curve
std::vector<int64_t> hilbertOrder(int dim, const std::vector<float> &coords, const float *min, const float *max);
}  // namespace reorder

}  // namespace gmshfem
