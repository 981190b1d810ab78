// Python binding of the host API (model, domain, functions, fields,
// equation DSL, Formulation<T>).  Used by the tests and bench.py; the C++
// API in gmshfem.hpp is the primary interface.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <memory>

#include <hip/hip_runtime_api.h>

#include "../csrc/refelem.h"
#include "gmshfem.hpp"
#include "serve.hpp"

namespace py = pybind11;
using namespace gmshfem;
using function::ScalarFunction;

namespace {

template <class V>
py::array_t<V> to_numpy(const std::vector<V> &v)
{
  py::array_t<V> a((py::ssize_t)v.size());
  if(!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(V));
  return a;
}

gfem_precision parse_precision(const std::string &s)
{
  if(s == "fp64") return GFEM_PREC_FP64;
  if(s == "mixed") return GFEM_PREC_MIXED;
  throw std::invalid_argument("precision must be 'fp64' or 'mixed'");
}

gfem_strategy parse_strategy(const std::string &s)
{
  if(s == "fused") return GFEM_STRATEGY_FUSED;
  if(s == "reference") return GFEM_STRATEGY_REFERENCE;
  if(s == "dense") return GFEM_STRATEGY_DENSE;
  throw std::invalid_argument("strategy must be 'fused', 'reference' or 'dense'");
}

problem::DofsSort parse_sort(const std::string &s)
{
  if(s == "default") return problem::DofsSort::Default;
  if(s == "hilbert") return problem::DofsSort::Hilbert;
  if(s == "none") return problem::DofsSort::None;
  throw std::invalid_argument("sort must be 'default', 'hilbert' or 'none'");
}

void check(gfem_status s, const char *what)
{
  if(s != GFEM_OK)
    throw std::runtime_error(std::string(what) + ": " + gfem_status_string(s) + ": " +
                             gfem_ctx_last_error(common::deviceContext()));
}

// One instance of gfem_assemble_many: its system, parameters and device
// outputs (allocated at the first assembly in a precision).
struct ManyInstance {
  gfem_system *sys = nullptr;
  int64_t nnz = 0, n = 0;
  int nrhs = 1;
  std::vector<gfem_term_param> params;
  std::vector<std::vector<double>> store;  // expression constants the params point to
  void *vals[2] = {nullptr, nullptr}, *rhs[2] = {nullptr, nullptr};
  int vb[2] = {0, 0};
  ManyInstance() = default;
  ManyInstance(const ManyInstance &) = delete;
  ~ManyInstance()
  {
    for(int p = 0; p < 2; ++p) {
      if(vals[p]) (void)hipFree(vals[p]);
      if(rhs[p]) (void)hipFree(rhs[p]);
    }
  }
  void ensure(int p)
  {
    if(vals[p]) return;
    int64_t nz = 0, nn = 0;
    check(gfem_system_info(sys, (gfem_precision)p, &nz, &nn, &vb[p]), "gfem_system_info");
    if(hipMalloc(&vals[p], (size_t)std::max<int64_t>(nz, 1) * vb[p]) != hipSuccess ||
       hipMalloc(&rhs[p], (size_t)std::max<int64_t>(nn * nrhs, 1) * vb[p]) != hipSuccess)
      throw std::runtime_error("assemble_many: output allocation failed");
  }
  // device values of precision p as complex128
  py::array_t<std::complex<double>> fetch(int p, bool rhs_side, int idx)
  {
    if(!vals[p]) throw std::runtime_error("instance not assembled in this precision");
    const int64_t cnt = rhs_side ? n : nnz;
    std::vector<unsigned char> h((size_t)cnt * vb[p]);
    const unsigned char *src = static_cast<const unsigned char *>(rhs_side ? rhs[p] : vals[p]);
    if(rhs_side) src += (size_t)idx * n * vb[p];
    if(cnt && hipMemcpy(h.data(), src, h.size(), hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("assemble_many: download failed");
    py::array_t<std::complex<double>> a((py::ssize_t)cnt);
    auto *o = a.mutable_data();
    for(int64_t i = 0; i < cnt; ++i) {
      const unsigned char *e = h.data() + (size_t)i * vb[p];
      if(vb[p] == 16) { double v[2]; std::memcpy(v, e, 16); o[i] = {v[0], v[1]}; }
      else if(vb[p] == 8 && p == 0) { double v; std::memcpy(&v, e, 8); o[i] = {v, 0.}; }
      else if(vb[p] == 8) { float v[2]; std::memcpy(v, e, 8); o[i] = {v[0], v[1]}; }
      else { float v; std::memcpy(&v, e, 4); o[i] = {v, 0.f}; }
    }
    return a;
  }
};

template <class T>
void bind_formulation(py::module_ &m, const char *name)
{
  using F = problem::Formulation<T>;
  py::class_<F>(m, name, py::dynamic_attr())
    .def(py::init<const std::string &>())
    .def(
      "integral",
      [](F &f, const equation::Unknown &l, const equation::Unknown &r, const domain::Domain &d, const std::string &g,
         bool herm) { return f.integral(l, r, d, g, herm ? problem::ProductType::Hermitian : problem::ProductType::Scalar); },
      py::arg("lhs"), py::arg("rhs"), py::arg("domain"), py::arg("integration"), py::arg("hermitian") = true)
    .def(
      "integral",
      [](F &f, const ScalarFunction &src, const equation::Unknown &r, const domain::Domain &d, const std::string &g,
         bool herm, unsigned idx) {
        return f.integral(src, r, d, g, herm ? problem::ProductType::Hermitian : problem::ProductType::Scalar, idx);
      },
      py::arg("source"), py::arg("rhs"), py::arg("domain"), py::arg("integration"), py::arg("hermitian") = true,
      py::arg("rhs_idx") = 0)
    .def("set_term_coefficient", &F::setTermCoefficient, py::arg("tag"), py::arg("coefficient"))
    .def("set_shifted_laplacian", &F::setShiftedLaplacian, py::arg("tag"), py::arg("beta") = 1.0)
    .def(
      "many_instance",
      [](F &f, const std::vector<std::pair<unsigned int, ScalarFunction>> &params) {
        auto inst = std::make_shared<ManyInstance>();
        f.instanceParams(params, inst->params, inst->store);
        inst->sys = f.system();
        int vb = 0;
        check(gfem_system_info(inst->sys, GFEM_PREC_FP64, &inst->nnz, &inst->n, &vb), "gfem_system_info");
        inst->nrhs = (int)f.numRHS();
        return inst;
      },
      py::arg("params") = std::vector<std::pair<unsigned int, ScalarFunction>>{})
    .def("set_symmetric", &F::setSymmetric)
    .def("set_angular_frequency", &F::setAngularFrequency)
    .def("num_rhs", &F::numRHS)
    .def(
      "pre",
      [](F &f, const std::string &s, bool device) {
        f.setPreOnDevice(device);
        double t;
        {
          py::gil_scoped_release rel;
          t = f.pre(parse_sort(s)).time();
        }
        return t;
      },
      py::arg("sort") = "default", py::arg("device") = true)
    .def(
      "assemble",
      [](F &f, const std::string &prec, const std::string &strat, bool download) {
        problem::AssembleOptions o;
        o.precision = parse_precision(prec);
        o.strategy = parse_strategy(strat);
        o.download = download;
        double t;
        {
          py::gil_scoped_release rel;
          t = f.assemble(o).time();
        }
        if(!f.lastError().empty()) throw std::runtime_error(f.lastError());
        return t;
      },
      py::arg("precision") = "fp64", py::arg("strategy") = "fused", py::arg("download") = true)
    .def(
      "solve",
      [](F &f, double tol, int restart, int max_iter, unsigned rhs_idx, bool download) {
        problem::SolveOptions o;
        o.tol = tol;
        o.restart = restart;
        o.maxIter = max_iter;
        o.rhsIdx = rhs_idx;
        o.download = download;
        double t;
        {
          py::gil_scoped_release rel;
          t = f.solve(o).time();
        }
        if(!f.lastError().empty()) throw std::runtime_error(f.lastError());
        const problem::SolveInfo &si = f.solveInfo();
        py::dict d;
        d["seconds"] = t;
        d["iterations"] = si.iterations;
        d["cycles"] = si.cycles;
        d["rel_residual"] = si.relResidual;
        d["converged"] = si.converged;
        return d;
      },
      py::arg("tol") = 1e-8, py::arg("restart") = 50, py::arg("max_iter") = 20000, py::arg("rhs_idx") = 0,
      py::arg("download") = true)
    .def("solution", [](const F &f) { return to_numpy(f.getSolution()); })
    .def("node_value", &F::nodeValue)
    .def(
      "assemble_async",
      [](F &f, const std::string &prec, const std::string &strat) {
        check(gfem_assemble(f.system(), parse_precision(prec), parse_strategy(strat), nullptr), "gfem_assemble");
      },
      py::arg("precision") = "fp64", py::arg("strategy") = "fused")
    .def("sync", [](F &) { check(gfem_stream_sync(common::deviceContext(), nullptr), "gfem_stream_sync"); })
    .def("set_gemm_mode",
         [](F &f, const std::string &m) {
           if(m != "fp32" && m != "bf16x3") throw std::invalid_argument("gemm mode: fp32 or bf16x3");
           if(!f.system()) throw std::runtime_error("set_gemm_mode: the system exists after the first assemble()");
           check(gfem_set_gemm_mode(f.system(), m == "fp32" ? GFEM_GEMM_FP32 : GFEM_GEMM_BF16X3), "set_gemm_mode");
         })
    .def("system_handle",
         [](F &f) {
           // the gfem_system* behind this formulation, for calls through the C ABI
           if(!f.system()) throw std::runtime_error("system_handle: the system exists after the first assemble()");
           return reinterpret_cast<uintptr_t>(f.system());
         })
    .def("set_profiling", [](F &f, bool on) { check(gfem_set_profiling(f.system(), on ? 1 : 0), "set_profiling"); })
    .def("kernel_times",
         [](F &f) {
           const char *names[256];
           double ms[256];
           int n = 0;
           check(gfem_last_kernel_times(f.system(), 256, names, ms, &n), "kernel_times");
           py::list out;
           for(int k = 0; k < n && k < 256; ++k) out.append(py::make_tuple(std::string(names[k]), ms[k]));
           return out;
         })
    .def("tuning",
         [](F &f, const std::string &prec) {
           int gs = 0, eb = 0, lanes = 0, rl = 0;
           check(gfem_tuning(f.system(), parse_precision(prec), &gs, &eb, &lanes, &rl), "gfem_tuning");
           py::dict d;
           d["gather_slots"] = gs;
           d["linear_eb_shift"] = eb;
           d["spmv_lanes"] = lanes;
           d["rhs_lanes"] = rl;
           int ow = 0;
           check(gfem_gather_one_wave(f.system(), parse_precision(prec), &ow), "gfem_gather_one_wave");
           d["gather_one_wave"] = ow != 0;
           return d;
         },
         py::arg("precision") = "fp64")
    .def("gather_bytes",
         [](F &f, const std::string &prec) {
           double b = 0.;
           check(gfem_gather_bytes(f.system(), parse_precision(prec), &b), "gather_bytes");
           return b;
         })
    .def_property_readonly("n_unknown", &F::getNumberOfUnknownDof)
    .def_property_readonly("n_fixed", &F::getNumberOfFixedDof)
    .def_property_readonly("nnz", &F::getNumberOfNonZeros)
    .def_property_readonly("last_error", &F::lastError)
    .def_property_readonly("pre_seconds", &F::preHostSeconds)
    .def("ai", [](const F &f) { return to_numpy(f.ai()); })
    .def("aj", [](const F &f) { return to_numpy(f.aj()); })
    .def("values", [](const F &f) { return to_numpy(f.getLHS()); })
    .def(
      "rhs", [](const F &f, unsigned i) { return to_numpy(f.getRHS(i)); }, py::arg("idx") = 0)
    .def("node_dof", &F::nodeDof)
    .def("unknown_keys",
         [](const F &f) {
           // hierarchical fields: one row per unknown dof, numDof order:
           // (entity vertex count, vertex tags ascending (0-padded), mode)
           const auto &k = f.unknownKeys();
           py::array_t<int64_t> a({(py::ssize_t)k.size(), (py::ssize_t)6});
           auto r = a.mutable_unchecked<2>();
           for(size_t i = 0; i < k.size(); ++i) {
             r(i, 0) = k[i].n;
             for(int j = 0; j < 4; ++j) r(i, 1 + j) = k[i].v[j];
             r(i, 5) = k[i].mode;
           }
           return a;
         })
    .def("system_maps",
         [](const F &f, int which) {
           int64_t n = 0;
           check(gfem_system_maps(f.system(), which, nullptr, 0, &n), "gfem_system_maps");
           const bool wide = which == 1 || which == 3;
           if(wide) {
             py::array_t<int64_t> a((py::ssize_t)n);
             check(gfem_system_maps(f.system(), which, a.mutable_data(), sizeof(int64_t) * (size_t)n, &n),
                   "gfem_system_maps");
             return py::object(a);
           }
           if(which == 5) {
             py::array_t<int32_t> a((py::ssize_t)n);
             check(gfem_system_maps(f.system(), which, a.mutable_data(), sizeof(int32_t) * (size_t)n, &n),
                   "gfem_system_maps");
             return py::object(a);
           }
           py::array_t<uint32_t> a((py::ssize_t)n);
           check(gfem_system_maps(f.system(), which, a.mutable_data(), sizeof(uint32_t) * (size_t)n, &n),
                 "gfem_system_maps");
           return py::object(a);
         })
    .def("num_groups",
         [](const F &f) {
           int64_t n = 0;
           check(gfem_group_maps(f.system(), -1, 0, nullptr, 0, &n), "gfem_group_maps");
           return n;
         })
    .def("group_maps",
         [](const F &f, int g, int which) {
           int64_t n = 0;
           check(gfem_group_maps(f.system(), g, which, nullptr, 0, &n), "gfem_group_maps");
           py::array_t<int32_t> a((py::ssize_t)n);
           check(gfem_group_maps(f.system(), g, which, a.mutable_data(), (size_t)n, &n), "gfem_group_maps");
           return a;
         })
    .def("block_maps", [](const F &f, int dim, int tag, int type) -> py::object {
      const auto *num = f.blockDofs(dim, tag, type);
      if(!num) return py::none();
      return py::make_tuple(to_numpy(*num), to_numpy(*f.blockTypes(dim, tag, type)),
                            to_numpy(*f.blockNonzeros(dim, tag, type)));
    });
}

}  // namespace

PYBIND11_MODULE(_gfem_host, m)
{
  m.doc() = "gfem host API: GmshFEM Formulation over the MI355X assembly engine";
  m.def("set_device", &common::setDevice);
  m.def("device_context_handle", []() {
    // the gfem_ctx* the formulations use, for calls through the C ABI
    return reinterpret_cast<uintptr_t>(common::deviceContext());
  });
  py::class_<ManyInstance, std::shared_ptr<ManyInstance>>(m, "ManyInstance")
    .def("values", [](ManyInstance &i, const std::string &p) { return i.fetch((int)parse_precision(p), false, 0); },
         py::arg("precision") = "fp64")
    .def("rhs", [](ManyInstance &i, const std::string &p, int idx) { return i.fetch((int)parse_precision(p), true, idx); },
         py::arg("precision") = "fp64", py::arg("idx") = 0);
  m.def(
    "assemble_many",
    [](const std::vector<std::shared_ptr<ManyInstance>> &inst, const std::string &prec, bool sync) {
      const gfem_precision p = parse_precision(prec);
      std::vector<gfem_instance> v(inst.size());
      for(size_t k = 0; k < inst.size(); ++k) {
        inst[k]->ensure((int)p);
        v[k].sys = inst[k]->sys;
        v[k].n_params = (int)inst[k]->params.size();
        v[k].params = inst[k]->params.data();
        v[k].vals = inst[k]->vals[(int)p];
        v[k].rhs = inst[k]->rhs[(int)p];
      }
      check(gfem_assemble_many((int)v.size(), v.data(), p, nullptr), "gfem_assemble_many");
      if(sync) check(gfem_stream_sync(common::deviceContext(), nullptr), "gfem_stream_sync");
    },
    py::arg("instances"), py::arg("precision") = "fp64", py::arg("sync") = true);

  py::class_<model::Model, std::shared_ptr<model::Model>>(m, "Model")
    .def_readonly("name", &model::Model::name)
    .def_readonly("dim", &model::Model::dim)
    .def_property_readonly("num_nodes", &model::Model::numNodes)
    .def("xyz",
         [](const model::Model &mo) {
           py::array_t<double> a({(py::ssize_t)mo.numNodes(), (py::ssize_t)3});
           std::memcpy(a.mutable_data(), mo.xyz.data(), mo.xyz.size() * sizeof(double));
           return a;
         })
    .def("entities", &model::Model::entities, py::arg("dim") = -1)
    .def_property_readonly("physicals",
                           [](const model::Model &mo) {
                             py::list out;
                             for(const auto &p : mo.physicals)
                               out.append(py::dict(py::arg("dim") = p.dim, py::arg("tag") = p.tag,
                                                   py::arg("name") = p.name, py::arg("entities") = p.entities));
                             return out;
                           })
    .def_property_readonly("file_node_tags", [](const model::Model &mo) { return mo.fileNodeTags; })
    .def("blocks", [](const model::Model &mo) {
      py::list out;
      for(const auto &b : mo.blocks) {
        py::array_t<int64_t> c({(py::ssize_t)b.size(), (py::ssize_t)b.npe});
        if(!b.conn.empty()) std::memcpy(c.mutable_data(), b.conn.data(), b.conn.size() * sizeof(int64_t));
        out.append(py::dict(py::arg("dim") = b.dim, py::arg("tag") = b.tag, py::arg("type") = b.type,
                            py::arg("order") = b.order, py::arg("conn") = c));
      }
      return out;
    });
  m.def("unit_square", &model::unitSquare, py::arg("n"), py::arg("order") = 1, py::arg("jitter") = 0.,
        py::arg("seed") = 0);
  m.def("unit_cube", &model::unitCube, py::arg("n"), py::arg("order") = 1, py::arg("jitter") = 0.,
        py::arg("seed") = 0);
  m.def("set_current_model", &model::setCurrent);
  m.def("read_msh", &model::readMsh, py::arg("path"));
  m.def("write_msh", &model::writeMsh, py::arg("model"), py::arg("path"));
  m.def("from_arrays", [](const std::string &name, int dim, py::array_t<double> xyz, py::list blocks) {
    std::vector<double> x(xyz.data(), xyz.data() + xyz.size());
    std::vector<model::Block> bl;
    for(auto item : blocks) {
      py::dict d = item.cast<py::dict>();
      model::Block b;
      b.tag = d["tag"].cast<int>();
      b.type = d["type"].cast<int>();
      py::array_t<int64_t> c = d["conn"].cast<py::array_t<int64_t>>();
      b.conn.assign(c.data(), c.data() + c.size());
      bl.push_back(std::move(b));
    }
    return model::fromArrays(name, dim, x, bl);
  });

  py::class_<domain::Domain>(m, "Domain")
    .def(py::init<>())
    .def(py::init<int, int>())
    .def(py::self | py::self)
    .def("entities", [](const domain::Domain &d) {
      return std::vector<std::pair<int, int>>(d.entities().begin(), d.entities().end());
    });

  py::class_<ScalarFunction>(m, "ScalarFunction")
    .def(py::init<double>())
    .def(py::init<std::complex<double>>())
    .def_property_readonly("is_constant", &ScalarFunction::isConstant)
    .def_property_readonly("value", &ScalarFunction::value)
    .def_property_readonly("ops", &ScalarFunction::ops)
    .def_property_readonly("consts", &ScalarFunction::consts)
    .def("evaluate", &ScalarFunction::evaluate)
    .def(py::self + py::self)
    .def(py::self - py::self)
    .def(py::self * py::self)
    .def(py::self / py::self)
    .def(-py::self)
    .def("__add__", [](const ScalarFunction &a, std::complex<double> b) { return a + ScalarFunction(b); })
    .def("__radd__", [](const ScalarFunction &a, std::complex<double> b) { return ScalarFunction(b) + a; })
    .def("__sub__", [](const ScalarFunction &a, std::complex<double> b) { return a - ScalarFunction(b); })
    .def("__rsub__", [](const ScalarFunction &a, std::complex<double> b) { return ScalarFunction(b) - a; })
    .def("__mul__", [](const ScalarFunction &a, std::complex<double> b) { return a * ScalarFunction(b); })
    .def("__rmul__", [](const ScalarFunction &a, std::complex<double> b) { return ScalarFunction(b) * a; })
    .def("__truediv__", [](const ScalarFunction &a, std::complex<double> b) { return a / ScalarFunction(b); })
    .def("__rtruediv__", [](const ScalarFunction &a, std::complex<double> b) { return ScalarFunction(b) / a; });
  py::implicitly_convertible<double, ScalarFunction>();
  py::implicitly_convertible<std::complex<double>, ScalarFunction>();
  m.def("x", &function::x);
  m.def("y", &function::y);
  m.def("z", &function::z);
  m.def("exp", &function::exp);
  m.def("sqrt", &function::sqrt);
  m.def("sin", &function::sin);
  m.def("cos", &function::cos);
  m.def("abs", &function::abs);
  m.def("pow", &function::pow);

  py::class_<field::Field>(m, "Field")
    .def(py::init([](const std::string &name, const domain::Domain &dom, int order, const std::string &basis) {
           if(basis == "lagrange") return new field::Field(name, dom, field::FunctionSpaceTypeForm0::Lagrange, order);
           if(basis == "hierarchical")
             return new field::Field(name, dom, field::FunctionSpaceTypeForm0::HierarchicalH1, order);
           throw std::invalid_argument("basis must be 'lagrange' or 'hierarchical'");
         }),
         py::arg("name"), py::arg("domain"), py::arg("order") = 1, py::arg("basis") = "lagrange")
    .def("add_constraint", &field::Field::addConstraint)
    .def_property_readonly("order", &field::Field::order)
    .def_property_readonly("basis", [](const field::Field &f) { return f.hierarchical() ? "hierarchical" : "lagrange"; });

  py::class_<equation::Unknown>(m, "Unknown")
    .def("__rmul__", [](const equation::Unknown &u, const ScalarFunction &c) { return c * u; })
    .def("__rmul__", [](const equation::Unknown &u, std::complex<double> c) { return ScalarFunction(c) * u; });
  m.def("dof", &equation::dof, py::keep_alive<0, 1>());
  m.def("tf", &equation::tf, py::keep_alive<0, 1>());
  m.def("dt_dof", &equation::dt_dof, py::keep_alive<0, 1>());
  m.def("dt2_dof", &equation::dt2_dof, py::keep_alive<0, 1>());
  m.def("grad", &equation::grad);

  bind_formulation<double>(m, "Formulation_f64");
  bind_formulation<float>(m, "Formulation_f32");
  bind_formulation<std::complex<double>>(m, "Formulation_c128");
  bind_formulation<std::complex<float>>(m, "Formulation_c64");

  py::class_<serve::Server>(m, "Server")
    .def(py::init<int, size_t, int>(), py::arg("restart") = 50, py::arg("hbm_bytes") = size_t(16) << 30,
         py::arg("max_requests") = 256)
    .def("add_mesh", &serve::Server::addMesh, py::arg("formulation"), py::keep_alive<1, 2>())
    .def("request_bytes", &serve::Server::requestBytes)
    .def("set_progress", &serve::Server::setProgress, py::arg("seconds"))
    .def(
      "run",
      [](serve::Server &sv, const py::list &schedule, bool keep) {
        std::vector<serve::Request> reqs;
        for(const auto &o : schedule) {
          const py::dict d = py::reinterpret_borrow<py::dict>(o);
          serve::Request r;
          r.mesh = d["mesh"].cast<int>();
          r.arrival = d["arrival"].cast<double>();
          if(d.contains("tol")) r.tol = d["tol"].cast<double>();
          if(d.contains("max_iter")) r.maxIter = d["max_iter"].cast<int>();
          if(d.contains("params"))
            for(const auto &p : d["params"]) {
              const py::tuple t = py::reinterpret_borrow<py::tuple>(p);
              r.params.emplace_back(t[0].cast<unsigned>(), t[1].cast<ScalarFunction>());
            }
          reqs.push_back(std::move(r));
        }
        std::vector<serve::Completed> done;
        {
          py::gil_scoped_release rel;
          done = sv.run(reqs, keep);
        }
        py::list out;
        for(const auto &c : done) {
          py::dict d;
          d["request"] = c.request;
          d["arrival"] = c.arrival;
          d["admitted"] = c.admitted;
          d["done"] = c.done;
          d["latency"] = c.latency();
          d["iterations"] = c.iterations;
          d["cycles"] = c.cycles;
          d["rel_residual"] = c.relResidual;
          d["converged"] = c.converged;
          if(keep) d["x"] = to_numpy(c.x);
          out.append(d);
        }
        return out;
      },
      py::arg("schedule"), py::arg("keep_solutions") = false)
    .def_property_readonly("stats", [](const serve::Server &sv) {
      py::dict d;
      d["seconds"] = sv.stats().seconds;
      d["steps"] = sv.stats().steps;
      d["max_live"] = sv.stats().maxLive;
      d["peak_bytes"] = sv.stats().peakBytes;
      d["max_waiting"] = sv.stats().maxWaiting;
      return d;
    });

  m.def("hilbert_key", [](int dim, std::vector<float> f, std::vector<float> mn, std::vector<float> mx) {
    return reorder::hilbertKey(dim, f.data(), mn.data(), mx.data());
  });
  m.def("hilbert_index", [](int dim, std::vector<uint32_t> u, int top) {
    return reorder::hilbertIndex(dim, u.data(), top);
  });
  m.def("lattice_nodes", [](int dim, int order) {
    std::vector<std::vector<int>> out;
    for(auto &l : gfem::lattice_nodes(dim, order)) out.push_back({l[0], l[1], l[2], l[3]});
    return out;
  });
  m.def("gauss_rule", [](int dim, int deg) {
    std::vector<double> p, w;
    gfem::gauss_rule(dim, deg, p, w);
    return py::make_tuple(p, w);
  });
  // hierarchical H1 basis (csrc/hbasis.h): values [npts, N] and reference
  // gradients [npts, N, 3] at reference points pts [npts, 3]
  m.def("hier_basis", [](int dim, int order, py::array_t<double, py::array::c_style | py::array::forcecast> pts) {
    if(dim < 1 || dim > 3 || order < 1 || order > GFEM_HIER_MAX_ORDER) throw std::invalid_argument("dim / order");
    if(pts.ndim() != 2 || pts.shape(1) != 3) throw std::invalid_argument("pts must be [npts, 3]");
    const int n = (int)pts.shape(0), N = gfem::hier::count(dim, order);
    py::array_t<double> v({(py::ssize_t)n, (py::ssize_t)N}), g({(py::ssize_t)n, (py::ssize_t)N, (py::ssize_t)3});
    gfem::hier::eval(dim, order, n, pts.data(), v.mutable_data(), g.mutable_data());
    return py::make_tuple(v, g);
  });
  // keys of the local functions of an element with ascending vertex tags:
  // rows (entity vertex count, vertex tags (0-padded), mode)
  m.def("hier_keys", [](int dim, int order, std::vector<int64_t> vt) {
    if(dim < 1 || dim > 3 || order < 1 || order > GFEM_HIER_MAX_ORDER || (int)vt.size() != dim + 1)
      throw std::invalid_argument("dim / order / vertices");
    std::vector<std::vector<int64_t>> out;
    for(const auto &k : gfem::hier::keys(dim, order, vt.data()))
      out.push_back({k.n, k.v[0], k.v[1], k.v[2], k.v[3], k.mode});
    return out;
  });
}
