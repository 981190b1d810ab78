"""Device GMRES (FP32 Krylov, FP64 iterative refinement) vs a direct solve.

The reference solves A x = -b with an exact sparse LU (PETSc/MUMPS,
fem-master/src/system/Solver.cpp:76-160, sign at :119-123).  Here the same
system, assembled by the CPU oracle, is solved by scipy's sparse LU, and the
device solution must agree within 1e-8 relative (2-norm) while the FP64
residual ||-b - A x|| / ||b|| reaches the requested tolerance."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

pytestmark = pytest.mark.gpu

SOL_RTOL = 1e-8


def _direct(spec):
    from oracle import oracle as O
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    ai, aj = P.csr()
    A = sp.csr_matrix((P.values(), aj, ai), shape=(P.n_unknown, P.n_unknown)).tocsc()
    b = P.rhs(0)
    return sla.spsolve(A, -b), A, b


def _solve(spec, scalar=None, **kw):
    from gfem import workloads as W
    F = W.build(spec, scalar)
    F.pre()
    F.assemble()
    info = F.solve(**kw)
    return F, info


@pytest.mark.parametrize("case", ["poisson2d_p1", "poisson2d_p2", "helmholtz3d_p1", "helmholtz3d_p2"])
def test_solution_matches_direct_solve(case):
    from gfem import workloads as W
    spec = {
        "poisson2d_p1": lambda: W.poisson2d(n=32, order=1),
        "poisson2d_p2": lambda: W.poisson2d(n=16, order=2),
        "helmholtz3d_p1": lambda: W.helmholtz3d(n=10, order=1),
        "helmholtz3d_p2": lambda: W.helmholtz3d(n=6, order=2),
    }[case]()
    print(f"case solve {case}", flush=True)
    x_ref, A, b = _direct(spec)
    F, info = _solve(spec, tol=1e-12)
    x = F.solution()
    assert info["converged"], info
    assert info["rel_residual"] <= 1e-12
    err = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
    assert err <= SOL_RTOL, (err, info)
    # the residual the solver reports is the true FP64 residual of its x
    r = -b - A @ x
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-11


def test_helmholtz_wavenumber_16pi():
    """k = 16 pi (the top of the serving range) on a P2 mesh: indefinite,
    still solved to the solution bar without a preconditioner."""
    from gfem import workloads as W
    print("case solve helmholtz k=16pi", flush=True)
    spec = W.helmholtz3d(n=6, order=2, k=16 * np.pi)
    x_ref, _, _ = _direct(spec)
    F, info = _solve(spec, tol=1e-12, max_iter=40000)
    assert info["converged"], info
    err = np.linalg.norm(F.solution() - x_ref) / np.linalg.norm(x_ref)
    assert err <= SOL_RTOL, (err, info)


def test_solve_is_deterministic_and_restartable():
    from gfem import workloads as W
    print("case solve determinism", flush=True)
    spec = W.helmholtz3d(n=5, order=2)
    F, i1 = _solve(spec, tol=1e-10, restart=20)
    x1 = F.solution().copy()
    i2 = F.solve(tol=1e-10, restart=20)
    assert np.array_equal(F.solution(), x1)
    assert i1["iterations"] == i2["iterations"]
    # a different restart length converges to the same solution
    F.solve(tol=1e-12, restart=50)
    assert np.linalg.norm(F.solution() - x1) / np.linalg.norm(x1) <= 1e-8


def test_single_precision_formulation_solves_in_fp64():
    """Formulation<float>: the engine assembles and refines in FP64 and
    returns the solution rounded to float."""
    from gfem import workloads as W
    print("case solve float formulation", flush=True)
    spec = W.poisson2d(n=24, order=1)
    x_ref, _, _ = _direct(spec)
    F, info = _solve(spec, "f32", tol=1e-10)
    assert info["converged"]
    x = F.solution()
    assert x.dtype == np.float32
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= 1e-6


def test_node_values_carry_dirichlet_data():
    """nodeValue: unknown dofs give the solution, fixed dofs the Dirichlet
    value (here a constant 1 everywhere with no source: u = 1)."""
    from gfem import workloads as W
    print("case node values", flush=True)
    spec = W.poisson2d(n=8, order=2)
    spec.dirichlet = [(spec.dirichlet[0][0], 1.0)]
    spec.terms = spec.terms[:1]
    F, info = _solve(spec, tol=1e-12)
    assert info["converged"]
    for tag in range(1, spec.model.num_nodes + 1):
        assert abs(F.node_value(tag) - 1.0) < 1e-9


def test_spmv_lanes_are_tuned_at_the_first_solve():
    from gfem import workloads as W
    print("case spmv autotune", flush=True)
    F, info = _solve(W.helmholtz3d(n=6, order=2), tol=1e-10)
    assert info["converged"]
    assert F.tuning("fp64")["spmv_lanes"] in (4, 8, 16)


@pytest.mark.parametrize("problem,restart,max_iter", [("helmholtz", 20, 100000), ("helmholtz", 23, 37),
                                                      ("poisson", 20, 100000), ("poisson", 23, 37)])
def test_graph_replayed_iterations_match_direct_launches(monkeypatch, problem, restart, max_iter):
    """Krylov iterations replayed from HIP graphs (blocks of 5) run the same
    kernels in the same order as direct launches: the same iteration count and
    a bitwise equal solution.  Also with a restart length that is not a
    multiple of the block and an iteration limit that ends inside a block.
    Both the complex (Helmholtz) and the real (Poisson) solver."""
    from gfem import workloads as W
    print(f"case solver graphs {problem} restart={restart} max_iter={max_iter}", flush=True)
    spec = W.helmholtz3d(n=5, order=2) if problem == "helmholtz" else W.poisson2d(n=24, order=2)
    monkeypatch.setenv("GFEM_SOLVER_GRAPHS", "0")
    F, i0 = _solve(spec, tol=1e-10, restart=restart, max_iter=max_iter)
    x0 = F.solution().copy()
    monkeypatch.setenv("GFEM_SOLVER_GRAPHS", "1")
    i1 = F.solve(tol=1e-10, restart=restart, max_iter=max_iter)
    assert np.array_equal(F.solution(), x0)
    assert i0["iterations"] == i1["iterations"] and i0["converged"] == i1["converged"]
    i2 = F.solve(tol=1e-10, restart=restart, max_iter=max_iter)  # the graphs are reused
    assert np.array_equal(F.solution(), x0) and i2["iterations"] == i0["iterations"]
