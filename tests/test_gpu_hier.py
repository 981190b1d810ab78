"""HierarchicalH1 fields assembled on the device, pinned by mathematics.

The reference's hierarchical space comes from gmsh (FunctionSpaceHierarchicalH1,
fem-master/src/field/FunctionSpaceHierarchicalH1.cpp:83-109), which is not part
of the reference tree: its tables and its numbering are parity unpinned.  The
space itself is P_p on each straight-sided simplex, which fixes:
  - polynomial reproduction: with u in P_p, Dirichlet data u (L2-projected on
    the boundary for p >= 2, as the reference does, DofsFactory.cpp:104-131)
    and source -Laplace(u), the discrete solution is u; checked at the
    vertices (the vertex dofs are point values) in 2-D (triangles) and 3-D
    (tetrahedra, some reversed by the ascending vertex order);
  - identity with Lagrange P_p on the same mesh: the two bases span the same
    space, so a Helmholtz problem with Robin faces (no constraint) has the same
    discrete solution; compared at the vertices;
  - the fused, reference and dense strategies and the mixed precision agree
    as they do for Lagrange fields; the device solver converges."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import gfem
from gfem import workloads as W

pytestmark = pytest.mark.gpu


def _mono(ex, ey, ez=0):
    f = gfem.ScalarFunction(1.0)
    for _ in range(ex):
        f = f * gfem.x()
    for _ in range(ey):
        f = f * gfem.y()
    for _ in range(ez):
        f = f * gfem.z()
    return f


def _direct(F):
    A = sp.csr_matrix((np.asarray(F.values()), F.aj(), F.ai()), shape=(F.n_unknown, F.n_unknown))
    return sla.spsolve(A.tocsc(), -np.asarray(F.rhs(0)))


def _vertex_values(F, model, x):
    """value at every node of a first-order mesh (or the vertex nodes of a
    higher-order one), keyed by rounded coordinates"""
    xyz = model.xyz()
    out = {}
    for tag in range(1, model.num_nodes + 1):
        num, t = F.node_dof(tag)
        if t == 1:
            v = x[num - 1]
        elif t == 2:
            v = F.node_value(tag)
        else:
            continue
        out[tuple(np.round(xyz[tag - 1], 10))] = v
    return out


def _poisson_spec(dim, p, n):
    if dim == 2:
        spec = W.poisson2d(n=n, order=p, basis="hierarchical", jitter=0.2)
        u = _mono(p, 0) + _mono(1, p - 1) + gfem.ScalarFunction(0.5)
        lap = p * (p - 1) * _mono(max(p - 2, 0), 0) if p >= 2 else gfem.ScalarFunction(0.0)
        if p >= 3:
            lap = lap + (p - 1) * (p - 2) * _mono(1, p - 3)
        ex = lambda X: X[0] ** p + X[0] * X[1] ** (p - 1) + 0.5
    else:
        m = gfem.unit_cube(n, 1, 0.15, 11)
        omega, gamma = [(3, 1)], [(2, s) for s in range(1, 7)]
        spec = W.Spec(name=f"poisson3d_n{n}_p{p}_h", model=m, order=p, field_domain=omega, scalar="f64",
                      basis="hierarchical", terms=[dict(kind="bilinear", domain=omega, gauss=2 * p, form=1)])
        u = _mono(p, 0) + _mono(0, 1, p - 1) + _mono(1, 1) + gfem.ScalarFunction(1.0)
        lap = p * (p - 1) * _mono(max(p - 2, 0), 0) if p >= 2 else gfem.ScalarFunction(0.0)
        if p >= 3:
            lap = lap + (p - 1) * (p - 2) * _mono(0, 1, p - 3)
        ex = lambda X: X[0] ** p + X[1] * X[2] ** (p - 1) + X[0] * X[1] + 1.0
    gamma = [(1, s) for s in range(1, 5)] if dim == 2 else [(2, s) for s in range(1, 7)]
    omega = spec.field_domain
    spec.dirichlet = [(gamma, u)]
    spec.terms = [dict(kind="bilinear", domain=omega, gauss=2 * p, form=1),
                  dict(kind="linear", domain=omega, gauss=2 * p, source=lap)]
    return spec, ex


@pytest.mark.parametrize("dim,p,n", [(2, 1, 6), (2, 2, 5), (2, 3, 4), (2, 4, 3), (3, 2, 3), (3, 3, 2), (3, 4, 2)])
def test_polynomial_solutions_are_reproduced(dim, p, n):
    print(f"case hierarchical poisson{dim}d p{p} n{n}", flush=True)
    spec, ex = _poisson_spec(dim, p, n)
    F = W.build(spec)
    F.pre()
    F.assemble()
    x = _direct(F)
    vals = _vertex_values(F, spec.model, x)
    assert len(vals) == spec.model.num_nodes
    err = max(abs(v - ex(np.array(k))) for k, v in vals.items())
    assert err < 1e-9, err


@pytest.mark.parametrize("order", [2, 3])
def test_same_solution_as_lagrange(order):
    print(f"case hierarchical vs lagrange helmholtz3d p{order}", flush=True)
    k = 2 * np.pi
    sl = W.helmholtz3d(n=3, order=order, k=k)
    sh = W.helmholtz3d(n=3, order=order, k=k, basis="hierarchical")
    # build() makes the spec's mesh current: pre() each formulation before the next build
    FL = W.build(sl)
    FL.pre()
    FL.assemble()
    vl = _vertex_values(FL, sl.model, _direct(FL))
    FH = W.build(sh)
    FH.pre()
    FH.assemble()
    vh = _vertex_values(FH, sh.model, _direct(FH))
    assert len(vh) == sh.model.num_nodes
    scale = max(abs(v) for v in vh.values())
    diff = max(abs(vh[c] - vl[c]) for c in vh)
    assert diff < 1e-9 * scale, (diff, scale)


def test_strategies_and_precisions_agree():
    from _helpers import assert_close
    print("case hierarchical strategies p3", flush=True)
    spec = W.helmholtz3d(n=2, order=3, basis="hierarchical")
    F = W.build(spec)
    F.pre()
    F.assemble()
    a, b = np.asarray(F.values()).copy(), np.asarray(F.rhs(0)).copy()
    for strat in ("reference", "dense"):
        F.assemble(strategy=strat)
        assert_close(F.values(), a, 1e-12, strat + " A")
        assert_close(F.rhs(0), b, 1e-12, strat + " b")
    F.assemble(precision="mixed")
    assert_close(F.values(), a, 2e-5, "mixed A")


def test_device_solve():
    print("case hierarchical device solve", flush=True)
    spec, ex = _poisson_spec(2, 3, 5)
    F = W.build(spec)
    F.pre()
    F.assemble()
    info = F.solve(tol=1e-11, restart=60, max_iter=5000)
    assert info["converged"]
    x = np.asarray(F.solution())
    assert np.max(np.abs(x - _direct(F))) < 1e-8 * max(1.0, np.max(np.abs(x)))
