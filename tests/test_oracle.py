"""Pinning the CPU oracle.

Each restated function of oracle/oracle.c is checked against something that
does not reuse its own formula: closed forms, invariants a plausible mistake
breaks, the gmsh manual's documented node positions, or the defining
property of a Hilbert curve.  (CPU only.)"""
import math
from itertools import product

import numpy as np
import pytest

from oracle import oracle as O


# ---------------------------------------------------------------- Gauss rules

def _simplex_monomial(dim, a):
    """exact integral of u^a0 v^a1 w^a2 over the gmsh reference element"""
    if dim == 1:
        k = a[0]
        return (1.0 - (-1.0) ** (k + 1)) / (k + 1)
    num = 1.0
    for ai in a[:dim]:
        num *= math.factorial(ai)
    return num / math.factorial(sum(a[:dim]) + dim)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 12])
def test_gauss_rule_exactness(dim, deg):
    pts, w = O.gauss_rule(dim, deg)
    assert np.all(w > 0)
    # symmetric table rules: every point strictly inside the element
    if dim > 1:
        assert np.all(pts[:, :dim] > 0) and np.all(pts[:, :dim].sum(axis=1) < 1)
    for a in product(range(deg + 1), repeat=dim):
        if sum(a) > deg:
            continue
        a3 = list(a) + [0] * (3 - dim)
        val = np.sum(w * pts[:, 0] ** a3[0] * pts[:, 1] ** a3[1] * pts[:, 2] ** a3[2])
        assert abs(val - _simplex_monomial(dim, a3)) < 1e-14, (dim, deg, a)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 12])
def test_engine_rules_are_the_oracle_rules(dim, deg):
    """the engine's GaussN (csrc/refelem.h) and the oracle's are the same
    points and weights in the same order (symmetric table rules: triangles to
    degree 8 with 1/3/6/7/12/16 points, tetrahedra to degree 6 with
    1/4/14/24 points; collapsed Gauss-Jacobi beyond)"""
    from gfem import gauss_rule
    pe, we = gauss_rule(dim, deg)
    po, wo = O.gauss_rule(dim, deg)
    pe, we = np.asarray(pe, dtype=np.float64).reshape(-1, 3), np.asarray(we, dtype=np.float64)
    assert pe.shape == po.shape
    expect = {2: [1, 1, 3, 6, 6, 7, 12, 16, 16], 3: [1, 1, 4, 14, 14, 14, 24]}
    if dim in expect and deg < len(expect[dim]):
        # the shared table: bit for bit
        assert len(wo) == expect[dim][deg]
        np.testing.assert_array_equal(pe, po)
        np.testing.assert_array_equal(we, wo)
    else:
        # computed rules (Gauss-Jacobi nodes by each side's own eigen-solver)
        np.testing.assert_allclose(pe, po, rtol=0, atol=1e-14)
        np.testing.assert_allclose(we, wo, rtol=1e-13, atol=0)


# ---------------------------------------------------------------- Lagrange basis

def _ref_point(dim, lat, p):
    """reference coordinates of a lattice node (gmsh reference elements)"""
    if dim == 1:
        return np.array([-1.0 + 2.0 * lat[1] / p, 0.0, 0.0])
    u = np.zeros(3)
    for j in range(dim):
        u[j] = lat[j + 1] / p
    return u


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_lagrange_kronecker_partition_reproduction(dim, p):
    nodes = O.ref_nodes(dim, p)
    N = len(nodes)
    assert N == math.comb(p + dim, dim)
    # nodes distinct, on the lattice
    assert len({tuple(n) for n in nodes}) == N and all(sum(n) == p for n in nodes)
    X = np.array([_ref_point(dim, n, p) for n in nodes])
    V = np.array([O.lagrange(dim, p, X[a])[0] for a in range(N)])
    assert np.allclose(V, np.eye(N), atol=1e-13)
    rng = np.random.default_rng(7)
    for _ in range(5):
        u = rng.random(3) * (2 if dim == 1 else 1)
        if dim == 1:
            u = np.array([u[0] - 1.0, 0, 0])
        else:
            u[dim:] = 0
            u[:dim] /= max(1.0, u[:dim].sum() * 1.05)
        val, grad = O.lagrange(dim, p, u)
        assert abs(val.sum() - 1.0) < 1e-13
        assert np.allclose(grad.sum(axis=0), 0.0, atol=1e-12)
        # linear reproduction: sum_a phi_a(u) X_a = u, sum_a X_a grad phi_a = I
        assert np.allclose(val @ X, u, atol=1e-13)
        G = X.T @ grad
        assert np.allclose(G[:dim, :dim], np.eye(dim), atol=1e-12)


def test_lagrange_p2_node_positions_gmsh_manual():
    """gmsh manual, node ordering: 6-node triangle edges (0,1),(1,2),(2,0);
    10-node tetrahedron edges (0,1),(1,2),(2,0),(3,0),(3,2),(3,1)."""
    tri = [tuple(_ref_point(2, n, 2)[:2]) for n in O.ref_nodes(2, 2)]
    assert tri == [(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)]
    tet = [tuple(_ref_point(3, n, 2)) for n in O.ref_nodes(3, 2)]
    assert tet == [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (0.5, 0, 0), (0.5, 0.5, 0), (0, 0.5, 0),
                   (0, 0, 0.5), (0, 0.5, 0.5), (0.5, 0, 0.5)]
    line = [_ref_point(1, n, 3)[0] for n in O.ref_nodes(1, 3)]
    assert np.allclose(line, [-1, 1, -1 / 3, 1 / 3])


# ---------------------------------------------------------------- Hilbert curve

@pytest.mark.parametrize("dim,k", [(2, 5), (3, 3)])
def test_hilbert_index_is_a_face_adjacent_path(dim, k):
    """Defining property of a Hilbert curve: the cells of a 2^k grid, sorted by
    index, form a path of face-adjacent cells starting at the origin."""
    side = 1 << k
    cells = list(product(range(side), repeat=dim))
    keys = [O.hilbert_index(dim, c, k - 1) for c in cells]
    assert sorted(keys) == list(range(side ** dim))
    path = [c for _, c in sorted(zip(keys, cells))]
    assert path[0] == (0,) * dim
    for a, b in zip(path, path[1:]):
        assert sum(abs(x - y) for x, y in zip(a, b)) == 1


def _numbering(spec):
    P = O.from_spec(spec)
    P.pre()
    m = spec.model
    xyz = m.xyz()
    order = sorted(range(1, m.num_nodes + 1), key=lambda t: P.node_dof(t)[0])
    assert [P.node_dof(t)[0] for t in order] == list(range(1, m.num_nodes + 1))
    return xyz[np.array(order) - 1]


@pytest.mark.parametrize("dim,cells", [(2, 3), (2, 7), (2, 15), (3, 3), (3, 7)])
def test_dof_numbering_walks_grid_neighbours(dim, cells):
    """Unjittered P1 grid of 2^k nodes per side, no constraint: each node falls
    in its own level-k Hilbert cell, so DofsManager::reorderWithHilbert must
    number them along a path of grid neighbours (checks quantisation, curve
    tables, stable sort and renumbering together)."""
    import gfem
    from gfem import workloads as W
    if dim == 2:
        spec = W.poisson2d(n=cells, order=1, jitter=0.0)
        spec.dirichlet = []
    else:
        spec = W.highorder3d(n=cells, order=1, jitter=0.0, gauss=2)
    pts = _numbering(spec)
    steps = np.abs(np.diff(pts, axis=0)) * cells
    assert np.allclose(steps.sum(axis=1), 1.0)
    assert np.allclose(pts[0], 0.0)


# ---------------------------------------------------------------- element terms

def _single(xyz, gmsh_type, conn, dim, terms, order=1):
    """oracle problem on one element, no constraint, dofs numbered by node"""
    P = O.Problem()
    P.set_nodes(np.asarray(xyz, float))
    P.add_block(dim, 1, gmsh_type, np.asarray([conn], np.int64))
    P.set_field(order, [(dim, 1)])
    for t in terms:
        if t[0] == "bil":
            P.add_bilinear([(dim, 1)], t[1], t[2], t[3] if len(t) > 3 else 1.0)
        else:
            P.add_linear([(dim, 1)], t[1], t[2])
    P.pre()
    P.assemble()
    n = len(conn)
    perm = [P.node_dof(t)[0] - 1 for t in conn]  # local node -> dof
    ai, aj = P.csr()
    v = P.values()
    A = np.zeros((n, n), complex)
    for r in range(n):
        A[r, aj[ai[r]:ai[r + 1]]] = v[ai[r]:ai[r + 1]]
    A = A[np.ix_(perm, perm)]
    b = P.rhs(0)[perm] if P.L.or_n_rhs(P.h) else None
    return A, b


def test_p1_triangle_cotangent_stiffness_mass_load():
    X = np.array([[0.1, 0.2, 0.0], [1.3, 0.1, 0.0], [0.4, 0.9, 0.0]])
    K, b = _single(X, 2, [1, 2, 3], 2, [("bil", 2, 1), ("lin", 2, -1.0)])
    area = 0.5 * abs(np.cross(X[1] - X[0], X[2] - X[0])[2])
    Kc = np.zeros((3, 3))
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        # cot of the angle at vertex i, opposite edge (j, k)
        e1, e2 = X[j] - X[i], X[k] - X[i]
        cot = np.dot(e1, e2) / np.linalg.norm(np.cross(e1, e2))
        Kc[j, k] = Kc[k, j] = -0.5 * cot
    for i in range(3):
        Kc[i, i] = -Kc[i].sum()
    assert np.allclose(K.real, Kc, rtol=1e-13, atol=1e-14)
    assert np.allclose(b.real, -area / 3, rtol=1e-13)
    M, _ = _single(X, 2, [1, 2, 3], 2, [("bil", 2, 0)])
    assert np.allclose(M.real, area / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-13)


def test_p1_tet_stiffness_and_mass():
    X = np.array([[0.0, 0.1, 0.0], [1.1, 0.0, 0.2], [0.2, 0.9, 0.1], [0.3, 0.2, 1.2]])
    K, _ = _single(X, 4, [1, 2, 3, 4], 3, [("bil", 2, 1)])
    D = np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]]).T
    vol = np.linalg.det(D) / 6
    G = np.linalg.inv(D)  # rows: gradients of lambda_1..3
    grads = np.vstack([-G.sum(axis=0), G])
    assert np.allclose(K.real, vol * grads @ grads.T, rtol=1e-12, atol=1e-14)
    M, _ = _single(X, 4, [1, 2, 3, 4], 3, [("bil", 2, 0)])
    assert np.allclose(M.real, vol / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-13)


def test_linear_term_expression_source():
    """b_i = -int x phi_i = -A/12 (x_i + sum_j x_j) on a P1 triangle."""
    import gfem
    X = np.array([[0.1, 0.2, 0.0], [1.3, 0.1, 0.0], [0.4, 0.9, 0.0]])
    _, b = _single(X, 2, [1, 2, 3], 2, [("bil", 2, 1), ("lin", 2, -gfem.x())])
    area = 0.5 * abs(np.cross(X[1] - X[0], X[2] - X[0])[2])
    assert np.allclose(b.real, -area / 12 * (X[:, 0] + X[:, 0].sum()), rtol=1e-13)


# ---------------------------------------------------------------- global assembly

def _dense(P):
    ai, aj = P.csr()
    v = P.values()
    n = P.n_unknown
    A = np.zeros((n, n), complex)
    for r in range(n):
        A[r, aj[ai[r]:ai[r + 1]]] = v[ai[r]:ai[r + 1]]
    return A


@pytest.mark.parametrize("dim,order,jit", [(2, 1, 0.2), (2, 2, 0.2), (2, 3, 0.1), (3, 1, 0.15), (3, 2, 0.15)])
def test_stiffness_patch_test(dim, order, jit):
    """no constraint: K 1 = 0 on every row, K u_lin = 0 on interior rows,
    K symmetric; the mass matrix sums to the volume."""
    from gfem import workloads as W
    if dim == 2:
        spec = W.poisson2d(n=5, order=order, jitter=jit)
        spec.dirichlet = []
        spec.terms = spec.terms[:1]
    else:
        spec = W.highorder3d(n=3, order=order, jitter=jit, gauss=2 * order)
        spec.terms = spec.terms[:1]
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    K = _dense(P).real
    xyz = spec.model.xyz()
    n = P.n_unknown
    pos = np.zeros((n, 3))
    for t in range(1, spec.model.num_nodes + 1):
        num, typ = P.node_dof(t)
        pos[num - 1] = xyz[t - 1]
    assert np.allclose(K, K.T, atol=1e-12)
    assert np.allclose(K.sum(axis=1), 0, atol=1e-11)
    lin = 0.3 + 1.7 * pos[:, 0] - 0.6 * pos[:, 1] + 0.9 * pos[:, 2]
    interior = np.all((pos > 1e-9) & (pos < 1 - 1e-9), axis=1)[:, None] if dim == 3 else \
        np.all((pos[:, :2] > 1e-9) & (pos[:, :2] < 1 - 1e-9), axis=1)[:, None]
    r = (K @ lin)[interior[:, 0]]
    assert np.abs(r).max() < 1e-11


def test_helmholtz_entry_sum_is_closed_form():
    """sum of all entries = -k^2 vol - i k area (K 1 = 0); source f = 1 gives
    sum(b) = -vol."""
    from gfem import workloads as W
    k = 3.0
    spec = W.helmholtz3d(n=3, order=2, k=k)
    spec.terms[3] = dict(kind="linear", domain=[(3, 1)], gauss=4, source=-1.0)
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    s = P.values().sum()
    assert abs(s - (-k * k - 6j * k)) < 1e-11
    assert abs(P.rhs(0).sum() - (-1.0)) < 1e-12


def test_dense_brute_force_p1_with_dirichlet_lift():
    """global P1 assembly rebuilt densely from per-element cotangent formulas,
    Dirichlet rows/cols removed and lifted: equals the oracle's CSR + RHS, and
    the pattern equals the brute-force nonzero structure."""
    from gfem import workloads as W
    spec = W.poisson2d(n=4, order=1, jitter=0.2)
    g = 0.75
    spec.dirichlet = [(spec.dirichlet[0][0], g)]
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    m = spec.model
    xyz = m.xyz()
    nn = m.num_nodes
    Kg = np.zeros((nn, nn))
    bg = np.zeros(nn)
    tri = [b for b in m.blocks() if b["dim"] == 2][0]["conn"]
    for c in tri:
        X = xyz[c - 1]
        area = 0.5 * abs(np.cross(X[1] - X[0], X[2] - X[0])[2])
        for i in range(3):
            j, k = (i + 1) % 3, (i + 2) % 3
            e1, e2 = X[j] - X[i], X[k] - X[i]
            cot = np.dot(e1, e2) / np.linalg.norm(np.cross(e1, e2))
            for a, b_ in ((j, k), (k, j)):
                Kg[c[a] - 1, c[b_] - 1] -= 0.5 * cot
                Kg[c[a] - 1, c[a] - 1] += 0.5 * cot
        for i in range(3):
            bg[c[i] - 1] -= area / 3
    num = np.array([P.node_dof(t)[0] for t in range(1, nn + 1)])
    typ = np.array([P.node_dof(t)[1] for t in range(1, nn + 1)])
    unk = np.where(typ == 2)[0]
    fix = np.where(typ == 1)[0]
    A = _dense(P)
    Kb = np.zeros_like(A)
    Kb[np.ix_(num[unk] - 1, num[unk] - 1)] = Kg[np.ix_(unk, unk)]
    assert np.allclose(A, Kb, atol=1e-13)
    ai, aj = P.csr()
    pattern = np.zeros(A.shape, bool)
    for r in range(A.shape[0]):
        pattern[r, aj[ai[r]:ai[r + 1]]] = True
    assert np.array_equal(pattern, np.abs(Kb) > 0)
    bb = np.zeros(len(unk))
    bb[num[unk] - 1] = bg[unk] + Kg[np.ix_(unk, fix)] @ np.full(len(fix), g)
    assert np.allclose(P.rhs(0), bb, atol=1e-13)


def test_constant_dirichlet_gives_constant_solution():
    """-Laplace u = 0 with u = g on the boundary: solving A x = -b gives x = g
    (sign convention of Solver.cpp:119-123 and the Dirichlet lifting)."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.linalg import spsolve
    from gfem import workloads as W
    spec = W.poisson2d(n=6, order=2, jitter=0.2)
    spec.dirichlet = [(spec.dirichlet[0][0], 2.5)]
    spec.terms = spec.terms[:1]
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    ai, aj = P.csr()
    A = csr_matrix((np.ascontiguousarray(P.values().real), aj, ai), shape=(P.n_unknown,) * 2)
    xsol = spsolve(A.tocsc(), np.ascontiguousarray(-P.rhs(0).real))
    assert np.allclose(xsol, 2.5, atol=1e-12)


def test_threaded_assembly_matches_sequential():
    """The OpenMP element loop (atomic scatter, used only to time the CPU
    baseline) assembles the same system as the sequential oracle loop, up to the
    order of the additions."""
    from gfem import workloads as W
    spec = W.helmholtz3d(n=5, order=2)
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    v1, r1 = P.values(), P.rhs(0)
    P.set_threads(4)
    P.assemble()
    v4, r4 = P.values(), P.rhs(0)
    assert np.abs(v4 - v1).max() <= 1e-13 * np.abs(v1).max()
    assert np.abs(r4 - r1).max() <= 1e-13 * np.abs(r1).max()
