"""Hierarchical H1 basis (fem-g_amd/csrc/hbasis.h) pinned by mathematics.

The reference takes this basis from gmsh ("H1Legendre", orientation
"HierarchicalSolin", fem-master/src/field/FunctionSpaceHierarchicalH1.cpp:
32-109), which is not part of the reference tree, so the tables cannot be
compared with gmsh's (parity unpinned against gmsh's normalisation and
orientation conventions).  What the space must satisfy is fixed by the
mathematics and checked here:
  - it is P_p: N = dim P_p functions, linearly independent, spanning every
    monomial of degree <= p;
  - vertex functions are the barycentric coordinates, every other function
    vanishes at the vertices;
  - an edge function restricted to its edge is the Lobatto function l_k =
    (L_k - L_{k-2}) / sqrt(2 (2k - 1)) of the published definition (Solin,
    Segeth, Dolezel 2004), computed here from numpy's Legendre series;
  - traces agree: on a face (edge) of a tetrahedron (triangle), the functions
    with the keys of the face's triangle (the edge's line) equal that
    element's functions and all others vanish, for any vertex tags (the
    conformity of the assembled space);
  - gradients are the derivatives of the values (central differences)."""
import itertools
import math

import numpy as np
import pytest

import gfem

ORDERS = [1, 2, 3, 4]


def _count(dim, p):
    return math.comb(p + dim, dim)


def _random_points(dim, n, rng):
    if dim == 1:
        x = rng.uniform(-1, 1, (n, 1))
    else:
        x = rng.dirichlet(np.ones(dim + 1), n)[:, 1:]
    pts = np.zeros((n, 3))
    pts[:, :dim] = x
    return pts


def _monomials(dim, p, pts):
    cols = []
    for e in itertools.product(range(p + 1), repeat=dim):
        if sum(e) <= p:
            cols.append(np.prod([pts[:, i] ** e[i] for i in range(dim)], axis=0))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("p", ORDERS)
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_basis_is_a_basis_of_P_p(dim, p):
    rng = np.random.default_rng(100 * dim + p)
    N = _count(dim, p)
    pts = _random_points(dim, 4 * N, rng)
    V, _ = gfem.hier_basis(dim, p, pts)
    assert V.shape == (4 * N, N)
    s = np.linalg.svd(V, compute_uv=False)
    assert s[-1] > 1e-8 * s[0], "functions are linearly dependent"
    Mo = _monomials(dim, p, pts)
    assert Mo.shape[1] == N
    coef, *_ = np.linalg.lstsq(V, Mo, rcond=None)
    assert np.max(np.abs(V @ coef - Mo)) < 1e-10, "a monomial of degree <= p is outside the span"
    # and nothing of higher degree: the span equals P_p, so V is in span(monomials)
    c2, *_ = np.linalg.lstsq(Mo, V, rcond=None)
    assert np.max(np.abs(Mo @ c2 - V)) < 1e-10


def _vertices(dim):
    if dim == 1:
        return np.array([[-1.0, 0, 0], [1.0, 0, 0]])
    v = np.zeros((dim + 1, 3))
    for j in range(dim):
        v[j + 1, j] = 1.0
    return v


@pytest.mark.parametrize("p", ORDERS)
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_vertex_functions_are_barycentric(dim, p):
    V, _ = gfem.hier_basis(dim, p, _vertices(dim))
    assert np.allclose(V[:, :dim + 1], np.eye(dim + 1), atol=1e-14)
    assert np.allclose(V[:, dim + 1:], 0.0, atol=1e-14)


def test_edge_functions_are_lobatto_functions():
    from numpy.polynomial import legendre as Lg
    x = np.linspace(-1, 1, 41)
    pts = np.zeros((x.size, 3))
    pts[:, 0] = x
    V, _ = gfem.hier_basis(1, 4, pts)
    for k in range(2, 5):
        ck = np.zeros(k + 1)
        ck[k] = 1.0
        ck2 = np.zeros(k + 1)
        ck2[k - 2] = 1.0
        lk = (Lg.legval(x, ck) - Lg.legval(x, ck2)) / math.sqrt(2 * (2 * k - 1))
        assert np.allclose(V[:, k], lk, atol=1e-13), k


def _trace_check(dim, p, tags, sub, rng):
    """functions of the dim-simplex with vertex tags `tags` (ascending) on its
    sub-simplex of local vertices `sub` against the sub-simplex's own."""
    sd = len(sub) - 1
    keys = [tuple(k) for k in gfem.hier_keys(dim, p, tags)]
    skeys = [tuple(k) for k in gfem.hier_keys(sd, p, [tags[i] for i in sub])]
    # points of the sub-simplex: barycentric weights on its vertices
    w = rng.dirichlet(np.ones(sd + 1), 25)
    lam = np.zeros((25, dim + 1))
    lam[:, sub] = w
    pts = np.zeros((25, 3))
    if dim == 1:
        pts[:, 0] = lam[:, 1] - lam[:, 0]
    else:
        pts[:, :dim] = lam[:, 1:]
    spts = np.zeros((25, 3))
    if sd == 1:
        spts[:, 0] = w[:, 1] - w[:, 0]
    else:
        spts[:, :sd] = w[:, 1:]
    V, _ = gfem.hier_basis(dim, p, pts)
    S, _ = gfem.hier_basis(sd, p, spts)
    for i, k in enumerate(keys):
        if k in skeys:
            assert np.allclose(V[:, i], S[:, skeys.index(k)], atol=1e-13), (k, sub)
        else:
            assert np.allclose(V[:, i], 0.0, atol=1e-13), (k, sub)


@pytest.mark.parametrize("p", ORDERS)
def test_traces_agree_on_faces_and_edges(p):
    rng = np.random.default_rng(7 + p)
    tet = [3, 8, 11, 20]
    for sub in itertools.combinations(range(4), 3):
        _trace_check(3, p, tet, list(sub), rng)
    for sub in itertools.combinations(range(4), 2):
        _trace_check(3, p, tet, list(sub), rng)
    tri = [2, 5, 9]
    for sub in itertools.combinations(range(3), 2):
        _trace_check(2, p, tri, list(sub), rng)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_gradients_are_derivatives(dim):
    rng = np.random.default_rng(3 + dim)
    pts = _random_points(dim, 8, rng) * 0.8 + (0.05 if dim > 1 else 0.0)
    _, G = gfem.hier_basis(dim, 4, pts)
    h = 1e-6
    for j in range(dim):
        d = np.zeros(3)
        d[j] = h
        Vp, _ = gfem.hier_basis(dim, 4, pts + d)
        Vm, _ = gfem.hier_basis(dim, 4, pts - d)
        assert np.allclose(G[:, :, j], (Vp - Vm) / (2 * h), atol=1e-7)
    assert np.all(G[:, :, dim:] == 0.0)


def test_keys_order_and_counts():
    for dim, p in itertools.product([1, 2, 3], ORDERS):
        k = gfem.hier_keys(dim, p, list(range(10, 10 + dim + 1)))
        assert len(k) == _count(dim, p)
        n = [r[0] for r in k]
        assert n == sorted(n), "vertices, edges, faces, interior"
        assert len(set(tuple(r) for r in k)) == len(k)
    with pytest.raises(ValueError):
        gfem.hier_basis(3, 5, np.zeros((1, 3)))


@pytest.mark.parametrize("order", ORDERS)
def test_hierarchical_pre_numbers_the_lagrange_space(order):
    """On the same mesh, hierarchical P_p and Lagrange P_p have one dof per
    entity mode: the same number of unknowns and the same sparsity (host
    pre(), no device; Robin boundary, so no constraint projection)."""
    from gfem import workloads as W
    FL = W.build(W.helmholtz3d(n=3, order=order))
    FL.pre(device=False)
    FH = W.build(W.helmholtz3d(n=3, order=order, basis="hierarchical"))
    FH.pre()
    assert FH.n_unknown == FL.n_unknown and FH.nnz == FL.nnz
    keys = FH.unknown_keys()
    assert keys.shape == (FH.n_unknown, 6)
    # every node of the first-order mesh is a vertex key (no Dirichlet here)
    spec = W.helmholtz3d(n=3, order=order, basis="hierarchical")
    assert int(np.sum(keys[:, 0] == 1)) == spec.model.num_nodes
    # entity vertex tags ascend; modes in range for the order
    for n in (2, 3, 4):
        sel = keys[keys[:, 0] == n]
        assert np.all(np.diff(sel[:, 1:1 + n], axis=1) > 0)
    e = keys[keys[:, 0] == 2]
    assert e.size == 0 or (e[:, 5].min() >= 2 and e[:, 5].max() <= order)
    assert len({tuple(r) for r in keys}) == len(keys)
