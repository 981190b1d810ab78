"""Golden fixtures (tests/golden/) pinning the oracle's Hilbert curve, its
DOF numbering and the AFrequency folding.  (CPU only.)

The Hilbert visiting orders in tests/golden/hilbert_orders.json were derived
by hand from the reference's table construction and index walk
(fem-master/src/reorder/Hilbert.cpp:18-158, :160-177) and its quantisation in
apply() (:188-219); tests/golden/hilbert_derivation.md gives the steps.  Each
fixture is checked against both the oracle (oracle/oracle.c) and the engine's
host tables (fem-g_amd/csrc/hilbert.h, shared with the device pre()), and
each is shown to reject the transposed and the reflected curve, the two
plausible mistakes that a face-adjacency property test lets through."""
import json
import math
import os
from itertools import product

import numpy as np
import pytest

import gfem
from gfem import workloads as W
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "hilbert_orders.json")) as _f:
    GOLD = json.load(_f)

INDEX_CASES = [("index_2d_begbit0", 2, 0), ("index_2d_begbit1", 2, 1), ("index_3d_begbit0", 3, 0),
               ("index_3d_begbit1", 3, 1)]
KEY_CASES = [("key_2d_unit_square_4x4", 2), ("key_3d_unit_cube_4x4x4", 3)]

# the two plausible mistakes: axes interleaved in the other order (the
# transposed curve) and one axis mirrored (the reflected curve)
MUTATIONS = {
    "transposed": lambda c, side: tuple(reversed(c)),
    "reflected": lambda c, side: (side - 1 - c[0],) + tuple(c[1:]),
}


def _index_order(index_fn, dim, top, mutate=None):
    side = 1 << (top + 1)
    cells = list(product(range(side), repeat=dim))
    keys = [index_fn(dim, list(mutate(c, side) if mutate else c), top) for c in cells]
    assert sorted(keys) == list(range(side ** dim))
    return [list(c) for _, c in sorted(zip(keys, cells))]


def _key_order(key_fn, dim, mutate=None):
    side = 4
    cells = list(product(range(side), repeat=dim))
    keys = []
    for c in cells:
        cc = mutate(c, side) if mutate else c
        keys.append(key_fn(dim, [(v + 0.5) / side for v in cc], [0.0] * dim, [1.0] * dim))
    assert len(set(keys)) == len(keys)
    return [list(c) for _, c in sorted(zip(keys, cells))]


SIDES = {"oracle": (O.hilbert_index, O.hilbert_key), "engine": (gfem.hilbert_index, gfem.hilbert_key)}


@pytest.mark.parametrize("side", list(SIDES))
@pytest.mark.parametrize("name,dim,top", INDEX_CASES)
def test_hilbert_index_fixture(side, name, dim, top):
    """getHilbertIndex from the root state ed = 0 (Hilbert.cpp:160-177)"""
    assert _index_order(SIDES[side][0], dim, top) == GOLD[name]["order"]


@pytest.mark.parametrize("side", list(SIDES))
@pytest.mark.parametrize("name,dim", KEY_CASES)
def test_hilbert_key_fixture(side, name, dim):
    """apply()'s float quantisation + index: the top digit is 0 for every point
    below the box maximum, so the curve of real points is rooted one state
    down (Hilbert.cpp:188-219)"""
    assert _key_order(SIDES[side][1], dim) == GOLD[name]["order"]


@pytest.mark.parametrize("mutation", list(MUTATIONS))
@pytest.mark.parametrize("name,dim,top", INDEX_CASES)
def test_index_fixture_rejects_planted_mistake(mutation, name, dim, top):
    """the oracle with a planted mistake (its inputs transposed or mirrored)
    fails the fixture, so the fixture can tell the curve from its images"""
    assert _index_order(O.hilbert_index, dim, top, MUTATIONS[mutation]) != GOLD[name]["order"]


@pytest.mark.parametrize("mutation", list(MUTATIONS))
@pytest.mark.parametrize("name,dim", KEY_CASES)
def test_key_fixture_rejects_planted_mistake(mutation, name, dim):
    assert _key_order(O.hilbert_key, dim, MUTATIONS[mutation]) != GOLD[name]["order"]


def test_key_fixture_differs_from_unshifted_root():
    """rooting the point curve at ed = 0 (forgetting that bit 31 / bit 20 is
    always 0) gives the transposed order: the key fixture pins the root"""
    assert GOLD["key_2d_unit_square_4x4"]["order"] != GOLD["index_2d_begbit1"]["order"]
    assert GOLD["key_3d_unit_cube_4x4x4"]["order"] != GOLD["index_3d_begbit1"]["order"]


def _dirichlet_grid(dim):
    if dim == 2:
        spec = W.poisson2d(n=4, order=1, jitter=0.0)
    else:
        spec = W.helmholtz3d(n=4, order=1, jitter=0.0)
        spec.dirichlet = [([(2, s) for s in range(1, 7)], 0.0)]
    return spec


def _numbered_nodes(spec, node_dof):
    """interior grid nodes (integer coordinates) in numDof order"""
    xyz = spec.model.xyz()
    dim = 2 if spec.name.startswith("poisson") else 3
    got = {}
    for t in range(1, spec.model.num_nodes + 1):
        num, unknown = node_dof(t)
        if unknown:
            got[num] = [int(round(v * 4)) for v in xyz[t - 1][:dim]]
    assert sorted(got) == list(range(1, len(got) + 1))
    return [got[k] for k in sorted(got)]


@pytest.mark.parametrize("dim,name", [(2, "dofs_p1_square_4x4_dirichlet"), (3, "dofs_p1_cube_4x4x4_dirichlet")])
def test_dof_numbering_fixture_oracle(dim, name):
    """DofsManager::reorderWithHilbert over the padded node box
    (DofsManager.cpp:118-239): the unknown nodes of a Dirichlet grid in the
    hand-derived order"""
    spec = _dirichlet_grid(dim)
    P = O.from_spec(spec)
    P.pre()
    order = _numbered_nodes(spec, lambda t: (P.node_dof(t)[0], P.node_dof(t)[1] == 2))
    assert order == GOLD[name]["nodes_in_numdof_order"]


@pytest.mark.parametrize("dim,name", [(2, "dofs_p1_square_4x4_dirichlet"), (3, "dofs_p1_cube_4x4x4_dirichlet")])
def test_dof_numbering_fixture_engine_host(dim, name):
    """the engine's host pre() gives the same numbering (the device pre() is
    held bitwise to the host pre() by the GPU set-up tests)"""
    spec = _dirichlet_grid(dim)
    F = W.build(spec)
    F.pre(device=False)
    order = _numbered_nodes(spec, lambda t: (F.node_dof(t)[0], F.node_dof(t)[1] == 1))
    assert order == GOLD[name]["nodes_in_numdof_order"]


# ---------------------------------------------------------------- AFrequency folding

def _tet_problem(terms, omega):
    """one jittered P1 tetrahedron, every node unknown, complex terms"""
    X = np.array([[0.1, 0.0, 0.05], [1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.15, 0.2, 1.1]])
    P = O.Problem()
    P.set_nodes(X)
    P.add_block(3, 1, 4, np.array([[1, 2, 3, 4]], np.int64))
    P.set_field(1, [(3, 1)])
    for form, coef, dt in terms:
        P.add_bilinear([(3, 1)], 2, form, coef, dt=dt)
    P.set_options(frequency=omega)
    P.pre()
    P.assemble()
    vol = abs(np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]]))) / 6.0
    perm = [P.node_dof(t)[0] - 1 for t in (1, 2, 3, 4)]
    ai, aj = P.csr()
    v = P.values()
    A = np.zeros((4, 4), complex)
    for r in range(4):
        A[r, aj[ai[r]:ai[r + 1]]] = v[ai[r]:ai[r + 1]]
    return A[np.ix_(perm, perm)], vol


@pytest.mark.parametrize("dt,factor", [(2, lambda w: -w * w), (1, lambda w: -1j * w)])
def test_frequency_folding_closed_form(dt, factor):
    """Assembler.cpp:262-305: a DtDt term enters as -w^2 times its matrix, a Dt
    term as -i w times it.  A lone mass term c * dt^k(u) * tf(u) on a P1 tet
    is factor * c * V/20 * (1 + delta_ij), whose entries sum to factor * c * V."""
    w, c = 3.7, 0.6
    A, vol = _tet_problem([(0, c, dt)], w)
    M = vol / 20.0 * (np.ones((4, 4)) + np.eye(4))
    np.testing.assert_allclose(A, factor(w) * c * M, rtol=1e-13, atol=1e-15)
    assert abs(A.sum() - factor(w) * c * vol) < 1e-13 * abs(factor(w) * c * vol)


def test_frequency_folding_mixed_terms():
    """NotDt terms are added unscaled beside the folded ones in the same
    element matrix: K + 0.5 M - w^2 (2 M) - i w (0.25 M)"""
    w = 2.0
    A, vol = _tet_problem([(1, 1.0, 0), (0, 0.5, 0), (0, 2.0, 2), (0, 0.25, 1)], w)
    K, _ = _tet_problem([(1, 1.0, 0)], 0.0)
    M = vol / 20.0 * (np.ones((4, 4)) + np.eye(4))
    expect = K + (0.5 - w * w * 2.0 - 1j * w * 0.25) * M
    np.testing.assert_allclose(A, expect, rtol=1e-13, atol=1e-14)
    assert abs(A.sum() - (0.5 - 2.0 * w * w - 0.25j * w) * vol) < 1e-12
    assert math.isclose(vol, abs(A.sum() / (0.5 - 2.0 * w * w - 0.25j * w)), rel_tol=1e-12)
