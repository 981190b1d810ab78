"""pre() on the device and the device map builder vs the host implementations.

The host pre() is pinned bit-exact to the oracle (tests/test_host.py).  The
device pre() must give the same DOF numbering (node -> numDof, type), the
same CSR pattern and the same element maps; the device map builder must give
the same gather records, RHS maps and element storage order as the host
builder (GFEM_HOST_FINALIZE=1), so the assembled values are bitwise equal.
Cases cover the reference's edge cases for this step: Dirichlet constraints
that overlap on shared nodes (first constraint wins, DofsFactory.cpp:39-74),
no renumbering, symmetric storage, boxes far from the origin (Hilbert
quantisation that leaves the curve's range, Hilbert.cpp:228-271), narrow and
wide gather records, and patterns long enough for many walk segments."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _translated(spec, shift):
    """the same spec on its mesh moved by `shift`"""
    from gfem import from_arrays
    m = spec.model
    xyz = m.xyz() + np.asarray(shift, dtype=np.float64)
    blocks = [dict(tag=b["tag"], type=b["type"], conn=b["conn"]) for b in m.blocks()]
    spec.model = from_arrays(m.name + "_moved", m.dim, xyz, blocks)
    return spec


def _two_constraints(spec):
    """u = 1 on curves 1-2, u = 2+i on curves 2-4: the nodes of curve 2 are
    fixed by the first constraint"""
    spec.scalar = "c128"
    spec.dirichlet = [([(1, 1), (1, 2)], 1.0), ([(1, 2), (1, 3), (1, 4)], 2.0 + 1.0j)]
    return spec


def _specs():
    from gfem import workloads as W

    def sym():
        s = W.poisson2d(n=9, order=2)
        s.symmetric = True
        return s

    def nosort():
        s = W.helmholtz3d(n=3, order=2)
        s.dofs_sort = "none"
        return s

    return {
        "poisson2d_p1": lambda: W.poisson2d(n=12, order=1),
        "poisson2d_p2": lambda: W.poisson2d(n=7, order=2),
        "poisson2d_p3": lambda: W.poisson2d(n=5, order=3),
        "helmholtz3d_p1": lambda: W.helmholtz3d(n=5, order=1),
        "helmholtz3d_p2": lambda: W.helmholtz3d(n=4, order=2),
        "helmholtz3d_p3": lambda: W.helmholtz3d(n=2, order=3),
        "highorder3d_p4": lambda: W.highorder3d(n=2, order=4, gauss=8),
        "symmetric": sym,
        "no_sort": nosort,
        "two_constraints": lambda: _two_constraints(W.poisson2d(n=8, order=2)),
        "far_box_2d": lambda: _translated(W.poisson2d(n=10, order=2), (1000.0, -500.0, 0.0)),
        "far_box_3d": lambda: _translated(W.helmholtz3d(n=4, order=2), (-3000.0, 20.0, 7000.0)),
        "helmholtz3d_p2_n12": lambda: W.helmholtz3d(n=12, order=2),
    }


def _blocks(spec):
    return [(b["dim"], b["tag"], b["type"]) for b in spec.model.blocks()]


PRE_CASES = ["poisson2d_p1", "poisson2d_p2", "poisson2d_p3", "helmholtz3d_p1", "helmholtz3d_p2", "helmholtz3d_p3",
             "highorder3d_p4", "symmetric", "no_sort", "two_constraints", "far_box_2d", "far_box_3d",
             "helmholtz3d_p2_n12"]


@pytest.mark.parametrize("case", PRE_CASES)
def test_device_pre_matches_host_pre(case):
    from gfem import workloads as W
    print(f"case device pre {case}", flush=True)
    spec = _specs()[case]()
    Fh = W.build(spec)
    Fh.pre(spec.dofs_sort, device=False)
    Fd = W.build(spec)
    Fd.pre(spec.dofs_sort, device=True)
    assert (Fd.n_unknown, Fd.n_fixed, Fd.nnz) == (Fh.n_unknown, Fh.n_fixed, Fh.nnz)
    assert np.array_equal(Fd.ai(), Fh.ai())
    assert np.array_equal(Fd.aj(), Fh.aj())
    nodes = range(1, spec.model.num_nodes + 1)
    assert [Fd.node_dof(t) for t in nodes] == [Fh.node_dof(t) for t in nodes]
    for key in _blocks(spec):
        mh, md = Fh.block_maps(*key), Fd.block_maps(*key)
        assert (mh is None) == (md is None), key
        if mh is not None:
            for a, b in zip(md, mh):
                assert np.array_equal(a, b), key


MAP_CASES = ["helmholtz3d_p2", "poisson2d_p3", "highorder3d_p4", "symmetric", "two_constraints",
             "helmholtz3d_p2_n12"]


def _maps(F):
    sm = [F.system_maps(w) for w in range(6)]
    gm = [[F.group_maps(g, w) for w in range(5)] for g in range(F.num_groups())]
    return sm, gm


@pytest.mark.parametrize("case", MAP_CASES)
def test_device_maps_match_host_builder(monkeypatch, case):
    """host pre() + host builder vs device pre() + device builder: every map
    and the assembled values bitwise equal (fixed tile choices, so both
    systems run the same kernels)"""
    from gfem import workloads as W
    print(f"case device maps {case}", flush=True)
    monkeypatch.setenv("GFEM_AUTOTUNE", "0")
    spec = _specs()[case]()
    monkeypatch.setenv("GFEM_HOST_FINALIZE", "1")
    Fh = W.build(spec)
    Fh.pre(spec.dofs_sort, device=False)
    Fh.assemble()
    monkeypatch.delenv("GFEM_HOST_FINALIZE")
    Fd = W.build(spec)
    Fd.pre(spec.dofs_sort)
    Fd.assemble()
    (sh, gh), (sd, gd) = _maps(Fh), _maps(Fd)
    names = ["records", "lin ptr", "lin words", "lift ptr", "lift words", "lift fixed"]
    for w in range(6):
        assert sh[w].dtype == sd[w].dtype and np.array_equal(sh[w], sd[w]), names[w]
    assert len(gh) == len(gd)
    for g in range(len(gh)):
        for w in range(5):
            assert np.array_equal(gh[g][w], gd[g][w]), (g, w)
    assert np.array_equal(Fh.values(), Fd.values())
    for r in range(Fh.num_rhs()):
        assert np.array_equal(Fh.rhs(r), Fd.rhs(r))
    # the mixed-precision assembly on the device maps too
    Fh.assemble(precision="mixed")
    Fd.assemble(precision="mixed")
    assert np.array_equal(Fh.values(), Fd.values())


def test_host_groups_on_the_device_builder():
    """host pre() (groups given as host arrays) + the device builder: the
    oracle's matrix, and records in the shared layout"""
    from gfem import workloads as W
    from oracle import oracle as O
    from _helpers import assert_close
    print("case host groups, device builder", flush=True)
    spec = W.helmholtz3d(n=5, order=2)
    F = W.build(spec)
    F.pre(device=False)
    F.assemble()
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    assert_close(F.values(), P.values(), 1e-12, "A")
    assert_close(F.rhs(0), P.rhs(0), 1e-12, "b")


def _curved(spec, amp):
    """the spec's mesh with every high-order (non-vertex) node moved by the
    smooth field amp * (sin 2 pi y, sin 2 pi x, sin 2 pi (x + y)): conforming
    curved elements, boundary edges included"""
    from gfem import from_arrays
    m = spec.model
    xyz = m.xyz().copy()
    blocks = m.blocks()
    ho = set()
    for b in blocks:
        nv = b["dim"] + 1
        if b["dim"] > 0 and b["conn"].shape[1] > nv:
            ho.update(int(t) for t in b["conn"][:, nv:].ravel())
    i = np.array(sorted(ho), dtype=np.int64) - 1
    x, y = xyz[i, 0].copy(), xyz[i, 1].copy()
    xyz[i, 0] += amp * np.sin(2 * np.pi * y)
    xyz[i, 1] += amp * np.sin(2 * np.pi * x)
    if m.dim == 3:
        xyz[i, 2] += amp * np.sin(2 * np.pi * (x + y))
    spec.model = from_arrays(m.name + "_curved", m.dim, xyz,
                             [dict(tag=b["tag"], type=b["type"], conn=b["conn"]) for b in blocks])
    return spec


def _curved_specs():
    from gfem import workloads as W, x, y
    p2 = W.poisson2d(n=6, order=2)
    p2.terms[0] = dict(kind="bilinear", domain=p2.field_domain, gauss=4, form=1, coef=1.0 + x() * y())
    p2.terms.append(dict(kind="bilinear", domain=p2.field_domain, gauss=4, form=0, coef=0.5))
    return {
        "poisson2d_p2": (p2, 0.008),
        "poisson2d_p3": (W.poisson2d(n=5, order=3), 0.004),
        "helmholtz3d_p2": (W.helmholtz3d(n=3, order=2), 0.01),
    }


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("case", ["poisson2d_p2", "poisson2d_p3", "helmholtz3d_p2"])
def test_curved_elements_match_oracle(case, device):
    """Curved high-order elements (a high-order node off the affine image of
    the vertices): J, det and x from all nodes at every Gauss point, as the
    reference's gmsh getJacobians (Assembler.cpp:130-145).  The engine
    integrates such groups point by point (per-point element matrices, no
    reference tensor).  FP64 within 1e-12 of the oracle for the fused, dense
    and reference strategies; mixed within the FP32 element-matrix bar.  The
    curving changes the oracle's matrix well beyond that bar."""
    from gfem import workloads as W
    from oracle import oracle as O
    from _helpers import assert_close
    print(f"case curved {case} device pre={device}", flush=True)
    spec, amp = _curved_specs()[case]
    # the curving matters: entrywise against the straight mesh, with the
    # coordinate-independent numbering (no Hilbert renumbering) on both
    spec.dofs_sort = "none"
    P0 = O.from_spec(spec)
    P0.pre()
    P0.assemble()
    spec = _curved(spec, amp)
    Pn = O.from_spec(spec)
    Pn.pre()
    Pn.assemble()
    d = np.abs(Pn.values() - P0.values()).max() / np.abs(P0.values()).max()
    assert 1e-3 < d < 0.5, d
    spec.dofs_sort = "default"
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    ref = P.values()
    F = W.build(spec)
    F.pre(device=device)
    for strat in ("fused", "dense", "reference"):
        F.assemble(precision="fp64", strategy=strat)
        assert_close(F.values(), ref, 1e-12, f"{case} A {strat}")
        assert_close(F.rhs(0), P.rhs(0), 1e-12, f"{case} b {strat}")
    F.assemble(precision="mixed")
    assert_close(F.values(), ref, 1e-4, f"{case} A mixed")
    assert_close(F.rhs(0), P.rhs(0), 1e-4, f"{case} b mixed")
