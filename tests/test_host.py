"""Host side on the CPU: Formulation::pre() of the engine against the oracle
(bit-exact DOF numbering, CSR pattern, element maps), the shared reference
element conventions, the expression compiler, the C ABI's exported symbols
and the error behaviour that needs no GPU."""
import ctypes

import numpy as np
import pytest

import gfem
from gfem import workloads as W
from oracle import oracle as O


def _specs():
    s_sym = W.poisson2d(n=6, order=2)
    s_sym.symmetric = True
    s_none = W.helmholtz3d(n=3, order=1)
    s_none.dofs_sort = "none"
    s_bnd = W.poisson2d(n=5, order=3, jitter=0.1)
    s_bnd.dirichlet = [([(1, 1), (1, 4)], 1.0), ([(1, 2)], 2.0)]
    return {
        "poisson_p1": W.poisson2d(n=9, order=1),
        "poisson_p2": W.poisson2d(n=6, order=2),
        "poisson_p4": W.poisson2d(n=3, order=4),
        "helm_p1": W.helmholtz3d(n=4, order=1),
        "helm_p2": W.helmholtz3d(n=3, order=2),
        "helm_p3": W.helmholtz3d(n=2, order=3),
        "high_p4": W.highorder3d(n=2, order=4),
        "symmetric": s_sym,
        "sort_none": s_none,
        "partial_dirichlet": s_bnd,
    }


@pytest.mark.parametrize("name", list(_specs()))
def test_pre_matches_oracle_bit_exact(name):
    spec = _specs()[name]
    F = W.build(spec)
    F.pre(spec.dofs_sort, device=False)
    P = O.from_spec(spec)
    P.pre()
    assert (F.n_unknown, F.n_fixed, F.nnz) == (P.n_unknown, P.n_fixed, P.nnz)
    ai, aj = P.csr()
    assert np.array_equal(F.ai(), ai) and np.array_equal(F.aj(), aj)
    for t in range(1, spec.model.num_nodes + 1):
        pn, pt = P.node_dof(t)
        fn, ft = F.node_dof(t)
        assert pn == fn and {0: 0, 1: 2, 2: 1}[ft] == pt, t
    for i, b in enumerate(spec.model.blocks()):
        maps = F.block_maps(b["dim"], b["tag"], b["type"])
        if maps is None:
            continue
        num, typ = P.block_maps(i)
        assert np.array_equal(maps[0], num.ravel()) and np.array_equal(maps[1], typ.ravel())
        e2n = maps[2]
        if e2n.size:
            # element -> nonzero map: position of (tf row i, dof col j) in the CSR
            N = b["conn"].shape[1]
            n2 = num.reshape(-1, N)
            t2 = typ.reshape(-1, N)
            e2n = e2n.reshape(-1, N, N)
            for e in range(0, n2.shape[0], max(1, n2.shape[0] // 7)):
                for a in range(N):
                    for c in range(N):
                        r, col = n2[e, a], n2[e, c]
                        if t2[e, a] != 2 or t2[e, c] != 2 or (spec.symmetric and col < r):
                            assert e2n[e, a, c] == -1
                        else:
                            assert aj[e2n[e, a, c]] == col and ai[r] <= e2n[e, a, c] < ai[r + 1]


def test_structured_nnz_closed_form():
    """P1 on the n x n two-triangle grid: interior vertex rows have 7 entries,
    so nnz = 7 (n-1)^2 - 4 (n-2)... counted exactly from the mesh graph."""
    n = 10
    spec = W.poisson2d(n=n, order=1, jitter=0.0)
    spec.dirichlet = []
    F = W.build(spec)
    F.pre(device=False)
    # graph of the diagonal split: each vertex connects to itself, 4 axis
    # neighbours and 2 diagonal neighbours (i+1,j+1), (i-1,j-1)
    cnt = 0
    for j in range(n + 1):
        for i in range(n + 1):
            for di, dj in ((0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1)):
                if 0 <= i + di <= n and 0 <= j + dj <= n:
                    cnt += 1
    assert F.nnz == cnt
    ai = F.ai()
    assert np.all(np.diff(ai) > 0)
    aj = F.aj()
    for r in range(F.n_unknown):
        row = aj[ai[r]:ai[r + 1]]
        assert np.all(np.diff(row) > 0)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_reference_node_convention_shared(dim, p):
    assert [list(l) for l in gfem.lattice_nodes(dim, p)] == O.ref_nodes(dim, p)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("deg", [0, 2, 4, 8])
def test_gauss_rules_agree(dim, deg):
    p, w = gfem.gauss_rule(dim, deg)
    po, wo = O.gauss_rule(dim, deg)
    assert np.allclose(np.array(p).reshape(-1, 3), po, atol=1e-15, rtol=0)
    assert np.allclose(w, wo, atol=1e-16, rtol=1e-14)


def test_hilbert_product_matches_oracle():
    rng = np.random.default_rng(3)
    for dim in (2, 3):
        mn, mx = [-0.25, 0.0, 1.0], [1.5, 2.0, 3.0]
        for _ in range(300):
            f = [float(np.float32(mn[i] + (mx[i] - mn[i]) * rng.random())) for i in range(3)]
            assert gfem.hilbert_key(dim, f, mn, mx) == O.hilbert_key(dim, f, mn, mx)
        for _ in range(100):
            u = [int(v) for v in rng.integers(0, 1 << 20, 3)]
            assert gfem.hilbert_index(dim, u, 19) == O.hilbert_index(dim, u, 19)


def test_expression_compiler():
    f = gfem.exp(-((gfem.x() - 0.5) * (gfem.x() - 0.5) + gfem.y() * 2.0) / 0.1) + 1j * gfem.z()
    for pt in [(0.1, 0.2, 0.3), (0.7, -0.4, 2.0)]:
        x, y, z = pt
        ref = np.exp(-((x - 0.5) ** 2 + 2 * y) / 0.1) + 1j * z
        assert abs(f.evaluate(*pt) - ref) < 1e-14
    c = (gfem.ScalarFunction(2.0) * 3.0 - 1.0) / 4.0
    assert c.is_constant and c.value == 1.25
    assert not f.is_constant and len(f.ops) % 2 == 0


def test_c_abi_exports_every_declared_symbol():
    lib = gfem.load_c_abi()
    names = gfem.c_abi_symbols()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n


def test_c_abi_errors_without_device():
    lib = gfem.load_c_abi()
    lib.gfem_status_string.restype = ctypes.c_char_p
    assert lib.gfem_status_string(5) == b"unsupported"
    ctx = ctypes.c_void_p()
    st = lib.gfem_ctx_create(0, ctypes.byref(ctx))
    if st == 0:
        pytest.skip("a HIP device is visible")
    assert st == 2  # GFEM_ERR_NO_DEVICE
    assert lib.gfem_ctx_create(0, None) == 1  # GFEM_ERR_INVALID_ARG


def test_formulation_error_paths():
    spec = W.poisson2d(n=3)
    F = gfem.Formulation("empty")
    gfem.set_current_model(spec.model)
    with pytest.raises(Exception):
        F.pre(device=False)  # no term
    u = gfem.Field("u", gfem.Domain(2, 1), 1)
    with pytest.raises(Exception):
        F.integral(gfem.tf(u), gfem.dof(u), gfem.Domain(2, 1), "Gauss2")  # tf on the dof side
    with pytest.raises(Exception):
        F.integral(gfem.grad(gfem.dof(u)), gfem.tf(u), gfem.Domain(2, 1), "Gauss2")  # form mismatch
    with pytest.raises(Exception):
        F.integral(gfem.dof(u), gfem.tf(u), gfem.Domain(2, 1), "Legendre3")
    with pytest.raises(ValueError):
        gfem.Formulation("x", "f16")


def test_solve_before_assemble_is_an_error():
    """solve() needs an assembled system: the error surfaces as the
    reference's exception (logged, solution left empty), without a device."""
    from gfem import workloads as W
    F = W.build(W.poisson2d(n=4, order=1))
    F.pre(device=False)
    with pytest.raises(RuntimeError, match="solve before assemble"):
        F.solve()


def test_term_coefficient_keeps_its_form():
    """set_term_coefficient: a constant stays a constant and an expression
    keeps its program (its constants may change); checked before any device
    work."""
    import gfem
    from gfem import workloads as W
    F = W.build(W.helmholtz3d(n=3, order=1))
    F.pre(device=False)
    with pytest.raises(Exception, match="keep its form"):
        F.set_term_coefficient(W.HELMHOLTZ_MASS, gfem.exp(gfem.x()))
    with pytest.raises(Exception, match="keep its form"):
        F.set_term_coefficient(W.HELMHOLTZ_SOURCE, gfem.ScalarFunction(1.0))
    for tag, f in W.helmholtz_params(3.0, (0.2, 0.3, 0.4)):
        F.set_term_coefficient(tag, f)
    with pytest.raises(Exception, match="unknown term"):
        F.set_term_coefficient(17, gfem.ScalarFunction(1.0))


@pytest.mark.parametrize("n,order", [(3, 1), (2, 2), (4, 2)])
def test_scatter_map_bytes_counts_the_element_maps(n, order):
    """the roofline's algorithmic bytes of the matrix scatter: 4 B per local
    DOF and 4 B per local pair of every element of a bilinear term, the
    Helmholtz cube having n^3 x 6 tetrahedra and 6 n^2 x 2 Robin triangles"""
    spec = W.helmholtz3d(n=n, order=order, k=5.0)
    Nt = (order + 1) * (order + 2) * (order + 3) // 6  # DOFs of a tetrahedron
    Nf = (order + 1) * (order + 2) // 2                 # of a triangle
    tets, tris = 6 * n ** 3, 12 * n ** 2
    assert spec.scatter_map_bytes() == tets * (4 * Nt + 4 * Nt * Nt) + tris * (4 * Nf + 4 * Nf * Nf)
    # the linear-only domains add nothing: the same mesh with the source term alone
    only_src = W.helmholtz3d(n=n, order=order, k=5.0)
    only_src.terms = [t for t in only_src.terms if t["kind"] == "linear"]
    assert only_src.scatter_map_bytes() == 0


def test_bench_defaults_measure_the_stated_workloads():
    """bench.py with no flags: one GPU, 20 timed assemblies after 3 warm-up
    steps, and the serving leg on the whole stated stream (no request cap,
    200 arrivals/s) with the multigrid preconditioner at shift 1, GMRES(30)"""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(os.path.dirname(__file__), "..",
                                                                               "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    a = bench.arg_parser().parse_args([])
    assert (a.gpus, a.steps, a.warmup) == (1, 20, 3)
    assert a.serve_requests is None and a.serve_rate == 200.0
    assert a.serve_mg == 1 and a.serve_mg_beta == 1.0 and a.serve_restart == 30
    assert a.batch_instances == 256 and a.batch_tol == 1e-8
