"""Shifted-Laplacian multigrid preconditioner of the batched solver
(gfem_mg_setup): preconditioned solutions against a direct sparse LU of the
oracle's system (the reference solves with MUMPS LU) to 1e-8 relative, fewer
iterations than the row-scaled iteration, and the batch guarantees kept: a
request's result bitwise the same alone and in a mix, and with graph replay or
kernel-by-kernel launches."""
import numpy as np
import pytest

from test_gpu_batch import SOL_RTOL, _direct, _resident, _schedule, _spec

pytestmark = pytest.mark.gpu


def _server(meshes, restart=30, mg=True):
    from gfem import Server, workloads as W
    Fs = [_resident(m) for m in meshes]
    sv = Server(restart=restart, hbm_bytes=1 << 31, max_requests=16)
    levels = []
    for F in Fs:
        sv.add_mesh(F)
        if mg:
            levels.append(F.set_shifted_laplacian(W.HELMHOLTZ_MASS))  # beta 1
    return sv, Fs, levels


def test_mg_hierarchy_levels():
    """rows per level: the DOFs, then uniform grids of twice the mean DOF
    spacing halved per level, down to at most 5 points per axis"""
    print("case mg levels", flush=True)
    _, Fs, levels = _server([(8, 2), (5, 1)])
    assert levels[0] == [17 ** 3, 10 ** 3, 6 ** 3, 4 ** 3], levels
    assert levels[1] == [6 ** 3, 4 ** 3], levels
    for lv, F in zip(levels, Fs):
        assert lv[0] == F.n_unknown


@pytest.mark.parametrize("meshes", [[(8, 2), (5, 1)], [(6, 1), (3, 3)]])
def test_mg_requests_match_direct_solves(meshes):
    """P1, P2 and P3 meshes, k from 2 pi to 16 pi and moving sources, all
    admitted together: every solution within 1e-8 of a direct LU"""
    print(f"case mg vs direct {meshes}", flush=True)
    sv, _, _ = _server(meshes)
    reqs = [(0, 2 * np.pi, (0.5, 0.5, 0.5)), (1, 9.0, (0.3, 0.3, 0.6)), (0, 16 * np.pi, (0.7, 0.4, 0.5)),
            (1, 30.0, (0.45, 0.6, 0.35)), (0, 6 * np.pi, (0.4, 0.6, 0.55))]
    out = sv.run(_schedule(reqs), keep_solutions=True)
    assert len(out) == len(reqs)
    for r in out:
        m, k, c = reqs[r["request"]]
        assert r["converged"], r
        x_ref, _ = _direct(_spec(meshes[m], k, c))
        err = np.linalg.norm(r["x"] - x_ref) / np.linalg.norm(x_ref)
        assert err <= SOL_RTOL, (err, r["request"], meshes[m], k, r["iterations"])


def test_mg_cuts_iterations():
    """the same requests with and without the preconditioner: at most a third
    of the row-scaled iteration count at k = 2 pi and 6 pi on a P2 mesh"""
    print("case mg iterations", flush=True)
    mesh = [(8, 2)]
    reqs = [(0, 2 * np.pi, (0.5, 0.5, 0.5)), (0, 6 * np.pi, (0.4, 0.5, 0.6))]
    its = {}
    for mg in (False, True):
        sv, _, _ = _server(mesh, mg=mg)
        its[mg] = {r["request"]: r["iterations"] for r in sv.run(_schedule(reqs), keep_solutions=False)}
    for i in range(len(reqs)):
        assert 3 * its[True][i] <= its[False][i], (i, its)


def test_mg_request_is_the_same_alone_and_in_a_mix():
    """a preconditioned request alone and among others (with and without a
    hierarchy of their own) arriving before and after it: bitwise equal
    solutions and iteration counts"""
    from gfem import Server, workloads as W
    print("case mg alone vs mix", flush=True)
    meshes = [(5, 1), (8, 2), (4, 2)]
    Fs = [_resident(m) for m in meshes]
    sv = Server(restart=25, hbm_bytes=1 << 31, max_requests=16)
    for F in Fs:
        sv.add_mesh(F)
    Fs[0].set_shifted_laplacian(W.HELMHOLTZ_MASS, 0.5)  # a smaller shift here
    Fs[1].set_shifted_laplacian(W.HELMHOLTZ_MASS, 0.5)  # mesh 2 stays row-scaled
    probe = (1, 20.0, (0.4, 0.55, 0.5))
    alone = sv.run(_schedule([probe]), keep_solutions=True)[0]
    mix = [(0, 10.0, (0.5, 0.5, 0.5)), (2, 7.0, (0.6, 0.3, 0.4)), probe, (0, 40.0, (0.3, 0.7, 0.5)),
           (1, 12.0, (0.5, 0.5, 0.2))]
    out = sv.run(_schedule(mix, gap=0.002), keep_solutions=True)
    inmix = [r for r in out if r["request"] == 2][0]
    assert inmix["iterations"] == alone["iterations"]
    assert np.array_equal(inmix["x"], alone["x"])
    for r in out:
        m, k, c = mix[r["request"]]
        x_ref, _ = _direct(_spec(meshes[m], k, c))
        assert np.linalg.norm(r["x"] - x_ref) / np.linalg.norm(x_ref) <= SOL_RTOL, r["request"]


def test_mg_graph_and_eager_batches_agree(monkeypatch):
    """the preconditioned iterations replayed as HIP graphs and launched one
    kernel at a time: bitwise equal solutions and iteration counts (restart
    23 runs both graphs)"""
    from gfem import Server
    print("case mg graph vs eager", flush=True)
    meshes = [(8, 2), (5, 1)]
    reqs = [(0, 9.0, (0.5, 0.5, 0.5)), (1, 14.0, (0.4, 0.6, 0.5)), (0, 25.0, (0.3, 0.5, 0.6))]
    _, Fs, _ = _server(meshes, restart=23)  # one set of systems and hierarchies for both runs
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GFEM_BATCH_GRAPH", mode)
        sv = Server(restart=23, hbm_bytes=1 << 31, max_requests=16)
        for F in Fs:
            sv.add_mesh(F)
        out[mode] = {r["request"]: r for r in sv.run(_schedule(reqs, gap=0.001), keep_solutions=True)}
    for i in range(len(reqs)):
        assert out["1"][i]["converged"]
        assert out["1"][i]["iterations"] == out["0"][i]["iterations"], i
        assert np.array_equal(out["1"][i]["x"], out["0"][i]["x"]), i


def test_mg_setup_keeps_the_parameters():
    """building the hierarchy assembles every term alone and restores the
    coefficients: the next assembly equals the oracle's for the same
    parameters"""
    from gfem import workloads as W
    from _helpers import assert_close
    print("case mg setup restores coefficients", flush=True)
    mesh = (4, 2)
    F = _resident(mesh)
    for tag, f in W.helmholtz_params(13.0, (0.4, 0.6, 0.5)):
        F.set_term_coefficient(tag, f)
    F.assemble()
    F.set_shifted_laplacian(W.HELMHOLTZ_MASS, 0.5)
    F.assemble()
    _, P = _direct(_spec(mesh, 13.0, (0.4, 0.6, 0.5)))
    assert_close(F.values(), P.values(), 1e-12, "A after mg setup")
    assert_close(F.rhs(0), P.rhs(0), 1e-12, "b after mg setup")


def test_mg_refuses_what_it_cannot_precondition():
    """real systems and non-bilinear or unknown terms are refused with an
    error, the system left usable"""
    from gfem import workloads as W
    print("case mg refusals", flush=True)
    F = W.build(W.poisson2d(n=6, order=1))
    F.pre()
    F.assemble()
    with pytest.raises(Exception):
        F.set_shifted_laplacian(0, 0.5)  # real system
    H = _resident((3, 1))
    H.assemble()
    with pytest.raises(Exception):
        H.set_shifted_laplacian(W.HELMHOLTZ_SOURCE, 0.5)  # a linear term
    with pytest.raises(Exception):
        H.set_shifted_laplacian(W.HELMHOLTZ_MASS, -1.0)  # beta must be positive
    H.assemble()


def test_mg_single_solve():
    """gfem_solve of a system with a hierarchy: within 1e-8 of a direct LU,
    at most a third of the row-scaled iterations, and bitwise the solution
    and iteration count of the same assembly served as a batched request"""
    from gfem import Server, workloads as W
    print("case mg single solve", flush=True)
    mesh, k, c = (8, 2), 6 * np.pi, (0.4, 0.5, 0.6)
    F = _resident(mesh)
    for tag, f in W.helmholtz_params(k, c):
        F.set_term_coefficient(tag, f)
    F.assemble(download=False)
    plain = F.solve(tol=1e-11, restart=30, max_iter=40000)
    F.set_shifted_laplacian(W.HELMHOLTZ_MASS)
    F.assemble(download=False)
    r = F.solve(tol=1e-11, restart=30, max_iter=40000)
    assert r["converged"], r
    x = F.solution()
    x_ref, _ = _direct(_spec(mesh, k, c))
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) <= SOL_RTOL
    assert 3 * r["iterations"] <= plain["iterations"], (r, plain)
    # a second solve reuses the system's batch: same result
    again = F.solve(tol=1e-11, restart=30, max_iter=40000)
    assert again["iterations"] == r["iterations"] and np.array_equal(F.solution(), x)
    sv = Server(restart=30, hbm_bytes=1 << 31, max_requests=4)
    sv.add_mesh(F)
    b = sv.run(_schedule([(0, k, c)]), keep_solutions=True)[0]
    assert b["iterations"] == r["iterations"]
    assert np.array_equal(b["x"], x)


def test_mg_setup_waits_for_batched_requests():
    """a new hierarchy is refused while a request of the system holds a batch
    slot (its cycle reads the hierarchy), and accepted once the result is
    fetched; the request itself converges"""
    import ctypes

    import gfem
    from gfem import workloads as W
    print("case mg set-up with a request in flight", flush=True)
    lib = gfem.load_c_abi()
    F = _resident((4, 2))
    for tag, f in W.helmholtz_params(9.0, (0.5, 0.5, 0.5)):
        F.set_term_coefficient(tag, f)
    F.assemble(download=False)
    F.set_shifted_laplacian(W.HELMHOLTZ_MASS)
    ctx = ctypes.c_void_p(gfem.device_context_handle())
    sysh = ctypes.c_void_p(F.system_handle())
    B = ctypes.c_void_p()
    assert lib.gfem_batch_create(ctx, 30, ctypes.c_size_t(1 << 30), 2, ctypes.byref(B)) == 0
    try:
        rid = ctypes.c_int64()
        assert lib.gfem_batch_submit(B, sysh, ctypes.c_double(1e-10), 10000, None, ctypes.byref(rid)) == 0
        with pytest.raises(Exception):
            F.set_shifted_laplacian(W.HELMHOLTZ_MASS)
        live = ctypes.c_int(1)
        steps = 0
        while live.value and steps < 1000:
            assert lib.gfem_batch_step(B, None, ctypes.byref(live)) == 0
            steps += 1
        assert live.value == 0

        class Info(ctypes.Structure):
            _fields_ = [("iterations", ctypes.c_int), ("cycles", ctypes.c_int), ("rel_residual", ctypes.c_double),
                        ("converged", ctypes.c_int)]
        info = Info()
        assert lib.gfem_batch_result(B, rid, None, ctypes.c_size_t(0), ctypes.byref(info)) == 0
        assert info.converged == 1 and info.rel_residual <= 1e-10
        F.set_shifted_laplacian(W.HELMHOLTZ_MASS)  # nothing in flight now
    finally:
        lib.gfem_batch_destroy(B)
