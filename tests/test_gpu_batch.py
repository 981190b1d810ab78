"""Parameter updates, batched solves and the serving loop.

Every solution is checked against a direct sparse LU of the oracle's system
for the same parameters (the reference solves with MUMPS LU), to 1e-8
relative; a request's result must be bitwise the same served alone and in a
mix of other requests."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from _helpers import assert_close

pytestmark = pytest.mark.gpu

SOL_RTOL = 1e-8


def _spec(mesh, k, center):
    from gfem import workloads as W
    n, order = mesh
    return W.helmholtz3d(n=n, order=order, k=k, center=center)


def _direct(spec):
    from oracle import oracle as O
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    ai, aj = P.csr()
    A = sp.csr_matrix((P.values(), aj, ai), shape=(P.n_unknown, P.n_unknown)).tocsc()
    return sla.spsolve(A, -P.rhs(0)), P


def _resident(mesh):
    """a formulation of the mesh, pre-processed, with placeholder parameters"""
    from gfem import workloads as W
    F = W.build(_spec(mesh, 1.0, (0.5, 0.5, 0.5)))
    F.pre()
    return F


def test_term_update_reassembles_with_new_parameters():
    """set_term_coefficient on an uploaded system: the next assembly equals
    the oracle's assembly with those parameters (k and source position)."""
    from gfem import workloads as W
    print("case term update", flush=True)
    mesh = (4, 2)
    F = _resident(mesh)
    F.assemble()
    for k, c in [(7.0, (0.3, 0.6, 0.4)), (31.0, (0.55, 0.45, 0.7))]:
        for tag, f in W.helmholtz_params(k, c):
            F.set_term_coefficient(tag, f)
        F.assemble()
        _, P = _direct(_spec(mesh, k, c))
        assert_close(F.values(), P.values(), 1e-12, f"A k={k}")
        assert_close(F.rhs(0), P.rhs(0), 1e-12, f"b k={k}")


def _schedule(reqs, gap=0.0):
    from gfem import workloads as W
    return [dict(mesh=m, arrival=i * gap, params=W.helmholtz_params(k, c), tol=1e-11, max_iter=40000)
            for i, (m, k, c) in enumerate(reqs)]


def test_batched_requests_match_direct_solves():
    """requests on two resident meshes with different wavenumbers and
    sources, all admitted together"""
    from gfem import Server
    print("case batch vs direct", flush=True)
    meshes = [(5, 1), (4, 2)]
    Fs = [_resident(m) for m in meshes]
    sv = Server(restart=30, hbm_bytes=1 << 30, max_requests=16)
    for F in Fs:
        sv.add_mesh(F)
    reqs = [(0, 2 * np.pi, (0.5, 0.5, 0.5)), (1, 9.0, (0.3, 0.3, 0.6)), (0, 16 * np.pi, (0.7, 0.4, 0.5)),
            (1, 30.0, (0.45, 0.6, 0.35))]
    out = sv.run(_schedule(reqs), keep_solutions=True)
    assert len(out) == len(reqs)
    for r in out:
        m, k, c = reqs[r["request"]]
        assert r["converged"], r
        x_ref, _ = _direct(_spec(meshes[m], k, c))
        err = np.linalg.norm(r["x"] - x_ref) / np.linalg.norm(x_ref)
        assert err <= SOL_RTOL, (err, r["request"], r["iterations"])
    assert sv.stats["max_live"] >= 2


def test_request_is_the_same_alone_and_in_a_mix():
    """the same request served alone and among others arriving before and
    after it: bitwise equal solutions and iteration counts"""
    from gfem import Server
    print("case alone vs mix", flush=True)
    meshes = [(5, 1), (4, 2)]
    Fs = [_resident(m) for m in meshes]
    sv = Server(restart=25, hbm_bytes=1 << 30, max_requests=16)
    for F in Fs:
        sv.add_mesh(F)
    probe = (1, 20.0, (0.4, 0.55, 0.5))
    alone = sv.run(_schedule([probe]), keep_solutions=True)[0]
    mix = [(0, 10.0, (0.5, 0.5, 0.5)), (1, 7.0, (0.6, 0.3, 0.4)), probe, (0, 40.0, (0.3, 0.7, 0.5)),
           (1, 12.0, (0.5, 0.5, 0.2))]
    out = sv.run(_schedule(mix, gap=0.002), keep_solutions=True)
    inmix = [r for r in out if r["request"] == 2][0]
    assert inmix["iterations"] == alone["iterations"]
    assert np.array_equal(inmix["x"], alone["x"])
    x_ref, _ = _direct(_spec(meshes[probe[0]], probe[1], probe[2]))
    assert np.linalg.norm(alone["x"] - x_ref) / np.linalg.norm(x_ref) <= SOL_RTOL


def test_admission_respects_the_hbm_budget():
    """a budget that holds two requests: five arrivals at once are admitted
    two at a time and all complete"""
    from gfem import Server
    print("case admission by budget", flush=True)
    mesh = (4, 2)
    F = _resident(mesh)
    probe = Server(restart=20, hbm_bytes=1 << 26, max_requests=8)
    probe.add_mesh(F)
    per = probe.request_bytes(0)
    sv = Server(restart=20, hbm_bytes=2 * per + 4096, max_requests=8)
    sv.add_mesh(F)
    reqs = [(0, 5.0 + i, (0.5, 0.5, 0.5)) for i in range(5)]
    out = sv.run(_schedule(reqs), keep_solutions=False)
    assert sorted(r["request"] for r in out) == list(range(5))
    assert all(r["converged"] for r in out)
    assert sv.stats["peak_bytes"] <= 2 * per + 4096
    assert sv.stats["max_live"] <= 2
    # later arrivals waited for memory: admitted after earlier ones ended
    first_done = min(r["done"] for r in out)
    assert max(r["admitted"] for r in out) >= first_done


def test_serving_reports_progress_and_queue(capfd):
    """a run with progress lines on: stderr names arrivals, admissions and
    completions; five arrivals at once on a budget of two requests leave
    five waiting before the first admission"""
    from gfem import Server
    print("case serving progress and queue", flush=True)
    mesh = (4, 2)
    F = _resident(mesh)
    probe = Server(restart=20, hbm_bytes=1 << 26, max_requests=8)
    probe.add_mesh(F)
    per = probe.request_bytes(0)
    sv = Server(restart=20, hbm_bytes=2 * per + 4096, max_requests=8)
    sv.add_mesh(F)
    sv.set_progress(1e-6)  # a line at every pass of the loop
    out = sv.run(_schedule([(0, 5.0 + i, (0.5, 0.5, 0.5)) for i in range(5)]), keep_solutions=False)
    assert len(out) == 5 and all(r["converged"] for r in out)
    assert sv.stats["max_waiting"] == 5
    lines = [ln for ln in capfd.readouterr().err.splitlines() if ln.startswith("serve:")]
    assert lines and "arrived 5" in lines[-1]
    assert any("waiting 3" in ln for ln in lines)  # two admitted, three held back by the budget


def test_graph_and_eager_batches_agree(monkeypatch):
    """The batched Krylov iterations replayed as HIP graphs (default) and
    launched one kernel at a time (GFEM_BATCH_GRAPH=0) give bitwise equal
    solutions and iteration counts.  A restart length of 23 runs both graphs
    (5 iterations between convergence checks, then single iterations); each is
    replayed many times with the host's checks and flag uploads between."""
    from gfem import Server
    print("case graph vs eager batch", flush=True)
    meshes = [(5, 1), (4, 2)]
    Fs = [_resident(m) for m in meshes]
    reqs = [(0, 9.0, (0.5, 0.5, 0.5)), (1, 14.0, (0.4, 0.6, 0.5)), (0, 25.0, (0.3, 0.5, 0.6))]
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("GFEM_BATCH_GRAPH", mode)
        sv = Server(restart=23, hbm_bytes=1 << 30, max_requests=8)
        for F in Fs:
            sv.add_mesh(F)
        out[mode] = {r["request"]: r for r in sv.run(_schedule(reqs, gap=0.001), keep_solutions=True)}
    for i in range(len(reqs)):
        assert out["1"][i]["converged"]
        assert out["1"][i]["iterations"] == out["0"][i]["iterations"], i
        assert np.array_equal(out["1"][i]["x"], out["0"][i]["x"]), i


def test_uneven_instances_orders_1_2_3():
    """Uneven instances in one batch: P1, P2 and P3 meshes of different sizes,
    k from 2 pi to 16 pi, GMRES(50) with the FP32 Krylov iteration and FP64
    refinement, all admitted at once: every solution within 1e-8 of a direct
    LU of the oracle's system for its parameters."""
    from gfem import Server
    print("case uneven instances P1/P2/P3", flush=True)
    meshes = [(6, 1), (3, 2), (2, 3), (3, 3)]
    Fs = [_resident(m) for m in meshes]
    sv = Server(restart=50, hbm_bytes=1 << 30, max_requests=16)
    for F in Fs:
        sv.add_mesh(F)
    reqs = [(0, 2 * np.pi, (0.5, 0.5, 0.5)), (1, 11.0, (0.35, 0.6, 0.45)), (2, 16 * np.pi, (0.6, 0.4, 0.5)),
            (3, 7.5, (0.4, 0.4, 0.65)), (2, 3 * np.pi, (0.5, 0.3, 0.5)), (3, 13 * np.pi, (0.55, 0.5, 0.35))]
    out = sv.run(_schedule(reqs), keep_solutions=True)
    assert len(out) == len(reqs)
    assert sv.stats["max_live"] == len(reqs)
    for r in out:
        m, k, c = reqs[r["request"]]
        assert r["converged"], r
        x_ref, _ = _direct(_spec(meshes[m], k, c))
        err = np.linalg.norm(r["x"] - x_ref) / np.linalg.norm(x_ref)
        assert err <= SOL_RTOL, (err, r["request"], meshes[m], k, r["iterations"])
