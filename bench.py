"""Benchmark: GDOF/s of Formulation::assemble() on one MI355X.

Workload (default): 3-D scalar Helmholtz, unit cube, 50^3 cubes x 6 jittered
P2 tetrahedra (750,000 tets, 1,030,301 DOFs) + 30,000 P2 Robin boundary
triangles, k = 10 pi, Gaussian source (sigma 0.05), complex128.  One step =
one full assembly (geometry, source evaluation, CSR matrix values and RHS)
with the mesh, pattern and maps resident in HBM.

value = element-DOFs assembled per second (sum over elements of DOFs per
element / assembly time) of the FP64 fused path; the mixed-precision path and
the plain reference-order GPU assembler are reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "fem-g_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gather_traffic.json")
GATHER_KERNEL = "k_gather<true, false, false, {}>"  # complex, FP64, 16-bit words (P2), slots per thread
GATHER_WAVE_KERNEL = "k_gather_wave<true, false, false>"  # the one-wave persistent form of 4 slots


def gather_kernel_name(args):
    return GATHER_WAVE_KERNEL if args.gather_slots == 4 and args.gather_wave else GATHER_KERNEL.format(args.gather_slots)


def measured_traffic(kernel):
    """HBM bytes per launch of the gather from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by a copy of known size), when it
    was taken on this kernel and workload; None otherwise."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    return t.get("traffic_bytes_per_launch") if t.get("kernel") == kernel else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_steps(F, steps, warmup, precision, strategy):
    for _ in range(warmup):
        F.assemble_async(precision, strategy)
    F.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        F.assemble_async(precision, strategy)
    F.sync()
    return (time.perf_counter() - t0) / steps


def kernel_profile(F, precision, strategy, reps=5):
    """average duration of each kernel over `reps` assemblies (HIP events)."""
    F.set_profiling(True)
    acc = {}
    for _ in range(reps):
        F.assemble_async(precision, strategy)
        F.sync()
        for name, ms in F.kernel_times():
            acc.setdefault(name, []).append(ms)
    F.set_profiling(False)
    return {k: sum(v) / len(v) * (len(v) / reps) / (len(v) / reps) for k, v in acc.items()}, acc


def copy_calibration(nbytes):
    """Device-to-device copy of nbytes (the runtime's copy kernel), used to
    scale the FETCH_SIZE / WRITE_SIZE counters: a plain copy moves exactly
    nbytes in each direction."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: 3 x {nbytes} bytes device-to-device")


def cpu_threads():
    """host cores this process may use (the OpenMP share on a shared box)"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() else n


def cpu_baseline(args):
    """The oracle (C restatement of the reference CPU assembler: per-Gauss-point
    dense A_e, binary-search scatter) on a bounded sample of the same workload:
    a smaller cube, same order, same terms.  Timed twice: with the element loop
    over all host cores (atomic scatter, as the reference's OpenMP loop) and
    with one thread."""
    from gfem import workloads as W
    from oracle import oracle as O
    n = args.cpu_n
    spec = W.helmholtz3d(n=n, order=args.order)
    P = O.from_spec(spec)
    P.pre()
    edofs = spec.element_dofs()
    res = {}
    for th in (cpu_threads(), 1):
        P.set_threads(th)
        t0 = time.perf_counter()
        P.assemble()
        res[th] = time.perf_counter() - t0
    nt = max(res)
    return {"value": edofs / res[nt] / 1e9, "unit": "GDOF/s", "cores": nt, "kind": "port",
            "single_thread_value": edofs / res[1] / 1e9,
            "sample": f"oracle assemble() of the same Helmholtz P{args.order} workload on a {n}^3x6-tet cube "
                      f"({edofs} element-DOFs): {res[nt]:.2f} s on {nt} threads, {res[1]:.2f} s on 1 thread"}


# resident meshes of the serving stream: unit cubes, (cells per side, order),
# 10k - 500k DOFs, orders 1 - 3
SERVE_MESHES = [(21, 1), (30, 1), (40, 1), (15, 2), (25, 2), (39, 2), (10, 3), (20, 3)]


def serving(args):
    """Helmholtz request stream: Poisson arrivals at `rate`/s, each a solve on
    one of the resident meshes with k ~ U[2pi, 16pi] and a random source
    position, batched on the GPU as requests arrive (continuous batching,
    admission by HBM budget).  Latency = arrival to solution on the host."""
    import numpy as np
    from gfem import Server, workloads as W
    t0 = time.perf_counter()
    Fs = []
    for n, order in SERVE_MESHES:
        F = W.build(W.helmholtz3d(n=n, order=order, k=10.0))
        F.pre()
        Fs.append(F)
    sv = Server(restart=args.serve_restart, hbm_bytes=args.serve_hbm_gb << 30, max_requests=args.serve_slots)
    for F in Fs:
        sv.add_mesh(F)
    log(f"serving: {len(Fs)} meshes resident ({time.perf_counter() - t0:.1f} s)")
    mg_levels, t_mg = [], 0.0
    if args.serve_mg:
        # shifted-Laplacian multigrid on every resident mesh (built once per mesh)
        t1 = time.perf_counter()
        mg_levels = [F.set_shifted_laplacian(W.HELMHOLTZ_MASS, args.serve_mg_beta) for F in Fs]
        t_mg = time.perf_counter() - t1
        log(f"serving: multigrid hierarchies in {t_mg:.1f} s, levels {[len(l) for l in mg_levels]}")
    rng = np.random.default_rng(2024)
    if args.serve_requests is None:
        # the whole stated stream: every arrival of 60 s at the rate
        arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, int(args.serve_rate * 66) + 100))
        arrivals = arrivals[arrivals < 60.0]
        sv.set_progress(20.0)
    else:
        arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, args.serve_requests))
    sched = []
    for i in range(len(arrivals)):
        m = int(rng.integers(0, len(Fs)))
        k = float(rng.uniform(2 * np.pi, 16 * np.pi))
        c = tuple(float(v) for v in rng.uniform(0.25, 0.75, 3))
        sched.append(dict(mesh=m, arrival=float(arrivals[i]), params=W.helmholtz_params(k, c), tol=args.serve_tol,
                          max_iter=30000))
    out = sv.run(sched)
    lat = np.array([r["latency"] for r in out])
    its = np.array([r["iterations"] for r in out])
    st = sv.stats
    res = {
        "requests": len(out), "arrival_rate_per_s": args.serve_rate, "tol": args.serve_tol,
        "restart": args.serve_restart, "slots": args.serve_slots, "hbm_budget_GB": args.serve_hbm_gb,
        "mesh_dofs": [int(F.n_unknown) for F in Fs],
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "solves_per_s": len(out) / st["seconds"], "seconds": st["seconds"],
        "mean_iterations": float(its.mean()), "converged": int(sum(r["converged"] for r in out)),
        "max_rel_residual": float(max(r["rel_residual"] for r in out)),
        "max_live": st["max_live"], "peak_hbm_GB": st["peak_bytes"] / 2**30,
        "preconditioner": (f"shifted-Laplacian multigrid V(1,1), beta {args.serve_mg_beta:g}, one V-cycle per "
                           "Krylov step" if args.serve_mg else "row-norm scaling"),
    }
    if args.serve_mg:
        res["multigrid"] = {"setup_s": t_mg, "rows_per_level": mg_levels}
    # the stated stream is Poisson arrivals at 200/s for 60 s (about 12,000
    # requests), served whole by default (or its first serve_requests
    # arrivals, a reduced configuration); report the steady-state rate and
    # the queue it built
    arr = np.array([r["arrival"] for r in out])
    adm = np.array([r["admitted"] for r in out])
    done = np.array([r["arrival"] + r["latency"] for r in out])
    t_lo, t_hi = np.percentile(done, 10), np.percentile(done, 90)
    inside = (done >= t_lo) & (done <= t_hi)
    events = np.sort(np.concatenate([arr, adm, done]))
    # arrived by t minus admitted (completed) by t: admission and completion follow arrival
    n_by = lambda a: np.searchsorted(np.sort(a), events, side="right")
    waiting = (n_by(arr) - n_by(adm)).tolist()
    in_system = (n_by(arr) - n_by(done)).tolist()
    stated = args.serve_rate * 60.0
    res.update({
        "config": "stated" if args.serve_requests is None else "reduced",
        "stated_config": {"arrival_rate_per_s": args.serve_rate, "seconds": 60.0, "requests": int(stated),
                          "measured": args.serve_requests is None,
                          "projected_seconds_at_sustained_rate": None},
        "sustained_solves_per_s": float(inside.sum() / max(t_hi - t_lo, 1e-9)),
        "sustained_window_s": [float(t_lo), float(t_hi)],
        "queue_depth_max": max(waiting), "queue_depth_mean": float(np.mean(waiting)),
        "in_system_max": max(in_system),
        "what": (f"{'all' if args.serve_requests is None else 'first'} {len(out)} requests of the stated stream (Poisson "
                 f"arrivals at {args.serve_rate:g}/s for 60 s); sustained rate = completions between the 10th and "
                 "90th percentile completion times"),
    })
    res["stated_config"]["projected_seconds_at_sustained_rate"] = stated / res["sustained_solves_per_s"]
    log(f"serving: {res['requests']} requests p50 {res['p50_latency_s']:.2f} s p99 {res['p99_latency_s']:.2f} s "
        f"{res['solves_per_s']:.2f} solves/s (sustained {res['sustained_solves_per_s']:.2f}), mean "
        f"{res['mean_iterations']:.0f} iterations, queue depth max {res['queue_depth_max']}")
    # the same requests one at a time (same solver, no batching), on a sample
    nb = min(args.serve_baseline, len(sched))
    t1 = time.perf_counter()
    for s in sched[:nb]:
        F = Fs[s["mesh"]]
        for tag, f in s["params"]:
            F.set_term_coefficient(tag, f)
        F.assemble(download=False)
        F.solve(tol=args.serve_tol, restart=args.serve_restart, max_iter=30000, download=False)
    dt = time.perf_counter() - t1
    res["one_at_a_time_solves_per_s"] = nb / dt
    res["one_at_a_time_sample"] = nb
    log(f"one at a time: {nb} requests in {dt:.2f} s")
    # solution error against a direct sparse LU of the FP64 system (the
    # reference solves with MUMPS LU): requests on the smallest mesh, served
    # through the same batch loop, to the stream's tolerance
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    small = min(range(len(Fs)), key=lambda i: Fs[i].n_unknown)
    F = Fs[small]
    chk = Server(restart=args.serve_restart, hbm_bytes=1 << 30, max_requests=4)
    chk.add_mesh(F)
    errs = []
    for s in [s for s in sched if s["mesh"] == small][:3]:
        r = chk.run([dict(s, mesh=0, arrival=0.0)], keep_solutions=True)[0]
        for tag, f in s["params"]:
            F.set_term_coefficient(tag, f)
        F.assemble()
        A = sp.csr_matrix((F.values(), F.aj(), F.ai()), shape=(F.n_unknown, F.n_unknown)).tocsc()
        xd = sla.spsolve(A, -F.rhs(0))
        errs.append(float(np.linalg.norm(r["x"] - xd) / np.linalg.norm(xd)))
    res["solution_check"] = {"mesh_dofs": int(F.n_unknown), "requests": len(errs),
                             "max_rel_err_vs_direct_lu": max(errs) if errs else None}
    log(f"solution check: {len(errs)} requests on {F.n_unknown} DOFs, max rel err vs direct LU "
        f"{max(errs) if errs else float('nan'):.2e}")
    return res


def poisson2d_leg(args):
    """The 2-D Poisson workload: -Laplace(u) = 1 on the unit square, 64x64
    cells x 2 jittered P1 triangles (8192 elements), Gauss2, Dirichlet u = 0,
    FP64; a correctness-size problem whose assembly is launch-bound."""
    from gfem import workloads as W
    spec = W.poisson2d(n=64, order=1)
    F = W.build(spec)
    F.pre()
    F.assemble(download=False)
    edofs = spec.element_dofs()
    steps = max(args.steps, 50)
    sec = timed_steps(F, steps, args.warmup, "fp64", "fused")
    prof, _ = kernel_profile(F, "fp64", "fused")
    log(f"poisson2d 64x64 P1 fp64: {sec * 1e3:.4f} ms/step  {edofs / sec / 1e9:.3f} GDOF/s")
    return {"workload": "poisson2d unit square 64x64x2 jittered P1 triangles, Gauss2, Dirichlet 0, real64",
            "elements": edofs // 3, "element_dofs": edofs, "n_unknown": int(F.n_unknown), "nnz": int(F.nnz),
            "value": edofs / sec / 1e9, "unit": "GDOF/s", "ms_per_step": sec * 1e3, "steps": steps,
            "kernel_ms": {k: round(v, 5) for k, v in prof.items()}}


def solver_leg(args):
    """One Helmholtz request on the largest serving mesh (39^3 x 6 P2 tets,
    493,039 unknowns, k = 10 pi) with the single-request solver: GMRES(30)
    in FP32 inside FP64 iterative refinement, to 1e-10.  SpMV bytes per
    product: nnz (complex64 value + 32-bit column) + row pointers (64-bit) +
    x read and y written (complex64)."""
    from gfem import workloads as W
    F = W.build(W.helmholtz3d(n=39, order=2, k=10 * 3.141592653589793))
    F.pre()
    F.assemble(download=False)
    F.solve(tol=1e-10, restart=30, max_iter=30000, download=False)  # workspace, graphs, SpMV lanes
    t0 = time.perf_counter()
    info = F.solve(tol=1e-10, restart=30, max_iter=30000, download=False)
    dt = time.perf_counter() - t0
    n, nnz = int(F.n_unknown), int(F.nnz)
    res = {"mesh": "helmholtz3d 39^3x6 P2, k=10pi", "n_unknown": n, "nnz": nnz, "iterations": info["iterations"],
           "cycles": info["cycles"], "seconds": dt, "ms_per_iteration": dt / max(info["iterations"], 1) * 1e3,
           "rel_residual": info["rel_residual"], "spmv_bytes": nnz * 12 + (n + 1) * 8 + 2 * n * 8}
    log(f"solver: {n} unknowns, {info['iterations']} iterations in {dt:.2f} s "
        f"({res['ms_per_iteration']:.3f} ms/iteration), rel residual {info['rel_residual']:.1e}")
    if args.solver_mg:
        # the same solve preconditioned by one shifted-Laplacian V-cycle per step
        t0 = time.perf_counter()
        levels = F.set_shifted_laplacian(W.HELMHOLTZ_MASS, args.serve_mg_beta)
        t_setup = time.perf_counter() - t0
        F.solve(tol=1e-10, restart=30, max_iter=30000, download=False)  # the system's one-request batch
        t0 = time.perf_counter()
        mi = F.solve(tol=1e-10, restart=30, max_iter=30000, download=False)
        dm = time.perf_counter() - t0
        res["multigrid"] = {"setup_s": t_setup, "rows_per_level": levels, "iterations": mi["iterations"],
                            "seconds": dm, "ms_per_iteration": dm / max(mi["iterations"], 1) * 1e3,
                            "rel_residual": mi["rel_residual"], "converged": bool(mi["converged"])}
        log(f"solver with multigrid: {mi['iterations']} iterations in {dm:.3f} s, rel residual "
            f"{mi['rel_residual']:.1e} (set-up {t_setup:.2f} s)")
    return res


def batched(args):
    """Uneven Helmholtz instances solved in one batch: tet meshes log-uniform
    in 2k-200k elements (n^3 x 6 cubes), order from {1, 2, 3}, k ~ U[2pi,
    16pi], random Gaussian source centre; all requests arrive at once.
    Requests on the same (n, order) share one resident mesh, pattern and maps.
    Assembly of every instance in mixed precision (complex64 values) against
    the plain FP64 GPU assembler; then GMRES(50) with the FP32 Krylov
    iteration and FP64 residuals / refinement to 1e-8 relative residual."""
    import math
    import numpy as np
    from gfem import Server, workloads as W
    rng = np.random.default_rng(7)
    inst = []
    for _ in range(args.batch_instances):
        ne = math.exp(rng.uniform(math.log(2000), math.log(200000)))
        n = max(2, int(round((ne / 6) ** (1 / 3))))
        order = int(rng.integers(1, 4))
        k = float(rng.uniform(2 * math.pi, 16 * math.pi))
        c = tuple(float(v) for v in rng.uniform(0.25, 0.75, 3))
        inst.append(((n, order), k, c))
    t0 = time.perf_counter()
    keys = sorted({m for m, _, _ in inst})
    Fs, edofs = {}, {}
    for m in keys:
        spec = W.helmholtz3d(n=m[0], order=m[1], k=10.0)
        F = W.build(spec)
        F.pre()
        F.assemble(download=False)  # resident system
        Fs[m], edofs[m] = F, spec.element_dofs()
    log(f"batched: {len(inst)} instances on {len(keys)} resident meshes ({time.perf_counter() - t0:.1f} s)")
    total_edofs = sum(edofs[m] for m, _, _ in inst)
    res = {"instances": len(inst), "resident_meshes": len(keys), "element_dofs": int(total_edofs),
           "unknowns": int(sum(Fs[m].n_unknown for m, _, _ in inst)),
           "nnz": int(sum(Fs[m].nnz for m, _, _ in inst))}
    # every instance assembled once with its own parameters: all together
    # (gfem_assemble_many: one launch per kernel kind per wave of distinct
    # meshes, planning and parameter upload included in the time) ...
    import gfem
    any_f = Fs[keys[0]]
    for prec in ("mixed", "fp64"):
        insts = [Fs[m].many_instance(W.helmholtz_params(k, c)) for m, k, c in inst]
        gfem.assemble_many(insts, prec)  # outputs allocated, geometry of each mesh in place
        reps = 3
        t1 = time.perf_counter()
        for _ in range(reps):
            gfem.assemble_many(insts, prec, sync=False)
        any_f.sync()
        dt = (time.perf_counter() - t1) / reps
        res[f"assembly_{prec}"] = {"seconds": dt, "GDOF_s": total_edofs / dt / 1e9,
                                   "what": "gfem_assemble_many over all instances (waves of distinct meshes)"}
        log(f"batched assembly_{prec} (assemble_many): {dt * 1e3:.1f} ms, {total_edofs / dt / 1e9:.2f} GDOF/s")
        del insts
    # ... and one instance at a time, back to back
    for name, prec, strat in (("assembly_mixed_one_at_a_time", "mixed", "fused"),
                              ("assembly_plain_fp64_gpu", "fp64", "reference")):
        any_f.sync()
        t1 = time.perf_counter()
        for m, k, c in inst:
            F = Fs[m]
            for tag, f in W.helmholtz_params(k, c):
                F.set_term_coefficient(tag, f)
            F.assemble_async(prec, strat)
        any_f.sync()
        dt = time.perf_counter() - t1
        res[name] = {"seconds": dt, "GDOF_s": total_edofs / dt / 1e9}
        log(f"batched {name}: {dt * 1e3:.1f} ms, {total_edofs / dt / 1e9:.2f} GDOF/s")
    res["mixed_vs_plain_fp64_gpu"] = res["assembly_plain_fp64_gpu"]["seconds"] / res["assembly_mixed"]["seconds"]
    # the batch solve: every request assembled at admission, GMRES(50)
    mesh_ids = {m: i for i, m in enumerate(keys)}
    sv = Server(restart=50, hbm_bytes=args.batch_hbm_gb << 30, max_requests=256)
    for m in keys:
        sv.add_mesh(Fs[m])
    if args.batch_mg:
        t1 = time.perf_counter()
        for m in keys:
            Fs[m].set_shifted_laplacian(W.HELMHOLTZ_MASS, args.serve_mg_beta)
        res["multigrid_setup_s"] = time.perf_counter() - t1
        log(f"batched: multigrid hierarchies of {len(keys)} meshes in {res['multigrid_setup_s']:.1f} s")
    res["preconditioner"] = ("shifted-Laplacian multigrid V(1,1), one V-cycle per Krylov step" if args.batch_mg
                             else "row-norm scaling")
    sched = [dict(mesh=mesh_ids[m], arrival=0.0, params=W.helmholtz_params(k, c), tol=args.batch_tol,
                  max_iter=60000) for m, k, c in inst]
    out = sv.run(sched)
    st = sv.stats
    its = np.array([r["iterations"] for r in out])
    lat = np.array([r["latency"] for r in out])
    res.update({
        "tol": args.batch_tol, "restart": 50, "solve_seconds": st["seconds"],
        "solves_per_s": len(out) / st["seconds"], "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)), "mean_iterations": float(its.mean()),
        "max_iterations": int(its.max()), "converged": int(sum(r["converged"] for r in out)),
        "max_rel_residual": float(max(r["rel_residual"] for r in out)), "max_live": st["max_live"],
        "peak_hbm_GB": st["peak_bytes"] / 2**30,
        "what": "assembly + GMRES(50) of every instance, all admitted at once; element DOFs per second of the "
                "whole batch in assembly_*",
    })
    log(f"batched solve: {len(out)} in {st['seconds']:.2f} s ({res['solves_per_s']:.1f} solves/s), "
        f"mean {its.mean():.0f} / max {its.max()} iterations, {res['converged']} converged")
    return res


def highorder(args):
    """Order-4 tetrahedra (32^3 x 6 = 196,608 tets, N = 35), mass + stiffness,
    Gauss8: element matrices as per-point B^T W B on the matrix cores (the
    hand-written MFMA kernel in FP64, FP32 and split-bf16, Jacobian inverses
    in FP64) against the tensor-form FP64 path; per-entry error of the FP32
    and bf16x3 matrix-core paths against FP64."""
    import numpy as np
    from gfem import gauss_rule, workloads as W
    t0 = time.perf_counter()
    spec = W.highorder3d(n=args.ho_n, order=4)
    F = W.build(spec)
    F.pre()
    F.assemble(download=False)  # creates the system
    log(f"highorder: {spec.name} n={F.n_unknown} nnz={F.nnz} set-up {time.perf_counter() - t0:.1f} s")
    edofs = spec.element_dofs()
    out = {"workload": f"unit cube {args.ho_n}^3x6 P4 tets, mass + stiffness, Gauss8", "elements": edofs // 35,
           "element_dofs": edofs, "n_unknown": int(F.n_unknown), "nnz": int(F.nnz)}
    ref = None
    for name, prec, strat in (("tensor_fp64", "fp64", "fused"), ("gemm_fp64", "fp64", "dense"),
                              ("gemm_fp32", "mixed", "dense"), ("gemm_bf16x3", "mixed", "dense")):
        split = name == "gemm_bf16x3"
        F.set_gemm_mode("bf16x3" if split else "fp32")
        sec = timed_steps(F, max(2, args.steps // 4), 1, prec, strat)
        prof, _ = kernel_profile(F, prec, strat, reps=3)
        rec = {"ms_per_step": sec * 1e3, "GDOF_s": edofs / sec / 1e9,
               "kernel_ms": {k: round(v, 4) for k, v in prof.items()}}
        if strat == "dense":
            K = 7 * len(gauss_rule(3, 8)[1])  # points of the Gauss8 tetrahedron rule
            flops = 2.0 * 35 * 35 * K * (edofs // 35)
            gemm_ms = prof.get("element_gemm", float("nan"))
            peak = 78.6 if prec == "fp64" else (2500.0 if split else 157.3)
            rec["gemm_tflops"] = flops / (gemm_ms * 1e-3) / 1e12
            # the split form runs 3 bf16 products per FP32 one
            rec["gemm_mfma_tflops"] = rec["gemm_tflops"] * (3 if split else 1)
            rec["gemm_peak_tflops"] = peak
            rec["gemm_frac"] = rec["gemm_mfma_tflops"] / peak
            rec["gemm_shape"] = (f"M=elements={edofs // 35}, N=N*N=1225, K=7Q={K} (component-major, each component's "
                                 f"points padded to a multiple of 32 or 64; padding not counted)" +
                                 (" x3 bf16 products" if split else ""))
            rec["gemm_kernel"] = "k_elem_mm (hand-written MFMA, csrc/elemmm.hip)"
        F.assemble(precision=prec, strategy=strat)
        vals = F.values()
        if ref is None:
            ref = vals
        else:
            err = np.abs(vals - ref)
            floor = 1e-2 * np.abs(ref).max()
            rec["max_rel_err_vs_tensor_fp64"] = float((err / (np.abs(ref) + floor)).max())
        out[name] = rec
        log(f"highorder {name}: {sec * 1e3:.2f} ms/step {edofs / sec / 1e9:.2f} GDOF/s "
            f"{'' if strat != 'dense' else '%.1f TFLOP/s GEMM' % rec['gemm_tflops']}")
    return out


def arg_parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=50)
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--cpu-n", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the mixed / reference-strategy lines")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-setup-host", action="store_true", help="skip the host set-up (pre + maps) timing")
    ap.add_argument("--no-highorder", action="store_true", help="skip the order-4 matrix-core comparison")
    ap.add_argument("--ho-n", type=int, default=32)
    ap.add_argument("--no-batched", action="store_true", help="skip the uneven-instance batch")
    ap.add_argument("--no-solver", action="store_true", help="skip the single-request solve")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-tol", type=float, default=1e-8)
    ap.add_argument("--batch-hbm-gb", type=int, default=96)
    ap.add_argument("--serve-requests", type=int, default=None,
                    help="serve only the first N arrivals (a reduced stream); default: the whole stated stream, "
                         "60 s of arrivals")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-tol", type=float, default=1e-10)
    ap.add_argument("--serve-hbm-gb", type=int, default=64)
    ap.add_argument("--serve-restart", type=int, default=30, help="GMRES restart length of the serving stream")
    ap.add_argument("--serve-slots", type=int, default=256, help="requests iterating together, at most")
    ap.add_argument("--serve-baseline", type=int, default=16, help="requests timed one at a time")
    ap.add_argument("--serve-mg", type=int, choices=[0, 1], default=1,
                    help="shifted-Laplacian multigrid preconditioner on the serving meshes (1) or row scaling (0)")
    ap.add_argument("--serve-mg-beta", type=float, default=1.0)
    ap.add_argument("--solver-mg", type=int, choices=[0, 1], default=1,
                    help="also time the single solve preconditioned by the shifted-Laplacian multigrid")
    ap.add_argument("--batch-mg", type=int, choices=[0, 1], default=1,
                    help="shifted-Laplacian multigrid preconditioner on the uneven-instance batch")
    ap.add_argument("--gather-slots", type=int, choices=[2, 4, 8], default=4,
                    help="contributions per gather thread (4: 128-thread blocks, 2: 256-thread blocks)")
    ap.add_argument("--gather-wave", type=int, choices=[0, 1], default=1,
                    help="4 slots: the one-wave persistent gather (1) or the 128-thread block (0); same values")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="BYTES",
                    help="run a device-to-device copy of BYTES first (counter calibration)")
    return ap


def main():
    args = arg_parser().parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build runs on a single "
              f"MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from gfem import workloads as W

    os.environ["GFEM_GATHER_SLOTS"] = str(args.gather_slots)
    os.environ["GFEM_GATHER_WAVE"] = str(args.gather_wave)
    if args.calib_copy:
        copy_calibration(args.calib_copy)

    t0 = time.perf_counter()
    spec = W.helmholtz3d(n=args.n, order=args.order)
    t_mesh = time.perf_counter() - t0
    log(f"mesh {spec.name}: {t_mesh:.2f} s")
    F = W.build(spec)
    t_pre = F.pre()  # on the device
    log(f"pre (device): {t_pre:.2f} s  n={F.n_unknown} nnz={F.nnz}")
    edofs = spec.element_dofs()
    t0 = time.perf_counter()
    F.assemble(download=False)  # first call builds the system's maps on the device
    t_first = time.perf_counter() - t0
    log(f"first assemble (system + maps, device): {t_first:.2f} s")
    setup = {"device": {"pre_s": t_pre, "first_assemble_s": t_first, "total_s": t_pre + t_first,
                        "what": "first set-up in the process (includes runtime start-up and the hipRTC compile "
                                "of the source program)"}}
    if not args.no_setup_host:
        # the same mesh set up again in the running process: what a service
        # pays per new mesh (the compiled source program is reused)
        Fw = W.build(spec)
        tw_pre = Fw.pre()
        t0 = time.perf_counter()
        Fw.assemble(download=False)
        tw_first = time.perf_counter() - t0
        del Fw
        setup["device_warm"] = {"pre_s": tw_pre, "first_assemble_s": tw_first, "total_s": tw_pre + tw_first}
        log(f"set-up device, warm: pre {tw_pre:.3f} s + first assemble {tw_first:.3f} s")
        # the same set-up on the host: pre() on the CPU, maps built on the CPU
        Fh = W.build(spec)
        th_pre = Fh.pre(device=False)
        os.environ["GFEM_HOST_FINALIZE"] = "1"
        t0 = time.perf_counter()
        Fh.assemble(download=False)
        th_first = time.perf_counter() - t0
        del os.environ["GFEM_HOST_FINALIZE"]
        del Fh
        setup["host"] = {"pre_s": th_pre, "first_assemble_s": th_first, "total_s": th_pre + th_first}
        setup["speedup"] = setup["host"]["total_s"] / setup["device_warm"]["total_s"]
        setup["speedup_first_in_process"] = setup["host"]["total_s"] / setup["device"]["total_s"]
        log(f"set-up host: pre {th_pre:.2f} s + first assemble {th_first:.2f} s; device "
            f"{setup['device_warm']['total_s']:.3f} s warm ({setup['speedup']:.1f}x), "
            f"{t_pre + t_first:.2f} s first in process")

    sec = timed_steps(F, args.steps, args.warmup, "fp64", "fused")
    value = edofs / sec / 1e9
    log(f"fused fp64: {sec * 1e3:.3f} ms/step  {value:.2f} GDOF/s")

    tuning = {"fp64": F.tuning("fp64")}
    prof, raw = kernel_profile(F, "fp64", "fused")
    gather_ms = prof.get("gather", float("nan"))
    # algorithmic bytes of the matrix scatter: the reference-format element
    # maps (element->DOF and element->nonzero, 4 B per entry) and the CSR
    # values written once; the engine reads its own compressed records and
    # element factor rows instead (engine_bytes, fewer)
    gbytes = spec.scatter_map_bytes() + 16.0 * F.nnz
    engine_bytes = F.gather_bytes("fp64")
    achieved = gbytes / (gather_ms * 1e-3) / 1e9
    log("kernel ms (fp64 fused): " + ", ".join(f"{k}={v:.4f}" for k, v in prof.items()))

    extra = {}
    if not args.no_extra:
        sec_m = timed_steps(F, args.steps, args.warmup, "mixed", "fused")
        prof_m, _ = kernel_profile(F, "mixed", "fused")
        gb_m = spec.scatter_map_bytes() + 8.0 * F.nnz  # complex64 values
        extra["mixed"] = {"value": edofs / sec_m / 1e9, "ms_per_step": sec_m * 1e3, "dtype": "c64 values, fp32 "
                          "contraction, fp64 geometry",
                          "gather_ms": prof_m.get("gather"),
                          "gather_GBs": gb_m / (prof_m.get("gather", float("nan")) * 1e-3) / 1e9}
        log(f"fused mixed: {sec_m * 1e3:.3f} ms/step  {edofs / sec_m / 1e9:.2f} GDOF/s")
        ref_steps = max(2, args.steps // 4)
        sec_r = timed_steps(F, ref_steps, 1, "fp64", "reference")
        extra["reference_gpu_fp64"] = {"value": edofs / sec_r / 1e9, "ms_per_step": sec_r * 1e3,
                                       "what": "per-Gauss-point dense element matrices + atomic CSR scatter"}
        # the mixed-precision engine against the plain FP64 GPU assembler
        extra["mixed_vs_plain_fp64_gpu"] = sec_r / sec_m
        log(f"reference-order fp64: {sec_r * 1e3:.3f} ms/step  {edofs / sec_r / 1e9:.2f} GDOF/s")

    if not args.no_extra:
        # a step that follows a change of the term parameters (new k and source
        # position, as every serving request makes): the reference tensors are
        # rebuilt on the stream inside the timed region
        import math
        ks = [2 * math.pi * (3 + 0.5 * (i % 7)) for i in range(args.steps + args.warmup)]

        def updated_step(i):
            for tag, f in W.helmholtz_params(ks[i], (0.5, 0.45 + 0.01 * (i % 5), 0.5)):
                F.set_term_coefficient(tag, f)
            F.assemble_async("fp64", "fused")

        for i in range(args.warmup):
            updated_step(i)
        F.sync()
        t1 = time.perf_counter()
        for i in range(args.steps):
            updated_step(args.warmup + i)
        F.sync()
        sec_u = (time.perf_counter() - t1) / args.steps
        extra["after_term_update_fp64"] = {"value": edofs / sec_u / 1e9, "ms_per_step": sec_u * 1e3,
                                           "what": "gfem_term_update of k, ik and the source centre, then the "
                                                   "FP64 assembly (reference tensors rebuilt each step)"}
        log(f"fp64 after term update: {sec_u * 1e3:.3f} ms/step  {edofs / sec_u / 1e9:.2f} GDOF/s")
        # restore the workload's own parameters
        for tag, f in W.helmholtz_params(2 * math.pi * 5):
            F.set_term_coefficient(tag, f)
        extra["poisson2d"] = poisson2d_leg(args)

    ho = None
    if not args.no_highorder:
        ho = highorder(args)

    sol = None
    if not args.no_solver:
        sol = solver_leg(args)

    bat = None
    if not args.no_batched:
        bat = batched(args)

    serve = None
    if not args.no_serve:
        serve = serving(args)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args)
            log(f"cpu baseline: {cpu['value']:.4f} GDOF/s")
        except Exception as e:  # the oracle is optional at run time
            log(f"cpu baseline skipped: {e}")

    line = {
        "metric": "GDOF/s assembled (FP64 vs mixed) at 1e-8 rel err; p50/p99 solve latency",
        "value": value,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": f"helmholtz3d unit cube {args.n}^3x6 jittered P{args.order} tets + Robin "
                               f"P{args.order} boundary triangles, k=10pi, Gaussian source, complex128",
                   "elements": sum(b["conn"].shape[0] for b in spec.model.blocks() if b["dim"] >= 2),
                   "element_dofs": edofs, "n_unknown": int(F.n_unknown), "nnz": int(F.nnz)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(gather_kernel_name(args)) if args.n == 50 and args.order == 2 else None,
                     "kernel": gather_kernel_name(args), "kernel_ms": gather_ms,
                     "algorithmic_bytes": gbytes,
                     "what": "algorithmic bytes = E (4 N + 4 N^2) of element maps over the bilinear-term elements "
                             "+ 16 B per nonzero written; engine_bytes = the engine's records + element factor "
                             "rows + values",
                     "engine_bytes": engine_bytes},
        "cpu_baseline": cpu,
        "autotuned": tuning,
    }
    line["setup"] = setup
    line.update(extra)
    if serve is not None:
        line["serving"] = serve
    if ho is not None:
        line["highorder"] = ho
    if bat is not None:
        line["batched"] = bat
    if sol is not None:
        line["solver"] = sol
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
