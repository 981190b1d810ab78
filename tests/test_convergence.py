"""H1 convergence of the L2 projection: the reference's known-answer check
(fem-master/unitTests/coverage/convergences.cpp:90-180).

Solve  int v tf = int f tf  for f = sin(10 x) + sin(10 y) on the unit square
(triangles, interior nodes jittered by +-0.2 h), Lagrange orders 1-3, rule
Gauss(5p), h = 1/4 ... 1/32 (the reference refines lc = 0.25 four times).
The relative L2 error |f - v| / |f|, integrated on the CPU with the oracle's
Gauss rule and Lagrange basis, must fall with slope p + 1 +- 0.5 in log h.
That is the reference's bar (convergences.cpp:176).

Two pipelines, both checked against that mathematics:
  - the oracle: CPU assembly and a direct sparse solve (pins the oracle's
    basis, rules, geometry and scatter together);
  - the engine: GPU assembly and the device solver (marked gpu)."""
import math

import numpy as np
import pytest

NS = (4, 8, 16, 32)


def _spec(n, p):
    from gfem import sin, x, y, workloads as W
    s = W.poisson2d(n=n, order=p, jitter=0.2, seed=77)
    s.name = f"l2proj_n{n}_p{p}"
    s.dirichlet = []
    f = sin(x() * 10.0) + sin(y() * 10.0)
    s.terms = [dict(kind="bilinear", domain=s.field_domain, gauss=5 * p, form=0),
               dict(kind="linear", domain=s.field_domain, gauss=5 * p, source=-f)]
    return s


def _rel_l2_error(spec, p, u_node):
    """|f - v|_L2 / |f|_L2 on the spec's triangles; u_node[tag - 1] = v at node tag"""
    from oracle import oracle as O
    m = spec.model
    xyz = m.xyz()
    conn = [b for b in m.blocks() if b["dim"] == 2][0]["conn"].astype(np.int64) - 1
    pts, w = O.gauss_rule(2, 5 * p)
    V = np.stack([O.lagrange(2, p, q)[0] for q in pts])  # [Q][N], gmsh node order
    X = xyz[conn[:, :3]]  # vertices first
    e1, e2 = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
    det = np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    xq = X[:, None, 0, :] + pts[None, :, 0:1] * e1[:, None, :] + pts[None, :, 1:2] * e2[:, None, :]
    f = np.sin(10.0 * xq[..., 0]) + np.sin(10.0 * xq[..., 1])
    vh = u_node[conn] @ V.T
    num = np.sum(w[None, :] * det[:, None] * (f - vh) ** 2)
    den = np.sum(w[None, :] * det[:, None] * f ** 2)
    return math.sqrt(num / den)


def _slope(errs):
    lh = np.log10([1.0 / n for n in NS])
    le = np.log10(errs)
    return float(np.polyfit(lh, le, 1)[0])


def _oracle_solution(spec):
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from oracle import oracle as O
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    ai, aj = P.csr()
    A = sp.csr_matrix((P.values(), aj, ai), shape=(P.n_unknown, P.n_unknown)).tocsc()
    x = np.real(sla.spsolve(A, -P.rhs(0)))  # a real system
    tags = range(1, spec.model.num_nodes + 1)
    return np.array([x[P.node_dof(t)[0] - 1] for t in tags])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_oracle_l2_projection_converges_at_p_plus_1(p):
    errs = [_rel_l2_error(s, p, _oracle_solution(s)) for s in (_spec(n, p) for n in NS)]
    slope = _slope(errs)
    assert abs(slope - (p + 1)) <= 0.5, (p, slope, errs)


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 2, 3])
def test_gpu_l2_projection_converges_at_p_plus_1(p):
    """the same on the device: assembly, pre() and the GMRES solve with FP64
    refinement; the errors also agree with the oracle pipeline's to 1e-8
    relative (both solve the same system to ~1e-12)"""
    from gfem import workloads as W
    print(f"case L2 projection convergence P{p}", flush=True)
    errs, errs_o = [], []
    for n in NS:
        s = _spec(n, p)
        F = W.build(s)
        F.pre()
        F.assemble()
        info = F.solve(tol=1e-12, restart=50, max_iter=20000)
        assert info["converged"], info
        u = np.real(np.array([F.node_value(t) for t in range(1, s.model.num_nodes + 1)]))
        errs.append(_rel_l2_error(s, p, u))
        errs_o.append(_rel_l2_error(s, p, _oracle_solution(s)))
    slope = _slope(errs)
    assert abs(slope - (p + 1)) <= 0.5, (p, slope, errs)
    np.testing.assert_allclose(errs, errs_o, rtol=1e-8)
