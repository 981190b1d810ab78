"""The reference's Dirichlet and Neumann tutorials as known-answer tests
(fem-master/tutorials/dirichlet/main.cpp:24-102,
fem-master/tutorials/neumann/main.cpp).

Both on the annulus 1 < r < 5 with v = 1 on the inner circle ("gammaSmall"),
terms integral(grad dof v, grad tf v, omega, Gauss6) and
integral(1, tf v, omega, Gauss6).  Their sum is zero, so Laplace(v) = 1, and
the exact solutions are radial:
  - dirichlet: v = 0 on the outer circle ("gammaBig"):
        v(r) = r^2 / 4 - 7 ln(r) / ln(5) + 3 / 4;
  - neumann: integral(-1, tf v, gammaBig, Gauss6), i.e. dv/dr = 1 at r = 5
    (a linear term on the boundary lines):
        v(r) = r^2 / 4 - 7.5 ln(r) + 3 / 4.
The tutorial meshes with gmsh; here a structured polar mesh stands in (nr
radial x nt angular cells, two triangles each).  Every node, high-order ones
included, is placed by the polar map, so the P2 edges on the circles (and the
angular and diagonal edges inside) are curved: the P2 cases run the engine's
isoparametric path.

  - CPU: the oracle pipeline (assembly, direct sparse solve) converges to
    v at the vertices with slope >= p + 1 - 0.5 in log h;
  - GPU: pattern bit-exact and A, b within 1e-12 of the oracle; the device
    solve agrees with the oracle's direct solve within 1e-8 relative."""
import math

import numpy as np
import pytest

R_IN, R_OUT = 1.0, 5.0
LOG_COEF = {"dirichlet": -7.0 / math.log(5.0), "neumann": -7.5}
PROBLEMS = ("dirichlet", "neumann")


def exact(r, problem):
    return r * r / 4.0 + LOG_COEF[problem] * np.log(r) + 0.75


def annulus(nr, nt, order):
    """gfem model of the annulus; node tag of lattice point (i, j) is
    i * (order * nt) + j + 1, i radial in [0, order * nr], j angular (periodic)"""
    from gfem import from_arrays
    T = order * nt
    ii, jj = np.meshgrid(np.arange(order * nr + 1), np.arange(T), indexing="ij")
    r = R_IN + (R_OUT - R_IN) * ii / (order * nr)
    th = 2.0 * np.pi * jj / T
    xyz = np.stack([r * np.cos(th), r * np.sin(th), np.zeros_like(r)], axis=-1).reshape(-1, 3)

    def tag(i, j):
        return i * T + (j % T) + 1

    tris = []
    for a in range(nr):
        for b in range(nt):
            i0, j0 = order * a, order * b
            for va, vb, vc in (((0, 0), (1, 0), (1, 1)), ((0, 0), (1, 1), (0, 1))):
                # counter-clockwise in (r, theta), and the polar map keeps orientation
                P = [(i0 + order * v[0], j0 + order * v[1]) for v in (va, vb, vc)]
                nodes = [tag(*p) for p in P]
                if order == 2:  # gmsh 6-node triangle: vertices, then edges 01, 12, 20
                    for s, e in ((0, 1), (1, 2), (2, 0)):
                        nodes.append(tag((P[s][0] + P[e][0]) // 2, (P[s][1] + P[e][1]) // 2))
                tris.append(nodes)

    def ring(i):
        lines = []
        for b in range(nt):
            j0 = order * b
            nodes = [tag(i, j0), tag(i, j0 + order)]
            if order == 2:  # gmsh 3-node line: ends, then middle
                nodes.append(tag(i, j0 + 1))
            lines.append(nodes)
        return np.array(lines, dtype=np.int64)

    tri_type, line_type = (2, 1) if order == 1 else (9, 8)
    return from_arrays(f"annulus_{nr}x{nt}_p{order}", 2, xyz, [
        dict(tag=1, type=tri_type, conn=np.array(tris, dtype=np.int64)),
        dict(tag=1, type=line_type, conn=ring(0)),
        dict(tag=2, type=line_type, conn=ring(order * nr)),
    ])


def tutorial_spec(nr, nt, order, problem):
    from gfem.workloads import Spec
    omega, small, big = [(2, 1)], [(1, 1)], [(1, 2)]
    terms = [dict(kind="bilinear", domain=omega, gauss=6, form=1),
             dict(kind="linear", domain=omega, gauss=6, source=1.0)]
    if problem == "dirichlet":
        dirichlet = [(small, 1.0), (big, 0.0)]
    else:
        dirichlet = [(small, 1.0)]
        terms.append(dict(kind="linear", domain=big, gauss=6, source=-1.0))
    return Spec(name=f"poisson_{problem}_{nr}x{nt}_p{order}", model=annulus(nr, nt, order), order=order,
                field_domain=omega, dirichlet=dirichlet, terms=terms)


def _vertex_tags(spec):
    """tags of the mesh vertices (lattice points of the P1 grid)"""
    m = spec.model
    conn = [b for b in m.blocks() if b["dim"] == 2][0]["conn"]
    return np.unique(conn[:, :3])


def _oracle_nodal(spec, tags):
    """oracle assembly + direct solve (A x = -b); Dirichlet nodes carry their value"""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    from oracle import oracle as O
    P = O.from_spec(spec)
    P.pre()
    P.assemble()
    ai, aj = P.csr()
    A = sp.csr_matrix((P.values(), aj, ai), shape=(P.n_unknown, P.n_unknown)).tocsc()
    x = np.real(sla.spsolve(A, -P.rhs(0)))
    xyz = spec.model.xyz()
    out = []
    for t in tags:
        num, typ = P.node_dof(int(t))
        if typ == 1:  # fixed: the constraint's value
            out.append(1.0 if np.linalg.norm(xyz[t - 1]) < 0.5 * (R_IN + R_OUT) else 0.0)
        else:
            out.append(x[num - 1])
    return P, np.array(out)


def _errors(spec, tags, u, problem):
    r = np.linalg.norm(spec.model.xyz()[tags - 1], axis=1)
    return np.abs(u - exact(r, problem)).max()


MESHES = ((4, 16), (8, 32), (16, 64))
# max vertex error bounds, just above the oracle's measured errors on the
# finest mesh (dirichlet 8.4e-3 / 4.0e-5, neumann 4.7e-2 / 7.5e-5 at P1 / P2)
# and on the middle one (3.3e-2 / 5.4e-4, 0.184 / 1.1e-3)
FINE_BOUND = {("dirichlet", 1): 1e-2, ("dirichlet", 2): 1e-4, ("neumann", 1): 6e-2, ("neumann", 2): 1e-4}
MID_BOUND = {("dirichlet", 1): 5e-2, ("dirichlet", 2): 1e-3, ("neumann", 1): 0.25, ("neumann", 2): 2e-3}


@pytest.mark.parametrize("problem", PROBLEMS)
@pytest.mark.parametrize("order", [1, 2])
def test_oracle_tutorial_converges_to_the_radial_solution(order, problem):
    errs, hs = [], []
    for nr, nt in MESHES:
        s = tutorial_spec(nr, nt, order, problem)
        tags = _vertex_tags(s)
        _, u = _oracle_nodal(s, tags)
        errs.append(_errors(s, tags, u, problem))
        hs.append(1.0 / nr)
    slope = float(np.polyfit(np.log10(hs), np.log10(errs), 1)[0])
    assert slope >= order + 1 - 0.5, (order, slope, errs)
    assert errs[-1] < FINE_BOUND[problem, order], errs


def test_tutorial_mesh_is_curved_at_p2():
    """the P2 mesh is not affine: a midpoint of an edge on the outer circle
    lies on the circle, off the chord of its two vertices"""
    m = annulus(2, 8, 2)
    xyz = m.xyz()
    outer = [b for b in m.blocks() if b["dim"] == 1 and b["tag"] == 2][0]["conn"]
    a, b, c = (xyz[outer[0, k] - 1] for k in range(3))
    assert abs(np.linalg.norm(c) - R_OUT) < 1e-14
    assert np.linalg.norm(c - 0.5 * (a + b)) > 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("problem", PROBLEMS)
@pytest.mark.parametrize("order", [1, 2])
def test_gpu_tutorial_matches_oracle_and_solution(order, problem):
    from gfem import workloads as W
    from _helpers import assert_close
    nr, nt = MESHES[1]
    print(f"case {problem} tutorial P{order} {nr}x{nt}", flush=True)
    s = tutorial_spec(nr, nt, order, problem)
    tags = _vertex_tags(s)
    P, u_o = _oracle_nodal(s, tags)
    F = W.build(s)
    F.pre()
    ai, aj = P.csr()
    assert np.array_equal(F.ai(), ai) and np.array_equal(F.aj(), aj)
    F.assemble()
    assert_close(F.values(), P.values(), 1e-12, f"tutorial P{order} A")
    assert_close(F.rhs(0), P.rhs(0), 1e-12, f"tutorial P{order} b")
    info = F.solve(tol=1e-12, restart=50, max_iter=20000)
    assert info["converged"], info
    u = np.real(np.array([F.node_value(int(t)) for t in tags]))
    np.testing.assert_allclose(u, u_o, rtol=1e-8, atol=1e-8 * np.abs(u_o).max())
    assert _errors(s, tags, u, problem) < MID_BOUND[problem, order]
