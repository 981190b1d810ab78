"""A mesh read from a Gmsh MSH 4.1 file through the device path: pre() on the
GPU and assemble() against the oracle on the same read model.

The file is written by this test, not by the engine's writer, with what gmsh
files carry and the generated meshes do not: sparse node and element tags
listed out of order over several node blocks, and the domains named by
physical groups ($PhysicalNames + entity physical tags).  The reader's key
service is the node tag -> node map behind the Lagrange keys
(fem-master/src/field/FunctionSpaceInterface.cpp:319-331)."""
import numpy as np
import pytest

import gfem
from gfem import workloads as W
from oracle import oracle as O
from _helpers import assert_close

pytestmark = pytest.mark.gpu

PHYS_OMEGA, PHYS_GAMMA = 5, 9


def _write_msh(model, path, seed=3):
    """MSH 4.1 ASCII: node tags 7*k + 13 in a shuffled order over three node
    blocks, element tags 1000 + 3*k, volume in physical 'omega', the six faces
    in physical 'gamma'"""
    rng = np.random.default_rng(seed)
    xyz = model.xyz()
    n = xyz.shape[0]
    tag = 7 * np.arange(1, n + 1) + 13  # file tag of node k+1
    order = rng.permutation(n)
    blocks = [b for b in model.blocks() if b["dim"] >= 2]
    ents = sorted({(b["dim"], b["tag"]) for b in blocks})
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$PhysicalNames", "2",
           f'3 {PHYS_OMEGA} "omega"', f'2 {PHYS_GAMMA} "gamma"', "$EndPhysicalNames", "$Entities"]
    nd = [sum(1 for d, _ in ents if d == k) for k in range(4)]
    out.append(f"0 0 {nd[2]} {nd[3]}")
    for d, t in ents:
        phys = PHYS_GAMMA if d == 2 else PHYS_OMEGA
        out.append(f"{t} 0 0 0 1 1 1 1 {phys} 0")
    out.append("$EndEntities")
    chunks = np.array_split(order, 3)
    out += ["$Nodes", f"3 {n} {tag.min()} {tag.max()}"]
    for i, c in enumerate(chunks):
        out.append(f"3 1 0 {len(c)}")
        out += [str(int(tag[k])) for k in c]
        out += [" ".join(repr(float(v)) for v in xyz[k]) for k in c]
    out.append("$EndNodes")
    ne = sum(b["conn"].shape[0] for b in blocks)
    out += ["$Elements", f"{len(blocks)} {ne} 1000 {1000 + 3 * (ne - 1)}"]
    k = 0
    for b in blocks:
        out.append(f"{b['dim']} {b['tag']} {b['type']} {b['conn'].shape[0]}")
        for row in b["conn"]:
            out.append(" ".join([str(1000 + 3 * k)] + [str(int(tag[t - 1])) for t in row]))
            k += 1
    out.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def test_msh_read_mesh_assembles_like_the_oracle(tmp_path):
    print("case msh-read helmholtz3d n=4 P2", flush=True)
    gen = W.helmholtz3d(n=4, order=2, jitter=0.1)
    p = str(tmp_path / "cube.msh")
    _write_msh(gen.model, p)
    m = gfem.read_msh(p)
    phys = {(q["dim"], q["tag"]): q for q in m.physicals}
    assert phys[(3, PHYS_OMEGA)]["name"] == "omega" and phys[(2, PHYS_GAMMA)]["name"] == "gamma"
    omega = [(3, t) for t in phys[(3, PHYS_OMEGA)]["entities"]]
    gamma = [(2, t) for t in phys[(2, PHYS_GAMMA)]["entities"]]
    assert sorted(gamma) == [(2, s) for s in range(1, 7)]
    # the same problem on the read model, domains taken from the physical groups
    spec = W.helmholtz3d(n=4, order=2, jitter=0.1)
    spec.model = m
    spec.field_domain = omega
    for t in spec.terms:
        t["domain"] = gamma if t["domain"][0][0] == 2 else omega
    F = W.build(spec)
    F.pre()  # on the device
    P = O.from_spec(spec)
    P.pre()
    ai, aj = P.csr()
    assert np.array_equal(F.ai(), ai) and np.array_equal(F.aj(), aj)
    F.assemble(precision="fp64")
    P.assemble()
    assert_close(F.values(), P.values(), 1e-12, "A")
    assert_close(F.rhs(0), P.rhs(0), 1e-12, "b")
    F.assemble(precision="mixed")
    assert_close(F.values(), P.values(), 2e-5, "A mixed")
    # node tags do not enter the numbering (first touch in element order,
    # Hilbert order of coordinates): the read mesh gives the generated
    # mesh's pattern and the same values
    G = W.build(gen)
    G.pre()
    assert np.array_equal(G.ai(), ai) and np.array_equal(G.aj(), aj)
    G.assemble(precision="fp64")
    F.assemble(precision="fp64")
    assert_close(F.values(), G.values(), 1e-12, "A read vs generated")
