"""Gmsh MSH 4.1 (ASCII) reading and writing of the host mesh model.

The format is gmsh's documented MSH 4.1 (sections $MeshFormat,
$PhysicalNames, $Entities, $Nodes, $Elements); the reference exchanges
meshes with gmsh in it (gmsh::open / gmsh::write, e.g.
fem-master/tutorials/nonLinear/main.cpp:72).  (CPU only.)"""
import numpy as np
import pytest

import gfem
from gfem import workloads as W


def _blocks(m):
    return {(b["dim"], b["tag"], b["type"]): b["conn"] for b in m.blocks()}


@pytest.mark.parametrize("make", [lambda: gfem.unit_cube(3, 2, 0.1, 5), lambda: gfem.unit_square(4, 3, 0.2, 7)])
def test_write_then_read_gives_the_same_model(tmp_path, make):
    m = make()
    p = str(tmp_path / "mesh.msh")
    gfem.write_msh(m, p)
    r = gfem.read_msh(p)
    assert r.num_nodes == m.num_nodes and r.dim == m.dim
    assert np.array_equal(r.xyz(), m.xyz())  # 17 significant digits round-trip exactly
    assert r.entities() == m.entities()
    a, b = _blocks(m), _blocks(r)
    assert a.keys() == b.keys()
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_read_back_mesh_preprocesses_identically(tmp_path):
    """DOF numbering and pattern of a formulation on the read-back mesh equal
    those on the generated mesh (bit-exact)."""
    spec = W.helmholtz3d(n=3, order=2)
    p = str(tmp_path / "cube.msh")
    gfem.write_msh(spec.model, p)
    F1 = W.build(spec)
    F1.pre(device=False)
    spec.model = gfem.read_msh(p)
    F2 = W.build(spec)
    F2.pre(device=False)
    assert np.array_equal(F1.ai(), F2.ai()) and np.array_equal(F1.aj(), F2.aj())


MSH = """$MeshFormat
4.1 0 8
$EndMeshFormat
$PhysicalNames
2
1 7 "boundary"
2 3 "domain"
$EndPhysicalNames
$Entities
0 2 1 0
11 0 0 0 1 0 0 1 7 0
12 1 0 0 1 1 0 1 7 0
5 0 0 0 1 1 0 1 3 0
$EndEntities
$Nodes
2 4 10 40
2 5 0 3
10
20
40
0 0 0
1 0 0
0 1 0
1 11 0 1
30
1 1 0
$EndNodes
$Elements
2 3 1 3
2 5 2 2
1 10 20 30
2 10 30 40
1 11 1 1
3 20 30
$EndElements
$NodeData
1
"u"
1
0
3
0
1
4
10 0.
20 0.
30 0.
40 0.
$EndNodeData
"""


def test_read_hand_written_file_with_sparse_tags_and_physicals(tmp_path):
    p = tmp_path / "two.msh"
    p.write_text(MSH)
    m = gfem.read_msh(str(p))
    assert m.num_nodes == 4 and m.dim == 2
    assert list(m.file_node_tags) == [10, 20, 30, 40]
    # node 30 (1,1) was listed last: compacted tags follow the file's tag order
    assert np.array_equal(m.xyz()[2], [1.0, 1.0, 0.0])
    b = _blocks(m)
    assert np.array_equal(b[(2, 5, 2)], [[1, 2, 3], [1, 3, 4]])
    assert np.array_equal(b[(1, 11, 1)], [[2, 3]])
    phys = {(p["dim"], p["tag"]): p for p in m.physicals}
    assert phys[(2, 3)]["name"] == "domain" and phys[(2, 3)]["entities"] == [5]
    assert phys[(1, 7)]["name"] == "boundary" and phys[(1, 7)]["entities"] == [11, 12]
    # written back with the file's own node tags
    q = str(tmp_path / "again.msh")
    gfem.write_msh(m, q)
    r = gfem.read_msh(q)
    assert list(r.file_node_tags) == [10, 20, 30, 40]
    assert np.array_equal(_blocks(r)[(2, 5, 2)], b[(2, 5, 2)])


@pytest.mark.parametrize("bad,msg", [
    ("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n", "version"),
    ("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n", "binary"),
    ("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Elements\n1 1 1 1\n2 1 99 1\n1 1 2 3\n$EndElements\n", "element type 99"),
    ("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 2 1 2\n0 1 0 2\n1\n2\n0 0 x\n", "number"),
])
def test_malformed_files_are_errors(tmp_path, bad, msg):
    p = tmp_path / "bad.msh"
    p.write_text(bad)
    with pytest.raises(Exception, match=msg):
        gfem.read_msh(str(p))
