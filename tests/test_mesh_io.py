"""spg::io::loadObj / loadMsh of the C++ mirror (source/geom/io/obj.cpp:10-35, msh.cpp:10-41). The reference
reads both formats through tinyobjloader and MshIO, submodules absent from its repository, and has no tests
or sample files for them; these fixtures are written here from the two formats' published specifications
(Wavefront OBJ; Gmsh MSH 4.1 ASCII) and pin what the reference keeps: OBJ vertices and the first shape's
triangles (polygons as fans, 1-based or negative indices, "i/t/n" corners), MSH nodes in file order and the
4-node tetrahedra with 1-based node tags."""
import numpy as np
import pytest

from spg_amd import native

OBJ = """# a quad and a triangle, then a second shape that is not read
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0.5
vt 0 0
vn 0 0 1
o first
f 1/1/1 2/1/1 3/1/1 4/1/1
f -4//1 -2//1 -1//1
g second
v 2 2 2
f 1 2 5
"""

MSH = """$MeshFormat
4.1 0 8
$EndMeshFormat
$Entities
0 0 0 1
1 0 0 0 1 1 1 0 0
$EndEntities
$Nodes
2 5 1 5
3 1 0 3
1
2
3
0 0 0
1 0 0
0 1 0
3 1 0 2
4
5
0 0 1
1 1 1
$EndNodes
$Elements
2 3 1 3
2 1 2 1
1 1 2 3
3 1 4 2
2 1 2 3 4
3 2 3 4 5
$EndElements
"""


def test_load_obj(tmp_path):
    p = tmp_path / "mesh.obj"
    p.write_text(OBJ)
    x, f = native.load_obj(str(p))
    assert np.array_equal(x, [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0.5], [2, 2, 2]])
    assert np.array_equal(f, [[0, 1, 2], [0, 2, 3], [0, 2, 3]])


def test_load_msh(tmp_path):
    p = tmp_path / "mesh.msh"
    p.write_text(MSH)
    x, t = native.load_msh(str(p))
    assert np.array_equal(x, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]])
    assert np.array_equal(t, [[0, 1, 2, 3], [1, 2, 3, 4]])


def test_loaded_meshes_build_energies(tmp_path):
    """A loaded tetrahedral mesh feeds a ParticleGroup's StableNeoHookeanEnergy like a generated one."""
    p = tmp_path / "mesh.msh"
    p.write_text(MSH)
    x, t = native.load_msh(str(p))
    pg = native.ParticleGroup()
    pg.add_particles(x, x, np.ones(len(x)))
    pg.add_energy("snh", t, np.tile([1e4, 0.3], (len(t), 1)))
    assert pg.energy_params(0).shape == (2, 12)


@pytest.mark.parametrize("text,msg", [
    ("v 0 0 0\nf 1 2 3\n", "index out of range"),
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", "fewer than 3 corners"),
])
def test_obj_errors(tmp_path, text, msg):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(native.NativeError, match=msg):
        native.load_obj(str(p))


def test_missing_files_and_versions(tmp_path):
    with pytest.raises(native.NativeError, match="cannot open"):
        native.load_obj(str(tmp_path / "none.obj"))
    p = tmp_path / "old.msh"
    p.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    with pytest.raises(native.NativeError, match="only MSH 4.1"):
        native.load_msh(str(p))
