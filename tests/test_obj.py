"""OBJ import (M/geometry/shape/TriangleMesh3F.java:534-960, 1200-1700):
fan triangulation, vertex sharing, mesh splitting by group and material,
generated normals, and the scene compiler accepting the result."""
import math

import numpy as np
import pytest

from dayflower_amd import obj
from dayflower_amd import scene as S

# regular tetrahedron, no normals: every vertex normal is the unit vector
# from the centroid (the normalised sum of its three unit face normals)
TETRA = """# tetrahedron
v 1 1 1
v 1 -1 -1
v -1 1 -1
v -1 -1 1
f 1 2 3
f 1 4 2
f 1 3 4
f 2 4 3
"""

PENTAGON = """v 0 0 0
v 1 0 0
v 1.3 1 0
v 0.5 1.6 0
v -0.3 1 0
vn 0 0 1
f 1//1 2//1 3//1 4//1 5//1
"""

GROUPS = """v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
vt 0 0
vt 1 0
vt 0 1
vt 0.25 0.75
g top
usemtl red
f 1/1 2/2 3/3
f 2/2 4/4 3/3
o other
f 1/1 2/2 4/4
usemtl blue
f 1/1 3/3 4/4
g bottom
f 1/2 3/3 4/4
"""


def test_generated_normals_of_a_regular_tetrahedron():
    (m,) = obj.parse_wavefront_object(TETRA)
    assert m.indices.shape == (4, 3) and m.positions.shape == (4, 3)
    expect = m.positions / np.linalg.norm(m.positions, axis=1, keepdims=True)  # centroid at the origin
    np.testing.assert_allclose(m.normals, expect, atol=1e-6)
    # winding of the file is kept: every face normal points outwards
    p = m.positions[m.indices]
    fn = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert ((fn * p.mean(1)).sum(1) > 0).all()


def test_polygon_is_a_fan_from_its_first_vertex():
    (m,) = obj.parse_wavefront_object(PENTAGON)
    assert m.indices.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4]]
    np.testing.assert_allclose(m.normals, np.tile([0, 0, 1], (5, 1)))
    # the fan covers the polygon: total area by the shoelace formula
    xy = m.positions[:, :2].astype(np.float64)
    shoelace = 0.5 * abs(np.dot(xy[:, 0], np.roll(xy[:, 1], -1)) - np.dot(xy[:, 1], np.roll(xy[:, 0], -1)))
    p = m.positions[m.indices].astype(np.float64)
    fan = 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1).sum()
    assert fan == pytest.approx(shoelace, rel=1e-6)


def test_meshes_split_on_group_or_material_not_object():
    """Names come from the first occurrence of each triangle's vertex A
    triple: the face after `usemtl blue` starts at 1/1, first read under
    (top, red), so it stays in the red mesh; the last face's vertex A 1/2 is
    new and opens (bottom, blue).  The red mesh carries the object name of
    its last triangle (vertex A 1/1, first read before `o other`)."""
    ms = obj.parse_wavefront_object(GROUPS, flip_texture_y=True)
    assert [(m.group_name, m.material_name, m.object_name, len(m.indices)) for m in ms] == [
        ("top", "red", "", 4), ("bottom", "blue", "other", 1)]
    # equal (position, texture) pairs share a vertex; a new pair does not
    assert len(ms[0].positions) == 4
    assert ms[1].uvs[0].tolist() == [1.0, 1.0]          # vt 2 = (1, 0), flipped


def test_mesh_object_name_is_the_last_triangles():
    src = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\no a\nf 1 2 3\no b\nf 4 3 2\n"
    (m,) = obj.parse_wavefront_object(src)
    assert (m.object_name, len(m.indices)) == ("b", 2)
    # a vertex A read again under a later group keeps its first group
    src = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\ng x\nf 1 2 3\ng y\nf 1 3 4\nf 4 3 2\n"
    ms = obj.parse_wavefront_object(src)
    assert [(m.group_name, len(m.indices)) for m in ms] == [("x", 2), ("y", 1)]


def test_scale_errors_and_compilation():
    (m,) = obj.parse_wavefront_object(TETRA, scale=2.0)
    np.testing.assert_allclose(np.abs(m.positions), 2.0)
    (mn,) = obj.parse_wavefront_object(TETRA, scale=-1.0)
    # mirrored positions -p, normals through the inverse transpose -n = -p/|p|
    np.testing.assert_allclose(mn.normals, mn.positions / np.linalg.norm(mn.positions, axis=1, keepdims=True), atol=1e-6)
    assert obj.parse_wavefront_object(TETRA, scale=0.0) == []
    with pytest.raises(ValueError):
        obj.parse_wavefront_object("v 0 0 0\nv 1 0 0\nv 0 1 0\nf -1 -2 -3\n")
    with pytest.raises(ValueError):
        obj.parse_wavefront_object("v 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvn 0 0 1\nf 1//1 2//1 3//1\nf 2 4 3\n")
    sc = S.Scene()
    mat = sc.add_material(S.matte((0.5, 0.5, 0.5)))
    prims = obj.add_to_scene(sc, obj.parse_wavefront_object(GROUPS), {"": mat})
    assert len(prims) == 2 and sc.indices.shape == (5, 3)
    with pytest.raises(KeyError):
        obj.add_to_scene(S.Scene(), obj.parse_wavefront_object(GROUPS), {"red": 0})
