"""OBJ import (no GPU), checked by geometry: a unit cube of quads becomes
12 outward-facing triangles; smooth normals at a cube corner point along
(+-1, +-1, +-1)/sqrt(3); given normals pass through; texture v is flipped;
tangents point along dP/du and are orthogonal to the normals."""
import numpy as np
import pytest

from colvillea_amd import imageio

CUBE = """# unit cube, quads, counter-clockwise seen from outside
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
f 1 4 3 2
f 5 6 7 8
f 1 2 6 5
f 2 3 7 6
f 3 4 8 7
f 4 1 5 8
"""


def test_cube_triangulated_with_smooth_normals(tmp_path):
    p = tmp_path / "cube.obj"
    p.write_text(CUBE)
    m = imageio.load_obj(p)
    assert m.triangles.shape == (12, 3) and m.vertices.shape == (8, 3)
    assert m.uvs is None and m.tangents is None
    v, t = m.vertices.astype(np.float64), m.triangles
    c = v.mean(axis=0)
    # every triangle faces outward (winding kept from the file)
    fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
    assert np.all(np.einsum("ij,ij->i", fn, v[t].mean(axis=1) - c) > 0)
    assert np.allclose(np.linalg.norm(m.normals, axis=1), 1.0, atol=1e-6)
    # a smooth normal lies inside the cone of its corner's face normals
    assert np.all(np.einsum("ij,ij->i", m.normals, v - c) > 0)


def test_octahedron_smooth_normals_by_symmetry(tmp_path):
    """Each vertex of a regular octahedron touches 4 congruent triangles, so
    by symmetry its smooth normal is the direction from the centre."""
    obj = "v 1 0 0\nv -1 0 0\nv 0 1 0\nv 0 -1 0\nv 0 0 1\nv 0 0 -1\n"
    obj += "f 1 3 5\nf 3 2 5\nf 2 4 5\nf 4 1 5\nf 3 1 6\nf 2 3 6\nf 4 2 6\nf 1 4 6\n"
    p = tmp_path / "oct.obj"
    p.write_text(obj)
    m = imageio.load_obj(p)
    assert np.allclose(m.normals, m.vertices, atol=1e-6)


def test_uv_flip_tangents_and_given_normals(tmp_path):
    obj = """v 0 0 0
v 2 0 0
v 2 0 -3
v 0 0 -3
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 1 0
f 1/1/1 2/2/1 3/3/1 4/4/1
"""
    p = tmp_path / "quad.obj"
    p.write_text(obj)
    m = imageio.load_obj(p)
    assert m.triangles.shape == (2, 3)
    assert np.allclose(m.normals, [0, 1, 0])
    # texture v flipped: the corner with vt (1, 1) has uv (1, 0)
    i = int(np.nonzero((m.vertices == [2, 0, -3]).all(axis=1))[0][0])
    assert np.allclose(m.uvs[i], [1.0, 0.0])
    # u grows along +x: tangent = +x, orthogonal to the normal
    assert np.allclose(m.tangents, [1, 0, 0], atol=1e-6)


def test_negative_indices_and_shared_positions(tmp_path):
    obj = """v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
vn 0 0 1
vn 0 0 -1
f -4//1 -3//1 -2//1
f -3//2 -1//2 -2//2
"""
    p = tmp_path / "neg.obj"
    p.write_text(obj)
    m = imageio.load_obj(p)
    # positions 2 and 3 appear with two different normals: 6 distinct corners
    assert m.vertices.shape == (6, 3) and m.triangles.shape == (2, 3)


def test_bad_obj_raises(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(OSError, match="out of range"):
        imageio.load_obj(p)
    p.write_text("v 0 0 0\n")
    with pytest.raises(OSError, match="no faces"):
        imageio.load_obj(p)
