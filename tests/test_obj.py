"""Mesh ingest on the host: the OBJ reader against the reference reader's
reading of each face form (ObjLoader.cpp:15-216, expectations worked out
from its code), the Mesh set-up steps (Mesh.cpp:17-131) against numpy
restatements, and -- where the reference's resources are present -- its
garment and body: counts from the files' own headers, mesh springs equal to
the oracle's (spring.cpp:82-391), and the committed scene fixture equal to a
fresh build."""
import os

import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Mesh, ObjLoader, X, Y, Z

# every test here runs in the CPU suite and again in the GPU suite
pytestmark = pytest.mark.usefixtures("both_suites")

REF = "/root/reference/resources"
U = np.uint32(0xFFFFFFFF)        # atoi("") - 1 stored unsigned


def write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_plain_faces_and_polygons(tmp_path):
    o = ObjLoader(write(tmp_path, "# c\nv 1 2 3\nv 4.5 -6 7e-1\nv 0 0 1\nvn 0 1 0\nvt 0.25 0.5\n"
                                  "f 1 2 3\nf 1 2 3 4\n"))
    np.testing.assert_array_equal(o.vertices, np.float32([[1, 2, 3, 1], [4.5, -6, 0.7, 1], [0, 0, 1, 1]]))
    np.testing.assert_array_equal(o.vertex_index, [[0, 1, 2], [0, 1, 2]])     # a quad keeps 3 corners
    np.testing.assert_array_equal(o.tex_index, [[0, 1, 2], [0, 1, 2]])        # default texture 0, 1, 2
    np.testing.assert_array_equal(o.normal_index, o.vertex_index)             # normal index = vertex index
    np.testing.assert_array_equal(o.normals, [[0, 1, 0]])
    np.testing.assert_array_equal(o.tex, [[0.25, 0.5]])


def test_slash_faces_as_the_reference_reads_them(tmp_path):
    o = ObjLoader(write(tmp_path, "f 1/2/3 4/5/6 12/345/7\nf 10//11 12//13 14//15\nf 16/17 18/19 20/21\n"))
    np.testing.assert_array_equal(o.vertex_index, [[0, 3, 11], [9, 11, 13], [15, 17, 19]])
    # "v//n": the first corner reads an empty texture index; the second and
    # third look for their second '/' one character late and read "/n"
    np.testing.assert_array_equal(o.tex_index, [[1, 4, 344], [U, U, U], [16, 18, 20]])
    # ... and then their normal index from the whole token ("12//13" -> 12);
    # "v/t" corners also read the normal from the whole token
    np.testing.assert_array_equal(o.normal_index, [[2, 5, 6], [10, 11, 13], [15, 17, 19]])


def test_corner_without_slash_in_a_slash_face_is_an_error(tmp_path):
    with pytest.raises(L.ClothCoreError, match="without '/'"):
        ObjLoader(write(tmp_path, "v 0 0 0\nf 1/1/1 2 3\n"))


def test_unified_normals_last_corner_wins(tmp_path):
    o = ObjLoader(write(tmp_path, "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 1 0 0\nvn 0 1 0\nvn 0 0 1\n"
                                  "f 1/1/1 2/1/2 3/1/3\nf 1/1/3 3/1/2 2/1/1\n"))
    # vertex 0: corners (f0, n0) then (f1, n2) -> n2; vertex 1: n1 then n0;
    # vertex 2: n2 then n1
    np.testing.assert_array_equal(o.unified_normals, np.float32([[0, 0, 1], [1, 0, 0], [0, 1, 0]]))


def np_center(v):
    s = np.zeros(3, np.float32)
    for row in v[:, :3]:                 # float sums in vertex order
        s = (s + row).astype(np.float32)
    return (s / np.float32(v.shape[0])).astype(np.float32)


def test_center_scale_translate_rotation():
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.normal(size=(57, 3)).astype(np.float32), np.ones((57, 1), np.float32)], 1)
    m = Mesh(v.copy(), np.zeros((0, 3), np.uint32))
    np.testing.assert_array_equal(m.get_center(), np_center(v))
    m.scale(3.33)
    want = v.copy()
    want[:, :3] = ((v[:, :3] - np_center(v)) * np.float32(3.33)).astype(np.float32)
    np.testing.assert_array_equal(m.vertices, want)
    m.translate(0, 1.98, 0.02)
    want[:, :3] = want[:, :3] + np.float32([0, 1.98, 0.02])
    np.testing.assert_array_equal(m.vertices, want)
    # rotation: glm columns, (c0 x + c1 y) + (c2 z + c3 w), angle * 3.1415 / 180;
    # cos / sin of the float angle may differ in the last bit between libms
    for deg, axis in ((90, X), (-4, Z), (30, Y)):
        c = np_center(want)
        a = np.float32(np.float64(np.float32(deg) / np.float32(180)) * 3.1415)
        co, si = np.float32(np.cos(np.float64(a))), np.float32(np.sin(np.float64(a)))
        cols = {X: [[1, 0, 0], [0, co, -si], [0, si, co]], Y: [[co, 0, -si], [0, 1, 0], [si, 0, co]],
                Z: [[co, -si, 0], [si, co, 0], [0, 0, 1]]}[axis]
        M = np.array(cols, np.float32).T                 # columns -> matrix
        d = (want[:, :3] - c).astype(np.float32)
        want[:, :3] = (d @ M.T + c).astype(np.float32)
        m.rotation(deg, axis)
        np.testing.assert_allclose(m.vertices, want, rtol=0, atol=2e-6 * max(1.0, float(np.abs(want).max())))
        want = m.vertices.copy()


ref = pytest.mark.skipif(not os.path.isdir(REF), reason="the reference's resources are not present")


@ref
@pytest.mark.parametrize("path, nv, nf", [("cloth/tshirt2/tshirt2.obj", 16773, 32589),
                                           ("pose/female.obj", 6890, 13776)])
def test_reference_files_read(path, nv, nf):
    o = ObjLoader(os.path.join(REF, path))
    assert o.vertices.shape == (nv, 4) and o.vertex_index.shape == (nf, 3)
    assert int(o.vertex_index.max()) == nv - 1
    np.testing.assert_array_equal(o.vertices[:, 3], 1.0)


@ref
def test_reference_scene_fixture_and_springs():
    from golden import make_reference_meshes as mk
    fresh = mk.build()
    fx = np.load(mk.OUT)
    for k, v in fresh.items():
        np.testing.assert_array_equal(fx[k], v, err_msg=k)
    # the mesh springs of the reference's garment equal the oracle's
    v4 = np.concatenate([fresh["cloth_verts"], np.ones((fresh["cloth_verts"].shape[0], 1), np.float32)], 1)
    got = S.springs_from_mesh(v4, fresh["cloth_faces"])
    want = O.springs_from_mesh(v4, fresh["cloth_faces"])
    for (gr, gc), (wr, wc) in zip(got, want):
        np.testing.assert_array_equal(gr, wr)
        np.testing.assert_array_equal(gc["end"], wc["end"])
        np.testing.assert_array_equal(gc["original"], wc["original"])
