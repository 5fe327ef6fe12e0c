"""The oracle's picking (find_first_intersect_triangle, Ray.cpp:8-77) and
VBO normals (verlet.cu:91-116, 163-181) pinned by closed forms: hits at known
t, misses, parallel rays, the reference's acceptance of hits behind the
origin and its tie rule (the later face), exact unit normals of axis-aligned
faces and the normalised sum at a crease."""
import numpy as np

import oracle as O
from simulator_amd import scene as S

TRI = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]])


def test_pick_closed_forms():
    f = np.uint32([[0, 1, 2]])
    assert O.pick(TRI, f, [0.25, 0.25, 2], [0, 0, -1]) == (0, 2.0)
    assert O.pick(TRI, f, [0.25, 0.25, -1], [0, 0, -1]) == (0, -1.0)     # behind the origin: still a hit
    assert O.pick(TRI, f, [2, 2, 1], [0, 0, -1])[0] == -1                # outside the triangle
    assert O.pick(TRI, f, [0.25, 0.25, 1], [1, 0, 0])[0] == -1           # parallel (|det| < FLT_EPSILON)
    assert O.pick(TRI, f, [0, 0, 1], [0, 0, -1]) == (0, 1.0)             # u = v = 0: on the corner, a hit
    assert O.pick(TRI, f, [0.5, 0.5, 1], [0, 0, -1]) == (0, 1.0)         # u + v = 1: on the edge, a hit


def test_pick_smallest_t_and_ties():
    v = np.concatenate([TRI, TRI + np.float32([0, 0, 1]), TRI + np.float32([0, 0, -3])])
    f = np.uint32([[0, 1, 2], [3, 4, 5], [6, 7, 8], [0, 1, 2]])
    # from z = 2 down: t = 1 (z = 1), 2 (z = 0, faces 0 and 3), 5 (z = -3)
    assert O.pick(v, f, [0.2, 0.2, 2], [0, 0, -1]) == (1, 1.0)
    # from z = 0.5 down: z = 1 is behind (t = -0.5), the smallest t
    assert O.pick(v, f, [0.2, 0.2, 0.5], [0, 0, -1]) == (1, -0.5)
    # upwards from z = 0.5: every face is behind, the smallest t is z = -3
    assert O.pick(v, f, [0.2, 0.2, 0.5], [0, 0, 1]) == (2, -3.5)
    # equal t: the later of two identical faces (map<float, int> overwrites)
    assert O.pick(v[:3], np.uint32([[0, 1, 2], [0, 1, 2]]), [0.2, 0.2, 1], [0, 0, -1]) == (1, 1.0)
    assert O.pick(v, f[[3, 2, 0]], [0.2, 0.2, 0.5], [0, 0, -1]) == (2, 0.5)     # z = -3 at t = 3.5 loses


def test_vertex_normals_closed_forms():
    # a flat grid: every face normal is exactly (0, 1, 0) or (0, -1, 0)
    nx = 5
    xs, zs = np.meshgrid(np.arange(nx, dtype=np.float32), np.arange(nx, dtype=np.float32))
    pos = np.stack([xs.ravel(), np.full(nx * nx, 0.7, np.float32), zs.ravel()], 1)
    n = O.vertex_normals(pos, S.grid_faces(nx, nx))
    np.testing.assert_array_equal(np.abs(n), np.tile(np.float32([0, 1, 0]), (nx * nx, 1)))
    assert len(set(np.sign(n[:, 1]))) == 1
    # a crease: normals cross((1,0,0), (0,1,0)) = (0, 0, 1) and
    # cross((0,1,0), (0,0,-1)) = (-1, 0, 0); the shared edge gets
    # normalize((-1, 0, 1)) = (-1, 0, 1) * (1 / sqrt(2)) in float
    pos = np.float32([[0, 0, 0], [0, 1, 0], [1, 0, 0], [0, 0, -1]])
    f = np.uint32([[0, 2, 1], [0, 1, 3]])
    n = O.vertex_normals(pos, f)
    s = np.float32(1) / np.sqrt(np.float32(2))
    np.testing.assert_array_equal(n[0], np.float32([-s, 0, s]))
    np.testing.assert_array_equal(n[1], np.float32([-s, 0, s]))
    np.testing.assert_array_equal(n[2], np.float32([0, 0, 1]))
    np.testing.assert_array_equal(n[3], np.float32([-1, 0, 0]))
    # a vertex on no face: normalize(0) is NaN, as glm's
    n = O.vertex_normals(np.concatenate([pos, np.float32([[5, 5, 5]])]), f)
    assert np.isnan(n[4]).all()
