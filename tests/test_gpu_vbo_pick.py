"""The VBO (cuda_update_vbo, simulation.cu:196-231) and picking
(Scene::OnMouseDown + find_first_intersect_triangle, scene.cpp:365-436,
Ray.cpp:8-77) on the GPU against the oracle, bit for bit: the layout
[vec4 pos][vec2 uv][vec3 normal] with the uv bytes left untouched, vertex
normals from the current or (lag, as the reference) the previous positions,
and the first face along a ray with its t and selected vertex."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import DeviceBuffer, say

pytestmark = pytest.mark.gpu


def draped(side=32, steps=40):
    sc = S.small(side=side, y0=0.806)
    faces = S.grid_faces(side, side)
    sim = Simulator.from_scene(sc)
    sim.set_faces(faces)
    sim.simulate(steps)
    return sc, faces, sim


def test_vbo_layout_and_normals(gpu):
    sc, faces, sim = draped()
    prev = sim.get_state()[0]
    sim.simulate(1)
    cur = sim.get_state()[0]
    n, n_tex = sc.n, sc.n
    nbytes = 16 * n + 8 * n_tex + 12 * n
    say(f"[vbo] {n} vertices, {faces.shape[0]} faces, {nbytes} bytes")
    for lag in (False, True):
        buf = DeviceBuffer(np.full(nbytes // 4, -7.25, np.float32))
        sim.cuda_update_vbo(buf.ptr.value, nbytes, n_tex, lag)
        h = buf.read()
        buf.free()
        pos = h[:4 * n].reshape(n, 4)
        uv = h[4 * n:4 * n + 2 * n_tex]
        nrm = h[4 * n + 2 * n_tex:].reshape(n, 3)
        np.testing.assert_array_equal(pos[:, :3], cur)
        np.testing.assert_array_equal(pos[:, 3], 1.0)
        np.testing.assert_array_equal(uv, -7.25)
        want = O.vertex_normals(prev if lag else cur, faces)
        np.testing.assert_array_equal(nrm, want)
        np.testing.assert_array_equal(sim.get_normals(), want)
    buf = DeviceBuffer(np.zeros(nbytes // 4, np.float32))
    with pytest.raises(Exception, match="too small"):
        sim.cuda_update_vbo(buf.ptr.value, nbytes - 4, n_tex)
    buf.free()
    sim.close()


def test_pick_matches_oracle(gpu):
    sc, faces, sim = draped()
    x = sim.get_state()[0]
    rng = np.random.default_rng(12)
    hits = 0
    for k in range(300):
        o = np.float32([rng.uniform(-0.2, 1.2), rng.choice([2.0, -1.0]), rng.uniform(-0.7, 0.7)])
        d = np.float32([rng.normal() * 0.2, -np.sign(o[1]), rng.normal() * 0.2])
        if k % 50 == 0:
            d = np.float32([1, 0, 0])                       # along the cloth: few or no faces
        d = (d / np.sqrt(np.float32((d * d).sum()))).astype(np.float32)
        f, v, t = sim.pick(o, d)
        of, ot = O.pick(x, faces, o, d)
        assert f == of, (k, o, d, f, of)
        if f >= 0:
            hits += 1
            assert t == ot and v == int(faces[f, 0]), (k, t, ot, v)
    say(f"[pick] {hits} of 300 rays hit the cloth")
    assert hits > 100
    sim.close()
