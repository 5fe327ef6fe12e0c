"""LBVH parity on the GPU: Morton keys, sorted order, Karras topology, refit
boxes and the first-hit collision query must equal the oracle bit for bit."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S

from helpers import (assert_bvh_equal, collider_sim, extend_like_library, oracle_bvh, say)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("level", [0, 2, 4])
def test_icosphere_build_bitexact(gpu, level):
    say(f"[lbvh] icosphere level {level}")
    v, t, nrm = S.icosphere(level, 0.3, (0.5, 0.5, 0.0))
    sim = collider_sim(v, t, nrm, extend=True)
    gv = sim.body_vertices()
    want_v = extend_like_library(v, t, nrm, 0.003)        # Mesh::vertex_extend(0.003)
    np.testing.assert_array_equal(gv, want_v)
    assert_bvh_equal(sim.bvh_export(), oracle_bvh(gv, t, 0.01))
    sim.close()


def test_capsule_body_build_bitexact(gpu):
    say("[lbvh] 12-capsule body, 50,688 triangles, depth 2 mm")
    body = S.capsules(64, 14, 6, L.CC_BODY_STATIC)
    sim = collider_sim(body.verts, body.tris, body.normals, extend=True, depth=0.002)
    gv = sim.body_vertices()
    assert_bvh_equal(sim.bvh_export(), oracle_bvh(gv, body.tris, 0.002))
    sim.close()


def test_large_body_build_bitexact(gpu):
    say("[lbvh] 12-capsule body, 501,120 triangles")
    body = S.capsules(144, 64, 18, L.CC_BODY_STATIC)
    assert body.tris.shape[0] == 501120
    sim = collider_sim(body.verts, body.tris, body.normals, extend=True, depth=0.002)
    gv = sim.body_vertices()
    assert_bvh_equal(sim.bvh_export(), oracle_bvh(gv, body.tris, 0.002))
    sim.close()


def test_single_and_two_triangles(gpu):
    say("[lbvh] m = 1 and m = 2")
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5]], np.float32)
    for tris in (np.array([[0, 1, 2]], np.uint32), np.array([[0, 1, 2], [1, 3, 2]], np.uint32)):
        sim = collider_sim(v, tris)
        e = sim.bvh_export()
        d = oracle_bvh(v, tris, 0.01)
        assert_bvh_equal(e, d)
        pts = np.array([[0.2, 0.2, -0.005], [0.2, 0.2, 0.3], [0.9, 0.9, 0.45]], np.float32)
        got = sim.collide_points(pts)
        want = O.collide_points(O.BvhHandle(d), pts)
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g, w)
        sim.close()


def test_duplicate_codes_and_flat_body(gpu):
    say("[lbvh] identical triangles (all codes equal) and a flat body with zero depth")
    rng = np.random.default_rng(4)
    # 300 copies of the same triangle: every Morton code is equal, order by index
    v = np.array([[0.1, 0.1, 0.1], [0.3, 0.1, 0.1], [0.1, 0.3, 0.1]], np.float32)
    tris = np.tile(np.array([[0, 1, 2]], np.uint32), (300, 1))
    sim = collider_sim(v, tris)
    assert_bvh_equal(sim.bvh_export(), oracle_bvh(v, tris, 0.01))
    sim.close()
    # flat grid in z = 0 with depth 0: zero root extent along z (0/0 -> cell 0)
    g = 20
    xs, ys = np.meshgrid(np.linspace(0, 1, g), np.linspace(0, 1, g))
    v = np.stack([xs.ravel(), ys.ravel(), np.zeros(g * g)], 1).astype(np.float32)
    v[:, :2] += rng.uniform(-1e-3, 1e-3, size=(g * g, 2)).astype(np.float32)
    q = []
    for j in range(g - 1):
        for i in range(g - 1):
            a = j * g + i
            q += [[a, a + 1, a + g], [a + 1, a + g + 1, a + g]]
    tris = np.array(q, np.uint32)
    sim = collider_sim(v, tris, depth=0.0)
    assert_bvh_equal(sim.bvh_export(), oracle_bvh(v, tris, 0.0))
    sim.close()


def test_collision_query_bitexact(gpu):
    say("[collide] 200k points around the extended icosphere")
    v, t, nrm = S.icosphere(4, 0.3, (0.5, 0.5, 0.0))
    sim = collider_sim(v, t, nrm, extend=True)
    gv = sim.body_vertices()
    d = oracle_bvh(gv, t, 0.01)
    rng = np.random.default_rng(5)
    n = 200_000
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pts = (np.array([0.5, 0.5, 0.0]) + dirs * rng.uniform(0.28, 0.32, size=(n, 1))).astype(np.float32)
    got = sim.collide_points(pts)
    want = O.collide_points(O.BvhHandle(d), pts)
    for name, g, w in zip(("hit", "inside", "pos", "normal"), got, want):
        assert np.array_equal(g, w), f"{name} differs at {np.where((g != w).reshape(n, -1).any(1))[0][:5]}"
    assert got[1].sum() > 10_000                     # plenty of projections exercised
    sim.close()


def test_refit_after_vertex_update_bitexact(gpu):
    say("[lbvh] set_vertices (rebuild) then refit of a moved body")
    v, t, nrm = S.icosphere(3, 0.3, (0.5, 0.5, 0.0))
    sim = collider_sim(v, t, nrm)
    rng = np.random.default_rng(6)
    v2 = (v + rng.uniform(-2e-3, 2e-3, size=v.shape)).astype(np.float32)
    sim.set_body_vertices(v2)
    d2 = oracle_bvh(v2, t, 0.01)
    assert_bvh_equal(sim.bvh_export(), d2)
    sim.close()
