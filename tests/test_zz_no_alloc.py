"""cc_collide_points_noalloc (no device allocation; its traversal spill area
is a fixed device array per GPU) against cc_collide_points, bit for bit:
two calls in a row on device buffers the test owns, the second with fewer
points, over a static body and over a refitted or rebuilt body between
steps; sizes beyond the maximum are rejected."""
import ctypes as C

import numpy as np
import pytest

from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import DeviceBuffer, say

pytestmark = pytest.mark.gpu

MAX_POINTS = 65536


def noalloc_query(sim, pts):
    """The query through cc_collide_points_noalloc on device buffers."""
    n = pts.shape[0]
    bufs = [DeviceBuffer(np.ascontiguousarray(pts, np.float32)), DeviceBuffer(np.full(n, -7, np.int32)),
            DeviceBuffer(np.full(n, 9, np.uint8)), DeviceBuffer(np.full((n, 3), 5, np.float32)),
            DeviceBuffer(np.full((n, 3), 5, np.float32)), DeviceBuffer(np.full(1, 77, np.uint32))]
    try:
        L.check(L.lib().cc_collide_points_noalloc(sim._h, bufs[0].ptr, n, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr,
                                                  bufs[4].ptr, bufs[5].ptr, None))
        return [b.read() for b in bufs[1:]]
    finally:
        for b in bufs:
            b.free()


def around_sphere(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (np.array([0.5, 0.5, 0.0]) + d * rng.uniform(0.29, 0.31, size=(n, 1))).astype(np.float32)


def same_as_existing(sim, pts, label):
    hit, ins, pos, nrm, status = noalloc_query(sim, pts)
    want = sim.collide_points(pts)
    for g, w, name in zip((hit, ins, pos, nrm), want, ("hit", "inside", "position", "normal")):
        np.testing.assert_array_equal(g, w, err_msg=f"{label}: {name}")
    assert int(status[0]) == 0
    say(f"[noalloc] {label}: {pts.shape[0]} points, {int(ins.sum())} inside, equal to cc_collide_points")
    return int(ins.sum())


def test_noalloc_query_equals_the_allocating_one_static_body(gpu):
    sim = Simulator.from_scene(S.small(side=16))
    assert same_as_existing(sim, around_sphere(MAX_POINTS, 1), "sphere, first call") > 1000
    assert same_as_existing(sim, around_sphere(3000, 2), "sphere, second call, fewer points") > 50
    # n = 0 is a no-op; beyond the maximum is an error, with nothing launched
    L.check(L.lib().cc_collide_points_noalloc(sim._h, None, 0, None, None, None, None, None, None))
    with pytest.raises(L.ClothCoreError) as e:
        L.check(L.lib().cc_collide_points_noalloc(sim._h, C.c_void_p(16), MAX_POINTS + 1, C.c_void_p(16),
                                                  C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None, None))
    assert e.value.code == L.CC_ERR_INVALID and "65536" in str(e.value)
    sim.close()


@pytest.mark.parametrize("mode", [L.CC_BODY_REFIT, L.CC_BODY_REBUILD])
def test_noalloc_query_equals_the_allocating_one_moving_body(gpu, mode):
    body = S.capsules(64, 14, 6, mode)
    sc = S.grid_scene("noalloc", 48, 1e-3, 0, body, box_depth=0.002, y0=0.93)
    sim = Simulator.from_scene(sc)
    sim.simulate(37)                       # single steps: poses issued ahead on the update lanes
    v = sim.body_vertices()
    rng = np.random.default_rng(5)
    pts = (v[rng.integers(0, v.shape[0], 20000)] + rng.normal(scale=3e-3, size=(20000, 3))).astype(np.float32)
    kind = "refitted" if mode == L.CC_BODY_REFIT else "rebuilt"
    assert same_as_existing(sim, pts, f"{kind} body, frame 37") > 1000
    sim.simulate(1)
    assert same_as_existing(sim, pts[:5000], f"{kind} body, frame 38, fewer points") > 200
    sim.close()
