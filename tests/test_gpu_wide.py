"""The four-wide traversal form (wide.hip) against the binary walk of
D_BVH::intersect (cudaBRT.cu:239-285): two simulators on the same scene, one
walking each form, stay bit-identical step after step (positions, previous
positions, contact normals), for a static body and for a body refitted every
frame; the per-point query gives the same leaf, inside flag and projection.
The binary walk is itself pinned to the oracle by the other GPU tests."""
import numpy as np
import pytest

from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say

pytestmark = pytest.mark.gpu


def pair(sc):
    a, b = Simulator.from_scene(sc), Simulator.from_scene(sc)
    b.set_bvh_width(2)
    return a, b


def assert_same_run(sc, blocks, per_block):
    a, b = pair(sc)
    wa, nw, bound = a.bvh_width()
    assert wa == 4 and nw > 0 and bound <= 64, (wa, nw, bound)
    assert b.bvh_width()[0] == 2
    contacts = 0
    for blk in range(blocks):
        a.simulate(per_block)
        b.simulate(per_block)
        xa, pa, ca = a.get_state()
        xb, pb, cb = b.get_state()
        np.testing.assert_array_equal(xa, xb, err_msg=f"block {blk}")
        np.testing.assert_array_equal(pa, pb, err_msg=f"block {blk}")
        np.testing.assert_array_equal(ca, cb, err_msg=f"block {blk}")
        contacts += int(np.any(ca != 0, axis=1).sum())
    say(f"[wide] {sc.name}: {nw} wide nodes, deferral bound {bound}, {contacts} contacts over {blocks} blocks")
    assert contacts > 0
    a.close()
    b.close()


def test_static_body_wide_equals_binary(gpu):
    assert_same_run(S.small(side=64, y0=0.806), blocks=6, per_block=25)


def test_refitted_body_wide_equals_binary(gpu):
    sc = S.grid_scene("refit128", 128, 5e-4, 200, S.capsules(24, 12, 6, L.CC_BODY_REFIT), box_depth=0.002,
                      y0=0.975)
    assert_same_run(sc, blocks=5, per_block=40)


def test_queries_wide_equals_binary(gpu):
    sc = S.small(side=16)
    a, b = pair(sc)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.28, 0.32, size=(20000, 1))          # around the sphere's surface
    pts = (np.array([0.5, 0.5, 0.0]) + d * r).astype(np.float32)
    ra, rb = a.collide_points(pts), b.collide_points(pts)
    for x, y in zip(ra, rb):
        np.testing.assert_array_equal(x, y)
    assert int((ra[0] >= 0).sum()) > 1000
    a.close()
    b.close()


def test_width_switch_and_rebuild(gpu):
    sc = S.small(side=16)
    a = Simulator.from_scene(sc)
    a.set_bvh_width(2)
    a.set_bvh_width(4)
    a.bvh_rebuild()                 # a new topology: the wide table is made again
    w, nw, _ = a.bvh_width()
    assert w == 4 and nw > 0
    with pytest.raises(L.ClothCoreError):
        a.set_bvh_width(3)
    a.close()
    g = Simulator.from_scene(S.garment(side=16))
    assert g.bvh_width()[0] == 2    # rebuilt every frame: always the binary walk
    g.close()
