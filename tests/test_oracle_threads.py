"""The oracle's threaded loops (triangle boxes, deformation, refit subtrees,
the step over vertices) give the same bits as one thread: the CPU baseline
at all cores runs the same algorithm."""
import numpy as np

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S


def run(nthreads):
    O.set_threads(nthreads)
    try:
        b = S.capsules(12, 8, 4, L.CC_BODY_REFIT)
        d = O.bvh_build(b.verts, b.tris, 0.002)
        caps = b.caps
        R = np.float32([0.99, 0, 0.1, 0, 1, 0, -0.1, 0, 0.99])
        dv = O.body_deform(b.verts, caps["vstart"], caps["p0"], caps["p1"],
                           np.linspace(0.97, 1.03, len(caps["phase"])).astype(np.float32), R,
                           np.float32([0.01, -0.02, 0.03]))
        lb = O.prim_boxes(dv, d["leaf_tri"], 0.002)
        bi = O.refit(d["childA"], d["childB"], d["parent_int"], lb)
        sc = S.small(side=32, y0=0.806)
        bb = O.bvh_build(sc.body.verts, sc.body.tris, sc.params.box_depth)
        pins = np.zeros(sc.n, np.uint8)
        pins[sc.pins] = 1
        x, xp, c = sc.pos.copy(), sc.pos.copy(), np.zeros_like(sc.pos)
        h = O.BvhHandle(bb)
        for _ in range(30):
            x, xp, c = O.step_f32(x, xp, c, sc.springs, O.params_from(sc.params), pins, h, nthreads=nthreads)
        return dv, lb, bi, x, xp, c
    finally:
        O.set_threads(1)


def test_threads_do_not_change_results():
    one, many = run(1), run(4)
    for a, b in zip(one, many):
        np.testing.assert_array_equal(a, b)
    # the refit over subtrees equals a fresh build's boxes of the same pose
    b = S.capsules(12, 8, 4, L.CC_BODY_REFIT)
    d = O.bvh_build(b.verts, b.tris, 0.002)
    O.set_threads(4)
    try:
        bi = O.refit(d["childA"], d["childB"], d["parent_int"], d["box_leaf"])
    finally:
        O.set_threads(1)
    np.testing.assert_array_equal(bi, d["box_int"])
