"""Batched instances on the GPU: every instance of a batch equals its own
Simulator bit for bit (same step, same body pose and LBVH every frame, same
pins), the batched body build equals the oracle's LBVH, and the batched step
stays within the per-step tolerance of the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import batch as B
from simulator_amd.simulator import Simulator

from helpers import BVH_FIELDS, assert_bvh_equal, say

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-6       # cloth extent is 1 m
MODES = {"static": [L.CC_BODY_STATIC], "refit": [L.CC_BODY_REFIT], "rebuild": [L.CC_BODY_REBUILD],
         "mixed": [L.CC_BODY_REBUILD, L.CC_BODY_REFIT, L.CC_BODY_STATIC, None]}


def scenes(count, seed, kind):
    scs = B.instances(count, seed=seed)
    for k, sc in enumerate(scs):
        mode = MODES[kind][k % len(MODES[kind])]
        if mode is None:
            sc.body = None
        else:
            sc.body.mode = mode
    return scs


@pytest.mark.parametrize("kind", list(MODES))
def test_batch_equals_simulators(gpu, kind):
    scs = scenes(8, 11, kind)
    say(f"[batch] {kind}: 8 instances, sides", [sc.nx for sc in scs])
    b = B.Batch(scs)
    nv, tile, order = b.info()
    say(f"[batch] {nv} vertices, tile {tile}, order {order}")
    assert nv == sum(sc.n for sc in scs)
    assert tile in (64, 128, 256) and order in (0, 1)
    sims = [Simulator.from_scene(sc) for sc in scs]
    done = 0
    for nsteps in (1, 16, 5, 32, 3, 199):                # single launches and 16-step graphs, both parities
        b.simulate(nsteps)
        for s in sims:
            s.simulate(nsteps)
        done += nsteps
        b.synchronize()
        for k, s in enumerate(sims):
            want = s.get_state()
            got = b.get_state(k)
            for g, w, name in zip(got, want, ("x_cur", "x_prev", "contact")):
                np.testing.assert_array_equal(g, w, err_msg=f"instance {k} {name} after {done} steps")
    assert b.frame == done
    touched = [int((np.abs(b.get_state(k)[2]).sum(axis=1) > 0).sum()) for k in range(len(scs))]
    say(f"[batch] contacts per instance after {done} steps: {touched}")
    assert sum(touched) > 1000
    for k, sc in enumerate(scs):
        if sc.body is None:
            assert touched[k] == 0
        np.testing.assert_array_equal(b.get_state(k)[0][sc.pins], sc.pos[sc.pins])
    for s in sims:
        s.close()
    b.close()


def refit_oracle(d0, verts, depth):
    """Topology of the frame-0 build, boxes recomputed from `verts`."""
    d = dict(d0)
    lb = O.prim_boxes(verts, d0["leaf_tri"], depth)
    d["box_leaf"] = lb
    d["box_int"] = O.refit(d0["childA"], d0["childB"], d0["parent_int"], lb)
    return d


def test_batch_body_trees(gpu):
    """Pose and LBVH of every body at several frames: equal to the single
    simulator's, and to the oracle's build (REBUILD, STATIC) or refit of the
    frame-0 topology (REFIT)."""
    scs = scenes(4, 23, "mixed")[:3]                     # rebuild, refit, static
    b = B.Batch(scs)
    sims = [Simulator.from_scene(sc) for sc in scs]
    d0 = {}
    for frame in (0, 1, 17, 40):
        if frame:
            step = frame - b.frame
            b.simulate(step)
            for s in sims:
                s.simulate(step)
            b.synchronize()
        for k, sc in enumerate(scs):
            say(f"[batch] body tree, frame {frame}, instance {k} mode {sc.body.mode}, {sc.body.tris.shape[0]} triangles")
            v, t = b.body_export(k)
            np.testing.assert_array_equal(v, sims[k].body_vertices())
            assert_bvh_equal(t, sims[k].bvh_export())
            np.testing.assert_array_equal(t["root"], sims[k].bvh_export()["root"])
            want = O.bvh_build(v, sc.body.tris, sc.params.box_depth)
            if frame == 0:
                d0[k] = want
            if sc.body.mode == L.CC_BODY_REFIT:
                want = refit_oracle(d0[k], v, sc.params.box_depth)
            assert_bvh_equal(t, want)
    for s in sims:
        s.close()
    b.close()


def test_batch_step_vs_fp64_oracle(gpu):
    scs = B.instances(4, seed=5)
    b = B.Batch(scs)
    say("[batch] per-step FP64 agreement, rebuilt bodies, sides", [sc.nx for sc in scs])
    pins = []
    for sc in scs:
        p = np.zeros(sc.n, np.uint8)
        p[sc.pins] = 1
        pins.append(p)
    worst, nflips = 0.0, 0
    for k in range(0, 300, 30):
        b.simulate(30 if k else 1)
        before = [b.get_state(i) for i in range(len(scs))]
        bodies = [b.body_export(i)[0] for i in range(len(scs))]   # the pose of the coming step
        b.simulate(1)
        b.synchronize()
        for i, sc in enumerate(scs):
            h = O.BvhHandle(O.bvh_build(bodies[i], sc.body.tris, sc.params.box_depth))
            x, xp, c = before[i]
            x64, _, _, ins64, _ = O.step_f64(x, xp, c, sc.springs, O.params_from(sc.params), pins[i], h,
                                             with_hit=True)
            xg, _, cg = b.get_state(i)
            # a vertex resting on the body (plane distance ~0) can test inside
            # in FP32 and outside in FP64 or the other way; those flips are
            # excluded from the tolerance and bounded below
            ins32 = np.abs(cg).sum(axis=1) > 0
            same = ins32 == (ins64 == 1)
            d = np.abs(xg - x64).max(axis=1)
            worst = max(worst, float(d[same].max()))
            flips = int((~same).sum())
            nflips += flips
            if flips:
                # bounded by one step's penetration depth
                assert d[~same].max() <= 1e-3, f"instance {i} step {k}: flip moved {d[~same].max():.3e}"
    say(f"[batch] worst one-step |gpu - fp64| = {worst:.3e}, {nflips} inside flips")
    assert worst <= TOL_STEP
    b.close()


def tiny_body(tris, mode, y=0.995, copies=1):
    """A flat body of a few triangles under the cloth, one capsule covering
    every vertex (so it can move), with the default sway and breathing."""
    from simulator_amd import scene as S
    v = np.array([[0.2, y, -0.3], [0.8, y, -0.3], [0.5, y, 0.3], [0.9, y - 0.05, 0.4]], np.float32)
    t = np.tile(np.asarray(tris, np.uint32).reshape(-1, 3), (copies, 1))
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    caps = {"vstart": np.array([0, 4], np.uint32), "p0": np.array([[0.3, y, 0.0]], np.float32),
            "p1": np.array([[0.7, y, 0.0]], np.float32), "radius": np.array([0.05], np.float32),
            "phase": np.array([0.5], np.float32)}
    return S.Body(v, t, nrm, mode, True, caps, S._motion(9))


def test_batch_edge_instances(gpu):
    """Bodies of 1, 2 and 3 triangles and of 300 identical triangles (every
    Morton code equal) in both moving modes, a one-vertex cloth without
    springs, and an instance without a body: each equals its own Simulator,
    state and body tree, bit for bit."""
    from simulator_amd import scene as S
    scs = []
    for mode in (L.CC_BODY_REBUILD, L.CC_BODY_REFIT):
        for tris, copies in (([0, 1, 2], 1), ([0, 1, 2, 1, 3, 2], 1), ([0, 1, 2, 1, 3, 2, 0, 3, 2], 1),
                             ([0, 1, 2], 300)):
            scs.append(S.grid_scene("edge", 12, 1e-3, 60, tiny_body(tris, mode, copies=copies), box_depth=0.01))
    one = S.Scene("one", 1, 1, np.array([[0.5, 0.999, 0.0]], np.float32),
                  [(np.zeros(2, np.uint32), np.zeros(0, L.SPRING_DTYPE))], np.zeros(0, np.uint32),
                  S.params(), tiny_body([0, 1, 2], L.CC_BODY_REBUILD), 60)
    scs.append(one)
    bare = S.grid_scene("bare", 9, 1e-3, 60, None)
    scs.append(bare)
    say(f"[batch] {len(scs)} edge instances")
    b = B.Batch(scs)
    sims = [Simulator.from_scene(sc) for sc in scs]
    for nsteps in (1, 16, 7, 20):
        b.simulate(nsteps)
        for s in sims:
            s.simulate(nsteps)
        b.synchronize()
        for k, (sc, s) in enumerate(zip(scs, sims)):
            for g, w, name in zip(b.get_state(k), s.get_state(), ("x_cur", "x_prev", "contact")):
                np.testing.assert_array_equal(g, w, err_msg=f"instance {k} {name} at frame {b.frame}")
            if sc.body is not None:
                v, t = b.body_export(k)
                np.testing.assert_array_equal(v, s.body_vertices(), err_msg=f"instance {k} body at {b.frame}")
                assert_bvh_equal(t, s.bvh_export())
    assert sum(int((np.abs(b.get_state(k)[2]).sum(axis=1) > 0).sum()) for k in range(len(scs))) > 0
    for s in sims:
        s.close()
    b.close()


def test_batch_sort_tile_sizes(gpu):
    """Bodies around the per-body sort's tile of 8192 keys: 8192 triangles
    (one full tile), 8193 and 16391 (a last tile of one key), 21000 (three
    tiles), rebuilt every frame; every tree equals the oracle's build and the
    single simulator's, bit for bit."""
    from simulator_amd import scene as S
    v, t, nrm, caps = S.capsule_body(12, 0.04, 0.15, *S.CAPSULE_BOX, 96, 24, 24, 3)
    assert t.shape[0] >= 21000
    scs = []
    for m in (8192, 8193, 16391, 21000):
        body = S.Body(v, np.ascontiguousarray(t[:m]), nrm, L.CC_BODY_REBUILD, True, caps, S._motion(5))
        scs.append(S.grid_scene("tiles", 24, 1e-3, 40, body, box_depth=0.002))
    b = B.Batch(scs)
    sims = [Simulator.from_scene(sc) for sc in scs]
    for frame in (0, 1, 17):
        if frame:
            step = frame - b.frame
            b.simulate(step)
            for s in sims:
                s.simulate(step)
            b.synchronize()
        for k, sc in enumerate(scs):
            say(f"[batch] sort tiles, frame {frame}, {sc.body.tris.shape[0]} triangles")
            pose, tree = b.body_export(k)
            np.testing.assert_array_equal(pose, sims[k].body_vertices())
            assert_bvh_equal(tree, O.bvh_build(pose, sc.body.tris, sc.params.box_depth))
            assert_bvh_equal(tree, sims[k].bvh_export())
    for s in sims:
        s.close()
    b.close()
