"""Per-frame body paths on the GPU: deformation, refit (chunk + top-tree
passes) and full rebuild must equal the oracle bit for bit."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S

from simulator_amd.simulator import Simulator

from helpers import assert_bvh_equal, collider_sim, extend_like_library, oracle_bvh, say

pytestmark = pytest.mark.gpu


def body_sim(body, mode, depth=0.002):
    return collider_sim(body.verts, body.tris, body.normals, mode=mode, depth=depth, extend=True,
                        caps=body.caps, motion=body.motion)


def check_deform(sim, body):
    R, T, breath = sim.body_motion()
    rest = extend_like_library(body.verts, body.tris, body.normals, 0.003)
    c = body.caps
    want = O.body_deform(rest, c["vstart"], c["p0"], c["p1"], breath, R, T)
    got = sim.body_vertices()
    np.testing.assert_array_equal(got, want)
    return got


def refit_oracle(d0, verts, depth):
    """Topology of the frame-0 build, boxes recomputed from `verts`."""
    d = dict(d0)
    lb = O.prim_boxes(verts, d0["leaf_tri"], depth)
    d["box_leaf"] = lb
    d["box_int"] = O.refit(d0["childA"], d0["childB"], d0["parent_int"], lb)
    d["verts"] = np.ascontiguousarray(verts, np.float32)
    return d


@pytest.mark.parametrize("shape", [(64, 14, 6), (144, 64, 18)])
def test_refit_mode_bitexact(gpu, shape):
    body = S.capsules(*shape, L.CC_BODY_REFIT)
    say(f"[body] refit mode, {body.tris.shape[0]} triangles")
    sim = body_sim(body, L.CC_BODY_REFIT)
    v0 = check_deform(sim, body)                       # frame 0 pose, topology source
    d0 = oracle_bvh(v0, body.tris, 0.002)
    assert_bvh_equal(sim.bvh_export(), d0)
    for k in (1, 7, 40):
        sim.simulate(k)
        v = check_deform(sim, body)
        assert_bvh_equal(sim.bvh_export(), refit_oracle(d0, v, 0.002))
    assert np.abs(v - v0).max() > 1e-3
    sim.close()


def test_rebuild_mode_bitexact(gpu):
    body = S.capsules(64, 14, 6, L.CC_BODY_REBUILD)
    say(f"[body] rebuild mode, {body.tris.shape[0]} triangles")
    sim = body_sim(body, L.CC_BODY_REBUILD)
    for k in (1, 3, 25):
        sim.simulate(k)
        v = check_deform(sim, body)
        assert_bvh_equal(sim.bvh_export(), oracle_bvh(v, body.tris, 0.002))
    sim.close()


def test_top_tree_global_fallback_bitexact(gpu):
    # ~657k triangles: more chunk-crossing nodes than the top pass holds in LDS
    body = S.capsules(160, 76, 20, L.CC_BODY_STATIC)
    say(f"[body] static build, {body.tris.shape[0]} triangles (top-tree fallback)")
    sim = collider_sim(body.verts, body.tris, body.normals, extend=True, depth=0.002)
    gv = sim.body_vertices()
    assert_bvh_equal(sim.bvh_export(), oracle_bvh(gv, body.tris, 0.002))
    sim.close()


def test_motion_parameters_follow_the_formula(gpu):
    body = S.capsules(16, 4, 2, L.CC_BODY_REFIT)
    sim = body_sim(body, L.CC_BODY_REFIT)
    m = body.motion
    dt = float(np.float32(sim.params.dt))
    for k in (0, 5, 30):
        if k:
            sim.simulate(k)
        f = sim.frame                    # the body holds the pose of the current frame
        R, T, b = sim.body_motion()
        t = f * dt
        w = np.sin(2 * np.pi * m["freq"] * t)
        th = m["rot_amp"] * w
        Rw = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
        P = np.array(m["pivot"])
        Tw = P - Rw @ P + m["amp"] * w * np.array(m["dir"])
        bw = 1 + m["breath_amp"] * np.sin(2 * np.pi * m["breath_freq"] * t + body.caps["phase"].astype(np.float64))
        np.testing.assert_allclose(R, Rw.ravel(), atol=1e-7)
        np.testing.assert_allclose(T, Tw, atol=1e-7)
        np.testing.assert_allclose(b, bw, atol=1e-7)
    sim.close()


def test_set_frame_recomputes_the_pose(gpu):
    body = S.capsules(16, 4, 2, L.CC_BODY_REFIT)
    sim = body_sim(body, L.CC_BODY_REFIT)
    sim.simulate(9)
    v9 = sim.body_vertices()
    R9 = sim.body_motion()[0]
    sim.set_frame(3)
    assert sim.frame == 3
    check_deform(sim, body)
    assert np.abs(sim.body_vertices() - v9).max() > 0
    sim.simulate(6)
    np.testing.assert_array_equal(sim.body_vertices(), v9)
    np.testing.assert_array_equal(sim.body_motion()[0], R9)
    sim.close()


def moving_scene(mode, side=48):
    body = S.capsules(64, 14, 6, mode)
    return S.grid_scene("moving", side, 1e-3, 0, body, box_depth=0.002, y0=0.93)


@pytest.mark.parametrize("mode", [L.CC_BODY_REFIT, L.CC_BODY_REBUILD])
def test_step_collides_with_the_pose_of_its_own_frame(gpu, mode):
    """The body update of frame f+1 runs beside the step of frame f; every
    step must still see the body of its own frame.  Per step: the pose read
    before the step, the oracle tree of that pose, and the contact outcome
    of the GPU's pre-projection positions, bit for bit."""
    sc = moving_scene(mode)
    name = "refit" if mode == L.CC_BODY_REFIT else "rebuild"
    say(f"[body] {name} body under a 48x48 cloth, 60 steps one by one")
    sim = Simulator.from_scene(sc)
    sim.set_debug(True)
    d0 = oracle_bvh(sim.body_vertices(), sc.body.tris, 0.002)
    free = np.ones(sc.n, bool)
    free[sc.pins] = False
    contacts = 0
    for k in range(60):
        v = sim.body_vertices()
        d = refit_oracle(d0, v, 0.002) if mode == L.CC_BODY_REFIT else oracle_bvh(v, sc.body.tris, 0.002)
        h = O.BvhHandle(d)
        sim.simulate(1)
        pre, hit, ins = sim.get_debug()
        xo, _, _ = sim.get_state()
        ohit, oins, opos, _ = O.collide_points(h, pre[free])
        np.testing.assert_array_equal(hit[free], ohit, err_msg=f"step {k}")
        np.testing.assert_array_equal(ins[free], oins, err_msg=f"step {k}")
        np.testing.assert_array_equal(xo[free], opos, err_msg=f"step {k}")
        contacts += int(ins.sum())
    say(f"[body] {name}: {contacts} contacts checked")
    assert contacts > 100
    sim.close()


@pytest.mark.parametrize("mode", [L.CC_BODY_REFIT, L.CC_BODY_REBUILD])
def test_multi_step_graph_equals_single_steps(gpu, mode):
    sc = moving_scene(mode)
    a, b = Simulator.from_scene(sc), Simulator.from_scene(sc)
    G = a.block_steps
    total = 3 + 2 * G + 5
    say(f"[body] 3 + {total - 3} steps (graphs of {G} from frame 3) vs {total} single steps (mode {mode})")
    a.simulate(3)
    a.simulate(total - 3)
    for _ in range(total):
        b.simulate(1)
    for x, y in zip(a.get_state(), b.get_state()):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(a.body_vertices(), b.body_vertices())
    assert a.frame == b.frame == total
    a.close()
    b.close()


def test_in_graph_timing_reports_every_step(gpu):
    sc = moving_scene(L.CC_BODY_REFIT, side=32)
    sim = Simulator.from_scene(sc)
    G = sim.block_steps
    sim.set_timing(True)
    sim.simulate(5 * G + 3)           # five graphs, then three single steps
    t = sim.timing()
    say(f"[body] in-graph timing: {t}")
    # every step of a graph; single steps of frames divisible by four
    assert t["steps"] == 5 * G + sum(1 for f in range(5 * G, 5 * G + 3) if f % 4 == 0)
    assert t["verlet_us"] > 0 and t["refit_us"] > 0
    x = sim.get_state()[0]
    sim.set_timing(False)
    ref = Simulator.from_scene(sc)
    ref.simulate(5 * G + 3)
    np.testing.assert_array_equal(ref.get_state()[0], x)     # timing changes no result
    sim.close()
    ref.close()


@pytest.mark.parametrize("mode", [L.CC_BODY_REFIT, L.CC_BODY_REBUILD])
def test_single_step_calls_are_timed(gpu, mode):
    """One call per frame (the reference's calling pattern): the step
    kernels of frames divisible by four and the body updates of frames
    divisible by eight are timed."""
    sc = moving_scene(mode, side=32)
    sim = Simulator.from_scene(sc)
    sim.simulate(3)
    sim.synchronize()
    sim.set_timing(True)
    for _ in range(20):
        sim.simulate(1)
    t = sim.timing()
    say(f"[body] single-step timing (mode {mode}): {t}")
    assert t["steps"] == sum(1 for f in range(3, 23) if f % 4 == 0)
    upd = t["refit_us"] if mode == L.CC_BODY_REFIT else t["build_us"]
    assert t["verlet_us"] > 0 and upd > 0
    sim.set_timing(False)
    sim.close()
