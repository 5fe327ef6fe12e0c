"""The configured workloads at their stated sizes on the GPU, against the
oracle.

  large    512x512 cloth (262,144 vertices) over the 501,120-triangle capsule
           body, deformed and refitted every frame, walked through its
           four-wide form, dt = 5e-4, 500 steps
  garment  111x111 cloth (12,321 vertices) over the 50,688-triangle body,
           rebuilt every frame, dt = 1e-3, 1000 steps
  batched  all 256 instances of the batched workload against one simulator
           per instance

Debug steps in contact (per step: the pose against the oracle's deformation,
the tree against the oracle's refit or build, the contact outcome of the
GPU's pre-projection positions bit for bit, one FP64 step from the GPU's
state within 1e-6 of the cloth extent, decision flips at resting contacts
excluded and bounded), then the whole run: finite, pins hold, graph blocks
equal one call per frame bit for bit."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import batch as B
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import assert_bvh_equal, extend_like_library, oracle_bvh, say

pytestmark = pytest.mark.gpu

EXTENT = 1.0            # the grid cloths span 1 m
TOL_STEP = 1e-6 * EXTENT


def pose_oracle(sim, sc):
    """The oracle's deformation of the rest pose with the current frame's
    motion parameters (rotation, translation, breathing per capsule)."""
    b = sc.body
    R, T, breath = sim.body_motion()
    rest = extend_like_library(b.verts, b.tris, b.normals, sc.params.body_extend)
    c = b.caps
    return O.body_deform(rest, c["vstart"], c["p0"], c["p1"], breath, R, T)


def refit_oracle(d0, verts, depth):
    """The frame-0 topology with boxes recomputed from `verts`."""
    d = dict(d0)
    d["box_leaf"] = O.prim_boxes(verts, d0["leaf_tri"], depth)
    d["box_int"] = O.refit(d0["childA"], d0["childB"], d0["parent_int"], d["box_leaf"])
    d["verts"] = np.ascontiguousarray(verts, np.float32)
    return d


def debug_steps(sim, sc, tree_of, nsteps, label, export_first=True):
    """nsteps single steps with debug outputs, each checked against the
    oracle; returns the number of contacts seen."""
    prm = O.params_from(sc.params)
    pins = np.zeros(sc.n, np.uint8)
    pins[sc.pins] = 1
    free = pins == 0
    sim.set_debug(True)
    contacts, flips, worst, worst_flip = 0, 0, 0.0, 0.0
    for k in range(nsteps):
        v = sim.body_vertices()
        np.testing.assert_array_equal(v, pose_oracle(sim, sc), err_msg=f"{label}: pose, frame {sim.frame}")
        d = tree_of(v)
        if export_first and k == 0:
            assert_bvh_equal(sim.bvh_export(), d)
            say(f"[{label}] frame {sim.frame}: LBVH of {d['m']} leaves bit-exact")
        h = O.BvhHandle(d)
        x, xp, c = sim.get_state()
        sim.simulate(1)
        pre, hit, ins = sim.get_debug()
        xo, xpo, co = sim.get_state()
        ohit, oins, opos, on = O.collide_points(h, pre[free])
        np.testing.assert_array_equal(hit[free], ohit, err_msg=f"{label}: hit, debug step {k}")
        np.testing.assert_array_equal(ins[free], oins, err_msg=f"{label}: inside, debug step {k}")
        np.testing.assert_array_equal(xo[free], opos, err_msg=f"{label}: position, debug step {k}")
        np.testing.assert_array_equal(co[free], on, err_msg=f"{label}: normal, debug step {k}")
        np.testing.assert_array_equal(xo[~free], x[~free])            # pins hold
        contacts += int(ins.sum())
        x64, _, _, ins64, hit64 = O.step_f64(x, xp, c, sc.springs, prm, pins, h, with_hit=True)
        diff = np.abs(xo - x64).max(1)
        flip = (ins64 != ins) | ((hit64 != hit) & (ins == 1))
        flips += int(flip.sum())
        worst = max(worst, float(diff[~flip].max()))
        if flip.any():
            worst_flip = max(worst_flip, float(diff[flip].max()))
    sim.set_debug(False)
    say(f"[{label}] {nsteps} debug steps: contacts {contacts}, FP32/FP64 flips {flips}, "
        f"worst |gpu - fp64| {worst:.3e} (tol {TOL_STEP:.0e}), worst at a flip {worst_flip:.3e}")
    assert worst <= TOL_STEP
    assert worst_flip <= 1e-3 * EXTENT       # at most one step's penetration
    return contacts


def whole_run(sc, nsteps, label):
    """The full run twice: one call (graph blocks) and one call per frame;
    bit for bit the same, finite, pins held, the cloth on the body."""
    a, b = Simulator.from_scene(sc), Simulator.from_scene(sc)
    G = a.block_steps
    say(f"[{label}] {nsteps} steps in one call ({nsteps // G} blocks of {G} + {nsteps % G}) "
        f"vs {nsteps} calls of one step")
    a.simulate(nsteps)
    for _ in range(nsteps):
        b.simulate(1)
    sa, sb = a.get_state(), b.get_state()
    for x, y, name in zip(sa, sb, ("x_cur", "x_prev", "contact")):
        np.testing.assert_array_equal(x, y, err_msg=f"{label}: {name}")
    np.testing.assert_array_equal(a.body_vertices(), b.body_vertices())
    assert a.frame == b.frame == nsteps
    a.check_finite()
    np.testing.assert_array_equal(sa[0][sc.pins], sc.pos[sc.pins])
    touching = int((np.abs(sa[2]).sum(1) > 0).sum())
    say(f"[{label}] after {nsteps} steps: {touching} vertices in contact, finite, pins held")
    a.close()
    b.close()
    return touching


@pytest.fixture(scope="module")
def large():
    return S.large()


def test_large_debug_steps_in_contact(gpu, large):
    sc = large
    say(f"[large] {sc.n} vertices, {sum(c.shape[0] for _, c in sc.springs)} springs, "
        f"body {sc.body.tris.shape[0]} triangles")
    sim = Simulator.from_scene(sc)
    width, nwide, bound = sim.bvh_width()
    say(f"[large] traversal width {width}, {nwide} wide nodes, deferral bound {bound}")
    assert width == 4
    v0 = sim.body_vertices()
    np.testing.assert_array_equal(v0, pose_oracle(sim, sc))
    d0 = oracle_bvh(v0, sc.body.tris, sc.params.box_depth)
    assert_bvh_equal(sim.bvh_export(), d0)
    sim.simulate(330)                    # the cloth reaches the body after ~160 steps
    c = debug_steps(sim, sc, lambda v: refit_oracle(d0, v, sc.params.box_depth), 3, "large")
    assert c > 3000
    sim.simulate(150)
    c = debug_steps(sim, sc, lambda v: refit_oracle(d0, v, sc.params.box_depth), 2, "large", export_first=False)
    assert c > 3000
    sim.close()


def test_large_whole_run(gpu, large):
    assert whole_run(large, 500, "large") > 1000


@pytest.fixture(scope="module")
def garment():
    return S.garment()


def test_garment_debug_steps_in_contact(gpu, garment):
    sc = garment
    say(f"[garment] {sc.n} vertices, body {sc.body.tris.shape[0]} triangles rebuilt every frame")
    sim = Simulator.from_scene(sc)
    tree = lambda v: oracle_bvh(v, sc.body.tris, sc.params.box_depth)   # noqa: E731
    sim.simulate(250)
    c = debug_steps(sim, sc, tree, 3, "garment")
    sim.simulate(600)
    c += debug_steps(sim, sc, tree, 3, "garment")
    assert c > 500
    sim.close()


def test_garment_whole_run(gpu, garment):
    assert whole_run(garment, 1000, "garment") > 200


def test_all_256_batched_instances_equal_their_simulators(gpu):
    scs = B.instances(256, seed=0)
    n = sum(sc.n for sc in scs)
    say(f"[batched] 256 instances, {n} vertices, {sum(sc.body.tris.shape[0] for sc in scs)} body triangles")
    b = B.Batch(scs)
    sims = [Simulator.from_scene(sc) for sc in scs]
    done = 0
    for nsteps in (1, 16, 7):            # a single launch, a 16-step graph, then singles
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
        say(f"[batched] all 256 instances bit-exact after {done} steps")
    v, t = b.body_export(255)
    np.testing.assert_array_equal(v, sims[255].body_vertices())
    assert_bvh_equal(t, oracle_bvh(v, scs[255].body.tris, scs[255].params.box_depth))
    for s in sims:
        s.close()
    b.close()
