"""Step parity on the GPU against the FP64 oracle (positions, by tolerance)
and the FP32 exact-order oracle (contact set, bit-exact for the GPU's
pre-projection positions)."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say

pytestmark = pytest.mark.gpu

EXTENT = 1.0          # cloth extent (m) of the 1 m x 1 m grid
TOL_STEP = 1e-6 * EXTENT
TOL_TRAJ = 1e-4 * EXTENT


def oracle_setup(sim, sc):
    gv = sim.body_vertices()
    d = O.bvh_build(gv, sc.body.tris, sc.params.box_depth)
    h = O.BvhHandle(d)
    prm = O.params_from(sc.params)
    pins = np.zeros(sc.n, np.uint8)
    pins[sc.pins] = 1
    return h, prm, pins


def test_small_trajectory_within_1e4(gpu):
    say("[step] 64x64 over icosphere, 200 steps vs FP64 oracle")
    sc = S.small()
    sim = Simulator.from_scene(sc)
    h, prm, pins = oracle_setup(sim, sc)
    sim.simulate(200)
    xg, xpg, cg = sim.get_state()
    x = sc.pos.astype(np.float64)
    xp, c = x.copy(), np.zeros_like(x)
    for _ in range(200):
        x, xp, c, _ = O.step_f64(x, xp, c, sc.springs, prm, pins, h)
    err = np.abs(xg - x).max()
    say(f"[step] max |x_gpu - x_fp64| after 200 steps = {err:.3e}")
    assert err <= TOL_TRAJ
    np.testing.assert_array_equal(xg[sc.pins], sc.pos[sc.pins])     # pins hold
    assert sim.frame == 200
    sim.close()


def test_small_per_step_within_1e6(gpu):
    say("[step] per-step FP64 agreement along the 64x64 trajectory")
    sc = S.small()
    sim = Simulator.from_scene(sc)
    h, prm, pins = oracle_setup(sim, sc)
    worst = 0.0
    for k in range(0, 200, 20):
        sim.simulate(20 if k else 1)
        x, xp, c = sim.get_state()
        sim.simulate(1)
        xg, _, _ = sim.get_state()
        x64, *_ = O.step_f64(x, xp, c, sc.springs, prm, pins, h)
        worst = max(worst, np.abs(xg - x64).max())
    say(f"[step] worst one-step |gpu - fp64| = {worst:.3e}")
    assert worst <= TOL_STEP
    sim.close()


def test_contact_set_bitexact_and_per_step(gpu):
    say("[step] contact-heavy 64x64 (cloth at y=0.806), 120 steps, debug outputs")
    sc = S.small(y0=0.806)
    sim = Simulator.from_scene(sc)
    h, prm, pins = oracle_setup(sim, sc)
    sim.set_debug(True)
    free = pins == 0
    total_contacts, flips, worst_flip = 0, 0, 0.0
    for k in range(120):
        x, xp, c = sim.get_state()
        sim.simulate(1)
        pre, hit, ins = sim.get_debug()
        xo, xpo, co = sim.get_state()
        # contact set given identical (GPU) pre-projection positions
        ohit, oins, opos, on = O.collide_points(h, pre[free])
        np.testing.assert_array_equal(hit[free], ohit)
        np.testing.assert_array_equal(ins[free], oins)
        np.testing.assert_array_equal(xo[free], opos)
        np.testing.assert_array_equal(co[free], on)
        np.testing.assert_array_equal(xpo[free & (ins == 1)], xo[free & (ins == 1)])
        total_contacts += int(ins.sum())
        # one FP64 step from the same state: agreement except at threshold flips
        x64, _, _, ins64, hit64 = O.step_f64(x, xp, c, sc.springs, prm, pins, h, with_hit=True)
        diff = np.abs(xo - x64).max(1)
        # A flip is a vertex whose inside/outside decision or first-hit leaf
        # differs between FP32 and FP64 pre-projection positions.  Resting
        # contacts sit on the plane (d ~ 0 after last step's projection), so
        # their decision is rounding noise; the positions then differ by |d|.
        flip = (ins64 != ins) | ((hit64 != hit) & (ins == 1))
        flips += int(flip.sum())
        assert diff[~flip].max() <= TOL_STEP, f"step {k}: {diff[~flip].max():.3e}"
        if flip.any():
            worst_flip = max(worst_flip, float(diff[flip].max()))
    say(f"[step] contacts {total_contacts}, FP32/FP64 decision flips {flips}, "
        f"largest position gap at a flip {worst_flip:.3e}")
    assert total_contacts > 10_000
    assert worst_flip <= 1e-3        # bounded by one step's penetration depth
    sim.close()


def test_update_vertex_and_state_roundtrip(gpu):
    say("[step] update_vertex / set_state / pins")
    sc = S.small(side=16)
    sim = Simulator.from_scene(sc)
    sim.simulate(5)
    x, xp, c = sim.get_state()
    sim.update_vertex([0.25, 0.9, 0.1], 37)               # simulation.cu:253-257
    x2, xp2, _ = sim.get_state()
    np.testing.assert_array_equal(x2[37], np.float32([0.25, 0.9, 0.1]))
    np.testing.assert_array_equal(xp2[37], np.float32([0.25, 0.9, 0.1]))
    sim.set_state(x, xp, c)
    x3, xp3, c3 = sim.get_state()
    np.testing.assert_array_equal(x3, x)
    np.testing.assert_array_equal(xp3, xp)
    sim.check_finite()
    sim.close()
