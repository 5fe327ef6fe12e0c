"""The reference's own scene on the GPU (main.cpp:24-35: tshirt2.obj over
female.obj with its constants, from the committed fixture
tests/golden/ref_tshirt_female.npz): the collision outcome bit for bit
against the FP32 oracle given the GPU's pre-projection positions, and every
step within the per-step tolerance of the FP64 oracle (decision flips at
resting contacts excluded and bounded)."""
import os

import numpy as np
import pytest

import oracle as O
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_tshirt_female.npz")


def reference_scene():
    return S.load_mesh_scene(FIXTURE)


def test_reference_scene_contacts_and_steps(gpu):
    sc = reference_scene()
    extent = float(np.ptp(sc.pos, axis=0).max())
    tol = 1e-6 * extent
    say(f"[reference] tshirt2 ({sc.n} vertices, {sum(c.shape[0] for _, c in sc.springs)} springs) "
        f"over female ({sc.body.tris.shape[0]} triangles), extent {extent:.2f}")
    sim = Simulator.from_scene(sc)
    gv = sim.body_vertices()
    h = O.BvhHandle(O.bvh_build(gv, sc.body.tris, sc.params.box_depth))
    prm = O.params_from(sc.params)
    pins = np.zeros(sc.n, np.uint8)
    sim.set_debug(True)
    contacts, flips, worst, worst_flip = 0, 0, 0.0, 0.0
    for k in range(60):
        x, xp, c = sim.get_state()
        sim.simulate(1)
        pre, hit, ins = sim.get_debug()
        xo, xpo, co = sim.get_state()
        ohit, oins, opos, on = O.collide_points(h, pre)
        np.testing.assert_array_equal(hit, ohit, err_msg=f"step {k}")
        np.testing.assert_array_equal(ins, oins, err_msg=f"step {k}")
        np.testing.assert_array_equal(xo, opos, err_msg=f"step {k}")
        np.testing.assert_array_equal(co, on, err_msg=f"step {k}")
        contacts += int(ins.sum())
        x64, _, _, ins64, hit64 = O.step_f64(x, xp, c, sc.springs, prm, pins, h, with_hit=True)
        diff = np.abs(xo - x64).max(1)
        flip = (ins64 != ins) | ((hit64 != hit) & (ins == 1))
        flips += int(flip.sum())
        worst = max(worst, float(diff[~flip].max()))
        if flip.any():
            worst_flip = max(worst_flip, float(diff[flip].max()))
    say(f"[reference] contacts {contacts}, flips {flips}, worst |gpu - fp64| {worst:.3e} (tol {tol:.1e}), "
        f"worst at a flip {worst_flip:.3e}")
    assert contacts > 1000
    assert worst <= tol
    assert worst_flip <= 0.01 * extent
    assert np.isfinite(xo).all()
    sim.close()
