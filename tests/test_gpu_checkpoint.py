"""Checkpoint and resume (cc_sim_save / cc_sim_load): a simulator restored
from a checkpoint into a fresh simulator of the same cloth steps exactly as
the saved one, bit for bit -- state, contact normals, pins, parameters and
frame (the moving body's pose follows from the frame).  Files of another
cloth, cut short or not checkpoints are refused.  The reference's
Simulator::save is empty (simulation.cu:233-235)."""
import numpy as np
import pytest

from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say

pytestmark = pytest.mark.gpu


def _scene(mode):
    if mode == L.CC_BODY_STATIC:
        return S.small(side=32)
    body = S.capsules(64, 14, 6, mode)
    return S.grid_scene("ckpt", 32, 1e-3, 0, body, box_depth=0.002, y0=0.93)


@pytest.mark.parametrize("mode", [L.CC_BODY_STATIC, L.CC_BODY_REFIT, L.CC_BODY_REBUILD])
def test_resume_equals_uninterrupted_run(gpu, tmp_path, mode):
    sc = _scene(mode)
    path = str(tmp_path / "cloth.ckpt")
    a = Simulator.from_scene(sc)
    a.set_pins(np.array([0, 5, 31], np.uint32))
    p = a.params                          # the scene's parameters, two of them changed
    p.k0, p.damp = 900.0, -0.03
    a.set_params(p)
    a.simulate(170)
    a.save(path)
    a.simulate(50)
    b = Simulator.from_scene(sc)          # its own pins and parameters, frame 0
    b.simulate(3)
    b.load(path)
    assert b.frame == 170
    b.simulate(50)
    for x, y, name in zip(a.get_state(), b.get_state(), ("x_cur", "x_prev", "contact")):
        np.testing.assert_array_equal(x, y, err_msg=name)
    in_contact = int((np.abs(a.get_state()[2]).sum(1) > 0).sum())
    say(f"[ckpt] mode {mode}: resumed at frame 170, equal after 50 more steps ({in_contact} in contact)")
    a.close()
    b.close()


def test_other_cloth_and_bad_files_are_refused(gpu, tmp_path):
    sc = _scene(L.CC_BODY_STATIC)
    a = Simulator.from_scene(sc)
    a.simulate(10)
    good = tmp_path / "a.ckpt"
    a.save(str(good))
    other = Simulator.from_scene(S.small(side=16))
    cut = tmp_path / "cut.ckpt"
    cut.write_bytes(good.read_bytes()[:-100])
    junk = tmp_path / "junk.ckpt"
    junk.write_bytes(b"not a checkpoint" * 64)
    for sim, path in ((other, good), (a, cut), (a, junk), (a, tmp_path / "missing.ckpt")):
        with pytest.raises(L.ClothCoreError) as e:
            sim.load(str(path))
        assert e.value.code == L.CC_ERR_INVALID, str(e.value)
    x = a.get_state()[0]
    a.load(str(good))                          # the refused loads changed nothing
    np.testing.assert_array_equal(a.get_state()[0], x)
    a.close()
    other.close()
