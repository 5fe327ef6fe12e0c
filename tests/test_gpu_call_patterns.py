"""Any pattern of calls gives the results of one call per frame, bit for
bit: runs of cc_sim_step of every length (single steps, whole step graphs,
runs that end inside a block), interleaved with synchronisation, state
reads, collision queries (allocating and allocation-free), timing switched
on and off, and the frame set back, against a simulator stepped once per
frame through the same frame changes.  Moving bodies, refitted and rebuilt:
the update lanes run up to a dozen frames (single steps) or a block (graphs)
ahead of the step, and every step must still see the pose of its own frame."""
import numpy as np
import pytest

from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say
from test_zz_no_alloc import noalloc_query

pytestmark = pytest.mark.gpu

LENGTHS = (1, 2, 3, 5, 7, 23, 24, 25, 30, 48, 49)


def _scene(mode):
    body = S.capsules(64, 14, 6, mode)
    return S.grid_scene("pattern", 32, 1e-3, 0, body, box_depth=0.002, y0=0.93)


def _same_state(a, b, label):
    for x, y, name in zip(a.get_state(), b.get_state(), ("x_cur", "x_prev", "contact")):
        np.testing.assert_array_equal(x, y, err_msg=f"{label}: {name}")


@pytest.mark.parametrize("mode", [L.CC_BODY_REFIT, L.CC_BODY_REBUILD])
def test_any_call_pattern_equals_one_call_per_frame(gpu, mode):
    sc = _scene(mode)
    rng = np.random.default_rng(20 + mode)
    a = Simulator.from_scene(sc)
    b = Simulator.from_scene(sc)
    timing = False
    done = []
    for it in range(24):
        if it == 12:                                       # the frame set back
            f = max(1, a.frame - 37)
            a.set_frame(f)
            b.set_frame(f)
            done.append(f"set_frame({f})")
        k = int(rng.choice(LENGTHS))
        a.simulate(k)
        for _ in range(k):
            b.simulate(1)
        act = int(rng.integers(0, 5))
        label = f"{mode} call {it}: simulate({k}) -> frame {a.frame}"
        if act == 0:
            a.synchronize()
        elif act == 1:
            _same_state(a, b, label)
        elif act in (2, 3):
            x = b.get_state()[0]
            pts = x[rng.choice(x.shape[0], 512, replace=False)]
            want = b.collide_points(pts)
            got = a.collide_points(pts) if act == 2 else noalloc_query(a, pts)[:4]
            for g, w, name in zip(got, want, ("hit", "inside", "position", "normal")):
                np.testing.assert_array_equal(g, w, err_msg=f"{label}: query {name}")
        else:
            timing = not timing
            a.set_timing(timing)
        done.append(f"{k}:{act}")
        assert a.frame == b.frame
    a.check_finite()
    _same_state(a, b, f"{mode} at the end")
    in_contact = int((np.abs(b.get_state()[2]).sum(1) > 0).sum())
    say(f"[patterns] mode {mode}: {' '.join(done)}; frame {a.frame}, {in_contact} vertices in contact")
    assert in_contact > 0
    a.close()
    b.close()


@pytest.mark.parametrize("mode", [L.CC_BODY_REFIT, L.CC_BODY_REBUILD])
@pytest.mark.parametrize("lanes,frames", [(1, 1), (2, 2), (3, 1), (4, 4)])
def test_lane_and_chain_settings_change_no_result(gpu, mode, lanes, frames, monkeypatch):
    # CC_BODY_LANES / CC_BUILD_FRAMES reshape the update pipeline (how many
    # streams, how many frames per chain of launches); graph blocks, runs
    # that end inside a block and single steps give the default's results
    sc = _scene(mode)
    ref = Simulator.from_scene(sc)
    monkeypatch.setenv("CC_BODY_LANES", str(lanes))
    monkeypatch.setenv("CC_BUILD_FRAMES", str(frames))
    sim = Simulator.from_scene(sc)
    monkeypatch.delenv("CC_BODY_LANES")
    monkeypatch.delenv("CC_BUILD_FRAMES")
    for n in (48, 7, 1, 1, 30, 24, 5):
        ref.simulate(n)
        sim.simulate(n)
        _same_state(sim, ref, f"lanes {lanes}, frames {frames}, mode {mode}, frame {sim.frame}")
    say(f"[patterns] {lanes} lanes of {frames} frames, mode {mode}: the default's {sim.frame} frames")
    ref.close()
    sim.close()
