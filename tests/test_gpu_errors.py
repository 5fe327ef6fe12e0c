"""Error paths on the GPU, every condition forced from the host (no kernel
faults, traps or hangs):

  * a NaN or an infinity written into the state is reported by the finite
    check (CC_ERR_NONFINITE), and a clean state passes again;
  * a traversal whose deferred entries exceed the stack limit -- lowered by
    the test hook from its default of 64, which no valid tree reaches -- is
    recorded on the device and reported by the next check
    (CC_ERR_TRAVERSAL), once; at the default limit nothing is recorded and
    the results equal a simulator that never had its limit lowered;
  * the batch's check passes after a run of every instance.

The reference prints HIP errors and continues (common.h:4-10) and has no
overflow test in its traversal (cudaBRT.cu:243-245)."""
import numpy as np
import pytest

from simulator_amd import _lib as L
from simulator_amd import batch as B
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say

pytestmark = pytest.mark.gpu


def expect_error(fn, code):
    with pytest.raises(L.ClothCoreError) as e:
        fn()
    assert e.value.code == code, str(e.value)
    return str(e.value)


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_nonfinite_state_is_reported(gpu, bad):
    sc = S.small(side=16)
    sim = Simulator.from_scene(sc)
    sim.simulate(3)
    sim.check_finite()
    x, xp, c = sim.get_state()
    y = x.copy()
    y[37, 1] = bad
    sim.set_state(y, xp, c)
    msg = expect_error(sim.check_finite, L.CC_ERR_NONFINITE)
    say(f"[errors] {bad}: {msg}")
    sim.set_state(x, xp, c)
    sim.check_finite()
    sim.close()


@pytest.mark.parametrize("width", [4, 2])
def test_traversal_overflow_is_recorded_and_reported(gpu, width):
    # a 48x48 cloth resting on the 50,688-triangle capsule body after 150
    # steps: a CPU replay of the binary walk over this state finds ~200
    # queries that defer two or more nodes at once
    body = S.capsules(64, 14, 6, L.CC_BODY_STATIC)
    sc = S.grid_scene("stack", 48, 1e-3, 0, body, box_depth=0.002, y0=0.93)
    sim = Simulator.from_scene(sc)
    ref = Simulator.from_scene(sc)
    for s in (sim, ref):
        s.set_bvh_width(width)
        s.simulate(150)
    sim.check_finite()                    # the default limit: nothing recorded
    sim.set_stack_limit(1)
    sim.simulate(5)
    msg = expect_error(sim.check_finite, L.CC_ERR_TRAVERSAL)
    say(f"[errors] width {width}, stack limit 1: {msg}")
    sim.check_finite()                    # reported once, then cleared
    sim.set_stack_limit(64)
    ref.simulate(5)
    # back at the default limit, the run continues without a record; from
    # the same state it equals the simulator that was never lowered
    sim.set_state(*ref.get_state())
    sim.simulate(30)
    ref.simulate(30)
    sim.check_finite()
    for a, b in zip(sim.get_state(), ref.get_state()):
        np.testing.assert_array_equal(a, b)
    assert int((np.abs(ref.get_state()[2]).sum(1) > 0).sum()) > 100     # in contact throughout
    sim.close()
    ref.close()


def test_stack_limit_is_checked(gpu):
    sim = Simulator.from_scene(S.small(side=8))
    for bad in (0, 65, -3):
        expect_error(lambda: sim.set_stack_limit(bad), L.CC_ERR_INVALID)
    sim.close()


def test_batch_check_after_a_run(gpu):
    scs = B.instances(6, seed=4)
    b = B.Batch(scs)
    b.simulate(40)
    b.check()
    b.close()
