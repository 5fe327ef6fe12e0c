"""Error paths on the GPU, every condition forced from the host (no kernel
faults, traps or hangs):

  * a NaN or an infinity written into the state is reported by the finite
    check (CC_ERR_NONFINITE), and a clean state passes again;
  * a traversal whose deferred entries exceed the stack limit -- lowered by
    the test hook from its default of 64, which no valid tree reaches -- is
    recorded on the device and reported by the next check
    (CC_ERR_TRAVERSAL), once; at the default limit nothing is recorded and
    the results equal a simulator that never had its limit lowered;
  * the batch's check passes after a run of every instance;
  * an allocation made to fail (cc_fault_inject) at any point of a creation
    is reported as CC_ERR_ALLOC and the failed creation frees what it had
    made; a launch check made to fail is reported as CC_ERR_HIP, the failed
    simulator can be destroyed, and a fresh one reproduces an undisturbed
    run bit for bit.

The reference prints HIP errors and continues (common.h:4-10) and has no
overflow test in its traversal (cudaBRT.cu:243-245)."""
import ctypes as C

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


@pytest.mark.parametrize("width,mode", [(4, L.CC_BODY_STATIC), (2, L.CC_BODY_STATIC), (4, L.CC_BODY_REBUILD)])
def test_traversal_overflow_is_recorded_and_reported(gpu, width, mode):
    # a 48x48 cloth resting on the 50,688-triangle capsule body after 150
    # steps: a CPU replay of the binary walk over this state finds ~200
    # queries that defer two or more nodes at once (a rebuilt body: the
    # wide records folded on the device every frame)
    body = S.capsules(64, 14, 6, mode)
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
    say(f"[errors] width {width}, body mode {mode}, stack limit 1: {msg}")
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


def _free_device_bytes():
    # the HIP runtime that the library itself loaded (another copy of the
    # runtime in the process would have its own state)
    L.lib()
    with open("/proc/self/maps") as f:
        paths = sorted({line.split()[-1] for line in f if "libamdhip64" in line},
                       key=lambda q: not q.startswith("/opt/rocm"))
    path = paths[0]
    hip = C.CDLL(path)
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _moving_scene():
    body = S.capsules(64, 14, 6, L.CC_BODY_REFIT)
    return S.grid_scene("moving", 24, 1e-3, 0, body, box_depth=0.002, y0=0.93)


def _fail_nth(kind, k, make):
    """make() with the k-th operation of `kind` failing; the object, or the
    error it raised."""
    L.fault_inject(kind, k)
    try:
        return make()
    except L.ClothCoreError as e:
        return e
    finally:
        L.fault_inject(kind, -1)


def test_allocation_failure_at_every_point_of_a_creation(gpu):
    sc = _moving_scene()
    Simulator.from_scene(sc).close()          # runtime set-up outside the count
    base = _free_device_bytes()
    k = 0
    while True:
        r = _fail_nth(L.CC_FAULT_ALLOC, k, lambda: Simulator.from_scene(sc))
        if not isinstance(r, L.ClothCoreError):
            break
        assert r.code == L.CC_ERR_ALLOC and "injected" in str(r), str(r)
        k += 1
        assert k < 2000, "the creation never completed"
    say(f"[errors] a moving-body creation makes {k} device allocations; each one failed in turn")
    sim = r
    sim.simulate(30)                          # the creation that got through works
    sim.check_finite()
    sim.close()
    lost = base - _free_device_bytes()
    say(f"[errors] device memory after {k} failed creations: {lost / 2**20:.1f} MiB below the start")
    assert lost < 64 * 2**20                  # a leaked creation would hold hundreds of MiB


def test_launch_failure_is_reported_and_a_fresh_simulator_recovers(gpu):
    sc = _moving_scene()
    r = _fail_nth(L.CC_FAULT_LAUNCH, 0, lambda: Simulator.from_scene(sc))
    assert isinstance(r, L.ClothCoreError) and r.code == L.CC_ERR_HIP and "injected" in str(r), r
    ref = Simulator.from_scene(sc)
    ref.simulate(40)
    for k in (0, 5, 40):                      # the 1st, 6th and 41st launch check of the call
        sim = Simulator.from_scene(sc)
        sim.simulate(2)
        r = _fail_nth(L.CC_FAULT_LAUNCH, k, lambda: sim.simulate(38))
        assert isinstance(r, L.ClothCoreError) and r.code == L.CC_ERR_HIP, r
        say(f"[errors] launch check {k} of a 38-step call failed: {r}")
        sim.close()
        again = Simulator.from_scene(sc)
        again.simulate(40)
        for a, b in zip(again.get_state(), ref.get_state()):
            np.testing.assert_array_equal(a, b)
        again.close()
    ref.close()


def test_batch_allocation_failure(gpu):
    scs = B.instances(4, seed=2)
    r = _fail_nth(L.CC_FAULT_ALLOC, 3, lambda: B.Batch(scs))
    assert isinstance(r, L.ClothCoreError) and r.code == L.CC_ERR_ALLOC, r
    b = B.Batch(scs)
    b.simulate(10)
    b.check()
    b.close()
