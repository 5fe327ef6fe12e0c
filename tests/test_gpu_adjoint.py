"""The GPU adjoint (cc_adj_*) on the differentiable workload: 128x128 cloth
over the 5120-triangle sphere, 200 steps, checkpoints every 20 steps; target
from hidden parameters (1200/450/80, damping 0.03, rest scale 0.98), loss
the MSE of the final positions from the start parameters (1000/500/100,
0.02, 1.0).  Checked against the FP64 CPU adjoint of the oracle (same
dL/dx) and FP64 central finite differences."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import scene as S
from simulator_amd.adjoint import Adjoint
from simulator_amd.simulator import Simulator

from helpers import oracle_bvh, say

pytestmark = pytest.mark.gpu

THETA0 = np.array([1000.0, 500.0, 100.0, -0.02, 1.0])
HIDDEN = np.array([1200.0, 450.0, 80.0, -0.03, 0.98])
NAMES = ("k_struct", "k_shear", "k_bend", "damp", "rest_scale")


@pytest.fixture(scope="module")
def c4(gpu):
    sc = S.adjoint_scene()
    sim = Simulator.from_scene(sc)
    pins = np.zeros(sc.n, np.uint8)
    pins[sc.pins] = 1
    h = O.BvhHandle(oracle_bvh(sim.body_vertices(), sc.body.tris, sc.params.box_depth))
    yield sc, sim, pins, h
    sim.close()


def fp64_loss(sc, pins, h, theta, target, v0=None):
    prm = O.params64_from(sc.params, k=theta[:3], damp=theta[3], rest_scale=theta[4])
    xf = O.rollout_f64(sc.pos, v0, sc.springs, prm, pins, h, 200)
    return float(np.mean((xf - target) ** 2)), xf


def linear(sc, pins, h, theta, g):
    prm = O.params64_from(sc.params, k=theta[:3], damp=theta[3], rest_scale=theta[4])
    return float(np.sum(g * O.rollout_f64(sc.pos, None, sc.springs, prm, pins, h, 200)))


def test_adjoint_matches_fp64_adjoint_and_central_differences(c4):
    sc, sim, pins, h = c4
    say("[adjoint] 128x128 cloth, 200 steps, checkpoints every 20")
    adj = Adjoint(sim, 200, 20)
    target = adj.forward(sc.pos, None, HIDDEN)
    xf = adj.forward(sc.pos, None, THETA0)
    g = (2.0 * (xf.astype(np.float64) - target) / xf.size).astype(np.float32)
    dth, dv0, dx0 = adj.backward(g)
    # FP64 oracle adjoint along the FP64 trajectory, fed the same dL/dx
    prm = O.params64_from(sc.params, k=THETA0[:3], damp=THETA0[3], rest_scale=THETA0[4])
    x64, tp, tq, tc, _ = O.rollout_f64(sc.pos, None, sc.springs, prm, pins, h, 200, trajectory=True)
    assert not np.any(tc), "this workload has no contacts (the cloth does not reach the sphere)"
    say(f"[adjoint]   max |x_gpu - x_fp64| = {np.abs(xf - x64).max():.2e}")
    dth64, dv064, _ = O.adjoint_f64(sc.springs, prm, pins, tp, tq, tc, g.astype(np.float64))
    rel = np.abs(dth - dth64) / np.abs(dth64)
    for k in range(5):
        say(f"[adjoint]   dL/d{NAMES[k]:10s} gpu {dth[k]: .9e}  fp64 adjoint {dth64[k]: .9e}  rel {rel[k]:.1e}")
    rv = np.linalg.norm(dv0 - dv064) / np.linalg.norm(dv064)
    say(f"[adjoint]   dL/dv0 relative error vs fp64 adjoint {rv:.1e}")
    assert np.all(rel <= 1e-3) and rv <= 1e-3
    # FP64 central differences (h = 1e-6 |theta|) of the vector-Jacobian
    # functional L_g(theta) = sum(g * x_fp64(theta)): the quantity both
    # adjoints compute for the dL/dx g
    for k in range(5):
        hk = 1e-6 * abs(THETA0[k])
        a, b = THETA0.copy(), THETA0.copy()
        a[k] += hk
        b[k] -= hk
        fd = (linear(sc, pins, h, a, g) - linear(sc, pins, h, b, g)) / (2 * hk)
        say(f"[adjoint]   d(g.x)/d{NAMES[k]:10s} FD {fd: .9e}  gpu {dth[k]: .9e}  rel {abs(dth[k] - fd) / abs(fd):.1e}")
        assert abs(dth[k] - fd) <= 1e-3 * abs(fd)
    # the MSE loss itself: FP64 adjoint vs FP64 central differences.  (The
    # GPU's MSE gradient differs from this FD by the difference of its FP32
    # residuals -- trajectories 8e-5 apart -- not by its adjoint; printed.)
    t64 = fp64_loss(sc, pins, h, HIDDEN, 0.0)[1]
    l0, x0_64 = fp64_loss(sc, pins, h, THETA0, t64)
    g64 = 2.0 * (x0_64 - t64) / x0_64.size
    dthc, _, _ = O.adjoint_f64(sc.springs, prm, pins, tp, tq, tc, g64)
    for k in range(5):
        hk = 1e-6 * abs(THETA0[k])
        a, b = THETA0.copy(), THETA0.copy()
        a[k] += hk
        b[k] -= hk
        fd = (fp64_loss(sc, pins, h, a, t64)[0] - fp64_loss(sc, pins, h, b, t64)[0]) / (2 * hk)
        say(f"[adjoint]   dMSE/d{NAMES[k]:10s} FD {fd: .9e}  fp64 adjoint {dthc[k]: .9e}  "
            f"(gpu, FP32 residuals: {dth[k]: .9e})")
        assert abs(dthc[k] - fd) <= 1e-3 * abs(fd)
    adj.close()


def test_checkpoint_interval_does_not_change_the_gradient(c4):
    sc, sim, pins, h = c4
    rng = np.random.default_rng(1)
    v0 = (0.01 * rng.standard_normal(sc.pos.shape)).astype(np.float32)
    g = rng.standard_normal(sc.pos.shape).astype(np.float32)
    out = []
    for ck in (20, 7, 200):
        adj = Adjoint(sim, 60, ck)
        adj.forward(sc.pos, v0, THETA0)
        out.append(adj.backward(g))
        adj.close()
    for o in out[1:]:
        for a, b in zip(out[0], o):
            np.testing.assert_array_equal(a, b)


def test_torch_function_gives_the_adjoint(c4):
    torch = pytest.importorskip("torch")
    from simulator_amd.adjoint import rollout
    sc, sim, pins, h = c4
    adj = Adjoint(sim, 40, 20)
    target = torch.from_numpy(adj.forward(sc.pos, None, HIDDEN))
    theta = torch.tensor(THETA0, dtype=torch.float64, requires_grad=True)
    v0 = torch.zeros(sc.n, 3, dtype=torch.float32, requires_grad=True)
    x = rollout(adj, theta, v0)
    loss = ((x - target) ** 2).mean()
    (gx,) = torch.autograd.grad(loss, x, retain_graph=True)     # what autograd feeds the Function
    loss.backward()
    dth, dv0, _ = adj.backward(gx.numpy())
    np.testing.assert_array_equal(theta.grad.numpy(), dth)
    np.testing.assert_array_equal(v0.grad.numpy(), dv0)
    assert theta.grad.dtype == torch.float64 and np.all(np.isfinite(dth)) and np.any(dth != 0)
    adj.close()
