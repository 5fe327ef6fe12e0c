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


def test_adjoint_with_active_contacts(gpu):
    """A rollout that falls onto the sphere (32x32 cloth starting 1 cm above
    its top, 120 steps): the contact-projection adjoint runs.  The GPU
    adjoint against (1) the FP64 adjoint along the GPU's own trajectory and
    contact record, and (2) FP64 central differences of the rollout with the
    GPU's contacts held fixed as planes (normal and anchor per step and
    vertex) -- the function the adjoint differentiates."""
    nsteps = 120
    sc = S.small(side=32, y0=0.81, steps=nsteps)
    pins = np.zeros(sc.n, np.uint8)
    pins[sc.pins] = 1
    # the GPU trajectory with its contact record, one step at a time
    sim = Simulator.from_scene(sc)
    body = sim.body_vertices()
    d = oracle_bvh(body, sc.body.tris, sc.params.box_depth)
    sim.set_debug(True)
    tp = np.zeros((nsteps + 1, sc.n, 3))
    tq, tc, ta = np.zeros_like(tp), np.zeros_like(tp), np.zeros_like(tp)
    x, xp, c = sim.get_state()
    tp[0], tq[0], tc[0] = x, xp, c
    inside = 0
    for t in range(nsteps):
        sim.simulate(1)
        _, hit, ins = sim.get_debug()
        x, xp, c = sim.get_state()
        tp[t + 1], tq[t + 1], tc[t + 1] = x, xp, c
        on = ins == 1
        tri = sc.body.tris[d["order"][hit[on]]]
        ta[t + 1][on] = body[tri[:, 0]]                  # the plane's anchor: the hit triangle's v0
        inside += int(on.sum())
    sim.close()
    say(f"[adjoint] contacts: {inside} vertex-steps in contact over {nsteps} steps")
    assert inside > 500
    # the adjoint's forward is the same rollout
    sim2 = Simulator.from_scene(sc)
    adj = Adjoint(sim2, nsteps, 20)
    xf = adj.forward(sc.pos, None, THETA0)
    np.testing.assert_array_equal(xf, tp[nsteps].astype(np.float32))
    w = np.random.default_rng(7).standard_normal((sc.n, 3)).astype(np.float32)
    dth, dv0, _ = adj.backward(w)
    prm = O.params64_from(sc.params, k=THETA0[:3], damp=THETA0[3], rest_scale=THETA0[4])
    dth64, dv064, _ = O.adjoint_f64(sc.springs, prm, pins, tp, tq, tc, w.astype(np.float64))
    rel = np.abs(dth - dth64) / np.abs(dth64)
    for k in range(5):
        say(f"[adjoint]   contacts: dL/d{NAMES[k]:10s} gpu {dth[k]: .9e}  fp64 adjoint {dth64[k]: .9e}  "
            f"rel {rel[k]:.1e}")
    rv = np.linalg.norm(dv0 - dv064) / np.linalg.norm(dv064)
    say(f"[adjoint]   contacts: dL/dv0 relative error vs fp64 adjoint {rv:.1e}")
    assert np.all(rel <= 1e-3) and rv <= 1e-3
    # The frozen-contact rollout (the GPU's planes held fixed) in FP64:
    # central differences of it pin the FP64 adjoint along its own trajectory
    # to 1e-5, and give the GPU adjoint within 1e-3 of the gradient's norm.
    # (Per component the GPU adjoint differs from these differences by how
    # much the derivative moves between the FP32 and the FP64 trajectories --
    # up to ~1e-2 for the structural stiffness, whose gradient is a small sum
    # of near-cancelling stretches; the FP64 adjoint evaluated along the GPU's
    # trajectory shows the same difference, printed.)
    def frozen(theta, v0=None, trajectory=False):
        p = O.params64_from(sc.params, k=theta[:3], damp=theta[3], rest_scale=theta[4])
        return O.rollout_f64_frozen(sc.pos, v0, sc.springs, p, pins, tc, ta, trajectory=trajectory)
    xz, fp, fq = frozen(THETA0, trajectory=True)
    assert np.abs(xz - tp[nsteps]).max() < 1e-4
    dthz, dv0z, _ = O.adjoint_f64(sc.springs, prm, pins, fp, fq, tc, w.astype(np.float64))
    fd = np.zeros(5)
    for k in range(5):
        # relative step 1e-4: smaller steps drown the stiffness gradients
        # (1e-6 of the others) in the rollout's FP64 roundoff
        hk = 1e-4 * abs(THETA0[k])
        a, b = THETA0.copy(), THETA0.copy()
        a[k] += hk
        b[k] -= hk
        fd[k] = float(np.sum(w * (frozen(a) - frozen(b)))) / (2 * hk)
        say(f"[adjoint]   contacts: d(w.x)/d{NAMES[k]:10s} FD {fd[k]: .9e}  fp64 adjoint (frozen) {dthz[k]: .9e} "
            f"rel {abs(dthz[k] - fd[k]) / abs(fd[k]):.1e}  gpu {dth[k]: .9e} rel {abs(dth[k] - fd[k]) / abs(fd[k]):.1e}"
            f"  (fp64 adjoint on the gpu trajectory: rel {abs(dth64[k] - fd[k]) / abs(fd[k]):.1e})")
        assert abs(dthz[k] - fd[k]) <= 1e-5 * abs(fd[k])
    gerr = np.linalg.norm(dth - fd) / np.linalg.norm(fd)
    say(f"[adjoint]   contacts: |gpu - FD| / |FD| over the five parameters {gerr:.1e}")
    assert gerr <= 1e-3
    r = np.random.default_rng(9).standard_normal((sc.n, 3))
    hv = 1e-4
    fdv = float(np.sum(w * (frozen(THETA0, hv * r) - frozen(THETA0, -hv * r)))) / (2 * hv)
    say(f"[adjoint]   contacts: dL/dv0 . r  FD {fdv: .9e}  fp64 adjoint (frozen) {float(np.sum(dv0z * r)): .9e}"
        f"  gpu {float(np.sum(dv0 * r)): .9e}")
    assert abs(float(np.sum(dv0z * r)) - fdv) <= 1e-5 * abs(fdv)
    assert abs(float(np.sum(dv0 * r)) - fdv) <= 1e-3 * abs(fdv)
    adj.close()
    sim2.close()
