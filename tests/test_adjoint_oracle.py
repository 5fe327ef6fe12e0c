"""The FP64 adjoint of the oracle against FP64 central finite differences.

This pins the adjoint's mathematics (spring-force, damping and Verlet
Jacobians, pins, the contact-force projection and the plane projection of
contacts held fixed) before the GPU adjoint is compared with it."""
import os
import sys

import numpy as np
import pytest

import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as G  # noqa: E402

THETA0 = (1000.0, 500.0, 100.0, -0.02, 1.0)     # k0, k1, k2, damp, rest_scale


def setup(side, y0, level, nsteps):
    pos, springs, pins = G.cloth(side, y0)
    bvh = None
    if level is not None:
        v, t = G.sphere(level)
        bvh = O.BvhHandle(O.bvh_build(v, t, G.DEPTH))
    rng = np.random.default_rng(3)
    w = rng.standard_normal((pos.shape[0], 3))       # L = sum(w * x_final)
    return pos.astype(np.float64), springs, pins, bvh, w


def loss(theta, v0, pos, springs, pins, bvh, w, nsteps, frozen=None):
    """L = sum(w * x_final); with `frozen` = (traj_c, traj_o) the contacts of
    a recorded rollout are held fixed (the function the adjoint
    differentiates).  Held fixed because a vertex resting on the body has its
    normal force removed and no velocity, so its next plane distance is zero
    up to rounding: the contact set of the true rollout changes under any
    perturbation, however small."""
    prm = O.params64_from(G.params(), k=theta[:3], damp=theta[3], rest_scale=theta[4])
    if frozen is not None:
        xf = O.rollout_f64_frozen(pos, v0, springs, prm, pins, *frozen)
    else:
        xf = O.rollout_f64(pos, v0, springs, prm, pins, bvh, nsteps)
    return float(np.sum(w * xf))


@pytest.mark.parametrize("case", ["springs only", "falling into contact"])
def test_fp64_adjoint_matches_central_differences(case):
    side, y0, level, nsteps = (12, 1.0, None, 60) if case == "springs only" else (16, 0.83, 2, 120)
    pos, springs, pins, bvh, w = setup(side, y0, level, nsteps)
    theta = np.array(THETA0)
    v0 = np.zeros_like(pos)
    prm = O.params64_from(G.params(), k=theta[:3], damp=theta[3], rest_scale=theta[4])
    xf, tp, tq, tc, to = O.rollout_f64(pos, v0, springs, prm, pins, bvh, nsteps, trajectory=True)
    ncontact = int((np.abs(tc).sum(2) > 0).sum())
    dp, dv0, dx0 = O.adjoint_f64(springs, prm, pins, tp, tq, tc, w)
    frozen = (tc, to) if level is not None else None
    # the frozen rollout at the recorded parameters is the rollout itself
    np.testing.assert_array_equal(O.rollout_f64_frozen(pos, v0, springs, prm, pins, tc, to), xf)
    # Relative steps of 1e-4 keep the FD truncation error below 1e-6 and its
    # roundoff (about 1e-16 |L| / h) far below the tolerance (smaller steps
    # drown gradients as small as the bend stiffness's in roundoff).
    for k in range(5):
        for rel, tol in ((1e-4, 1e-5),):
            h = rel * abs(theta[k])
            tp_, tm_ = theta.copy(), theta.copy()
            tp_[k] += h
            tm_[k] -= h
            fd = (loss(tp_, v0, pos, springs, pins, bvh, w, nsteps, frozen) -
                  loss(tm_, v0, pos, springs, pins, bvh, w, nsteps, frozen)) / (2 * h)
            assert abs(dp[k] - fd) <= tol * abs(fd), (case, k, rel, dp[k], fd)
    rng = np.random.default_rng(5)
    # pinned vertices included: their initial x_prev enters their
    # neighbours' spring damping in the first step
    r = rng.standard_normal(pos.shape)
    h = 1e-4
    fd = (loss(theta, v0 + h * r, pos, springs, pins, bvh, w, nsteps, frozen) -
          loss(theta, v0 - h * r, pos, springs, pins, bvh, w, nsteps, frozen)) / (2 * h)
    assert abs(np.sum(dv0 * r) - fd) <= 1e-5 * abs(fd), (case, np.sum(dv0 * r), fd)
    if level is not None:
        assert ncontact > 100            # the contact path was exercised
