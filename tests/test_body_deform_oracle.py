"""The oracle's per-frame body deformation (or_body_deform) pinned by what
the mathematics fixes, independently of its arithmetic:

  identity   R = I, T = 0, breath = 1: the rest pose, to rounding
  breathing  R = I, T = 0, breath[k] per capsule: every vertex keeps its
             closest point on its capsule's axis segment and its distance to
             that segment scales by its own capsule's breath[k]
  rigid      breath = 1, a rotation about +y and a translation: pairwise
             distances are kept, and the result is R @ X + T

A vertex assigned to the wrong capsule, a dropped or sign-flipped term, an
unclamped axis parameter or a transposed rotation fails one of them."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S


@pytest.fixture(scope="module")
def body():
    return S.capsules(24, 8, 4, L.CC_BODY_REFIT, seed=3)


def deform(b, breath, R, T):
    c = b.caps
    return O.body_deform(b.verts, c["vstart"], c["p0"], c["p1"], np.asarray(breath, np.float32),
                         np.asarray(R, np.float32).ravel(), np.asarray(T, np.float32))


def capsule_of(b):
    vs = b.caps["vstart"]
    return np.searchsorted(vs, np.arange(b.verts.shape[0]), side="right") - 1


def axis_point(X, p0, p1):
    """Closest point of segment [p0, p1] to each row of X (float64)."""
    d = p1 - p0
    t = np.clip(np.einsum("ij,ij->i", X - p0, d) / np.einsum("ij,ij->i", d, d), 0.0, 1.0)
    return p0 + t[:, None] * d, t


def test_identity_motion_returns_the_rest_pose(body):
    nc = len(body.caps["radius"])
    out = deform(body, np.ones(nc), np.eye(3), np.zeros(3))
    # a + (X - a) rounds twice: at most one unit in the last place of X
    ulp = np.spacing(np.abs(body.verts).max())
    assert np.abs(out - body.verts).max() <= ulp
    assert np.mean(out == body.verts) > 0.9


def test_breathing_scales_the_distance_to_the_capsule_axis(body):
    c = body.caps
    nc = len(c["radius"])
    breath = np.float32(1.0) + np.float32(0.02) * np.linspace(-1, 1, nc).astype(np.float32)
    out = deform(body, breath, np.eye(3), np.zeros(3)).astype(np.float64)
    k = capsule_of(body)
    p0 = c["p0"].reshape(-1, 3)[k].astype(np.float64)
    p1 = c["p1"].reshape(-1, 3)[k].astype(np.float64)
    X = body.verts.astype(np.float64)
    a_in, t_in = axis_point(X, p0, p1)
    a_out, t_out = axis_point(out, p0, p1)
    r_in = np.linalg.norm(X - a_in, axis=1)
    r_out = np.linalg.norm(out - a_out, axis=1)
    np.testing.assert_allclose(a_out, a_in, atol=2e-6)              # the same axis point
    np.testing.assert_allclose(r_out, breath[k].astype(np.float64) * r_in, rtol=2e-5, atol=2e-7)
    # the rest pose lies on each capsule's surface, and the breathing differs
    # from capsule to capsule, so a vertex given another capsule's factor fails
    np.testing.assert_allclose(r_in, c["radius"][k], rtol=1e-4)
    assert np.ptp(breath) > 0.03


@pytest.mark.parametrize("theta", [0.3, -1.1])
def test_rigid_motion_keeps_distances(body, theta):
    nc = len(body.caps["radius"])
    R = np.array([[np.cos(theta), 0, np.sin(theta)], [0, 1, 0], [-np.sin(theta), 0, np.cos(theta)]])
    T = np.array([0.05, -0.02, 0.11])
    out = deform(body, np.ones(nc), R, T).astype(np.float64)
    X = body.verts.astype(np.float64)
    rng = np.random.default_rng(4)
    i, j = rng.integers(0, X.shape[0], size=(2, 4000))
    np.testing.assert_allclose(np.linalg.norm(out[i] - out[j], axis=1), np.linalg.norm(X[i] - X[j], axis=1),
                               atol=1e-6)
    R32 = R.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(out, X @ R32.T + T.astype(np.float32), atol=1e-6)
    # a transposed rotation moves points by far more than that
    assert np.abs(out - (X @ R32 + T.astype(np.float32))).max() > 1e-2


def test_empty_capsule_ranges_are_skipped():
    """A capsule with no vertices (vstart[k] == vstart[k+1]): the vertices
    after it belong to the next non-empty capsule."""
    X = np.array([[0, 1, 0], [0, 2, 0], [5, 0.5, 0]], np.float32)
    vstart = np.array([0, 2, 2, 3], np.uint32)          # capsule 1 is empty
    p0 = np.array([[0, 0, 0], [9, 9, 9], [5, 0, 0]], np.float32)
    p1 = np.array([[0, 0, 1], [9, 9, 8], [5, 0, 1]], np.float32)
    breath = np.array([2.0, 7.0, 3.0], np.float32)
    out = O.body_deform(X, vstart, p0, p1, breath, np.eye(3, dtype=np.float32).ravel(), np.zeros(3, np.float32))
    np.testing.assert_array_equal(out, np.float32([[0, 2, 0], [0, 4, 0], [5, 1.5, 0]]))
