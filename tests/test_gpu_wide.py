"""The four-wide traversal form (wide.hip) against the binary walk of
D_BVH::intersect (cudaBRT.cu:239-285): two simulators on the same scene, one
walking each form, stay bit-identical step after step (positions, previous
positions, contact normals), for a static body, a body refitted every frame
and a body rebuilt every frame (its wide records folded on the device); the
per-point query gives the same leaf, inside flag and projection.  The binary
walk is itself pinned to the oracle by the other GPU tests."""
import numpy as np
import pytest

from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say

pytestmark = pytest.mark.gpu


def pair(sc):
    a, b = Simulator.from_scene(sc), Simulator.from_scene(sc)
    b.set_bvh_width(2)
    return a, b


def assert_same_run(sc, blocks, per_block):
    a, b = pair(sc)
    wa, nw, bound = a.bvh_width()
    assert wa == 4 and nw > 0 and bound <= 64, (wa, nw, bound)
    assert b.bvh_width()[0] == 2
    contacts = 0
    for blk in range(blocks):
        a.simulate(per_block)
        b.simulate(per_block)
        xa, pa, ca = a.get_state()
        xb, pb, cb = b.get_state()
        np.testing.assert_array_equal(xa, xb, err_msg=f"block {blk}")
        np.testing.assert_array_equal(pa, pb, err_msg=f"block {blk}")
        np.testing.assert_array_equal(ca, cb, err_msg=f"block {blk}")
        contacts += int(np.any(ca != 0, axis=1).sum())
    say(f"[wide] {sc.name}: {nw} wide nodes, deferral bound {bound}, {contacts} contacts over {blocks} blocks")
    assert contacts > 0
    a.close()
    b.close()


def test_static_body_wide_equals_binary(gpu):
    assert_same_run(S.small(side=64, y0=0.806), blocks=6, per_block=25)


def test_refitted_body_wide_equals_binary(gpu):
    sc = S.grid_scene("refit128", 128, 5e-4, 200, S.capsules(24, 12, 6, L.CC_BODY_REFIT), box_depth=0.002,
                      y0=0.975)
    assert_same_run(sc, blocks=5, per_block=40)


def test_rebuilt_body_wide_equals_binary(gpu):
    sc = S.grid_scene("rebuild128", 128, 5e-4, 200, S.capsules(24, 12, 6, L.CC_BODY_REBUILD), box_depth=0.002,
                      y0=0.975)
    m = sc.body.tris.shape[0]
    a = Simulator.from_scene(sc)
    assert a.bvh_width() == (4, m - 1, 64), a.bvh_width()    # one record per internal node
    a.close()
    assert_same_run(sc, blocks=5, per_block=40)


def test_karras_rebuild_path_equals_bottom_up(gpu, monkeypatch):
    # A rebuilt body beyond 2^21 triangles takes the Karras searches, the
    # refit and then the fold; CC_REBUILD_BOTTOM_UP=0 forces that path on a
    # small body: the same trees and the same run as the bottom-up build,
    # bit for bit, both walked through their folded records
    sc = S.grid_scene("rebuild96", 96, 5e-4, 200, S.capsules(24, 12, 6, L.CC_BODY_REBUILD), box_depth=0.002,
                      y0=0.975)
    monkeypatch.setenv("CC_REBUILD_BOTTOM_UP", "0")
    k = Simulator.from_scene(sc)
    monkeypatch.delenv("CC_REBUILD_BOTTOM_UP")
    b = Simulator.from_scene(sc)
    assert k.bvh_width() == b.bvh_width() == (4, sc.body.tris.shape[0] - 1, 64)
    contacts = 0
    for blk in range(4):
        k.simulate(40)
        b.simulate(40)
        for x, y in zip(k.get_state(), b.get_state()):
            np.testing.assert_array_equal(x, y, err_msg=f"block {blk}")
        tk, tb = k.bvh_export(), b.bvh_export()
        for key in tk:
            np.testing.assert_array_equal(tk[key], tb[key], err_msg=f"block {blk}: {key}")
        contacts += int(np.any(k.get_state()[2] != 0, axis=1).sum())
    say(f"[wide] Karras rebuild path: {contacts} contacts over 4 blocks, equal to the bottom-up build")
    assert contacts > 0
    k.close()
    b.close()


def test_queries_wide_equals_binary(gpu):
    sc = S.small(side=16)
    a, b = pair(sc)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.28, 0.32, size=(20000, 1))          # around the sphere's surface
    pts = (np.array([0.5, 0.5, 0.0]) + d * r).astype(np.float32)
    ra, rb = a.collide_points(pts), b.collide_points(pts)
    for x, y in zip(ra, rb):
        np.testing.assert_array_equal(x, y)
    assert int((ra[0] >= 0).sum()) > 1000
    a.close()
    b.close()


def test_width_switch_and_rebuild(gpu):
    sc = S.small(side=16)
    a = Simulator.from_scene(sc)
    a.set_bvh_width(2)
    a.set_bvh_width(4)
    a.bvh_rebuild()                 # a new topology: the wide table is made again
    w, nw, _ = a.bvh_width()
    assert w == 4 and nw > 0
    with pytest.raises(L.ClothCoreError):
        a.set_bvh_width(3)
    a.close()
    g = Simulator.from_scene(S.garment(side=16))
    w, nw, bound = g.bvh_width()    # rebuilt every frame: records folded on the device
    assert w == 4 and nw == g.bvh_size - 1 and bound == 64, (w, nw, bound)
    g.set_bvh_width(2)
    assert g.bvh_width()[0] == 2
    g.close()


def crossed_body(pairs, mode=L.CC_BODY_STATIC):
    """2 * pairs + 1 triangles whose boxes share the centre (0, 0, 0), so one
    Morton code and a tree interleaved by index: per pair one triangle flat in
    y and one flat in x, then one whose box is the cube [-1, 1]^3.  A point
    such as (0.5, 0.5, z) is in the box of every node that holds both kinds
    and in no leaf box but the last one's: the walk visits the whole tree,
    deferring three entries at most levels.  With 3000 pairs (6,001 leaves)
    the last leaf lies at the bottom of the rightmost path; an emulation of
    the four-wide walk over the oracle's tree visits 3,393 records for such a
    point, with up to 18 entries deferred at once.  (With 2^k + 1 leaves the
    last one would be the root's own leaf child, found at once.)  A rebuilt
    body gets one capsule over all its vertices and no motion."""
    v = []
    for _ in range(pairs):
        v += [(-1.0, -0.1, -1.0), (1.0, 0.1, -1.0), (0.0, 0.0, 1.0)]
        v += [(-0.1, -1.0, -1.0), (0.1, 1.0, -1.0), (0.0, 0.0, 1.0)]
    v += [(-1.0, -1.0, -1.0), (1.0, 1.0, -1.0), (0.0, 0.0, 1.0)]
    verts = np.array(v, np.float32)
    tris = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    caps = None
    if mode == L.CC_BODY_REBUILD:
        caps = {"vstart": np.array([0, len(v)], np.uint32), "p0": np.zeros((1, 3), np.float32),
                "p1": np.array([[0.0, 0.0, 1.0]], np.float32), "radius": np.ones(1, np.float32),
                "phase": np.zeros(1, np.float32)}
    return S.Body(verts, tris, np.zeros_like(verts), mode, False, caps)


@pytest.mark.parametrize("mode", [L.CC_BODY_STATIC, L.CC_BODY_REBUILD])
def test_queries_deep_stack_wide_equals_binary(gpu, mode):
    # the box depth is 0 so that a box's centre is its triangle's centroid
    sc = S.grid_scene("crossed", 8, 1e-3, 10, crossed_body(3000, mode), box_depth=0.0)
    a, b = pair(sc)
    w, nw, bound = a.bvh_width()
    # more deferrals than the LDS part of the stack holds (16 entries); a
    # rebuilt body reports the stack limit
    assert w == 4 and 16 < bound <= 64, (w, nw, bound)
    rng = np.random.default_rng(11)
    z = rng.uniform(-0.9, 0.9, 96)
    s = rng.choice([-0.5, 0.5], size=(96, 2))
    deep = np.stack([s[:, 0], s[:, 1], z], axis=1)            # the whole tree, then the last leaf
    near = np.stack([rng.uniform(-0.9, 0.9, 64), rng.uniform(-0.05, 0.05, 64),
                     rng.uniform(-0.9, 0.9, 64)], axis=1)     # in the flat-in-y boxes
    out = np.array([[2.0, 0.0, 0.0], [0.0, 0.0, 1.5], [-1.5, 0.5, 0.0]])
    pts = np.concatenate([deep, near, out]).astype(np.float32)
    say(f"[wide] crossed body: {nw} wide nodes, deferral bound {bound}, {pts.shape[0]} queries")
    ra, rb = a.collide_points(pts), b.collide_points(pts)
    for x, y in zip(ra, rb):
        np.testing.assert_array_equal(x, y)
    assert np.all(ra[0][:96] == ra[0][0]) and ra[0][0] >= 0    # every deep query ends on the same leaf
    assert np.all(ra[0][-3:] == -1)
    a.close()
    b.close()


def one_capsule(nv):
    """One capsule over all nv vertices and no motion: a rebuilt body that
    keeps its rest pose."""
    return {"vstart": np.array([0, nv], np.uint32), "p0": np.zeros((1, 3), np.float32),
            "p1": np.array([[0.0, 0.0, 1.0]], np.float32), "radius": np.ones(1, np.float32),
            "phase": np.zeros(1, np.float32)}


@pytest.mark.parametrize("case", ["one", "two", "three", "same300"])
def test_small_and_repeated_rebuilt_bodies(gpu, case):
    # The device fold at its edges: one triangle (no record), two and three
    # (records of leaves only, one split-free record with a pair), and 300
    # copies of one triangle (all Morton codes equal, the tree ordered by
    # index): the tree against the oracle's build, and the point queries
    # against the oracle and the binary walk, bit for bit
    from helpers import assert_bvh_equal, collider_sim, oracle_bvh
    import oracle as O
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.5], [0.5, 1.5, 0.2]], np.float32)
    tris = np.array({"one": [[0, 1, 2]], "two": [[0, 1, 2], [1, 3, 2]],
                     "three": [[0, 1, 2], [1, 3, 2], [2, 3, 4]], "same300": [[0, 1, 2]] * 300}[case], np.uint32)
    a = collider_sim(v, tris, mode=L.CC_BODY_REBUILD, caps=one_capsule(len(v)))
    b = collider_sim(v, tris, mode=L.CC_BODY_REBUILD, caps=one_capsule(len(v)))
    b.set_bvh_width(2)
    m = tris.shape[0]
    assert a.bvh_width()[:2] == ((4, m - 1) if m >= 2 else (2, 0)), a.bvh_width()
    d = oracle_bvh(a.body_vertices(), tris, 0.01)
    assert_bvh_equal(a.bvh_export(), d)
    rng = np.random.default_rng(17)
    pts = np.concatenate([rng.uniform([-0.2, -0.2, -0.1], [1.2, 1.6, 0.6], size=(3000, 3)),
                          np.array([[0.2, 0.2, -0.005], [0.9, 0.9, 0.45], [0.3, 0.3, 0.0]])]).astype(np.float32)
    got, alt = a.collide_points(pts), b.collide_points(pts)
    want = O.collide_points(O.BvhHandle(d), pts)
    for g, w, x in zip(got, want, alt):
        np.testing.assert_array_equal(g, w)
        np.testing.assert_array_equal(g, x)
    say(f"[wide] rebuilt body '{case}': {m} triangles, {int((got[0] >= 0).sum())} of {pts.shape[0]} queries hit")
    a.close()
    b.close()
