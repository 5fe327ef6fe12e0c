"""The four-wide form of a rebuilt body (k_fold_grouped, wide.hip, and the
split records of wide_visit, cc_device.h), restated step for step in Python
over the oracle's tree, against the oracle's binary first-hit walk
(or_query, which restates D_BVH::intersect, cudaBRT.cu:239-285): the same
leaf for every query point, and never more deferred entries than the tree
is deep (the binary walk's own bound), on the capsule body, on random
triangle soups with repeated Morton codes, and on a body built to make the
walk visit the whole tree.  The kernels themselves are checked on the GPU
(test_gpu_wide.py: the wide walk equals the binary walk step for step on a
rebuilt body; test_gpu_configs.py: the garment's contacts equal the
oracle's)."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S

EMPTY = (np.array([np.inf] * 3 + [-np.inf] * 3, np.float32), 0)


def decode(c):
    """reference encoding -> library link (internal index >= 0, leaf ~l)"""
    return ~(-c - 1) if c < 0 else c - 1


def fold(d):
    """k_fold_grouped: one record per internal node x, (entries, split)."""
    mi = d["m"] - 1
    cA = [decode(int(c)) for c in d["childA"]]
    cB = [decode(int(c)) for c in d["childB"]]

    def box(link):
        return d["box_leaf"][~link] if link < 0 else d["box_int"][link]

    def pair(y):                      # y's children, a leaf child first, A before B
        a, b = cA[y], cB[y]
        if a >= 0 and b < 0:
            return [(box(b), b), (box(a), a)]
        return [(box(a), a), (box(b), b)]

    recs = []
    for x in range(mi):
        a, b = cA[x], cB[x]
        if a < 0 and b < 0:
            recs.append(([(box(a), a), (box(b), b), EMPTY, EMPTY], 0))
        elif a < 0:
            recs.append(([(box(a), a)] + pair(b) + [EMPTY], 0))
        elif b < 0:
            recs.append(([(box(b), b)] + pair(a) + [EMPTY], 0))
        else:
            recs.append((pair(a) + pair(b), b))
    return recs


def in_box(b, p):
    return b[0] <= p[0] <= b[3] and b[1] <= p[1] <= b[4] and b[2] <= p[2] <= b[5]


def walk(recs, p):
    """wide_first_hit over split records: (leaf or -1, visits, most deferred)."""
    stack, node, visits, most = [], 0, 0, 0
    while True:
        visits += 1
        ents, split = recs[node]
        ins = [in_box(b, p) for b, _ in ents]
        links = [l for _, l in ents]
        if split != 0 and (ins[0] or ins[1]):
            # B's record when both of B's children hold p, else the one that does
            b_next = split if (ins[2] and ins[3]) else (links[2] if ins[2] else links[3])
            ins[2], links[2], ins[3] = ins[2] or ins[3], b_next, False
        cand = [e for e in range(4) if ins[e]]
        if not cand:
            if not stack:
                return -1, visits, most
            node = stack.pop()
            if node < 0:
                return ~node, visits, most
            continue
        c = links[cand[0]]
        if c < 0:
            return ~c, visits, most
        for e in reversed(cand[1:]):  # the lowest later entry on top
            stack.append(links[e])
        most = max(most, len(stack))
        node = c


def depth_of(d):
    cA = [decode(int(c)) for c in d["childA"]]
    cB = [decode(int(c)) for c in d["childB"]]
    best, todo = 0, [(0, 1)]
    while todo:
        x, h = todo.pop()
        best = max(best, h)
        for c in (cA[x], cB[x]):
            if c >= 0:
                todo.append((c, h + 1))
    return best


def check(d, pts):
    recs = fold(d)
    h = O.BvhHandle(d)
    depth = depth_of(d)
    hits = vis_w = vis_b = most = 0
    for p in pts.astype(np.float32):
        want, vb = O.query(h, p)
        got, vw, mw = walk(recs, p)
        assert got == want, (p, got, want)
        assert mw <= depth - 1, (p, mw, depth)
        hits += got >= 0
        vis_w += vw
        vis_b += vb
        most = max(most, mw)
    return hits, vis_w, vis_b, most, depth


def surface_points(v, t, n, rng, spread):
    tri = rng.integers(0, t.shape[0], n)
    w = rng.dirichlet([1.0, 1.0, 1.0], n)
    p = np.einsum("nk,nkc->nc", w, v[t[tri]])
    return p + rng.normal(scale=spread, size=p.shape)


def test_capsule_body_first_hits_equal_the_binary_walk():
    body = S.capsules(64, 14, 6, L.CC_BODY_STATIC)
    d = O.bvh_build(body.verts, body.tris, 0.002)
    rng = np.random.default_rng(3)
    pts = np.concatenate([surface_points(body.verts, body.tris, 3000, rng, 0.003),
                          rng.uniform(body.verts.min(0), body.verts.max(0), size=(500, 3))])
    hits, vw, vb, most, depth = check(d, pts)
    assert hits > 1500
    # two binary levels per record: about half the visits of the binary walk
    assert vw < 0.7 * vb, (vw, vb)
    print(f"capsules: {hits} hits, visits {vw} wide / {vb} binary, deferred <= {most}, depth {depth}")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_triangle_soup_first_hits_equal_the_binary_walk(seed):
    rng = np.random.default_rng(seed)
    n = 1500
    c = rng.uniform(0.0, 1.0, size=(n, 3))
    c[: n // 3] = np.round(c[: n // 3] * 4) / 4      # clusters: repeated Morton codes
    v = (c[:, None, :] + rng.normal(scale=0.03, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    t = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    d = O.bvh_build(v, t, 0.01)
    pts = np.concatenate([rng.uniform(-0.05, 1.05, size=(1500, 3)), c[:500] + rng.normal(scale=0.01, size=(500, 3))])
    hits, vw, vb, most, depth = check(d, pts)
    assert hits > 300


def test_whole_tree_walk_stays_within_the_depth():
    # 2 * 600 + 1 triangles whose boxes share one centre, interleaved flat in
    # y and flat in x, then a cube: a point such as (0.5, 0.5, z) lies in the
    # box of every node that holds both kinds and in no leaf box but the last
    # one's, so the walk visits the whole tree with the most deferrals
    v = []
    for _ in range(600):
        v += [(-1.0, -0.1, -1.0), (1.0, 0.1, -1.0), (0.0, 0.0, 1.0)]
        v += [(-0.1, -1.0, -1.0), (0.1, 1.0, -1.0), (0.0, 0.0, 1.0)]
    v += [(-1.0, -1.0, -1.0), (1.0, 1.0, -1.0), (0.0, 0.0, 1.0)]
    v = np.array(v, np.float32)
    t = np.arange(v.shape[0], dtype=np.uint32).reshape(-1, 3)
    d = O.bvh_build(v, t, 0.0)
    rng = np.random.default_rng(11)
    z = rng.uniform(-0.9, 0.9, 24)
    s = rng.choice([-0.5, 0.5], size=(24, 2))
    pts = np.stack([s[:, 0], s[:, 1], z], axis=1)
    hits, vw, vb, most, depth = check(d, pts)
    assert hits == 24
    assert most > 4, most                            # deep stacks were exercised
    print(f"crossed: visits {vw} wide / {vb} binary, deferred <= {most}, depth {depth}")


def test_a_planted_mistake_is_caught():
    # the split word left out (entries 2-3 deferred one by one): the first
    # hits stay the same but the deferrals exceed the binary walk's bound on
    # the whole-tree body, which is what the split is for
    v = []
    for _ in range(300):
        v += [(-1.0, -0.1, -1.0), (1.0, 0.1, -1.0), (0.0, 0.0, 1.0)]
        v += [(-0.1, -1.0, -1.0), (0.1, 1.0, -1.0), (0.0, 0.0, 1.0)]
    v += [(-1.0, -1.0, -1.0), (1.0, 1.0, -1.0), (0.0, 0.0, 1.0)]
    v = np.array(v, np.float32)
    t = np.arange(v.shape[0], dtype=np.uint32).reshape(-1, 3)
    d = O.bvh_build(v, t, 0.0)
    recs = [(ents, 0) for ents, _ in fold(d)]
    depth = depth_of(d)
    worst = max(walk(recs, np.array([0.5, 0.5, z], np.float32))[2] for z in np.linspace(-0.9, 0.9, 9))
    assert worst > depth - 1, (worst, depth)
