"""The chunked bottom-up LBVH construction of the batched bodies, restated
step for step in Python, against the oracle's Karras build
(processInternalNode, cudaBRT.cu:9-161): the same children and parents at
the same node indices, for random keys (duplicate Morton codes included),
small chunks (several levels) and any arrival order.  The kernel itself is
checked bit for bit on the GPU (test_gpu_batch.py); this pins the algorithm
it runs (chunks of items, arrivals at splits, unfinished nodes handed to the
next level in leaf order, left child = its last leaf, right child = its
first)."""
import numpy as np
import pytest

import oracle as O


def clz64(x):
    return 64 - int(x).bit_length() if x else 64


def level_build(keys, C, G, rng):
    m = len(keys)
    K = [int(k) for k in keys]
    cA = [None] * (m - 1)
    cB = [None] * (m - 1)
    par = [None] * (m - 1)
    lpar = [None] * m
    pend_max = 0

    def delta(a):                      # delta(a, a + 1)
        return clz64(K[a] ^ K[a + 1])

    def link(c, p):
        if c < 0:
            lpar[~c] = p
        else:
            par[c] = p

    chunks = [[(l, l, ~l) for l in range(a, min(m, a + C))] for a in range(0, m, C)]
    while True:
        outs = []
        for ch in chunks:
            n = len(ch)
            sD = [delta(ch[0][0] - 1) if ch[0][0] > 0 else -1] + \
                 [delta(r) if r + 1 < m else -1 for (_, r, _) in ch]
            half, pend = {}, []
            for t in rng.permutation(n):
                kl = kr = int(t)
                gl, gr, me = ch[t]
                merged = False
                for _ in range(70):
                    dl, dr = sD[kl], sD[kr + 1]
                    if dl < 0 and dr < 0:
                        assert merged and (gl, gr) == (0, m - 1)
                        cA[0], cB[0], par[0] = ca, cb, -1
                        link(ca, 0)
                        link(cb, 0)
                        break
                    assert dl != dr              # unique keys never tie
                    right = dr > dl
                    if merged:
                        me = gr if right else gl
                        cA[me], cB[me] = ca, cb
                        link(ca, me)
                        link(cb, me)
                    if (kr == n - 1) if right else (kl == 0):
                        pend.append((gl, gr, me))
                        break
                    sp = kr if right else kl - 1
                    side = 0 if right else 1
                    half[(sp, side)] = (kl if right else kr, gl if right else gr)
                    if (sp, 1 - side) not in half:
                        break                    # first child at this split
                    fk, fg = half[(sp, 1 - side)]
                    if right:
                        s, gr, kr = gr, fg, fk
                    else:
                        s, gl, kl = gl - 1, fg, fk
                    ca = ~s if gl == s else s
                    cb = ~(s + 1) if s + 1 == gr else s + 1
                    merged = True
            for (sp, side), (_, fg) in half.items():     # children still waiting at a split
                if (sp, 1 - side) in half:
                    continue
                s = ch[sp][1]
                l, r = (fg, s) if side == 0 else (s + 1, fg)
                pend.append((l, r, (r if side == 0 else l) if l < r else ~l))
            pend.sort()
            pend_max = max(pend_max, len(pend))
            outs.append(pend)
        if len(chunks) == 1:
            assert outs[0] == []
            return cA, cB, par, lpar, pend_max
        chunks = [sum(outs[i:i + G], []) for i in range(0, len(outs), G)]


@pytest.mark.parametrize("seed", range(6))
def test_level_build_equals_karras(seed):
    rng = np.random.default_rng(seed)
    for _ in range(8):
        m = int(rng.integers(2, 1500))
        codes = rng.integers(0, 2 ** int(rng.integers(2, 31)), size=m).astype(np.uint64)
        keys = np.sort((codes << np.uint64(32)) | np.arange(m, dtype=np.uint64))
        C = int(rng.choice([4, 8, 16, 64]))
        G = int(rng.choice([2, 4, 8]))
        cA, cB, par, lpar, pend_max = level_build(keys, C, G, rng)
        oA, oB, oP, oL = O.karras(keys)
        enc = np.vectorize(lambda c: c + 1 if c >= 0 else c)
        np.testing.assert_array_equal(enc(np.array(cA)), oA)
        np.testing.assert_array_equal(enc(np.array(cB)), oB)
        np.testing.assert_array_equal(np.array(par) + 1, oP)
        np.testing.assert_array_equal(np.array(lpar) + 1, oL)
        assert pend_max <= 2 * 63
