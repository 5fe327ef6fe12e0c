"""Pins the CPU oracle itself against closed forms, brute force and
invariants (the reference ships no tests or fixtures).  CPU only."""
import numpy as np
import pytest

import oracle as O


# ------------------------------------------------------------------ Morton
def interleave_bits(x, y, z):
    """Textbook Morton interleave: bit b of x lands at 3b+2, y at 3b+1, z at 3b."""
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def test_expand_bits_bruteforce():
    for v in range(1024):
        want = sum(((v >> b) & 1) << (3 * b) for b in range(10))
        assert O.expand_bits(v) == want


def test_morton_closed_forms():
    assert O.morton3d(0.0, 0.0, 0.0) == 0
    assert O.morton3d(1.0, 1.0, 1.0) == 0x3FFFFFFF
    assert O.morton3d(1.0 / 1024, 0.0, 0.0) == 4
    assert O.morton3d(0.0, 1.0 / 1024, 0.0) == 2
    assert O.morton3d(0.0, 0.0, 1.0 / 1024) == 1
    # clamping: below 0 and above 1
    assert O.morton3d(-5.0, 2.0, 0.5) == interleave_bits(0, 1023, 512)
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x, y, z = rng.random(3).astype(np.float32)
        ix, iy, iz = (min(max(int(np.float32(c) * np.float32(1024.0)), 0), 1023) for c in (x, y, z))
        assert O.morton3d(x, y, z) == interleave_bits(ix, iy, iz)


def test_morton_nan_and_flat_root():
    # a zero-extent root (all boxes coincide) gives the (0,0,0) cell
    box = np.array([0.2, 0.3, 0.4, 0.2, 0.3, 0.4], np.float32)
    codes = O.morton_codes(box[None, :], box)
    assert codes[0] == 0


# ------------------------------------------------------------------ boxes
def test_prim_box_axis_aligned():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    b = O.prim_boxes(v, np.array([[0, 1, 2]], np.uint32), 0.01)[0]
    # normal (0,0,1): box grows 0.01 towards -z only
    np.testing.assert_array_equal(b, np.array([0, 0, -0.01, 1, 1, 0], np.float32))
    r = O.root_box(b[None, :])
    np.testing.assert_array_equal(r, b)   # sentinel (+-100) loses to real coordinates


def test_root_box_keeps_sentinel_semantics():
    # coordinates beyond the +-100 sentinel are clamped by it (bbox.h:17-22)
    b = np.array([[150, 150, 150, 160, 160, 160]], np.float32)
    np.testing.assert_array_equal(O.root_box(b), np.array([100, 100, 100, 160, 160, 160], np.float32))


# ------------------------------------------------------------------ sort
def test_sort_is_stable_by_code():
    rng = np.random.default_rng(2)
    codes = rng.integers(0, 50, size=3000).astype(np.uint32)   # many duplicates
    keys, order = O.sort_keys(codes)
    ref = np.argsort(codes, kind="stable")
    np.testing.assert_array_equal(order, ref)
    np.testing.assert_array_equal(keys, (codes[ref].astype(np.uint64) << 32) | ref.astype(np.uint64))


def test_dedupe_keeps_lowest_index():
    codes = np.array([5, 3, 5, 3, 9, 3], np.uint32)
    keys, order = O.sort_keys(codes)
    k, o = O.dedupe_sorted(keys, order)
    np.testing.assert_array_equal(o, [1, 0, 4])


# ------------------------------------------------------------------ Karras
def topdown_build(keys):
    """Brute-force Karras tree: node over [f, l] splits after the last key
    that shares the first differing bit of keys[f] and keys[l]; child A is
    leaf/internal `split`, child B leaf/internal `split + 1`; root is 0."""
    m = len(keys)
    cA = np.zeros(m - 1, np.int32)
    cB = np.zeros(m - 1, np.int32)
    pI = np.zeros(m - 1, np.int32)
    pL = np.zeros(m, np.int32)
    ks = [int(k) for k in keys]
    stack = [(0, 0, m - 1)]
    while stack:
        idx, f, l = stack.pop()
        x = ks[f] ^ ks[l]
        top = x.bit_length() - 1          # highest differing bit
        split = f
        for i in range(f, l):
            if (ks[i] >> top) == (ks[f] >> top):
                split = i
        if split == f:
            cA[idx] = -(split + 1)
            pL[split] = idx + 1
        else:
            cA[idx] = split + 1
            pI[split] = idx + 1
            stack.append((split, f, split))
        if split + 1 == l:
            cB[idx] = -(split + 2)
            pL[split + 1] = idx + 1
        else:
            cB[idx] = split + 2
            pI[split + 1] = idx + 1
            stack.append((split + 1, split + 1, l))
    return cA, cB, pI, pL


@pytest.mark.parametrize("m,dup", [(2, False), (3, False), (7, True), (33, True), (64, False), (64, True)])
def test_karras_matches_topdown(m, dup):
    rng = np.random.default_rng(m + (100 if dup else 0))
    hi = 8 if dup else 1 << 30
    codes = rng.integers(0, hi, size=m).astype(np.uint32)
    keys, _ = O.sort_keys(codes)
    got = O.karras(keys)
    want = topdown_build(keys)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


def check_tree_invariants(cA, cB, pI, pL):
    m = len(pL)
    assert pI[0] == 0                        # root has no parent
    assert (pI[1:] > 0).all() and (pL > 0).all()
    # every node except the root is the child of exactly one node
    seen_i = np.zeros(m - 1, int)
    seen_l = np.zeros(m, int)
    for x in range(m - 1):
        for c in (cA[x], cB[x]):
            if c < 0:
                seen_l[-c - 1] += 1
                assert pL[-c - 1] == x + 1
            else:
                seen_i[c - 1] += 1
                assert pI[c - 1] == x + 1
    assert (seen_l == 1).all() and seen_i[0] == 0 and (seen_i[1:] == 1).all()


def leaf_range(cA, cB, x):
    lo, hi, stack = 1 << 62, -1, [x]
    while stack:
        y = stack.pop()
        for c in (cA[y], cB[y]):
            if c < 0:
                lo, hi = min(lo, -c - 1), max(hi, -c - 1)
            else:
                stack.append(c - 1)
    return lo, hi


def test_karras_invariants_icosphere():
    v, t = O.icosphere(3, 0.3, (0.5, 0.5, 0.0))
    d = O.bvh_build(v, t, 0.01)
    check_tree_invariants(d["childA"], d["childB"], d["parent_int"], d["parent_leaf"])
    # children partition the parent's contiguous leaf range
    for x in range(0, d["m"] - 1, 37):
        lo, hi = leaf_range(d["childA"], d["childB"], x)
        a = d["childA"][x]
        alo, ahi = (-a - 1, -a - 1) if a < 0 else leaf_range(d["childA"], d["childB"], a - 1)
        b = d["childB"][x]
        blo, bhi = (-b - 1, -b - 1) if b < 0 else leaf_range(d["childA"], d["childB"], b - 1)
        assert (alo, bhi) == (lo, hi) and ahi + 1 == blo


# ------------------------------------------------------------------ refit + traversal
def test_refit_equals_union_of_leaves():
    v, t = O.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    d = O.bvh_build(v, t, 0.01)
    for x in range(d["m"] - 1):
        lo, hi = leaf_range(d["childA"], d["childB"], x)
        bl = d["box_leaf"][lo:hi + 1]
        want = np.concatenate([np.minimum(bl[:, :3].min(0), 100), np.maximum(bl[:, 3:].max(0), -100)])
        np.testing.assert_array_equal(d["box_int"][x], want.astype(np.float32))


def replay_first_hit(d, p):
    """Recursive restatement of the first-hit order on the exported tree."""
    def has(b):
        return all(b[i] <= p[i] <= b[3 + i] for i in range(3))

    def box(c):
        return d["box_leaf"][-c - 1] if c < 0 else d["box_int"][c - 1]

    def visit(x):
        a, b = d["childA"][x], d["childB"][x]
        oa, ob = has(box(a)), has(box(b))
        if oa and a < 0:
            return -a - 1
        if ob and b < 0:
            return -b - 1
        if oa and a > 0:
            r = visit(a - 1)
            if r >= 0:
                return r
        if ob and b > 0:
            r = visit(b - 1)
            if r >= 0:
                return r
        return -1
    return visit(0)


def test_query_matches_replay_and_containment():
    v, t = O.icosphere(3, 0.3, (0.5, 0.5, 0.0))
    d = O.bvh_build(v, t, 0.01)
    h = O.BvhHandle(d)
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(600, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    pts = (np.array([0.5, 0.5, 0.0]) + dirs * rng.uniform(0.27, 0.31, size=(600, 1))).astype(np.float32)
    hits = 0
    for p in pts:
        r, _ = O.query(h, p)
        assert r == replay_first_hit(d, p)
        contains = [(d["box_leaf"][l, :3] <= p).all() and (p <= d["box_leaf"][l, 3:]).all()
                    for l in range(d["m"])]
        if r < 0:
            assert not any(contains)
        else:
            assert contains[r]
            hits += 1
    assert hits > 100


def test_single_leaf_tree():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    d = O.bvh_build(v, np.array([[0, 1, 2]], np.uint32), 0.01)
    h = O.BvhHandle(d)
    assert O.query(h, np.array([0.2, 0.2, -0.005], np.float32))[0] == 0
    assert O.query(h, np.array([0.2, 0.2, 0.5], np.float32))[0] == -1


# ------------------------------------------------------------------ plane test
def test_plane_test_cases():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([0, 1, 2], np.uint32)
    ins, d, n = O.plane_test(v, t, np.array([0.2, 0.2, 0.5], np.float32))
    assert not ins and d == np.float32(0.5)
    np.testing.assert_array_equal(n, [0, 0, 1])
    ins, d, n = O.plane_test(v, t, np.array([0.2, 0.2, -0.25], np.float32))
    assert ins and d == np.float32(-0.25)
    ins, d, _ = O.plane_test(v, t, np.array([0.2, 0.2, 0.0], np.float32))
    assert ins and d == 0.0                           # on the plane counts as inside
    # reversed winding flips the normal
    ins, d, n = O.plane_test(v, np.array([0, 2, 1], np.uint32), np.array([0.2, 0.2, 0.5], np.float32))
    assert ins and d == np.float32(-0.5)


# ------------------------------------------------------------------ Verlet
def make_params(dt=1e-3, mass=0.3, g=-9.8, damp=0.0, kd=0.0, k=(0, 0, 0), eps=0.0):
    p = O.or_params()
    p.dt, p.mass, p.damp, p.kd, p.eps, p.rest_scale = dt, mass, damp, kd, eps, 1.0
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, g, 0.0
    for i in range(3):
        p.k[i] = k[i]
    return p


def no_springs(n):
    return [(np.zeros(n + 1, np.uint32), np.zeros(0, O.SPRING_DTYPE))]


@pytest.mark.parametrize("damp", [0.0, -0.02])
def test_free_fall_closed_form(damp):
    prm = make_params(damp=damp)
    x0 = np.array([[0.1, 1.0, -0.2]])
    x, xp, c = x0.copy(), x0.copy(), np.zeros_like(x0)
    N = 300
    for _ in range(N):
        x, xp, c, _ = O.step_f64(x, xp, c, no_springs(1), prm)
    # the parameters are FP32 values widened to FP64
    dt, g = np.float64(np.float32(1e-3)), np.float64(np.float32(-9.8))
    dmp, mf = np.float64(np.float32(damp)), np.float64(np.float32(0.3))
    if damp == 0.0:
        want = 1.0 + g * dt * dt * N * (N + 1) / 2
    else:
        r = 1.0 + dmp * dt / mf
        dd = dt ** 2 * g
        # displacement d_n = dd (1 - r^n)/(1 - r); x_N = x_0 + sum_{n=1..N} d_n
        want = 1.0 + sum(dd * (1 - r ** n) / (1 - r) for n in range(1, N + 1))
    assert abs(x[0, 1] - want) < 1e-10     # a wrong term would be >= g dt^2 = 1e-5
    assert x[0, 0] == 0.1 and x[0, 2] == -0.2


def two_mass(k=50.0, L0=0.125, stretch=0.02):   # L0 exact in FP32
    pos = np.array([[0.0, 0.0, 0.0], [L0 + stretch, 0.0, 0.0]])
    row = np.array([0, 1, 2], np.uint32)
    col = np.zeros(2, O.SPRING_DTYPE)
    col["end"] = [1, 0]
    col["original"] = [L0, L0]
    return pos, [(row, col)]


def test_two_mass_spring_frequency_and_momentum():
    k, m, dt = 50.0, 0.3, 1e-3
    prm = make_params(dt=dt, mass=m, g=0.0, k=(k, 0, 0))
    pos, springs = two_mass(k)
    x, xp, c = pos.copy(), pos.copy(), np.zeros_like(pos)
    sep = []
    for _ in range(2000):
        x, xp, c, _ = O.step_f64(x, xp, c, springs, prm)
        sep.append(x[1, 0] - x[0, 0])
        assert abs(x[:, 0].sum() - pos[:, 0].sum()) < 1e-12     # momentum (centre fixed)
    sep = np.array(sep) - 0.125
    # relative displacement oscillates at omega = sqrt(2k/m); find zero crossings
    zc = np.where(np.diff(np.sign(sep)) != 0)[0]
    period = 2 * np.mean(np.diff(zc)) * dt
    omega = np.sqrt(2 * k / m)
    assert abs(period - 2 * np.pi / omega) / (2 * np.pi / omega) < 2e-3
    assert abs(np.abs(sep).max() - 0.02) < 1e-3          # amplitude kept (no damping)


def test_spring_at_rest_force_is_eps_bias_only():
    k = 100.0
    prm = make_params(dt=1e-3, g=0.0, k=(k, 0, 0), eps=np.float32(2.0 ** -23))
    pos, springs = two_mass(k, stretch=0.0)
    x, xp, c, _ = O.step_f64(pos.copy(), pos.copy(), np.zeros_like(pos), springs, prm)
    # dP = (-L0 + eps, eps, eps): |F| ~ ks * (|dP| - L0) ~ ks * eps along -x of dP
    disp = (x - pos)[0]
    acc = disp / 1e-6
    F = acc * 0.3
    eps = 2.0 ** -23
    assert abs(abs(F[0]) - k * eps) < 0.1 * k * eps
    prm0 = make_params(dt=1e-3, g=0.0, k=(k, 0, 0), eps=0.0)
    x0, *_ = O.step_f64(pos.copy(), pos.copy(), np.zeros_like(pos), springs, prm0)
    assert np.abs(x0 - pos).max() < 1e-15


def test_fp32_matches_fp64_per_step():
    pos = O.grid_cloth(16, 16, 1 / 15, 1.0, 1e-4, 5)
    sp = O.grid_springs(16, 16, pos)
    prm = make_params(k=(1000.0, 500.0, 100.0), damp=-0.02, kd=-0.5, eps=np.float32(2.0 ** -23))
    pins = np.zeros(256, np.uint8)
    pins[[0, 15]] = 1
    x, xp, c = pos.copy(), pos.copy(), np.zeros_like(pos)
    for _ in range(50):
        x, xp, c = O.step_f32(x, xp, c, sp, prm, pins)
    a32, *_ = O.step_f32(x, xp, c, sp, prm, pins)
    a64, *_ = O.step_f64(x, xp, c, sp, prm, pins)
    assert np.abs(a32 - a64).max() < 1e-6 * 1.0        # 1e-6 of cloth extent (1 m)
    np.testing.assert_array_equal(a32[[0, 15]], x[[0, 15]])   # pins hold
