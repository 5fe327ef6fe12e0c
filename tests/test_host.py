"""CPU tests of the library's host code (spring builders, generators,
vertex_extend) against the oracle, and of the C ABI surface.  No GPU."""
import os
import re

import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S

# every test here runs in the CPU suite and again in the GPU suite
pytestmark = pytest.mark.usefixtures("both_suites")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "clothcore.h")).read()
    names = set(re.findall(r"\b(cc_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 30
    lib = L.lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes prototype and vice versa
    assert names == set(L.PROTOTYPES), names ^ set(L.PROTOTYPES)


def test_errors_are_reported_not_printed():
    # an invalid call returns an error code with a message (common.h:4-10 printed and went on)
    with pytest.raises(L.ClothCoreError) as e:
        L.check(L.lib().cc_springs_grid(0, 0, None, None))
    assert e.value.code == L.CC_ERR_INVALID and "cc_springs_grid" in str(e.value)


def test_params_default_are_reference_constants():
    p = L.default_params()
    assert np.float32(p.dt) == np.float32(1 / 40.0) and np.float32(p.mass) == np.float32(0.3)
    assert np.float32(p.gravity[1]) == np.float32(-0.00981) and np.float32(p.damp) == np.float32(-0.02)
    assert p.kd == -0.5 and p.k[0] == 100.0 and p.k[1] == 2.0
    assert p.eps == np.float32(2.0 ** -23) and np.float32(p.box_depth) == np.float32(0.01)


@pytest.mark.parametrize("side,seed", [(16, 0), (64, 0), (111, 3)])
def test_grid_cloth_and_springs_match_oracle(side, seed):
    sp = 1.0 / (side - 1)
    pos = S.grid_cloth(side, side, sp, 1.0, 1e-4, seed)
    np.testing.assert_array_equal(pos.view(np.uint32), O.grid_cloth(side, side, sp, 1.0, 1e-4, seed).view(np.uint32))
    got = S.grid_springs(side, side, pos)
    want = O.grid_springs(side, side, pos)
    for (gr, gc), (wr, wc) in zip(got, want):
        np.testing.assert_array_equal(gr, wr)
        np.testing.assert_array_equal(gc.view(np.uint64), wc.view(np.uint64))


def test_grid_spring_counts_closed_form():
    # undirected structural / shear / bend counts 2n(n-1), 2(n-1)^2, 2n(n-2)
    for n in (64, 111, 512):
        cnt = O.grid_spring_counts(n, n)
        assert list(cnt // 2) == [2 * n * (n - 1), 2 * (n - 1) ** 2, 2 * n * (n - 2)]
    pos = S.grid_cloth(64, 64, 1 / 63, 1.0, 1e-4, 0)
    assert sum(c.shape[0] for _, c in S.grid_springs(64, 64, pos)) == 47876


def test_icosphere_matches_oracle_and_is_closed_outward():
    for lv in (0, 2, 4):
        v, t, nrm = S.icosphere(lv, 0.3, (0.5, 0.5, 0.0))
        ov, ot = O.icosphere(lv, 0.3, (0.5, 0.5, 0.0))
        np.testing.assert_array_equal(v.view(np.uint32), ov.view(np.uint32))
        np.testing.assert_array_equal(t, ot)
        assert v.shape[0] == 10 * 4 ** lv + 2 and t.shape[0] == 20 * 4 ** lv
        c = v[t].mean(1) - np.float32([0.5, 0.5, 0.0])
        fn = np.cross(v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]])
        assert (np.einsum("ij,ij->i", fn, c) > 0).all()
        # closed 2-manifold: every undirected edge in exactly two faces
        e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert (cnt == 2).all()


def test_capsule_body_closed_outward_and_counts():
    body = S.capsules(64, 14, 6, L.CC_BODY_REBUILD)
    v, t = body.verts, body.tris
    assert t.shape[0] == 50688 and v.shape[0] == 12 * (2 + 64 * 33)
    caps = body.caps
    for k in range(12):
        a, b = caps["vstart"][k], caps["vstart"][k + 1]
        tk = t[(t >= a).all(1) & (t < b).all(1)]
        assert tk.shape[0] == 50688 // 12
        e = np.sort(np.concatenate([tk[:, [0, 1]], tk[:, [1, 2]], tk[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert (cnt == 2).all()
        # outward: face normal points away from the capsule axis
        p0, p1 = caps["p0"][k].astype(np.float64), caps["p1"][k].astype(np.float64)
        cen = v[tk].mean(1).astype(np.float64)
        d = p1 - p0
        s = np.clip(((cen - p0) @ d) / (d @ d), 0, 1)
        axis_pt = p0 + s[:, None] * d
        fn = np.cross(v[tk[:, 1]] - v[tk[:, 0]], v[tk[:, 2]] - v[tk[:, 0]])
        assert (np.einsum("ij,ij->i", fn, cen - axis_pt) > 0).all()
        # vertices lie at the capsule radius
        vv = v[a:b].astype(np.float64)
        s = np.clip(((vv - p0) @ d) / (d @ d), 0, 1)
        r = np.linalg.norm(vv - (p0 + s[:, None] * d), axis=1)
        np.testing.assert_allclose(r, caps["radius"][k], rtol=1e-5)


def two_triangle_square():
    v = np.array([[0, 0, 0, 1], [1, 0, 0, 1], [1, 1, 0, 1], [0, 1, 0, 1]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return v, f


def test_springs_from_mesh_known_answer():
    v, f = two_triangle_square()
    (rs, cs), (rb, cb) = S.springs_from_mesh(v, f)
    # structure: 1-ring from face edges in first-seen order (spring.cpp:290-312)
    assert list(rs) == [0, 3, 5, 8, 10]
    assert list(cs["end"]) == [1, 2, 3, 0, 2, 0, 1, 3, 0, 2]
    # bend: the shared edge (0,2) pairs its opposite vertices 1 and 3
    assert list(rb) == [0, 0, 1, 1, 2]
    assert list(cb["end"]) == [3, 1]
    np.testing.assert_array_equal(cb["original"], np.float32([np.sqrt(np.float32(2))] * 2))


def test_springs_from_mesh_matches_oracle_on_meshes():
    v, t, _ = S.icosphere(3, 0.3, (0.5, 0.5, 0.0))
    v4 = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], 1)
    body = S.capsules(16, 4, 2, L.CC_BODY_STATIC)
    b4 = np.concatenate([body.verts, np.ones((body.verts.shape[0], 1), np.float32)], 1)
    for verts4, faces in ((v4, t), (b4, body.tris)):
        got = S.springs_from_mesh(verts4, faces)
        want = O.springs_from_mesh(verts4, faces)
        for (gr, gc), (wr, wc) in zip(got, want):
            np.testing.assert_array_equal(gr, wr)
            np.testing.assert_array_equal(gc.view(np.uint64), wc.view(np.uint64))


def test_vertex_extend_matches_oracle():
    v, t, nrm = S.icosphere(2, 0.3, (0.5, 0.5, 0.0))
    from simulator_amd.simulator import Mesh
    m = Mesh(v, t, nrm)
    m.vertex_extend(0.003)
    v4 = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], 1)
    want = O.vertex_extend(v4, t, t, nrm, 0.003)
    np.testing.assert_array_equal(m.vertices, want)
    r = np.linalg.norm(m.vertices[:, :3] - np.float32([0.5, 0.5, 0.0]), axis=1)
    np.testing.assert_allclose(r, 0.303, rtol=1e-5)


def test_fault_inject_arguments():
    """cc_fault_inject takes CC_FAULT_ALLOC / CC_FAULT_LAUNCH and any count
    (negative: off); other kinds are CC_ERR_INVALID.  No device work."""
    lib = L.lib()
    assert lib.cc_fault_inject(0, 1) == L.CC_ERR_INVALID
    assert lib.cc_fault_inject(3, 1) == L.CC_ERR_INVALID
    for kind in (L.CC_FAULT_ALLOC, L.CC_FAULT_LAUNCH):
        assert lib.cc_fault_inject(kind, 5) == L.CC_OK
        assert lib.cc_fault_inject(kind, -1) == L.CC_OK


def test_checkpoint_arguments():
    """cc_sim_save / cc_sim_load refuse a null simulator or path (no device
    work)."""
    lib = L.lib()
    assert lib.cc_sim_save(None, b"/tmp/x.ckpt") == L.CC_ERR_INVALID
    assert lib.cc_sim_load(None, b"/tmp/x.ckpt") == L.CC_ERR_INVALID
