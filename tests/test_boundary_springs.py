"""Boundary springs of SINGLE_LAYER_BOUNDARY meshes (Springs::Springs,
spring.cpp:82-91, 98-209, 314-323; kdtree.cpp) and the OBJ vertex groups
they read (ObjLoader.cpp:34-41).

The oracle's nearest-neighbour search is pinned by brute force on tie-free
points and by hand-worked tie cases of the reference's kd-tree order; the
library's springs equal the oracle's on meshes whose boundary points sit at
edge midpoints (every nearest-neighbour question a tie) and on jittered
meshes (no ties).  No shipped reference file has such groups (its scene
builds the garment as SINGLE_LAYER_NOB, main.cpp:24), so the meshes are
synthetic."""
import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import ObjLoader

# every test here runs in the CPU suite and again in the GPU suite
pytestmark = pytest.mark.usefixtures("both_suites")


def test_kd_nearest_is_the_nearest_without_ties():
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1, 1, size=(700, 3)).astype(np.float32)
    q = rng.uniform(-1.3, 1.3, size=(400, 3)).astype(np.float32)
    d = ((q[:, None, :].astype(np.float64) - pts[None].astype(np.float64)) ** 2).sum(-1)
    np.testing.assert_array_equal(O.kd_nearest(pts, q), d.argmin(1))


def test_kd_nearest_ties_follow_the_walk():
    # root R far away; P = (1, 1, 0) is inserted before Q = (1, -1, 0) and
    # both are at distance 1 from (1, 0, 0): the walk reaches Q first (Q is
    # P's left child, the nearer side for y = 0 <= 1) and P does not beat it
    pts = np.float32([[10, 0, 0], [1, 1, 0], [1, -1, 0]])
    assert O.kd_nearest(pts, np.float32([[1, 0, 0]]))[0] == 2
    # the root, the first best, keeps any tie
    assert O.kd_nearest(np.float32([[0, 0, 0], [2, 0, 0]]), np.float32([[1, 0, 0]]))[0] == 0
    # equal coordinate goes right on insertion; a query at the split goes left first
    pts = np.float32([[0, 0, 0], [0, 5, 0], [-1, 0, 0]])
    assert O.kd_nearest(pts, np.float32([[-0.5, 0, 0]]))[0] == 0   # R and (-1,0,0) tie: R first


def grid_piece(nx, x0, h):
    i, j = np.meshgrid(np.arange(nx), np.arange(nx), indexing="xy")
    v = np.stack([x0 + i.ravel() * h, np.zeros(nx * nx), j.ravel() * h], 1)
    f = []
    for jj in range(nx - 1):
        for ii in range(nx - 1):
            a = jj * nx + ii
            f += [(a, a + nx, a + 1), (a + 1, a + nx, a + nx + 1)]
    return v, np.asarray(f, np.int64)


def border_midpoints(nx, x0, h, lift):
    m = []
    for k in range(nx - 1):
        for (x, z) in ((x0 + (k + 0.5) * h, 0.0), (x0 + (k + 0.5) * h, (nx - 1) * h),
                       (x0, (k + 0.5) * h), (x0 + (nx - 1) * h, (k + 0.5) * h)):
            m.append((x, lift, z))
    return np.asarray(m)


def two_piece_mesh(jitter=0.0, seed=0):
    """[piece A][boundary A][piece B][boundary B]: 17x17 grids 1/16 apart
    (dyadic, so squared distances are exact), boundary points on the
    border-edge midpoints lifted by 1/64, pieces 1/8 apart in x."""
    nx, h = 17, 1.0 / 16
    va, fa = grid_piece(nx, 0.0, h)
    ba = border_midpoints(nx, 0.0, h, 1.0 / 64)
    x1 = (nx - 1) * h + 0.125
    vb, fb = grid_piece(nx, x1, h)
    bb = border_midpoints(nx, x1, h, 1.0 / 64)
    v = np.concatenate([va, ba, vb, bb])
    if jitter:
        v = v + np.random.default_rng(seed).uniform(-jitter, jitter, size=v.shape)
    v4 = np.concatenate([v, np.ones((v.shape[0], 1))], 1).astype(np.float32)
    off_b = len(va) + len(ba)
    faces = np.concatenate([fa, fb + off_b]).astype(np.uint32)
    groups = np.array([len(va), len(ba), len(vb), len(bb)], np.uint32)
    return v4, faces, groups


@pytest.mark.parametrize("jitter", [0.0, 2e-3])
def test_library_boundary_springs_equal_the_oracle(jitter):
    v4, faces, groups = two_piece_mesh(jitter)
    got = S.springs_from_mesh(v4, faces, groups)
    want = O.springs_from_mesh(v4, faces, groups)
    for (gr, gc), (wr, wc) in zip(got, want):
        np.testing.assert_array_equal(gr, wr)
        np.testing.assert_array_equal(gc["end"], wc["end"])
        np.testing.assert_array_equal(gc["original"], wc["original"])
    plain = S.springs_from_mesh(v4, faces)
    cb, bd = O.boundary_springs(v4, groups)
    # the structural rows grow by 2 per boundary tie and 1 per boundary pair
    assert got[0][0][-1] == plain[0][0][-1] + 2 * len(cb) + len(bd)
    np.testing.assert_array_equal(got[1][0], plain[1][0])           # bend rows unchanged
    assert len(cb) == int(groups[1] + groups[3]) and len(bd) > 0
    # every boundary point on a midpoint ties its edge's two endpoints: the
    # kd-tree's order, not the lower index, picks one of them
    if not jitter:
        pa = v4[:groups[0], :3].astype(np.float64)
        lowest = [int(np.argmin(((pa - v4[k, :3]) ** 2).sum(1))) for k, _ in cb[:groups[1]]]
        assert any(int(r) != lo for (_, r), lo in zip(cb[:groups[1]], lowest))


def test_rejected_groups():
    v4, faces, groups = two_piece_mesh()
    bad = {
        "odd": groups[:3],
        "past": np.array([groups[0], groups[1], groups[2], groups[3] + 1], np.uint32),
        "empty piece": np.array([0, groups[1], groups[2], groups[3]], np.uint32),
        "few": np.array([groups[0], 40, groups[2], 40], np.uint32),
        "alone": groups[:2],
    }
    for name, g in bad.items():
        with pytest.raises(L.ClothCoreError):
            S.springs_from_mesh(v4, faces, g)
        with pytest.raises(ValueError):
            O.springs_from_mesh(v4, faces, g)
    with pytest.raises(L.ClothCoreError):
        S.springs_from_mesh(v4, faces, np.zeros(0, np.uint32))   # BOUNDARY without groups


def test_obj_vertex_groups(tmp_path):
    p = tmp_path / "groups.obj"
    p.write_text("# object piece1\nv 0 0 0\nv 1 0 0\nv 0 1 0\n# 3 vertices\n"
                 "# object piece1_boundary\nv 0 0 1\n# 1 vertices\n"
                 "# Object meshlab_style\n# 4 vertices, 0 vertices normals\n"
                 "f 1 2 3\n")
    o = ObjLoader(str(p))
    np.testing.assert_array_equal(o.vertex_object, [3, 1])
    q = tmp_path / "early.obj"
    q.write_text("v 0 0 0\n# 1 vertices\n")
    with pytest.raises(L.ClothCoreError, match="before any object"):
        ObjLoader(str(q))
    r = tmp_path / "none.obj"
    r.write_text("# Object tshirt2.obj\nv 0 0 0\n# 1 vertices, 0 vertices normals\n")
    assert ObjLoader(str(r)).vertex_object.size == 0


def test_obj_groups_to_springs(tmp_path):
    v4, faces, groups = two_piece_mesh(1e-3, seed=3)
    lines = []
    start = 0
    for k, cnt in enumerate(groups):
        lines.append(f"# object group{k}")
        for v in v4[start:start + cnt]:
            lines.append("v %.9g %.9g %.9g" % tuple(v[:3]))
        lines.append(f"# {cnt} vertices")
        start += cnt
    lines += ["f %d %d %d" % tuple(t + 1) for t in faces]
    p = tmp_path / "pieces.obj"
    p.write_text("\n".join(lines) + "\n")
    o = ObjLoader(str(p))
    np.testing.assert_array_equal(o.vertex_object, groups)
    got = S.springs_from_mesh(o.vertices, o.vertex_index, o.vertex_object)
    want = O.springs_from_mesh(o.vertices, o.vertex_index, o.vertex_object)
    for (gr, gc), (wr, wc) in zip(got, want):
        np.testing.assert_array_equal(gr, wr)
        np.testing.assert_array_equal(gc["end"], wc["end"])
        np.testing.assert_array_equal(gc["original"], wc["original"])
