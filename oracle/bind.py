"""numpy-facing ctypes binding of oracle/liboracle.so (TEST INFRASTRUCTURE ONLY).

Each function names the reference code it restates; see oracle.h for the
behaviour the oracle defines where the reference is undefined.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SPRING_DTYPE = np.dtype([("end", "<i4"), ("original", "<f4")])


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class or_params(C.Structure):
    _fields_ = [("dt", C.c_float), ("mass", C.c_float), ("gravity", C.c_float * 3),
                ("damp", C.c_float), ("kd", C.c_float), ("k", C.c_float * 3),
                ("rest_scale", C.c_float), ("eps", C.c_float)]


class or_params64(C.Structure):
    _fields_ = [("dt", C.c_double), ("mass", C.c_double), ("gravity", C.c_double * 3),
                ("damp", C.c_double), ("kd", C.c_double), ("k", C.c_double * 3),
                ("rest_scale", C.c_double), ("eps", C.c_double)]


class or_bvh(C.Structure):
    _fields_ = [("m", C.c_uint32), ("verts", C.c_void_p), ("leaf_tri", C.c_void_p),
                ("childA", C.c_void_p), ("childB", C.c_void_p), ("box_int", C.c_void_p),
                ("box_leaf", C.c_void_p)]


_lib = None


def olib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        V, u32, u64, f32, f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_double
        sig = {
            "or_splitmix64": (u64, [V]),
            "or_uniform01": (f64, [V]),
            "or_grid_cloth": (None, [u32, u32, f64, f64, f64, u64, V]),
            "or_grid_spring_counts": (None, [u32, u32, V]),
            "or_grid_springs": (None, [u32, u32, V, V, V, V, V, V, V]),
            "or_icosphere": (u32, [u32, f64, V, V, V, V]),
            "or_springs_from_mesh": (C.c_int, [u32, V, u32, V, V, V, V, V, V, V]),
            "or_springs_from_mesh_ex": (C.c_int, [u32, V, u32, V, u32, V, u32, V, V, V, V, V, V, V]),
            "or_kd_nearest": (C.c_int32, [u32, V, u32, V, V]),
            "or_boundary_springs": (C.c_int32, [u32, V, u32, V, V, V, V]),
            "or_vertex_extend": (None, [u32, V, u32, V, V, V, f32]),
            "or_prim_box": (None, [V, V, f32, V]),
            "or_prim_boxes": (None, [V, u32, V, f32, V]),
            "or_set_threads": (None, [C.c_int]),
            "or_root_box": (None, [u32, V, V]),
            "or_expand_bits": (u32, [u32]),
            "or_morton3d": (u32, [f32, f32, f32]),
            "or_box_morton": (u32, [V, V]),
            "or_morton_codes": (None, [u32, V, V, V]),
            "or_sort_keys": (None, [u32, V, V, V]),
            "or_dedupe_sorted": (u32, [u32, V, V]),
            "or_karras": (None, [u32, V, V, V, V, V]),
            "or_refit": (None, [u32, V, V, V, V, V]),
            "or_bvh_build": (u32, [V, u32, V, f32, C.c_int, V, V, V, V, V, V, V, V, V]),
            "or_query": (C.c_int32, [V, V, V]),
            "or_query_f64": (C.c_int32, [V, V]),
            "or_plane_test": (C.c_int, [V, V, V, V, V]),
            "or_step_f32": (None, [u32, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V,
                                   C.c_int]),
            "or_step_f64": (None, [u32, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V, V]),
            "or_collide_points": (None, [u32, V, V, V, V, V, V]),
            "or_body_deform": (None, [u32, V, u32, V, V, V, V, V, V, V]),
            "or_pick": (C.c_int32, [u32, V, V, V, V, V]),
            "or_vertex_normals": (None, [u32, V, u32, V, V]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------------ scenes
def grid_cloth(nx, nz, spacing, y0, jitter, seed):
    pos = np.zeros((nx * nz, 3), np.float32)
    olib().or_grid_cloth(nx, nz, spacing, y0, jitter, seed, _p(pos))
    return pos


def grid_spring_counts(nx, nz):
    cnt = np.zeros(3, np.uint32)
    olib().or_grid_spring_counts(nx, nz, _p(cnt))
    return cnt


def grid_springs(nx, nz, pos):
    cnt = np.zeros(3, np.uint32)
    olib().or_grid_spring_counts(nx, nz, _p(cnt))
    n = nx * nz
    rows = [np.zeros(n + 1, np.uint32) for _ in range(3)]
    cols = [np.zeros(int(c), SPRING_DTYPE) for c in cnt]
    olib().or_grid_springs(nx, nz, _p(_c(pos, np.float32)), _p(rows[0]), _p(cols[0]),
                           _p(rows[1]), _p(cols[1]), _p(rows[2]), _p(cols[2]))
    return list(zip(rows, cols))


def icosphere(level, radius, center):
    nv = 10 * 4 ** level + 2
    nf = 20 * 4 ** level
    v = np.zeros((nv, 3), np.float32)
    t = np.zeros((nf, 3), np.uint32)
    c = np.asarray(center, np.float64)
    k = C.c_uint32()
    got = olib().or_icosphere(level, radius, _p(c), _p(v), _p(t), C.byref(k))
    assert got == nv and k.value == nf
    return v, t


BOUNDARY_ERRORS = {-1: "odd group count", -2: "groups past the vertex count", -3: "an empty piece",
                   -4: "fewer than 100 boundary vertices", -5: "a boundary with no other boundary"}


def kd_nearest(pts, queries):
    """Index in pts of each query's nearest point, as the reference's
    kd-tree (kdtree.cpp) finds it, ties included."""
    p, q = _c(pts, np.float32), _c(queries, np.float32)
    out = np.zeros(q.shape[0], np.uint32)
    if olib().or_kd_nearest(p.shape[0], _p(p), q.shape[0], _p(q), _p(out)) < 0:
        raise ValueError("kd_nearest: no points")
    return out


def boundary_springs(verts4, vertex_object):
    """(cloth-boundary pairs (k, 2), boundary set pairs (j, 2)) of
    Springs::get_cloth_boundary_spring / get_boundary_boundary_spring
    (spring.cpp:98-209); ValueError for the inputs the reference reads out
    of range."""
    v = _c(verts4, np.float32)
    g = _c(vertex_object, np.uint32)
    nv = v.shape[0]
    cb = np.zeros((nv + 1, 2), np.uint32)
    bd = np.zeros((nv + 1, 2), np.uint32)
    nbd = C.c_uint32()
    r = olib().or_boundary_springs(nv, _p(v), g.size, _p(g), _p(cb), _p(bd), C.byref(nbd))
    if r < 0:
        raise ValueError("boundary springs: " + BOUNDARY_ERRORS[r])
    return cb[:r].copy(), bd[:nbd.value].copy()


def springs_from_mesh(verts4, faces, vertex_object=None):
    """Springs::Springs + create_neigh + CSR_spring (spring.cpp:82-391):
    SINGLE_LAYER_NOB, or SINGLE_LAYER_BOUNDARY given the group counts."""
    v = _c(verts4, np.float32)
    f = _c(faces, np.uint32)
    nv, nf = v.shape[0], f.size // 3
    cbp = np.zeros((0, 2), np.uint32)
    bdp = np.zeros((0, 2), np.uint32)
    if vertex_object is not None:
        cbp, bdp = boundary_springs(v, vertex_object)
    cbp, bdp = _c(cbp, np.uint32), _c(bdp, np.uint32)
    ns, nb = C.c_uint32(), C.c_uint32()
    L = olib()
    L.or_springs_from_mesh_ex(nv, _p(v), nf, _p(f), cbp.shape[0], _p(cbp), bdp.shape[0], _p(bdp), None, None,
                              C.byref(ns), None, None, C.byref(nb))
    rs, rb = np.zeros(nv + 1, np.uint32), np.zeros(nv + 1, np.uint32)
    cs, cb = np.zeros(ns.value, SPRING_DTYPE), np.zeros(nb.value, SPRING_DTYPE)
    L.or_springs_from_mesh_ex(nv, _p(v), nf, _p(f), cbp.shape[0], _p(cbp), bdp.shape[0], _p(bdp), _p(rs),
                              _p(cs), C.byref(ns), _p(rb), _p(cb), C.byref(nb))
    return (rs, cs), (rb, cb)


def vertex_extend(verts4, faces, normal_index, normals3, dist):
    v = _c(verts4, np.float32).copy()
    f = _c(faces, np.uint32)
    ni = _c(normal_index, np.uint32)
    nrm = _c(normals3, np.float32)
    olib().or_vertex_extend(v.shape[0], _p(v), f.size // 3, _p(f), _p(ni), _p(nrm), dist)
    return v


# ------------------------------------------------------------------ LBVH
def set_threads(n):
    """Threads of the oracle's per-item loops (boxes, deformation, refit
    subtrees); results do not depend on it."""
    olib().or_set_threads(int(n))


def prim_boxes(verts, tris, depth):
    v, t = _c(verts, np.float32), _c(tris, np.uint32)
    m = t.size // 3
    out = np.zeros((m, 6), np.float32)
    olib().or_prim_boxes(_p(v), m, _p(t), depth, _p(out))
    return out


def root_box(boxes):
    b = _c(boxes, np.float32)
    out = np.zeros(6, np.float32)
    olib().or_root_box(b.shape[0], _p(b), _p(out))
    return out


def expand_bits(v):
    return olib().or_expand_bits(int(v))


def morton3d(x, y, z):
    return olib().or_morton3d(x, y, z)


def morton_codes(boxes, root):
    b = _c(boxes, np.float32)
    out = np.zeros(b.shape[0], np.uint32)
    olib().or_morton_codes(b.shape[0], _p(b), _p(_c(root, np.float32)), _p(out))
    return out


def sort_keys(codes):
    c = _c(codes, np.uint32)
    keys = np.zeros(c.size, np.uint64)
    order = np.zeros(c.size, np.uint32)
    olib().or_sort_keys(c.size, _p(c), _p(keys), _p(order))
    return keys, order


def dedupe_sorted(keys, order):
    k, o = keys.copy(), order.copy()
    n = olib().or_dedupe_sorted(k.size, _p(k), _p(o))
    return k[:n], o[:n]


def karras(keys):
    k = _c(keys, np.uint64)
    m = k.size
    mi = max(m - 1, 0)
    cA, cB, pI = (np.zeros(mi, np.int32) for _ in range(3))
    pL = np.zeros(m, np.int32)
    olib().or_karras(m, _p(k), _p(cA), _p(cB), _p(pI), _p(pL))
    return cA, cB, pI, pL


def refit(childA, childB, parent_int, box_leaf):
    m = box_leaf.shape[0]
    out = np.zeros((max(m - 1, 0), 6), np.float32)
    olib().or_refit(m, _p(_c(childA, np.int32)), _p(_c(childB, np.int32)),
                    _p(_c(parent_int, np.int32)), _p(_c(box_leaf, np.float32)), _p(out))
    return out


def bvh_build(verts, tris, depth, compat_dedupe=False):
    """BVHAccel construction (bvh.cu:264-297) with the oracle's defined root box
    and 64-bit keys; returns a dict of numpy arrays in the reference encoding."""
    v, t = _c(verts, np.float32), _c(tris, np.uint32)
    m = t.size // 3
    mi = max(m - 1, 0)
    d = dict(keys=np.zeros(m, np.uint64), order=np.zeros(m, np.uint32),
             leaf_tri=np.zeros((m, 3), np.uint32), childA=np.zeros(mi, np.int32),
             childB=np.zeros(mi, np.int32), parent_int=np.zeros(mi, np.int32),
             parent_leaf=np.zeros(m, np.int32), box_leaf=np.zeros((m, 6), np.float32),
             box_int=np.zeros((mi, 6), np.float32))
    mk = olib().or_bvh_build(_p(v), m, _p(t), depth, 1 if compat_dedupe else 0,
                             *(_p(d[k]) for k in ("keys", "order", "leaf_tri", "childA", "childB",
                                                  "parent_int", "parent_leaf", "box_leaf",
                                                  "box_int")))
    if mk != m:
        for k in ("keys", "order", "leaf_tri", "parent_leaf", "box_leaf"):
            d[k] = d[k][:mk].copy()
        for k in ("childA", "childB", "parent_int", "box_int"):
            d[k] = d[k][:max(mk - 1, 0)].copy()
    d["m"] = mk
    d["verts"] = v
    return d


class BvhHandle:
    """Keeps the arrays of an or_bvh view alive."""

    def __init__(self, d):
        self.d = d
        self.s = or_bvh(d["m"], _p(d["verts"]), _p(d["leaf_tri"]), _p(d["childA"]), _p(d["childB"]),
                        _p(d["box_int"]), _p(d["box_leaf"]))

    @property
    def ref(self):
        return C.byref(self.s)


def query(h, p):
    pp = np.asarray(p, np.float32)
    vis = C.c_int32()
    r = olib().or_query(h.ref, _p(pp), C.byref(vis))
    return r, vis.value


def plane_test(verts, tri, p):
    d = C.c_float()
    n = np.zeros(3, np.float32)
    r = olib().or_plane_test(_p(_c(verts, np.float32)), _p(_c(tri, np.uint32)),
                             _p(np.asarray(p, np.float32)), C.byref(d), _p(n))
    return bool(r), d.value, n


def collide_points(h, pts):
    p = _c(pts, np.float32)
    n = p.shape[0]
    hit = np.zeros(n, np.int32)
    ins = np.zeros(n, np.uint8)
    op = np.zeros((n, 3), np.float32)
    on = np.zeros((n, 3), np.float32)
    olib().or_collide_points(n, _p(p), h.ref, _p(hit), _p(ins), _p(op), _p(on))
    return hit, ins, op, on


# ------------------------------------------------------------------ step
def params_from(p):
    """Copy the physics fields of a cc_params-like object."""
    q = or_params()
    q.dt, q.mass, q.damp, q.kd = p.dt, p.mass, p.damp, p.kd
    for i in range(3):
        q.gravity[i] = p.gravity[i]
        q.k[i] = p.k[i]
    q.rest_scale, q.eps = p.rest_scale, p.eps
    return q


def _springs_args(springs):
    out = []
    for t in range(3):
        if t < len(springs):
            out += [_p(springs[t][0]), _p(springs[t][1])]
        else:
            out += [None, None]
    return out


def step_f32(pos, prev, contact, springs, prm, pins=None, bvh=None, nthreads=1, debug=False):
    n = pos.shape[0]
    pos, prev, contact = (_c(a, np.float32) for a in (pos, prev, contact))
    op, oq, oc = (np.zeros((n, 3), np.float32) for _ in range(3))
    pre = np.zeros((n, 3), np.float32) if debug else None
    hit = np.zeros(n, np.int32) if debug else None
    ins = np.zeros(n, np.uint8) if debug else None
    sp = [(_c(r, np.uint32), _c(c, SPRING_DTYPE)) for r, c in springs]
    pins_a = None if pins is None else _c(pins, np.uint8)
    olib().or_step_f32(n, _p(pos), _p(prev), _p(contact), *_springs_args(sp), C.byref(prm),
                       _p(pins_a), bvh.ref if bvh is not None else None, _p(op), _p(oq), _p(oc),
                       _p(pre), _p(hit), _p(ins), nthreads)
    if debug:
        return op, oq, oc, pre, hit, ins
    return op, oq, oc


def step_f64(pos, prev, contact, springs, prm, pins=None, bvh=None, with_hit=False):
    n = pos.shape[0]
    pos, prev, contact = (_c(a, np.float64) for a in (pos, prev, contact))
    op, oq, oc = (np.zeros((n, 3), np.float64) for _ in range(3))
    ins = np.zeros(n, np.uint8)
    hit = np.zeros(n, np.int32)
    sp = [(_c(r, np.uint32), _c(c, SPRING_DTYPE)) for r, c in springs]
    pins_a = None if pins is None else _c(pins, np.uint8)
    olib().or_step_f64(n, _p(pos), _p(prev), _p(contact), *_springs_args(sp), C.byref(prm),
                       _p(pins_a), bvh.ref if bvh is not None else None, _p(op), _p(oq), _p(oc),
                       _p(ins), _p(hit))
    if with_hit:
        return op, oq, oc, ins, hit
    return op, oq, oc, ins


def body_deform(rest, vstart, p0, p1, breath, R, T):
    r = _c(rest, np.float32)
    out = np.zeros_like(r)
    olib().or_body_deform(r.shape[0], _p(r), len(breath), _p(_c(vstart, np.uint32)),
                          _p(_c(p0, np.float32)), _p(_c(p1, np.float32)),
                          _p(_c(breath, np.float32)), _p(_c(R, np.float32)), _p(_c(T, np.float32)),
                          _p(out))
    return out


# ------------------------------------------------------------------ adjoint
def params64_from(p, k=None, damp=None, rest_scale=None):
    """Double-precision copy of cc_params-like p, with optional overrides of
    the differentiable parameters."""
    q = or_params64()
    q.dt, q.mass, q.damp, q.kd = float(p.dt), float(p.mass), float(p.damp), float(p.kd)
    for i in range(3):
        q.gravity[i] = float(p.gravity[i])
        q.k[i] = float(p.k[i])
    q.rest_scale, q.eps = float(p.rest_scale), float(p.eps)
    if k is not None:
        for i in range(3):
            q.k[i] = float(k[i])
    if damp is not None:
        q.damp = float(damp)
    if rest_scale is not None:
        q.rest_scale = float(rest_scale)
    return q


def rollout_f64(x0, v0, springs, prm64, pins=None, bvh=None, nsteps=1, trajectory=False):
    """FP64 rollout; returns x_final, or (x_final, traj_p, traj_q, traj_c, traj_a)."""
    n = x0.shape[0]
    x0 = _c(x0, np.float64)
    v0 = _c(np.zeros_like(x0) if v0 is None else v0, np.float64)
    xf = np.zeros((n, 3), np.float64)
    tp = tq = tc = to = None
    if trajectory:
        tp, tq, tc, to = (np.zeros((nsteps + 1, n, 3), np.float64) for _ in range(4))
    sp = [(_c(r, np.uint32), _c(c, SPRING_DTYPE)) for r, c in springs]
    pins_a = None if pins is None else _c(pins, np.uint8)
    olib().or_rollout_f64(n, _p(x0), _p(v0), *_springs_args(sp), C.byref(prm64), _p(pins_a),
                          bvh.ref if bvh is not None else None, nsteps, _p(xf), _p(tp), _p(tq), _p(tc), _p(to))
    return (xf, tp, tq, tc, to) if trajectory else xf


def rollout_f64_frozen(x0, v0, springs, prm64, pins, traj_c, traj_a, trajectory=False):
    """FP64 rollout with the recorded contacts held fixed; returns x_final,
    or (x_final, traj_p, traj_q)."""
    n = x0.shape[0]
    nsteps = traj_c.shape[0] - 1
    x0 = _c(x0, np.float64)
    v0 = _c(np.zeros_like(x0) if v0 is None else v0, np.float64)
    xf = np.zeros((n, 3), np.float64)
    tp = tq = None
    if trajectory:
        tp, tq = (np.zeros((nsteps + 1, n, 3), np.float64) for _ in range(2))
    sp = [(_c(r, np.uint32), _c(c, SPRING_DTYPE)) for r, c in springs]
    pins_a = None if pins is None else _c(pins, np.uint8)
    olib().or_rollout_f64_frozen(n, _p(x0), _p(v0), *_springs_args(sp), C.byref(prm64), _p(pins_a), nsteps,
                                 _p(_c(traj_c, np.float64)), _p(_c(traj_a, np.float64)), _p(xf), _p(tp), _p(tq))
    return (xf, tp, tq) if trajectory else xf


def adjoint_f64(springs, prm64, pins, traj_p, traj_q, traj_c, dl_dxf):
    """FP64 adjoint along a recorded trajectory (contacts held fixed);
    returns (dL/d(k0, k1, k2, damp, rest_scale), dL/dv0, dL/dx0)."""
    nsteps = traj_p.shape[0] - 1
    n = traj_p.shape[1]
    sp = [(_c(r, np.uint32), _c(c, SPRING_DTYPE)) for r, c in springs]
    pins_a = None if pins is None else _c(pins, np.uint8)
    dp = np.zeros(5, np.float64)
    dv0 = np.zeros((n, 3), np.float64)
    dx0 = np.zeros((n, 3), np.float64)
    olib().or_adjoint_f64(n, *_springs_args(sp), C.byref(prm64), _p(pins_a), nsteps,
                          _p(_c(traj_p, np.float64)), _p(_c(traj_q, np.float64)), _p(_c(traj_c, np.float64)),
                          _p(_c(dl_dxf, np.float64)), _p(dp), _p(dv0), _p(dx0))
    return dp, dv0, dx0


def pick(verts3, faces, origin, direction):
    """find_first_intersect_triangle (Ray.cpp:45-77): (face or -1, t)."""
    v = _c(verts3, np.float32)
    f = _c(faces, np.uint32)
    o = _c(origin, np.float32)
    d = _c(direction, np.float32)
    t = C.c_float()
    face = olib().or_pick(f.size // 3, _p(f), _p(v), _p(o), _p(d), C.byref(t))
    return int(face), float(t.value)


def vertex_normals(pos3, faces):
    """The VBO's vertex normals (verlet.cu:91-116, 163-181)."""
    p = _c(pos3, np.float32)
    f = _c(faces, np.uint32)
    out = np.zeros((p.shape[0], 3), np.float32)
    olib().or_vertex_normals(p.shape[0], _p(p), f.size // 3, _p(f), _p(out))
    return out
