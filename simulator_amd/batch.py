"""Batched instances over the C ABI (cc_batch_*).

`instances(count, seed)` generates the batched workload: `count` independent
grid cloths of side log-uniform in [16, 192] over their own moving
capsule-union bodies of 2k-40k triangles (6 capsules, rigid sway and
breathing, LBVH rebuilt every frame), each with spring constants drawn
within +/-30% of the small workload's, dt = 1e-3.  `Batch(scenes)` advances
all of them together (bodies posed and rebuilt in a few launches, then one
step launch); it equals one `Simulator` per scene stepped in turn, bit for
bit.

The reference holds a single Simulator per Scene (simulator.h:9-59,
scene.cpp:176-186); a batch is that object repeated, with the per-vertex step
of verlet.cu:129-161 shared by every instance.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from . import scene as S
from .simulator import body_desc

SIDE_RANGE = (16, 192)
CAPSULE_S = (16, 64)        # capsule tessellation: nseg = s, nring = ncyl = s / 4
BATCH_CAPS = 6
K_BASE = (1000.0, 500.0, 100.0)


def capsule_triangles(s, ncaps=BATCH_CAPS):
    """Triangle count of a capsule body of tessellation s (two pole fans and
    2 * nring + ncyl - 2 quad rows per capsule)."""
    nring = ncyl = s // 4
    return ncaps * s * 2 * (2 * nring + ncyl - 1)


def instances(count=256, seed=0, steps=300, mode=L.CC_BODY_REBUILD):
    """Seeded scenes of the batched workload (see the module docstring);
    `mode` is the bodies' per-frame update (REBUILD: posed and rebuilt every
    frame, REFIT: posed and refitted, STATIC: fixed)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(SIDE_RANGE[0]), np.log(SIDE_RANGE[1])
    out = []
    for k in range(count):
        side = int(round(float(np.exp(rng.uniform(lo, hi)))))
        s = 4 * int(rng.integers(CAPSULE_S[0] // 4, CAPSULE_S[1] // 4 + 1))
        ks = tuple(float(b * rng.uniform(0.7, 1.3)) for b in K_BASE)
        bseed = int(rng.integers(1, 2**31 - 1))
        v, t, nrm, caps = S.capsule_body(BATCH_CAPS, 0.04, 0.15, *S.CAPSULE_BOX, s, s // 4, s // 4, bseed)
        body = S.Body(v, t, nrm, mode, True, caps, S._motion(bseed))
        sc = S.grid_scene("batched", side, 1e-3, steps, body, k=ks, box_depth=0.002, seed=1000 + k)
        out.append(sc)
    return out


def _cloth_desc(sc, keep):
    v = np.concatenate([sc.pos.astype(np.float32), np.ones((sc.n, 1), np.float32)], axis=1)
    v = np.ascontiguousarray(v)
    sp = [(np.ascontiguousarray(r, np.uint32), np.ascontiguousarray(c, L.SPRING_DTYPE)) for r, c in sc.springs]
    keep += [v, sp]
    cd = L.cc_cloth_desc()
    cd.n = sc.n
    cd.verts4 = C.cast(L.ptr(v), C.POINTER(C.c_float))
    cd.ntypes = len(sp)
    for t, (r, c) in enumerate(sp):
        cd.row[t] = C.cast(L.ptr(r), C.POINTER(C.c_uint32))
        cd.col[t] = C.cast(L.ptr(c), C.POINTER(L.cc_spring))
    return cd


def _body_desc(body, keep):
    if body is None or body.mode == L.CC_BODY_NONE:
        bd = L.cc_body_desc()
        bd.mode = L.CC_BODY_NONE
        return bd
    return body_desc(body, keep)


class Batch:
    """All scenes advanced together; instance k is scenes[k]."""

    def __init__(self, scenes):
        lib = L.lib()
        self.count = len(scenes)
        if self.count < 1:
            raise ValueError("Batch needs at least one scene")
        keep = []
        cds = (L.cc_cloth_desc * self.count)(*[_cloth_desc(sc, keep) for sc in scenes])
        bds = (L.cc_body_desc * self.count)(*[_body_desc(sc.body, keep) for sc in scenes])
        prm = (L.cc_params * self.count)(*[sc.params for sc in scenes])
        self.sizes = [sc.n for sc in scenes]
        self._body_m = [0 if sc.body is None else sc.body.tris.shape[0] for sc in scenes]
        self._body_nv = [0 if sc.body is None else sc.body.verts.shape[0] for sc in scenes]
        h = C.c_void_p()
        L.check(lib.cc_batch_create(cds, bds, prm, self.count, C.byref(h)))
        self._h = h
        for k, sc in enumerate(scenes):
            if sc.pins is not None and len(sc.pins):
                self.set_pins(k, sc.pins)

    def close(self):
        if getattr(self, "_h", None):
            L.lib().cc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def simulate(self, nsteps=1, stream=None):
        L.check(L.lib().cc_batch_step(self._h, int(nsteps), stream))

    def check(self):
        """cc_sim_check_finite over every instance (cc_batch_check)."""
        L.check(L.lib().cc_batch_check(self._h))

    def synchronize(self):
        L.check(L.lib().cc_batch_synchronize(self._h))

    def set_pins(self, inst, idx):
        a = np.ascontiguousarray(idx, np.uint32)
        L.check(L.lib().cc_batch_set_pins(self._h, int(inst), L.ptr(a), a.size))

    def get_state(self, inst):
        n = self.sizes[inst]
        x, xp, c = (np.zeros((n, 3), np.float32) for _ in range(3))
        L.check(L.lib().cc_batch_get_state(self._h, int(inst), L.ptr(x), L.ptr(xp), L.ptr(c)))
        return x, xp, c

    def tuning(self):
        """The tuner's ms per step for tiles 64 / 128 / 256 x order 0 / 1
        (a (3, 2) array) and whether it timed the fused sort + step launch."""
        ms = np.zeros(6, np.float32)
        fused = C.c_uint32()
        L.check(L.lib().cc_batch_tuning(self._h, L.ptr(ms), C.byref(fused)))
        return ms.reshape(3, 2), bool(fused.value)

    def info(self):
        """(total vertices, workgroup size, workgroup order)."""
        v, t, o = C.c_uint32(), C.c_uint32(), C.c_uint32()
        L.check(L.lib().cc_batch_info(self._h, C.byref(v), C.byref(t), C.byref(o)))
        return v.value, t.value, o.value

    @property
    def frame(self):
        f = C.c_uint64()
        L.check(L.lib().cc_batch_frame(self._h, C.byref(f)))
        return f.value

    def body_export(self, inst):
        """(vertices, LBVH dict in the keys of Simulator.bvh_export) of
        instance inst's body at the current frame."""
        m, nv = self._body_m[inst], self._body_nv[inst]
        mi = max(m - 1, 0)
        verts = np.zeros((nv, 3), np.float32)
        d = {"keys": np.zeros(m, np.uint64), "order": np.zeros(m, np.uint32),
             "childA": np.zeros(mi, np.int32), "childB": np.zeros(mi, np.int32),
             "parent_int": np.zeros(mi, np.int32), "parent_leaf": np.zeros(m, np.int32),
             "box_int": np.zeros((mi, 6), np.float32), "box_leaf": np.zeros((m, 6), np.float32),
             "root": np.zeros(6, np.float32)}
        L.check(L.lib().cc_batch_body_export(self._h, int(inst), L.ptr(verts), L.ptr(d["keys"]), L.ptr(d["order"]),
                                             L.ptr(d["childA"]), L.ptr(d["childB"]), L.ptr(d["parent_int"]),
                                             L.ptr(d["parent_leaf"]), L.ptr(d["box_int"]), L.ptr(d["box_leaf"]),
                                             L.ptr(d["root"])))
        return verts, d
