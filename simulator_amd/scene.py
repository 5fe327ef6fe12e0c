"""Seeded synthetic scenes for the five workloads (the reference reads OBJ
files; these inputs are generated, identically on every machine, by the
library's splitmix64-seeded host generators).

  small    64x64 grid cloth over a level-4 icosphere (static LBVH), 200 steps
  garment  111x111 cloth over 12 capsules (~50k triangles), LBVH rebuilt per frame
  large    512x512 cloth over a 12-capsule body (~500k triangles), refit per frame
  batched  256 uneven cloth / body instances (see simulator_amd.batch)
  adjoint  128x128 cloth over the icosphere, 200 steps, differentiable
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


@dataclass
class Body:
    verts: np.ndarray            # (nv, 3) float32
    tris: np.ndarray             # (nf, 3) uint32
    normals: np.ndarray          # (nv, 3) float32
    mode: int = L.CC_BODY_STATIC
    extend: bool = True
    caps: dict = None            # vstart, p0, p1, radius, phase
    motion: dict = field(default_factory=dict)


@dataclass
class Scene:
    name: str
    nx: int
    nz: int
    pos: np.ndarray              # (n, 3) float32
    springs: list                # [(row uint32[n+1], col SPRING_DTYPE[nnz])] x 3
    pins: np.ndarray             # uint32
    params: object               # L.cc_params
    body: Body = None
    steps: int = 200

    @property
    def n(self):
        return self.pos.shape[0]


def grid_cloth(nx, nz, spacing, y0=1.0, jitter=1e-4, seed=0):
    pos = np.zeros((nx * nz, 3), np.float32)
    L.check(L.lib().cc_scene_grid_cloth(nx, nz, spacing, y0, jitter, seed, L.ptr(pos)))
    return pos


def _csr_to_numpy(c):
    n, nnz = c.n, c.nnz
    row = np.ctypeslib.as_array(c.row, shape=(n + 1,)).copy()
    col = np.frombuffer(C.string_at(C.cast(c.col, C.c_void_p), nnz * 8), dtype=L.SPRING_DTYPE).copy()
    return row, col


def grid_springs(nx, nz, pos):
    """Structural, shear and bend CSR sets of a grid cloth."""
    out = (L.cc_csr * 3)()
    L.check(L.lib().cc_springs_grid(nx, nz, L.ptr(np.ascontiguousarray(pos, np.float32)), out))
    res = [_csr_to_numpy(out[t]) for t in range(3)]
    for t in range(3):
        L.lib().cc_csr_free(C.byref(out[t]))
    return res


def springs_from_mesh(verts4, faces, vertex_object=None):
    """Springs::Springs + CSR_structure_spring / CSR_bend_spring
    (spring.cpp:82-391): a SINGLE_LAYER_NOB mesh, or, given the vertex
    counts of its groups (ObjLoader.vertex_object, pairs {piece, boundary}),
    a SINGLE_LAYER_BOUNDARY mesh with its boundary springs (:98-209)."""
    verts4 = np.ascontiguousarray(verts4, np.float32)
    faces = np.ascontiguousarray(faces, np.uint32)
    s, b = L.cc_csr(), L.cc_csr()
    if vertex_object is None:
        L.check(L.lib().cc_springs_from_mesh(L.ptr(verts4), verts4.shape[0], L.ptr(faces),
                                             faces.size // 3, C.byref(s), C.byref(b)))
    else:
        g = np.ascontiguousarray(vertex_object, np.uint32)
        L.check(L.lib().cc_springs_from_mesh_groups(L.ptr(verts4), verts4.shape[0], L.ptr(faces),
                                                    faces.size // 3, L.ptr(g), g.size, C.byref(s),
                                                    C.byref(b)))
    res = (_csr_to_numpy(s), _csr_to_numpy(b))
    L.lib().cc_csr_free(C.byref(s))
    L.lib().cc_csr_free(C.byref(b))
    return res


def icosphere(level, radius, center):
    nv = 10 * 4 ** level + 2
    nf = 20 * 4 ** level
    v = np.zeros((nv, 3), np.float32)
    t = np.zeros((nf, 3), np.uint32)
    nrm = np.zeros((nv, 3), np.float32)
    c = (C.c_double * 3)(*center)
    L.check(L.lib().cc_scene_icosphere(level, radius, c, L.ptr(v), L.ptr(t), L.ptr(nrm)))
    return v, t, nrm


def capsule_body(ncaps, rmin, rmax, lo, hi, nseg, nring, ncyl, seed):
    lo_c, hi_c = (C.c_double * 3)(*lo), (C.c_double * 3)(*hi)
    nv, nf = C.c_uint32(), C.c_uint32()
    lib = L.lib()
    L.check(lib.cc_scene_capsule_body(ncaps, rmin, rmax, lo_c, hi_c, nseg, nring, ncyl, seed,
                                      None, None, None, C.byref(nv), C.byref(nf), None))
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nf.value, 3), np.uint32)
    nrm = np.zeros((nv.value, 3), np.float32)
    caps = {"vstart": np.zeros(ncaps + 1, np.uint32), "p0": np.zeros((ncaps, 3), np.float32),
            "p1": np.zeros((ncaps, 3), np.float32), "radius": np.zeros(ncaps, np.float32),
            "phase": np.zeros(ncaps, np.float32)}
    cs = L.cc_capsules(ncaps, *(C.cast(L.ptr(caps[k]), C.POINTER(T)) for k, T in
                                (("vstart", C.c_uint32), ("p0", C.c_float), ("p1", C.c_float),
                                 ("radius", C.c_float), ("phase", C.c_float))))
    L.check(lib.cc_scene_capsule_body(ncaps, rmin, rmax, lo_c, hi_c, nseg, nring, ncyl, seed,
                                      L.ptr(v), L.ptr(t), L.ptr(nrm), C.byref(nv), C.byref(nf),
                                      C.byref(cs)))
    return v, t, nrm, caps


def params(dt=1e-3, k=(1000.0, 500.0, 100.0), damp=-0.02, mass=0.3, gravity=-9.8,
           box_depth=0.01, rest_scale=1.0):
    p = L.default_params()
    p.dt = dt
    p.mass = mass
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, gravity, 0.0
    p.damp = damp
    p.k[0], p.k[1], p.k[2] = k
    p.box_depth = box_depth
    p.rest_scale = rest_scale
    return p


DEFAULT_MOTION = dict(amp=0.05, freq=1.0, rot_amp=0.1, breath_amp=0.02, breath_freq=2.0)
CAPSULE_BOX = ((0.15, 0.45, -0.35), (0.85, 0.8, 0.35))


def _motion(seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    m = dict(DEFAULT_MOTION)
    m["dir"] = tuple(float(x) for x in d)
    m["pivot"] = (0.5, 0.6, 0.0)
    return m


def grid_scene(name, side, dt, steps, body, k=(1000.0, 500.0, 100.0), damp=-0.02,
               rest_scale=1.0, box_depth=0.01, seed=0, y0=1.0):
    spacing = 1.0 / (side - 1)
    pos = grid_cloth(side, side, spacing, y0, 1e-4, seed)
    springs = grid_springs(side, side, pos)
    pins = np.array([0, side - 1], np.uint32)      # the two corners at z = -0.5
    prm = params(dt=dt, k=k, damp=damp, box_depth=box_depth, rest_scale=rest_scale)
    return Scene(name, side, side, pos, springs, pins, prm, body, steps)


def sphere_body(level=4):
    v, t, nrm = icosphere(level, 0.3, (0.5, 0.5, 0.0))
    return Body(v, t, nrm, L.CC_BODY_STATIC, True)


def capsules(nseg, nring, ncyl, mode, seed=7):
    lo, hi = CAPSULE_BOX
    v, t, nrm, caps = capsule_body(12, 0.04, 0.15, lo, hi, nseg, nring, ncyl, seed)
    return Body(v, t, nrm, mode, True, caps, _motion(seed))


def small(side=64, steps=200, y0=1.0):
    """64x64 cloth over a 2562-vertex / 5120-triangle icosphere, static LBVH."""
    return grid_scene("small", side, 1e-3, steps, sphere_body(), y0=y0)


def garment(side=111, steps=1000):
    """111x111 cloth over 12 capsules (50,688 triangles), LBVH rebuilt every frame."""
    return grid_scene("garment", side, 1e-3, steps, capsules(64, 14, 6, L.CC_BODY_REBUILD),
                      box_depth=0.002)


def large(side=512, steps=500):
    """512x512 cloth over 12 capsules (501,120 triangles), LBVH refitted every frame."""
    return grid_scene("large", side, 5e-4, steps, capsules(144, 64, 18, L.CC_BODY_REFIT),
                      box_depth=0.002)


def adjoint_scene(side=128, steps=200, k=(1000.0, 500.0, 100.0), damp=-0.02, rest_scale=1.0):
    """128x128 cloth over the 5120-triangle icosphere (differentiable workload)."""
    return grid_scene("adjoint", side, 1e-3, steps, sphere_body(), k=k, damp=damp,
                      rest_scale=rest_scale)


def reference_scene(cloth_obj, body_obj, steps=200):
    """The reference's own scene (main.cpp:24-35): tshirt2.obj rotated 90
    degrees about x and -4 about z, scaled 3.33 about its centre and moved by
    (0, 1.98, 0.02), over female.obj scaled 3.226 and moved by (0, 1.8, 0);
    structure and bend springs from the mesh (Springs(Mesh*), spring.cpp:
    82-91), the reference's constants (verlet.cu:13-21), a static body
    extended by 0.003 along its normals (simulation.cu:127-138), no pins."""
    from .simulator import Mesh, X, Z
    cloth = Mesh.from_obj(cloth_obj)
    cloth.rotation(90, X)
    cloth.rotation(-4, Z)
    cloth.scale(3.33)
    cloth.translate(0, 1.98, 0.02)
    body = Mesh.from_obj(body_obj)
    body.scale(3.226)
    body.translate(0, 1.8, 0)
    return mesh_scene("reference", cloth.vertices[:, :3], cloth.faces, body.vertices[:, :3], body.faces,
                      body.normals, steps)


def mesh_scene(name, cloth_verts, cloth_faces, body_verts, body_faces, body_normals, steps=200):
    """A cloth mesh over a static body mesh with the reference's constants and
    its mesh springs (structure, then bend)."""
    pos = np.ascontiguousarray(cloth_verts, np.float32)
    v4 = np.concatenate([pos, np.ones((pos.shape[0], 1), np.float32)], axis=1)
    s, b = springs_from_mesh(v4, np.asarray(cloth_faces, np.uint32))
    body = Body(np.ascontiguousarray(body_verts, np.float32), np.asarray(body_faces, np.uint32),
                np.ascontiguousarray(body_normals, np.float32), L.CC_BODY_STATIC, True)
    sc = Scene(name, 0, 0, pos, [s, b], np.zeros(0, np.uint32), L.default_params(), body, steps)
    sc.faces = np.asarray(cloth_faces, np.uint32)
    return sc


def load_mesh_scene(npz_path, name="reference", steps=200):
    """mesh_scene from arrays saved as cloth_verts / cloth_faces / body_verts /
    body_faces / body_normals (e.g. tests/golden/ref_tshirt_female.npz, the
    reference's own scene)."""
    f = np.load(npz_path)
    return mesh_scene(name, f["cloth_verts"], f["cloth_faces"], f["body_verts"], f["body_faces"],
                      f["body_normals"], steps)


def grid_faces(nx, nz):
    """Two triangles per cell of an nx x nz grid cloth (vertex (i, j) =
    j * nx + i), as the faces of its VBO / picking mesh."""
    f = []
    for j in range(nz - 1):
        for i in range(nx - 1):
            a = j * nx + i
            f += [(a, a + nx, a + 1), (a + 1, a + nx, a + nx + 1)]
    return np.asarray(f, np.uint32).reshape(-1, 3)
