"""Host-side mirror of the reference's simulator interface over the C ABI.

Reference classes (sutongkui/simulator):
  Simulator  src/simulator.h:9-59, src/simulation.cu:47-257
  Springs    src/spring.h:14-36,   src/spring.cpp:82-391
  BVHAccel   src/bvh/bvh.h:20-101, src/bvh/bvh.cu:264-297
Method names and argument meaning follow the reference; errors raise
ClothCoreError instead of printing and continuing (common.h:4-10).
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L
from . import scene as S


class ObjLoader:
    """ObjLoader (src/ObjLoader.h:19-38, ObjLoader.cpp:15-216), geometry
    only, through cc_obj_read: vertices (n, 4), normals (nvn, 3), tex
    (nvt, 2) and faces as vertex / tex / normal index triples (0-based,
    unsigned as the reference stores them), and the vertex counts of its
    "# object" groups (vertex_object)."""

    def __init__(self, file):
        o = L.cc_obj()
        L.check(L.lib().cc_obj_read(str(file).encode(), C.byref(o)))
        try:
            def arr(ptr, n, k, dt):
                return np.ctypeslib.as_array(ptr, shape=(n * k,)).astype(dt).reshape(n, k) if n else \
                    np.zeros((0, k), dt)
            self.obj_file = str(file)
            self.vertices = arr(o.verts4, o.nv, 4, np.float32)
            self.normals = arr(o.normals3, o.nvn, 3, np.float32)
            self.tex = arr(o.tex2, o.nvt, 2, np.float32)
            self.vertex_index = arr(o.vidx, o.nf, 3, np.uint32)
            self.tex_index = arr(o.tidx, o.nf, 3, np.uint32)
            self.normal_index = arr(o.nidx, o.nf, 3, np.uint32)
            # vertex counts of the "# object" groups (vertex_object, without
            # the names), in file order
            self.vertex_object = (np.ctypeslib.as_array(o.group_counts, shape=(o.ngroups,)).astype(np.uint32)
                                  if o.ngroups else np.zeros(0, np.uint32))
            # per-vertex normals / uv as Mesh::unified makes them (None when a
            # face names a vertex the file does not have)
            self.unified_normals = self.unified_tex = None
            if not o.nf or int(self.vertex_index.max()) < o.nv:
                self.unified_normals = np.zeros((o.nv, 3), np.float32)
                self.unified_tex = np.zeros((o.nv, 2), np.float32)
                L.check(L.lib().cc_mesh_unify(C.byref(o), L.ptr(self.unified_normals),
                                              L.ptr(self.unified_tex)))
        finally:
            L.lib().cc_obj_free(C.byref(o))


X, Y, Z = 0, 1, 2      # direction (Mesh.h:6)


class Mesh:
    """The parts of Mesh (src/Mesh.h:8-51) the simulation path reads:
    vertices as vec4 (w = 1), faces as vertex index triples, normals; built
    from arrays or, as Mesh(string) does, from an OBJ file (Mesh.from_obj)."""

    def __init__(self, vertices, faces, normals=None, normal_index=None):
        v = np.asarray(vertices, np.float32)
        if v.shape[1] == 3:
            v = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], axis=1)
        self.vertices = np.ascontiguousarray(v)
        self.faces = np.ascontiguousarray(np.asarray(faces, np.uint32).reshape(-1, 3))
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float32)
        self.normal_index = None if normal_index is None else np.ascontiguousarray(normal_index, np.uint32)

    @classmethod
    def from_obj(cls, file):
        """Mesh(string file) = ObjLoader + Mesh::unified (Mesh.cpp:11-38)."""
        o = ObjLoader(file)
        if o.unified_normals is None:
            raise ValueError(f"{file}: a face names a vertex the file does not have")
        m = cls(o.vertices, o.vertex_index, o.unified_normals, o.normal_index)
        m.tex = o.unified_tex
        return m

    def get_center(self):
        c = np.zeros(3, np.float32)
        L.check(L.lib().cc_mesh_center(L.ptr(self.vertices), self.vertices.shape[0], L.ptr(c)))
        return c

    def scale(self, S):
        """Mesh::scale (Mesh.cpp:40-50): about the center."""
        L.check(L.lib().cc_mesh_scale(L.ptr(self.vertices), self.vertices.shape[0], float(S)))

    def translate(self, x, y, z):
        L.check(L.lib().cc_mesh_translate(L.ptr(self.vertices), self.vertices.shape[0], float(x), float(y),
                                          float(z)))

    def rotation(self, angle, direction):
        """Mesh::rotation (Mesh.cpp:61-92): degrees about the center."""
        L.check(L.lib().cc_mesh_rotate(L.ptr(self.vertices), self.vertices.shape[0], float(angle),
                                       int(direction)))

    @property
    def vertex_indices(self):
        return self.faces.reshape(-1)

    def vertex_extend(self, dist):
        """Mesh::vertex_extend (Mesh.cpp:101-117): the normal of the first
        corner visiting each vertex (its normal index into the per-vertex
        normals, as the reference indexes them)."""
        if self.normals is None:
            raise ValueError("vertex_extend needs normals")
        ni = None if self.normal_index is None else L.ptr(self.normal_index)
        if self.normal_index is not None and self.normal_index.size and \
                int(self.normal_index.max()) >= self.normals.shape[0]:
            raise ValueError("vertex_extend: a normal index exceeds the per-vertex normals")
        L.check(L.lib().cc_mesh_vertex_extend(L.ptr(self.vertices), 4, self.vertices.shape[0],
                                              L.ptr(self.faces), self.faces.shape[0], ni,
                                              L.ptr(self.normals), float(dist)))


class Springs:
    """Springs(Mesh*) (spring.cpp:82-91): structure and bend CSR sets."""

    def __init__(self, mesh):
        self._s, self._b = S.springs_from_mesh(mesh.vertices, mesh.faces)

    def CSR_structure_spring(self):
        return self._s

    def CSR_bend_spring(self):
        return self._b


def _params_copy(p):
    q = L.cc_params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(L.cc_params))
    return q


def body_desc(body, keep):
    """cc_body_desc of a scene.Body; the arrays it points to are appended to
    `keep`, which must outlive the create call."""
    bv = np.ascontiguousarray(body.verts, np.float32)
    bt = np.ascontiguousarray(body.tris, np.uint32)
    bn = np.ascontiguousarray(body.normals, np.float32)
    keep += [bv, bt, bn]
    bd = L.cc_body_desc()
    bd.nv, bd.stride, bd.nf = bv.shape[0], 3, bt.shape[0]
    bd.verts = C.cast(L.ptr(bv), C.POINTER(C.c_float))
    bd.tris = C.cast(L.ptr(bt), C.POINTER(C.c_uint32))
    bd.normals3 = C.cast(L.ptr(bn), C.POINTER(C.c_float))
    bd.mode = body.mode
    bd.extend = 1 if body.extend else 0
    if body.caps is not None:
        caps = {k: np.ascontiguousarray(a) for k, a in body.caps.items()}
        keep.append(caps)
        cs = L.cc_capsules(len(caps["radius"]),
                           *(C.cast(L.ptr(caps[k]), C.POINTER(T)) for k, T in
                             (("vstart", C.c_uint32), ("p0", C.c_float), ("p1", C.c_float),
                              ("radius", C.c_float), ("phase", C.c_float))))
        keep.append(cs)
        bd.caps = C.pointer(cs)
    m = body.motion or {}
    bd.amp = m.get("amp", 0.0)
    bd.freq = m.get("freq", 0.0)
    bd.rot_amp = m.get("rot_amp", 0.0)
    bd.breath_amp = m.get("breath_amp", 0.0)
    bd.breath_freq = m.get("breath_freq", 0.0)
    for c in range(3):
        bd.dir[c] = m.get("dir", (0, 0, 0))[c]
        bd.pivot[c] = m.get("pivot", (0, 0, 0))[c]
    return bd


class Simulator:
    """Simulator(Mesh& cloth, Mesh& body) (simulation.cu:47-52).

    cloth_vertices: (n, 3|4) float32; springs: list of 1..3 (row, col) CSR
    sets summed in that order; body: scene.Body or None; params: cc_params.
    """

    def __init__(self, cloth_vertices, springs, body=None, params=None, pins=None):
        lib = L.lib()
        v = np.asarray(cloth_vertices, np.float32)
        if v.shape[1] == 3:
            v = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], axis=1)
        self._v4 = np.ascontiguousarray(v)
        self.n = self._v4.shape[0]
        self._springs = [(np.ascontiguousarray(r, np.uint32), np.ascontiguousarray(c, L.SPRING_DTYPE))
                         for r, c in springs]
        self.params = _params_copy(params) if params is not None else L.default_params()
        cd = L.cc_cloth_desc()
        cd.n = self.n
        cd.verts4 = C.cast(L.ptr(self._v4), C.POINTER(C.c_float))
        cd.ntypes = len(self._springs)
        for t, (r, c) in enumerate(self._springs):
            cd.row[t] = C.cast(L.ptr(r), C.POINTER(C.c_uint32))
            cd.col[t] = C.cast(L.ptr(c), C.POINTER(L.cc_spring))
        self.body = body
        bd_ptr = None
        keep = []
        if body is not None and body.mode != L.CC_BODY_NONE:
            bd = body_desc(body, keep)
            bd_ptr = C.byref(bd)
        h = C.c_void_p()
        L.check(lib.cc_sim_create(C.byref(cd), bd_ptr, C.byref(self.params), C.byref(h)))
        self._h = h
        if pins is not None and len(pins):
            self.set_pins(pins)

    @classmethod
    def from_scene(cls, sc):
        return cls(sc.pos, sc.springs, sc.body, sc.params, sc.pins)

    def close(self):
        if getattr(self, "_h", None):
            L.lib().cc_sim_destroy(self._h)
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

    # ---- reference interface
    def simulate(self, nsteps=1, stream=None):
        """Simulator::simulate (simulation.cu:141-150); asynchronous."""
        L.check(L.lib().cc_sim_step(self._h, int(nsteps), stream))

    def update_vertex(self, new_value, idx):
        """Simulator::update_vertex (simulation.cu:253-257)."""
        v = np.ascontiguousarray(np.asarray(new_value, np.float32).reshape(3))
        L.check(L.lib().cc_sim_update_vertex(self._h, int(idx), L.ptr(v)))

    def cuda_update_vbo(self, dst_ptr, nbytes, n_tex=0, lag=False):
        """cuda_update_vbo (simulation.cu:196-231) into a device buffer with
        the reference layout [vec4 pos][vec2 uv][vec3 normal]."""
        L.check(L.lib().cc_sim_write_vbo(self._h, C.c_void_p(dst_ptr), int(nbytes), int(n_tex),
                                         1 if lag else 0))

    # ---- extensions
    def synchronize(self):
        L.check(L.lib().cc_sim_synchronize(self._h))

    def set_pins(self, idx):
        idx = np.ascontiguousarray(idx, np.uint32)
        L.check(L.lib().cc_sim_set_pins(self._h, L.ptr(idx), idx.size))

    def set_params(self, params):
        self.params = _params_copy(params)
        L.check(L.lib().cc_sim_set_params(self._h, C.byref(self.params)))

    def get_state(self):
        x = np.zeros((self.n, 3), np.float32)
        xp = np.zeros((self.n, 3), np.float32)
        c = np.zeros((self.n, 3), np.float32)
        L.check(L.lib().cc_sim_get_state(self._h, L.ptr(x), L.ptr(xp), L.ptr(c)))
        return x, xp, c

    def set_state(self, x=None, xp=None, c=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
        x, xp, c = f(x), f(xp), f(c)
        L.check(L.lib().cc_sim_set_state(self._h, L.ptr(x), L.ptr(xp), L.ptr(c)))

    @property
    def frame(self):
        f = C.c_uint64()
        L.check(L.lib().cc_sim_frame(self._h, C.byref(f)))
        return f.value

    @property
    def block_steps(self):
        """Steps per captured block: simulate() replays whole blocks from
        graphs (a moving body's poses computed a block ahead on the update
        lanes) and launches the remaining steps one by one."""
        g = C.c_int()
        L.check(L.lib().cc_sim_block_steps(self._h, C.byref(g)))
        return g.value

    def set_frame(self, frame):
        L.check(L.lib().cc_sim_set_frame(self._h, int(frame)))

    def save(self, path):
        """Checkpoint (cc_sim_save): state, contact normals, pins,
        parameters and frame to a binary file."""
        L.check(L.lib().cc_sim_save(self._h, os.fsencode(path)))

    def load(self, path):
        """Restore a checkpoint written by save() for this cloth (cc_sim_load)."""
        L.check(L.lib().cc_sim_load(self._h, os.fsencode(path)))

    def set_debug(self, on=True):
        L.check(L.lib().cc_sim_set_debug(self._h, 1 if on else 0))

    def get_debug(self):
        pre = np.zeros((self.n, 3), np.float32)
        hit = np.zeros(self.n, np.int32)
        ins = np.zeros(self.n, np.uint8)
        L.check(L.lib().cc_sim_get_debug(self._h, L.ptr(pre), L.ptr(hit), L.ptr(ins)))
        return pre, hit, ins

    def set_counting(self, on=True):
        L.check(L.lib().cc_sim_set_counting(self._h, 1 if on else 0))

    def counters(self):
        a, b = C.c_uint64(), C.c_uint64()
        L.check(L.lib().cc_sim_get_counters(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def check_finite(self):
        """Raise on a non-finite position (CC_ERR_NONFINITE) or a traversal
        fault since the last check (CC_ERR_TRAVERSAL)."""
        L.check(L.lib().cc_sim_check_finite(self._h))

    def set_stack_limit(self, limit):
        """Test hook: the traversal stack limit (1..64, default 64)."""
        L.check(L.lib().cc_sim_set_stack_limit(self._h, int(limit)))

    # ---- body / LBVH
    @property
    def bvh_size(self):
        m = C.c_uint32()
        L.check(L.lib().cc_bvh_size(self._h, C.byref(m)))
        return m.value

    def body_vertices(self):
        nv = self.body.verts.shape[0]
        v = np.zeros((nv, 3), np.float32)
        L.check(L.lib().cc_body_get_vertices(self._h, L.ptr(v)))
        return v

    def body_motion(self):
        nc = len(self.body.caps["radius"])
        R = np.zeros(9, np.float32)
        T = np.zeros(3, np.float32)
        b = np.zeros(nc, np.float32)
        L.check(L.lib().cc_body_get_motion(self._h, L.ptr(R), L.ptr(T), L.ptr(b)))
        return R, T, b

    def set_body_vertices(self, verts3):
        v = np.ascontiguousarray(verts3, np.float32)
        L.check(L.lib().cc_body_set_vertices(self._h, L.ptr(v)))

    def bvh_rebuild(self):
        L.check(L.lib().cc_bvh_rebuild(self._h))

    def bvh_refit(self):
        L.check(L.lib().cc_bvh_refit(self._h))

    def set_bvh_width(self, width):
        """Traversal through the four-wide form (4, default) or the binary
        tree (2); the contacts are the same."""
        L.check(L.lib().cc_sim_set_bvh_width(self._h, int(width)))

    def bvh_width(self):
        """(width in use, wide node count, deferral bound of the tree)."""
        w, nw, sb = C.c_int(), C.c_uint32(), C.c_uint32()
        L.check(L.lib().cc_bvh_width(self._h, C.byref(w), C.byref(nw), C.byref(sb)))
        return w.value, nw.value, sb.value

    def bvh_export(self):
        m = self.bvh_size
        mi = max(m - 1, 0)
        out = dict(keys=np.zeros(m, np.uint64), order=np.zeros(m, np.uint32),
                   childA=np.zeros(mi, np.int32), childB=np.zeros(mi, np.int32),
                   parent_int=np.zeros(mi, np.int32), parent_leaf=np.zeros(m, np.int32),
                   box_int=np.zeros((mi, 6), np.float32), box_leaf=np.zeros((m, 6), np.float32),
                   root=np.zeros(6, np.float32))
        L.check(L.lib().cc_bvh_export(self._h, *(L.ptr(out[k]) for k in
                                                 ("keys", "order", "childA", "childB", "parent_int",
                                                  "parent_leaf", "box_int", "box_leaf", "root"))))
        return out

    def collide_points(self, pts):
        pts = np.ascontiguousarray(pts, np.float32)
        n = pts.shape[0]
        hit = np.zeros(n, np.int32)
        ins = np.zeros(n, np.uint8)
        op = np.zeros((n, 3), np.float32)
        on = np.zeros((n, 3), np.float32)
        L.check(L.lib().cc_collide_points(self._h, L.ptr(pts), n, L.ptr(hit), L.ptr(ins), L.ptr(op),
                                          L.ptr(on)))
        return hit, ins, op, on

    def set_faces(self, faces):
        f = np.ascontiguousarray(faces, np.uint32)
        L.check(L.lib().cc_sim_set_faces(self._h, L.ptr(f), f.size // 3))

    def pick(self, origin, direction):
        """(face, vertex, t) of the first cloth face along the ray
        (Scene::OnMouseDown + find_first_intersect_triangle), face = -1 if
        none; needs set_faces."""
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        f, v, t = C.c_int32(), C.c_int32(), C.c_float()
        L.check(L.lib().cc_sim_pick(self._h, L.ptr(o), L.ptr(d), C.byref(f), C.byref(v), C.byref(t)))
        return f.value, v.value, t.value

    def get_normals(self):
        nrm = np.zeros((self.n, 3), np.float32)
        L.check(L.lib().cc_sim_get_normals(self._h, L.ptr(nrm)))
        return nrm

    def time_kernels(self, nsteps):
        """Per-kernel-group times with nothing overlapped (cc_sim_time_kernels)."""
        t = L.cc_timing()
        L.check(L.lib().cc_sim_time_kernels(self._h, int(nsteps), C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.cc_timing._fields_}

    def set_timing(self, enable=True):
        """Event timing inside the step graphs (cc_sim_set_timing)."""
        L.check(L.lib().cc_sim_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        """Averages of the in-graph event timing since it was enabled."""
        t = L.cc_timing()
        L.check(L.lib().cc_sim_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in L.cc_timing._fields_}


class BVHAccel:
    """BVHAccel(Mesh& body) (bvh.cu:264-297): an LBVH over a triangle body,
    built on the device.  Backed by a collider-only simulator."""

    def __init__(self, verts3, tris, box_depth=0.01):
        body = S.Body(np.asarray(verts3, np.float32), np.asarray(tris, np.uint32),
                      np.zeros_like(np.asarray(verts3, np.float32)), L.CC_BODY_STATIC, False)
        p = L.default_params()
        p.box_depth = box_depth
        one = np.zeros((1, 3), np.float32)
        row = np.zeros(2, np.uint32)
        col = np.zeros(0, L.SPRING_DTYPE)
        self._sim = Simulator(one, [(row, col)], body, p)

    def export(self):
        return self._sim.bvh_export()

    def intersect(self, pts):
        """First-hit leaf per point plus projection (verlet.cu:24-50)."""
        return self._sim.collide_points(pts)

    def set_vertices(self, verts3):
        self._sim.set_body_vertices(verts3)

    def close(self):
        self._sim.close()
