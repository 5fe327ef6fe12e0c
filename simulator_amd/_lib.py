"""ctypes binding of libclothcore.so (the C ABI declared in include/clothcore.h).

The library is built in-tree by simulator_amd.build; importing this module
fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclothcore.so")

CC_OK = 0
CC_ERR_INVALID = 1
CC_ERR_HIP = 2
CC_ERR_ALLOC = 3
CC_ERR_NONFINITE = 4
CC_ERR_UNSUPPORTED = 5
CC_ERR_TRAVERSAL = 6

CC_FAULT_ALLOC = 1
CC_FAULT_LAUNCH = 2

CC_BODY_NONE = 0
CC_BODY_STATIC = 1
CC_BODY_REFIT = 2
CC_BODY_REBUILD = 3


class ClothCoreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"clothcore error {code}: {msg}")
        self.code = code


class cc_spring(C.Structure):
    _fields_ = [("end", C.c_int32), ("original", C.c_float)]


SPRING_DTYPE = np.dtype([("end", "<i4"), ("original", "<f4")])


class cc_csr(C.Structure):
    _fields_ = [("row", C.POINTER(C.c_uint32)), ("col", C.POINTER(cc_spring)),
                ("n", C.c_uint32), ("nnz", C.c_uint32)]


class cc_params(C.Structure):
    _fields_ = [("dt", C.c_float), ("mass", C.c_float), ("gravity", C.c_float * 3),
                ("damp", C.c_float), ("kd", C.c_float), ("k", C.c_float * 3),
                ("rest_scale", C.c_float), ("eps", C.c_float), ("box_depth", C.c_float),
                ("body_extend", C.c_float)]


class cc_capsules(C.Structure):
    _fields_ = [("ncaps", C.c_uint32), ("vstart", C.POINTER(C.c_uint32)),
                ("p0", C.POINTER(C.c_float)), ("p1", C.POINTER(C.c_float)),
                ("radius", C.POINTER(C.c_float)), ("phase", C.POINTER(C.c_float))]


class cc_cloth_desc(C.Structure):
    _fields_ = [("n", C.c_uint32), ("verts4", C.POINTER(C.c_float)), ("ntypes", C.c_uint32),
                ("row", C.POINTER(C.c_uint32) * 3), ("col", C.POINTER(cc_spring) * 3)]


class cc_obj(C.Structure):
    _fields_ = [("nv", C.c_uint32), ("nvn", C.c_uint32), ("nvt", C.c_uint32), ("nf", C.c_uint32),
                ("verts4", C.POINTER(C.c_float)), ("normals3", C.POINTER(C.c_float)),
                ("tex2", C.POINTER(C.c_float)), ("vidx", C.POINTER(C.c_uint32)),
                ("tidx", C.POINTER(C.c_uint32)), ("nidx", C.POINTER(C.c_uint32)),
                ("ngroups", C.c_uint32), ("group_counts", C.POINTER(C.c_uint32))]


class cc_body_desc(C.Structure):
    _fields_ = [("nv", C.c_uint32), ("verts", C.POINTER(C.c_float)), ("stride", C.c_uint32),
                ("nf", C.c_uint32), ("tris", C.POINTER(C.c_uint32)), ("mode", C.c_uint32),
                ("extend", C.c_uint32), ("normals3", C.POINTER(C.c_float)),
                ("caps", C.POINTER(cc_capsules)),
                ("amp", C.c_double), ("freq", C.c_double), ("dir", C.c_double * 3),
                ("rot_amp", C.c_double), ("pivot", C.c_double * 3), ("breath_amp", C.c_double),
                ("breath_freq", C.c_double)]


class cc_timing(C.Structure):
    _fields_ = [("step_us", C.c_double), ("verlet_us", C.c_double), ("body_us", C.c_double),
                ("refit_us", C.c_double), ("build_us", C.c_double), ("steps", C.c_uint32)]


P = C.c_void_p
u32, i32, u64, f32, f64 = C.c_uint32, C.c_int32, C.c_uint64, C.c_float, C.c_double

# name -> (restype, argtypes); every declaration of include/clothcore.h
PROTOTYPES = {
    "cc_params_default": (None, [C.POINTER(cc_params)]),
    "cc_version": (C.c_int, []),
    "cc_last_error": (C.c_char_p, []),
    "cc_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "cc_springs_from_mesh": (C.c_int, [P, u32, P, u32, C.POINTER(cc_csr), C.POINTER(cc_csr)]),
    "cc_springs_from_mesh_groups": (C.c_int, [P, u32, P, u32, P, u32, C.POINTER(cc_csr), C.POINTER(cc_csr)]),
    "cc_springs_grid": (C.c_int, [u32, u32, P, C.POINTER(cc_csr)]),
    "cc_csr_free": (None, [C.POINTER(cc_csr)]),
    "cc_scene_grid_cloth": (C.c_int, [u32, u32, f64, f64, f64, u64, P]),
    "cc_scene_icosphere": (C.c_int, [u32, f64, P, P, P, P]),
    "cc_scene_capsule_body": (C.c_int, [u32, f64, f64, P, P, u32, u32, u32, u64, P, P, P,
                                        C.POINTER(u32), C.POINTER(u32), C.POINTER(cc_capsules)]),
    "cc_mesh_vertex_extend": (C.c_int, [P, u32, u32, P, u32, P, P, f32]),
    "cc_sim_create": (C.c_int, [C.POINTER(cc_cloth_desc), C.POINTER(cc_body_desc),
                                C.POINTER(cc_params), C.POINTER(P)]),
    "cc_sim_destroy": (C.c_int, [P]),
    "cc_fault_inject": (C.c_int, [C.c_int, C.c_int64]),
    "cc_sim_save": (C.c_int, [P, C.c_char_p]),
    "cc_sim_load": (C.c_int, [P, C.c_char_p]),
    "cc_sim_step": (C.c_int, [P, C.c_int, P]),
    "cc_sim_synchronize": (C.c_int, [P]),
    "cc_sim_update_vertex": (C.c_int, [P, u32, P]),
    "cc_sim_set_pins": (C.c_int, [P, P, u32]),
    "cc_sim_set_params": (C.c_int, [P, C.POINTER(cc_params)]),
    "cc_sim_get_state": (C.c_int, [P, P, P, P]),
    "cc_sim_set_state": (C.c_int, [P, P, P, P]),
    "cc_sim_frame": (C.c_int, [P, C.POINTER(u64)]),
    "cc_sim_block_steps": (C.c_int, [P, C.POINTER(C.c_int)]),
    "cc_sim_set_frame": (C.c_int, [P, u64]),
    "cc_sim_set_debug": (C.c_int, [P, C.c_int]),
    "cc_sim_get_debug": (C.c_int, [P, P, P, P]),
    "cc_sim_set_counting": (C.c_int, [P, C.c_int]),
    "cc_sim_get_counters": (C.c_int, [P, C.POINTER(u64), C.POINTER(u64)]),
    "cc_sim_check_finite": (C.c_int, [P]),
    "cc_sim_set_stack_limit": (C.c_int, [P, C.c_int]),
    "cc_body_set_vertices": (C.c_int, [P, P]),
    "cc_body_get_vertices": (C.c_int, [P, P]),
    "cc_body_get_motion": (C.c_int, [P, P, P, P]),
    "cc_bvh_rebuild": (C.c_int, [P]),
    "cc_bvh_refit": (C.c_int, [P]),
    "cc_bvh_size": (C.c_int, [P, C.POINTER(u32)]),
    "cc_sim_set_bvh_width": (C.c_int, [P, C.c_int]),
    "cc_bvh_width": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "cc_bvh_export": (C.c_int, [P, P, P, P, P, P, P, P, P, P]),
    "cc_collide_points": (C.c_int, [P, P, u32, P, P, P, P]),
    "cc_collide_points_noalloc": (C.c_int, [P, P, u32, P, P, P, P, P, P]),
    "cc_sim_set_faces": (C.c_int, [P, P, u32]),
    "cc_sim_write_vbo": (C.c_int, [P, P, C.c_size_t, u32, C.c_int]),
    "cc_sim_get_normals": (C.c_int, [P, P]),
    "cc_sim_time_kernels": (C.c_int, [P, C.c_int, C.POINTER(cc_timing)]),
    "cc_sim_set_timing": (C.c_int, [P, C.c_int]),
    "cc_adj_create": (C.c_int, [P, C.c_uint32, C.c_uint32, C.POINTER(P)]),
    "cc_adj_destroy": (C.c_int, [P]),
    "cc_adj_forward": (C.c_int, [P, P, P, P, P]),
    "cc_adj_backward": (C.c_int, [P, P, P, P, P]),
    "cc_obj_read": (C.c_int, [C.c_char_p, P]),
    "cc_obj_free": (None, [P]),
    "cc_mesh_unify": (C.c_int, [P, P, P]),
    "cc_mesh_center": (C.c_int, [P, C.c_uint32, P]),
    "cc_mesh_scale": (C.c_int, [P, C.c_uint32, C.c_float]),
    "cc_mesh_translate": (C.c_int, [P, C.c_uint32, C.c_float, C.c_float, C.c_float]),
    "cc_mesh_rotate": (C.c_int, [P, C.c_uint32, C.c_float, C.c_int]),
    "cc_sim_pick": (C.c_int, [P, P, P, P, P, P]),
    "cc_batch_create": (C.c_int, [P, P, P, C.c_uint32, C.POINTER(P)]),
    "cc_batch_destroy": (C.c_int, [P]),
    "cc_batch_step": (C.c_int, [P, C.c_int, P]),
    "cc_batch_synchronize": (C.c_int, [P]),
    "cc_batch_check": (C.c_int, [P]),
    "cc_batch_set_pins": (C.c_int, [P, C.c_uint32, P, C.c_uint32]),
    "cc_batch_get_state": (C.c_int, [P, C.c_uint32, P, P, P]),
    "cc_batch_info": (C.c_int, [P, P, P, P]),
    "cc_batch_tuning": (C.c_int, [P, P, P]),
    "cc_batch_body_export": (C.c_int, [P, C.c_uint32, P, P, P, P, P, P, P, P, P, P]),
    "cc_batch_frame": (C.c_int, [P, P]),
    "cc_sim_get_timing": (C.c_int, [P, C.POINTER(cc_timing)]),
}

_lib = None


def lib():
    """Load libclothcore.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build it with `python -m simulator_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != CC_OK:
        msg = lib().cc_last_error()
        raise ClothCoreError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Raw data pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def device_count():
    n = C.c_int(0)
    check(lib().cc_device_count(C.byref(n)))
    return n.value


def fault_inject(kind, count):
    """cc_fault_inject: fail the allocation (CC_FAULT_ALLOC) or launch check
    (CC_FAULT_LAUNCH) after `count` further successful ones; count < 0 turns
    the kind off."""
    check(lib().cc_fault_inject(int(kind), int(count)))


def default_params():
    p = cc_params()
    lib().cc_params_default(C.byref(p))
    return p
