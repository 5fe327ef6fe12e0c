"""Shared helpers for the parity tests (oracle on the CPU vs the HIP path)."""
import ctypes as C
import sys

import numpy as np

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

BVH_FIELDS = ("keys", "order", "childA", "childB", "parent_int", "parent_leaf", "box_int", "box_leaf")


class DeviceBuffer:
    """Device memory for a test, from the HIP runtime the library links
    (hipMalloc / hipMemcpy / hipFree); filled from and read into numpy."""

    def __init__(self, host):
        L.lib()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), C.c_size_t(self.nbytes), 1) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def collider_sim(verts, tris, normals=None, mode=L.CC_BODY_STATIC, depth=0.01, extend=False,
                 caps=None, motion=None):
    """A simulator with a one-vertex cloth, used to exercise the body / LBVH."""
    nrm = np.zeros_like(verts) if normals is None else normals
    body = S.Body(np.asarray(verts, np.float32), np.asarray(tris, np.uint32), nrm, mode, extend,
                  caps, motion or {})
    p = L.default_params()
    p.box_depth = depth
    one = np.zeros((1, 3), np.float32)
    return Simulator(one, [(np.zeros(2, np.uint32), np.zeros(0, L.SPRING_DTYPE))], body, p)


def oracle_bvh(verts, tris, depth):
    return O.bvh_build(np.asarray(verts, np.float32), np.asarray(tris, np.uint32), depth)


def assert_bvh_equal(got, want):
    for k in BVH_FIELDS:
        g, w = got[k], want[k]
        assert g.shape == w.shape, (k, g.shape, w.shape)
        if g.dtype.kind == "f":
            bad = np.where(~(g == w).all(axis=-1) if g.ndim > 1 else ~(g == w))[0]
        else:
            bad = np.where(g != w)[0]
        assert bad.size == 0, f"{k}: {bad.size} mismatches, first at {bad[:5]}: {g[bad[:3]]} vs {w[bad[:3]]}"


def extend_like_library(verts, tris, normals, dist):
    """Mesh::vertex_extend with per-vertex normals, through the oracle."""
    v4 = np.concatenate([verts, np.ones((verts.shape[0], 1), np.float32)], axis=1)
    out = O.vertex_extend(v4, tris, tris, normals, dist)
    return np.ascontiguousarray(out[:, :3])
