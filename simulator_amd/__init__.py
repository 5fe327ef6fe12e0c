"""simulator_amd -- MI355X (gfx950) cloth simulation core.

CSR mass-spring Verlet step, Karras LBVH over a triangle body and
point-vs-LBVH collision projection as hand-written HIP kernels behind the C
ABI of include/clothcore.h.  The classes mirror the reference's Simulator,
Springs and BVHAccel.  There is no CPU fallback: the library must be built
(python -m simulator_amd.build).
"""
from ._lib import ClothCoreError, lib  # noqa: F401


def __getattr__(name):
    if name in ("Simulator", "Springs", "BVHAccel", "Mesh"):
        from . import simulator
        return getattr(simulator, name)
    raise AttributeError(name)
