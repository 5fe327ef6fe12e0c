"""Golden fixtures for the cloth step and the body LBVH, written by the oracle.

The reference ships no test vectors, so these fixtures are outputs of the
oracle (the CPU restatement of the reference path) on fixed, seeded inputs.
Each fixture stores its inputs too, so the checks need no generator:

  lbvh_ico2.npz, lbvh_ico4.npz   icosphere (level 2: 320 triangles, level 4:
      5120 triangles), radius 0.3 at (0.5, 0.5, 0), vertices moved 0.003 along
      their normals (Mesh::vertex_extend); the Karras LBVH with expanded boxes
      of depth 0.01: sorted keys, order, links, parent links, boxes.
  cloth64.npz    64x64 grid cloth at y = 1.0 over the level-4 sphere (the
      small configuration): FP64 positions after 1, 10 and 200 steps.
  cloth16.npz    16x16 grid cloth at y = 0.83 over the level-2 sphere, falling
      into contact (first contact in step 82): FP64 positions after 80 and
      200 steps and the per-step contact flags (bit i of row s: vertex i
      projected in step s + 1).

Run `python tests/golden/make_golden.py` to rewrite them (only after a
deliberate change of the oracle).  tests/test_golden.py checks that the
oracle still reproduces every fixture bit for bit; tests/test_gpu_golden.py
checks the HIP path against them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

FLT_EPS = float(np.finfo(np.float32).eps)
CENTER = (0.5, 0.5, 0.0)
RADIUS = 0.3
EXTEND = 0.003
DEPTH = 0.01


def params():
    """The physics of the small configuration (dt 1e-3, k 1000/500/100)."""
    p = O.or_params()
    p.dt, p.mass, p.damp, p.kd = 1e-3, 0.3, -0.02, -0.5
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.8, 0.0
    p.k[0], p.k[1], p.k[2] = 1000.0, 500.0, 100.0
    p.rest_scale, p.eps = 1.0, FLT_EPS
    return p


def params_array(p):
    return np.array([p.dt, p.mass, p.gravity[0], p.gravity[1], p.gravity[2], p.damp, p.kd,
                     p.k[0], p.k[1], p.k[2], p.rest_scale, p.eps], np.float32)


def sphere(level):
    """Icosphere vertices extended along their (radial) normals."""
    v, t = O.icosphere(level, RADIUS, CENTER)
    nrm = (v.astype(np.float64) - np.array(CENTER)) / RADIUS
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    v4 = np.concatenate([v, np.ones((v.shape[0], 1), np.float32)], 1)
    ext = O.vertex_extend(v4, t, t, nrm, EXTEND)[:, :3]
    return np.ascontiguousarray(ext), t


def lbvh_fixture(level):
    v, t = sphere(level)
    d = O.bvh_build(v, t, DEPTH)
    out = {"verts": v, "tris": t, "depth": np.float32(DEPTH)}
    for k in ("keys", "order", "childA", "childB", "parent_int", "parent_leaf", "box_int", "box_leaf"):
        out[k] = d[k]
    return out


def cloth(side, y0):
    pos = O.grid_cloth(side, side, 1.0 / (side - 1), y0, 1e-4, 0)
    springs = O.grid_springs(side, side, pos)
    pins = np.zeros(side * side, np.uint8)
    pins[[0, side - 1]] = 1
    return pos, springs, pins


def run_f64(pos, springs, pins, level, checkpoints, flags=False):
    v, t = sphere(level)
    h = O.BvhHandle(O.bvh_build(v, t, DEPTH))
    prm = params()
    x = pos.astype(np.float64)
    xp, c = x.copy(), np.zeros_like(x)
    out, bits = {}, []
    for s in range(1, max(checkpoints) + 1):
        x, xp, c, ins = O.step_f64(x, xp, c, springs, prm, pins, h)
        bits.append(np.packbits(ins.astype(bool), bitorder="little"))
        if s in checkpoints:
            out[f"x_{s}"] = x.copy()
    if flags:
        out["contact_bits"] = np.stack(bits)
    return v, t, out


def cloth_fixture(side, y0, level, checkpoints, flags):
    pos, springs, pins = cloth(side, y0)
    v, t, res = run_f64(pos, springs, pins, level, checkpoints, flags)
    out = {"pos0": pos, "pins": pins, "params": params_array(params()), "body_verts": v,
           "body_tris": t, "depth": np.float32(DEPTH)}
    for k, (row, col) in enumerate(springs):
        out[f"row{k}"] = row
        out[f"end{k}"] = np.ascontiguousarray(col["end"])
        out[f"rest{k}"] = np.ascontiguousarray(col["original"])
    out.update(res)
    return out


FIXTURES = {
    "lbvh_ico2": lambda: lbvh_fixture(2),
    "lbvh_ico4": lambda: lbvh_fixture(4),
    "cloth64": lambda: cloth_fixture(64, 1.0, 4, (1, 10, 200), False),
    "cloth16": lambda: cloth_fixture(16, 0.83, 2, (80, 200), True),
}


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def springs_of(fx):
    """The fixture's three CSR spring sets in the library's record layout."""
    from oracle.bind import SPRING_DTYPE
    out = []
    for k in range(3):
        col = np.zeros(fx[f"end{k}"].shape[0], SPRING_DTYPE)
        col["end"] = fx[f"end{k}"]
        col["original"] = fx[f"rest{k}"]
        out.append((fx[f"row{k}"], col))
    return out


if __name__ == "__main__":
    O.build()
    for name, make in FIXTURES.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **make())
        print(f"wrote {path} ({os.path.getsize(path)} bytes)")
