"""Host-side rows end to end on the GPU: the library's spring builders and
OBJ reader produce the springs and cloth that the GPU then steps.

Per case: the library's springs equal the oracle's bit for bit
(Springs::create_neigh / CSR_spring, spring.cpp:82-391; boundary springs,
spring.cpp:98-252), and the GPU, stepping the cloth built from the
library's springs, stays within 1e-6 of the cloth extent per step of an
FP64 step over the oracle's own springs, with the contact outcome bit for
bit against the oracle for the GPU's pre-projection positions."""
import os

import numpy as np
import pytest

import oracle as O
from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import ObjLoader, Simulator

from helpers import say
from test_boundary_springs import two_piece_mesh

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_tshirt_female.npz")


def same_springs(got, want):
    assert len(got) == len(want)
    for (gr, gc), (wr, wc) in zip(got, want):
        np.testing.assert_array_equal(gr, wr)
        np.testing.assert_array_equal(gc["end"], wc["end"])
        np.testing.assert_array_equal(gc["original"].view(np.uint32), wc["original"].view(np.uint32))


def step_against_oracle(sc, oracle_springs, nsteps, label):
    """nsteps GPU steps of sc (library springs) against FP64 steps over the
    oracle's springs from the GPU's state; returns the contacts seen."""
    sim = Simulator.from_scene(sc)
    h = None
    if sc.body is not None:
        h = O.BvhHandle(O.bvh_build(sim.body_vertices(), sc.body.tris, sc.params.box_depth))
    prm = O.params_from(sc.params)
    pins = np.zeros(sc.n, np.uint8)
    pins[sc.pins] = 1
    free = pins == 0
    extent = float(np.ptp(sc.pos, axis=0).max())
    sim.set_debug(True)
    worst, contacts, flips = 0.0, 0, 0
    for k in range(nsteps):
        x, xp, c = sim.get_state()
        sim.simulate(1)
        pre, hit, ins = sim.get_debug()
        xo, _, co = sim.get_state()
        if h is not None:
            ohit, oins, opos, on = O.collide_points(h, pre[free])
            np.testing.assert_array_equal(hit[free], ohit, err_msg=f"{label}: step {k}")
            np.testing.assert_array_equal(ins[free], oins, err_msg=f"{label}: step {k}")
            np.testing.assert_array_equal(xo[free], opos, err_msg=f"{label}: step {k}")
            np.testing.assert_array_equal(co[free], on, err_msg=f"{label}: step {k}")
        contacts += int(ins.sum())
        x64, _, _, ins64, hit64 = O.step_f64(x, xp, c, oracle_springs, prm, pins, h, with_hit=True)
        flip = (ins64 != ins) | ((hit64 != hit) & (ins == 1))
        flips += int(flip.sum())
        worst = max(worst, float(np.abs(xo - x64).max(1)[~flip].max()))
    say(f"[{label}] {nsteps} GPU steps over library springs vs FP64 over oracle springs: worst "
        f"{worst:.2e} (extent {extent:.2f}), contacts {contacts}, flips {flips}")
    assert worst <= 1e-6 * extent
    sim.close()
    return contacts


def test_reference_garment_mesh_springs_on_the_gpu(gpu):
    """The reference's own garment (tshirt2.obj) over its body (female.obj):
    mesh springs from the library, stepped on the GPU."""
    sc = S.load_mesh_scene(FIXTURE)
    v4 = np.concatenate([sc.pos, np.ones((sc.n, 1), np.float32)], 1)
    want = O.springs_from_mesh(v4, sc.faces)
    same_springs(sc.springs, want)
    assert step_against_oracle(sc, want, 25, "garment mesh springs") > 100


def test_obj_file_to_gpu_step(gpu, tmp_path):
    """The garment written as an OBJ file (v / vn / f v//vn lines), read back
    by the library's reader, its springs built and stepped on the GPU."""
    f = np.load(FIXTURE)
    cv, cf = f["cloth_verts"], f["cloth_faces"]
    lines = ["# garment from the fixture"]
    lines += ["v %.9g %.9g %.9g" % tuple(p) for p in cv.astype(np.float64)]
    lines += ["vn 0 1 0"]
    lines += ["f %d//1 %d//1 %d//1" % tuple(t + 1) for t in cf]
    path = tmp_path / "garment.obj"
    path.write_text("\n".join(lines) + "\n")
    o = ObjLoader(str(path))
    np.testing.assert_array_equal(o.vertices[:, :3], cv)
    # "v//n" corners: the vertex index is read right (only the texture and
    # normal indices of the second and third corners are read as the
    # reference's reader reads them, ObjLoader.cpp:129-152)
    np.testing.assert_array_equal(o.vertex_index, cf)
    faces = np.ascontiguousarray(o.vertex_index, np.uint32)
    sc = S.mesh_scene("obj", o.vertices[:, :3], faces, f["body_verts"], f["body_faces"], f["body_normals"])
    v4 = np.concatenate([sc.pos, np.ones((sc.n, 1), np.float32)], 1)
    want = O.springs_from_mesh(v4, faces)
    same_springs(sc.springs, want)
    step_against_oracle(sc, want, 10, "OBJ garment")


@pytest.mark.parametrize("jitter", [0.0, 2e-3])
def test_boundary_springs_on_the_gpu(gpu, jitter):
    """A two-piece garment with boundary springs between its pieces: the
    library's springs equal the oracle's and the GPU steps the cloth they
    make (no collider)."""
    v4, faces, groups = two_piece_mesh(jitter)
    got = S.springs_from_mesh(v4, faces, groups)
    want = O.springs_from_mesh(v4, faces, groups)
    same_springs(got, want)
    pos = np.ascontiguousarray(v4[:, :3] + np.float32([0.0, 1.0, 0.0]))
    sc = S.Scene("boundary", 0, 0, pos, list(got), np.array([0, 16], np.uint32), L.default_params(), None, 20)
    step_against_oracle(sc, want, 20, f"boundary springs (jitter {jitter})")


def test_grid_springs_at_the_configured_sizes_on_the_gpu(gpu):
    """The grid generator's springs at 111x111 and 512x512 (the garment and
    large cloths) against the oracle's, then a few GPU steps of the 512x512
    cloth against FP64 steps over the oracle's springs."""
    for sc in (S.garment(), S.large()):
        want = O.grid_springs(sc.nx, sc.nz, sc.pos)
        same_springs(sc.springs, want)
        say(f"[grid] {sc.nx}x{sc.nz}: {sum(c.shape[0] for _, c in want)} springs equal the oracle's")
    step_against_oracle(sc, want, 3, "512x512 grid springs")
