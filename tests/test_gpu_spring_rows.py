"""The step kernel's spring layouts (the merged CSR rows of CSR_spring /
CSR_bend_spring, spring.cpp:363-391, and the fixed-width copy a cloth of at
most 12, or 24, springs per vertex also gets) sum the same springs in the
same order: two simulators, one per layout, stay bit-identical, with pinned
vertices moved by set_pins mid-run (grid cloths: 12 wide; the reference's
garment: 24 wide)."""
import os

import numpy as np
import pytest

from simulator_amd import _lib as L
from simulator_amd import scene as S
from simulator_amd.simulator import Simulator

from helpers import say

pytestmark = pytest.mark.gpu


def sim_with_layout(sc, fixed_width):
    old = os.environ.get("CC_SPRINGS_ELL")
    os.environ["CC_SPRINGS_ELL"] = "1" if fixed_width else "0"
    try:
        return Simulator.from_scene(sc)
    finally:
        if old is None:
            del os.environ["CC_SPRINGS_ELL"]
        else:
            os.environ["CC_SPRINGS_ELL"] = old


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("make", [lambda: S.small(side=48, y0=0.806),
                                  lambda: S.grid_scene("refit96", 96, 5e-4, 120,
                                                       S.capsules(24, 12, 6, L.CC_BODY_REFIT),
                                                       box_depth=0.002, y0=0.975),
                                  # the reference's garment: mesh springs, up to 24 per vertex
                                  lambda: S.load_mesh_scene(os.path.join(GOLDEN, "ref_tshirt_female.npz"))],
                         ids=["grid-static", "grid-refit", "reference-mesh"])
def test_fixed_width_rows_equal_csr_rows(gpu, make):
    sc = make()
    a, b = sim_with_layout(sc, True), sim_with_layout(sc, False)
    contacts = 0
    for blk, nsteps in enumerate((1, 16, 7, 40, 33)):
        if blk == 2:                       # new pins on both
            pins = np.array([0, max(sc.nx - 1, 1), sc.n // 2], np.uint32)
            a.set_pins(pins)
            b.set_pins(pins)
        a.simulate(nsteps)
        b.simulate(nsteps)
        for g, w, name in zip(a.get_state(), b.get_state(), ("x_cur", "x_prev", "contact")):
            np.testing.assert_array_equal(g, w, err_msg=f"{name} after block {blk}")
        contacts += int(np.any(a.get_state()[2] != 0, axis=1).sum())
    say(f"[springs] {sc.name}: {sc.n} vertices, {contacts} contacts over the run")
    assert contacts > 0
    a.close()
    b.close()
