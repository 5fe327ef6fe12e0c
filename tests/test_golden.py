"""The oracle reproduces every committed golden fixture bit for bit (the
fixtures pin the oracle; tests/test_gpu_golden.py pins the HIP path)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_golden as G  # noqa: E402


@pytest.mark.parametrize("name", sorted(G.FIXTURES))
def test_oracle_reproduces_fixture(name):
    want = G.load(name)
    got = G.FIXTURES[name]()
    assert sorted(got) == sorted(want)
    for k in want:
        np.testing.assert_array_equal(np.asarray(got[k]), want[k], err_msg=f"{name}:{k}")


def test_fixtures_are_nontrivial():
    c16 = G.load("cloth16")
    flags = np.unpackbits(c16["contact_bits"], axis=1, bitorder="little")[:, :256]
    assert flags.sum() > 1000 and (flags.sum(1) > 0).sum() > 100     # contact in most steps
    c64 = G.load("cloth64")
    assert np.abs(c64["x_200"] - c64["pos0"]).max() > 0.1           # the cloth fell
    l4 = G.load("lbvh_ico4")
    assert l4["keys"].shape == (5120,) and np.all(np.diff(l4["keys"].astype(np.float64)) > 0)
