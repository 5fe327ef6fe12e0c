"""The environment switches the library reads at cc_sim_create (getenv in
simulator_amd/csrc) and the table of them in INTEGRATION.md name the same
variables, so neither drifts from the other."""
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_in_sources():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "simulator_amd", "csrc", "*")):
        with open(path) as f:
            names |= set(re.findall(r'getenv\("(CC_[A-Z0-9_]+)"\)', f.read()))
    return names


def documented():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        return set(re.findall(r"^\| `(CC_[A-Z0-9_]+)` \|", f.read(), re.M))


def test_every_switch_is_documented_and_every_documented_switch_is_read():
    src, doc = read_in_sources(), documented()
    assert src, "no getenv switches found in the sources"
    assert src == doc, {"read but undocumented": sorted(src - doc), "documented but not read": sorted(doc - src)}
