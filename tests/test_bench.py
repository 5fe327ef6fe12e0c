"""The bench's arithmetic and its one-GPU contract, on the CPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

AB = {"verlet": 47_104_032, "deform": 1, "refit": 1, "springs": 3_135_492}


def timing(verlet=0.0, refit=0.0, build=0.0, body=0.0, steps=0):
    return {"step_us": 0.0, "verlet_us": verlet, "body_us": body, "refit_us": refit, "build_us": build,
            "steps": steps}


def test_live_step_time_is_used_when_recorded():
    kern, iso, dom, ach = bench.step_kernel_figures(timing(40.0, 60.0, steps=20), timing(37.5, 50.0, body=5.0),
                                                    AB, rebuild=False)
    assert dom == "live"
    assert ach == pytest.approx(47_104_032 / 40e-6 / 1e9)
    assert kern == {"verlet": 40.0, "body_update": 60.0}
    assert iso == {"verlet": 37.5, "refit": 50.0, "deform": 5.0}


def test_zero_live_time_falls_back_to_the_isolated_time():
    # no step kernel timed live (verlet_us == 0): no division by zero
    kern, iso, dom, ach = bench.step_kernel_figures(timing(), timing(37.5, 50.0), AB, rebuild=False)
    assert dom == "isolated"
    assert ach == pytest.approx(47_104_032 / 37.5e-6 / 1e9)


def test_no_step_time_at_all_gives_no_bandwidth():
    kern, iso, dom, ach = bench.step_kernel_figures(timing(), timing(), AB, rebuild=True)
    assert dom == "none" and ach is None


def test_rebuilt_body_reports_the_build_time():
    kern, iso, _, _ = bench.step_kernel_figures(timing(30.0, build=80.0, steps=5), timing(29.0, build=70.0),
                                                AB, rebuild=True)
    assert kern["body_update"] == 80.0
    assert iso["build"] == 70.0 and "refit" not in iso


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""
    with pytest.raises(json.JSONDecodeError):
        json.loads(r.stdout or "x")
