"""The C++ facade that mirrors the reference's H_psi API compiles against
include/effm_hpsi.h and links libeffm_hip.so; on a GPU it runs the
closed-form checks of tests/cpp/facade_check.cpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build(tmp_path):
    exe = str(tmp_path / "facade_check")
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "effmass_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "facade_check.cpp"),
                    "-L", os.path.join(ROOT, "effmass_amd"), "-leffm_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "effmass_amd"), "-o", exe], check=True)
    return exe


def test_facade_compiles(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_facade_runs(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
