"""The C++ facade that mirrors the reference's H_psi API compiles against
include/effm_hpsi.h and links libeffm_hip.so, both as a header (C++20, g++)
and through the C++23 module unit `import H_psi.hip;`
(effmass_amd/host/H_psi_hip.cppm, clang++); on a GPU both builds run the
closed-form checks of tests/cpp/facade_check.cpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "effmass_amd", "host"),
       "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__"]
LINK = ["-L", os.path.join(ROOT, "effmass_amd"), "-leffm_hip", "-L", "/opt/rocm/lib", "-lamdhip64",
        "-Wl,-rpath," + os.path.join(ROOT, "effmass_amd"), "-Wl,-rpath,/opt/rocm/lib"]
CLANG = "/opt/rocm/llvm/bin/clang++"


def _build(tmp_path):
    exe = str(tmp_path / "facade_check")
    subprocess.run(["g++", "-std=c++20", "-O1", *INC, os.path.join(ROOT, "tests", "cpp", "facade_check.cpp"),
                    *LINK, "-o", exe], check=True)
    return exe


def _build_module(tmp_path):
    pcm = str(tmp_path / "H_psi.hip.pcm")
    exe = str(tmp_path / "facade_check_module")
    subprocess.run([CLANG, "-std=c++23", *INC, "--precompile",
                    os.path.join(ROOT, "effmass_amd", "host", "H_psi_hip.cppm"), "-o", pcm], check=True)
    subprocess.run([CLANG, "-std=c++23", "-O1", *INC, "-DEFFM_USE_MODULE", f"-fmodule-file=H_psi.hip={pcm}",
                    os.path.join(ROOT, "tests", "cpp", "facade_check.cpp"), pcm, *LINK, "-o", exe], check=True)
    return exe


def test_facade_compiles(tmp_path):
    assert os.path.exists(_build(tmp_path))


def test_module_unit_compiles(tmp_path):
    assert os.path.exists(_build_module(tmp_path))


@pytest.mark.gpu
@pytest.mark.parametrize("build", [_build, _build_module], ids=["header", "module"])
def test_facade_runs(tmp_path, build):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
