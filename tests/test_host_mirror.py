"""C++ host mirror of the reference's call surface (dare_amd/include/dare).

The driver tests/cpp/test_host_mirror.cpp restates reference tests on the mirror:
MatrixSystem/test/test_Trilinos.cpp:60-191, MatrixSystem/test/test_TrilinosSolver.cpp:98-205,
test/test_integration_Cartesian_Trilinos.cpp:82-160 and MPI/test/test_HaloBuffer.cpp:33-115,
plus a 3-D diffusion run on the device assembly checked against the heat kernel.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "test_host_mirror")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_headers_compile_standalone():
    """Every mirror header compiles on its own (self-contained includes)."""
    hdr = os.path.join(ROOT, "dare_amd", "include", "dare")
    for name in sorted(os.listdir(hdr)):
        if not name.endswith(".h"):
            continue
        r = subprocess.run(
            ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
             "-I", os.path.join(ROOT, "dare_amd", "include"), "-x", "c++", os.path.join(hdr, name)],
            capture_output=True, text=True)
        assert r.returncode == 0, f"{name}:\n{r.stderr}"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_driver_builds():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", CPP], check=True)
    assert os.access(BIN, os.X_OK)


@pytest.mark.gpu
def test_host_mirror_driver():
    assert os.access(BIN, os.X_OK), "tests/cpp/test_host_mirror is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    for case in ("TrilinosBuild", "SolveLaplace", "SolveScalar1D", "HaloExchange", "DeviceDiffusion3D"):
        assert f"[ OK ] {case}" in r.stdout
