"""C++ host mirror of the reference's call surface (dare_amd/include/dare).

The C++ program tests/cpp/test_host_mirror.cpp restates reference tests on the mirror:
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
def test_host_mirror_driver(tmp_path):
    assert os.access(BIN, os.X_OK), "tests/cpp/test_host_mirror is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    for case in ("TrilinosBuild", "SolveLaplace", "SolveScalar1D", "HaloExchange", "DeviceDiffusion3D", "VTKWrite"):
        assert f"[ OK ] {case}" in r.stdout
    # the VTK file of a field whose periodic ghosts were filled on the device: internal cells
    # and face ghosts hold the internal cell they wrap to; edge and corner ghosts (two or more
    # indices outside) are not exchanged (HaloBuffer face ghosts only) and keep their 0
    from test_vtk_writer import read_vts
    import numpy as np
    head, arrays, ext, _ = read_vts(str(tmp_path / "mirror3D_4" / "mirror3D_4_0.vts"))
    n, g = [6, 5, 4], 2
    assert ext == [0, 10, 0, 9, 0, 8]
    # the Field names its time level 0 "conc_0"
    vals = arrays["conc_0"][1].reshape(n[2] + 2 * g, n[1] + 2 * g, n[0] + 2 * g, 2)
    for k in range(n[2] + 2 * g):
        for j in range(n[1] + 2 * g):
            for i in range(n[0] + 2 * g):
                outside = sum(not (g <= q < g + m) for q, m in ((i, n[0]), (j, n[1]), (k, n[2])))
                wi, wj, wk = (i - g) % n[0], (j - g) % n[1], (k - g) % n[2]
                for c in range(2):
                    exp = 1e6 * wi + 1e3 * wj + wk + 0.125 * c if outside <= 1 else 0.0
                    assert vals[k, j, i, c] == exp, (i, j, k, c)
    assert arrays["TimeValue"][1].tolist() == [0.5]
