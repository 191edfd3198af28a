"""The operator classes of the C++ host mirror (dare_amd/include/dare/Operators.h, Stencils.h).

tests/cpp/test_operators.cpp evaluates reference-style functors -- the patterns of
test/Verification/counter_diffusion2D.cpp:106-162 and advection_periodic2D.cpp:108-125, a 3-D
functor with harmonic-mean face diffusivity, TVD from cell-centred velocity fields and a
source, and a two-component reaction functor with explicit cross-component coefficients --
through the generic MatrixBlock path on several OpenMP threads, and compares the assembled
CSR, right-hand side and initial guess bit for bit with the oracle (CPU) and with the device
assembly of the same terms (GPU).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_operators")
CASES = ("CounterDiffusion2D", "AdvectionPeriodic2D", "General3D", "Reaction3D")


def _run(mode):
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    for case in CASES:
        assert f"[ OK ] {case}" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_operator_classes_match_oracle():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "test_operators"], check=True)
    _run("cpu")


@pytest.mark.gpu
def test_operator_classes_match_device_assembly():
    assert os.access(BIN, os.X_OK), "tests/cpp/test_operators is not built (run __graft_entry__.build())"
    _run("gpu")
