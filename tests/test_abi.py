"""CPU checks of the drop-in boundary: the library loads and exports every declared symbol."""
import os
import re
import subprocess

import pytest

import dare_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:int|int32_t|void\s*\*|const char\s*\*)\s*(dare_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["dare_ctx_create", "dare_ghost_fill", "dare_grad_faces", "dare_div_faces",
                     "dare_tvd_face_flux", "dare_csr_build_structure", "dare_csr_fill_values",
                     "dare_csr_upload_host", "dare_spmv", "dare_dot", "dare_axpby",
                     "dare_bicgstab2_solve", "dare_refine_solve_mixed"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(dare_amd.LIB_PATH):
        pytest.skip("libdare_hip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", dare_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_ctypes_load_and_abi_version():
    if not os.path.exists(dare_amd.LIB_PATH):
        pytest.skip("libdare_hip.so not built")
    L = dare_amd.lib()
    assert L.dare_abi_version() == 1
    for n in declared_functions():
        assert hasattr(L, n)


def test_struct_layouts_match_header():
    import ctypes
    # dare_grid: 3 int32 + 3 int32 + 3 int32 + pad + 3 double
    assert ctypes.sizeof(dare_amd.Grid) == 4 * 9 + 4 + 24
    assert ctypes.sizeof(dare_amd.Terms) % 8 == 0
    assert dare_amd.Terms.bc_value.offset == dare_amd.Terms.bc_type.offset + 24


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(dare_amd, "_lib", None)
    monkeypatch.setattr(dare_amd, "LIB_PATH", "/nonexistent/libdare_hip.so")
    with pytest.raises(dare_amd.DareError):
        dare_amd.lib()
