"""VTK output of the host mirror (dare_amd/include/dare/IO.h), checked on the CPU.

tests/cpp/test_vtk_writer.cpp writes 1-D, 2-D and 3-D grid vectors whose cell values encode
their local index; this test reads the files back (XML header + raw appended blocks) and
checks what IO/VTKWriter.inl:38-95 and Grid/Cartesian/VTKOptions_Cartesian.h:144-204,264-322
fix: file layout and names (VTKWriter.cpp:28-46), extents, ghost level, points at the cell
vertices offset by -g*dx in VTK's x-fastest order, one cell tuple per local cell (ghosts
included) in that order, array and component names, and the TimeValue field. A staggered
representation (CartesianRepresentation.inl:50-77) is written as its own grid "<grid>_Y":
one more internal cell along the staggered axis and points shifted by -dx/2 there, with dx
the grid's cell width.
"""
import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "test_vtk_writer")


def read_vts(path):
    raw = open(path, "rb").read()
    head_end = raw.index(b'<AppendedData encoding="raw">')
    head = raw[:head_end].decode()
    base = raw.index(b"_", head_end) + 1
    arrays = {}
    for tag in re.finditer(r"<DataArray ([^>]*)/>", head):
        attrs = dict(re.findall(r'(\w+)="([^"]*)"', tag.group(1)))
        off = base + int(attrs["offset"])
        nbytes = struct.unpack_from("<Q", raw, off)[0]
        arrays[attrs["Name"]] = (attrs, np.frombuffer(raw, dtype="<f8", count=nbytes // 8, offset=off + 8))
    ext = [int(v) for v in re.search(r'<Piece Extent="([^"]*)"', head).group(1).split()]
    whole = [int(v) for v in re.search(r'WholeExtent="([^"]*)"', head).group(1).split()]
    return head, arrays, ext, whole


def code(i, j, k, c):
    return 1e6 * i + 1e3 * j + k + 0.125 * c


@pytest.fixture(scope="module")
def out(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not installed")
    subprocess.run(["make", "-s", "-C", CPP, "test_vtk_writer"], check=True)
    d = str(tmp_path_factory.mktemp("vtk"))
    r = subprocess.run([BIN, d], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return d


@pytest.mark.parametrize("name,step,grid_res,size,g,dim,time,stag", [
    ("Cartesian3D", 7, [5, 6, 7], [1.0, 1.2, 1.4], 2, 3, 0.25, [0, 0, 0]),
    ("plane", 3, [4, 3], [2.0, 1.5], 1, 2, None, [0, 0]),
    ("Cartesian1D", 0, [9], [0.9], 2, 1, 1.5, [0]),
    ("stag", 2, [5, 6, 7], [1.0, 1.2, 1.4], 2, 3, 0.5, [0, 0, 0]),
    ("stag_Y", 2, [5, 6, 7], [1.0, 1.2, 1.4], 2, 3, 0.5, [0, 1, 0]),
    ("line_X", 1, [6], [1.2], 1, 1, None, [1]),
])
def test_vtk_files(out, name, step, grid_res, size, g, dim, time, stag):
    res = [n + s for n, s in zip(grid_res, stag)]      # internal cells of the representation
    sub = f"{name}_{step}"
    pvts = os.path.join(out, sub + ".pvts")
    vts = os.path.join(out, sub, sub + "_0.vts")
    assert os.path.exists(pvts) and os.path.exists(vts)
    loc = [n + 2 * g for n in res]
    # extents: whole = internal cells, piece = local cells from offset 0; 1-D -> one cell in y
    whole = [0, 0, 0, 0, 0, 0]
    piece = [0, 0, 0, 0, 0, 0]
    for d in range(dim):
        whole[2 * d + 1] = res[d]
        piece[2 * d + 1] = loc[d]
    if dim == 1:
        whole[3] = piece[3] = 1
    ptext = open(pvts).read()
    assert f'WholeExtent="{" ".join(map(str, whole))}"' in ptext
    assert f'GhostLevel="{g}"' in ptext
    assert f'Source="{sub}/{sub}_0.vts"' in ptext
    head, arrays, ext, wext = read_vts(vts)
    assert ext == piece and wext == piece
    # points: vertex index times dx minus g dx, x fastest
    npt = [1, 1, 1]
    dx = [0.0, 0.0, 0.0]
    for d in range(dim):
        npt[d] = loc[d] + 1
        dx[d] = size[d] / grid_res[d]
    if dim == 1:
        npt[1], dx[1] = 2, dx[0]
    pts = arrays["Points"][1].reshape(npt[2], npt[1], npt[0], 3)
    kk, jj, ii = np.meshgrid(np.arange(npt[2]), np.arange(npt[1]), np.arange(npt[0]), indexing="ij")
    off = [(-0.5 * dx[d] if stag[d] else 0.0) - g * dx[d] if d < dim else 0.0 for d in range(3)]
    np.testing.assert_array_equal(pts[..., 0], dx[0] * ii + off[0])
    np.testing.assert_array_equal(pts[..., 1], dx[1] * jj + off[1])
    np.testing.assert_array_equal(pts[..., 2], dx[2] * kk + off[2])
    # cell data: one tuple per local cell, VTK order (x fastest)
    ncell = int(np.prod(loc))
    for vname, (attrs, vals) in arrays.items():
        if vname in ("Points", "TimeValue"):
            continue
        nc = int(attrs["NumberOfComponents"])
        assert vals.size == ncell * nc
        cshape = [loc[d] if d < dim else 1 for d in range(3)]
        v = vals.reshape(cshape[2], cshape[1], cshape[0], nc)
        for c in range(nc):
            k3, j3, i3 = np.meshgrid(np.arange(cshape[2]), np.arange(cshape[1]), np.arange(cshape[0]), indexing="ij")
            exp = code(i3, j3 if dim > 1 else 0, k3 if dim > 2 else 0, c)
            np.testing.assert_array_equal(v[..., c], exp)
    if time is None:
        assert "TimeValue" not in arrays and "TimeValue" not in ptext
    else:
        assert arrays["TimeValue"][1].tolist() == [time]


def test_vtk_names(out):
    head, arrays, _, _ = read_vts(os.path.join(out, "Cartesian3D_7", "Cartesian3D_7_0.vts"))
    assert set(arrays) == {"TimeValue", "phi", "vel", "Points"}
    attrs = arrays["vel"][0]
    # empty component names are replaced by the component index (VTKWriter.inl:177-184)
    assert (attrs["ComponentName0"], attrs["ComponentName1"], attrs["ComponentName2"]) == ("u", "1", "w")
    assert "ComponentName0" not in arrays["phi"][0]
    head2, arrays2, _, _ = read_vts(os.path.join(out, "plane_3", "plane_3_0.vts"))
    assert (arrays2["c"][0]["ComponentName0"], arrays2["c"][0]["ComponentName1"]) == ("0", "1")


def test_vtk_staggered_grouping(out):
    """A cell-centred and a staggered vector of one grid go to two grids named after their
    representations (VTKWriter.inl:124-137 groups by GetName())."""
    _, a, _, _ = read_vts(os.path.join(out, "stag_2", "stag_2_0.vts"))
    _, b, _, _ = read_vts(os.path.join(out, "stag_Y_2", "stag_Y_2_0.vts"))
    assert set(a) == {"TimeValue", "p", "Points"}
    assert set(b) == {"TimeValue", "v", "Points"}
