"""CPU oracle of the DaRe FV hot path — TEST INFRASTRUCTURE ONLY.

ctypes wrapper over oracle/liboracle.so (oracle.cpp, a C++ restatement of the reference) and,
where it was built, oracle/_ref/libdare_ref.so (the reference's own standalone headers). Only
tests/ and __graft_entry__.smoke() (to check the HIP results against it) and bench.py's
cpu_baseline (to time it on the host) use this package.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libdare_ref.so")

_lib = None
_ref = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OGrid(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("ncomp", ctypes.c_int32), ("ghost", ctypes.c_int32),
                ("n", ctypes.c_int32 * 3), ("periodic", ctypes.c_int32 * 3),
                ("size", ctypes.c_double * 3)]


class OTerms(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("diff_const", ctypes.c_double),
                ("diff_cell", ctypes.c_void_p), ("diff_comp", ctypes.c_double * 8),
                ("use_diff_comp", ctypes.c_int32), ("conv_limiter", ctypes.c_int32),
                ("vel", ctypes.c_void_p * 3), ("source", ctypes.c_void_p),
                ("react", ctypes.c_void_p), ("bc_type", ctypes.c_int32 * 6),
                ("bc_value", ctypes.c_double * 6), ("div_weights", ctypes.c_double * 3),
                ("div_levels", ctypes.c_int32)]


def grid(n, size=None, ghost=2, ncomp=1, periodic=None):
    n = list(n)
    dim = len(n)
    g = OGrid()
    g.dim, g.ncomp, g.ghost = dim, ncomp, ghost
    size = list(size) if size is not None else [1.0] * dim
    periodic = list(periodic) if periodic is not None else [0] * dim
    for d in range(3):
        g.n[d] = n[d] if d < dim else 1
        g.periodic[d] = int(periodic[d]) if d < dim else 0
        g.size[d] = float(size[d]) if d < dim else 1.0
    return g


def _p(a):
    return None if a is None else a.ctypes.data


@contextlib.contextmanager
def using(path):
    """Routes every wrapper of this module to another build of oracle.cpp for the duration of
    the block (the planted-mistake checks of tests/ load mutant builds this way)."""
    global _lib
    saved = _lib
    _lib = _load(path)
    try:
        yield _lib
    finally:
        _lib = saved


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _load(LIB)
    return _lib


def _load(path):
    L = ctypes.CDLL(path)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    G = ctypes.POINTER(OGrid)
    T = ctypes.POINTER(OTerms)
    sig = {
        "orc_limiter": (dbl, [i32, dbl]),
        "orc_diffusion_n0d1": (dbl, [dbl, dbl, i32]),
        "orc_num_cells_padded": (i64, [G]),
        "orc_num_cells_internal": (i64, [G]),
        "orc_map_internal_to_local": (i64, [G, i64]),
        "orc_num_faces": (i64, [G, i32]),
        "orc_cell_center": (None, [G, vp, vp]),
        "orc_grad_faces": (ctypes.c_int, [G, vp, vp, vp, vp]),
        "orc_div_faces": (ctypes.c_int, [G, vp, vp, vp, vp]),
        "orc_interpolate_to_faces": (ctypes.c_int, [G, vp, vp, vp, vp, vp, vp]),
        "orc_tvd_face_flux": (ctypes.c_int, [G, i32, vp, vp, vp, vp, vp, vp, vp]),
        "orc_tvd_explicit_step": (ctypes.c_int, [G, i32, vp, vp, vp, vp, dbl, vp]),
        "orc_tvd_matrix_cell": (ctypes.c_int, [G, i32, vp, vp, vp, vp, i64, i32, vp, vp, vp]),
        "orc_ghost_fill": (ctypes.c_int, [G, vp, vp, vp]),
        "orc_assemble": (i64, [G, T, vp, vp, vp, vp, i64, vp, vp]),
        "orc_spmv": (None, [i64, vp, vp, vp, vp, vp]),
        "orc_dot": (dbl, [i64, vp, vp]),
        "orc_bicgstab2": (ctypes.c_int, [i64, vp, vp, vp, vp, vp, dbl, i32,
                                         ctypes.POINTER(i32), ctypes.POINTER(dbl), vp, i32]),
        "orc_copy_to_field": (None, [G, vp, vp]),
        "orc_set_threads": (None, [i32]),
        "orc_get_threads": (i32, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def ref():
    """The reference's own headers compiled into oracle/_ref, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        R = ctypes.CDLL(REF_LIB)
        dbl = ctypes.c_double
        R.ref_limiter.restype = dbl
        R.ref_limiter.argtypes = [ctypes.c_int, dbl]
        R.ref_diffusion_n0d1.restype = dbl
        R.ref_diffusion_n0d1.argtypes = [dbl, dbl, ctypes.c_int]
        R.ref_tvd_face_value.restype = dbl
        R.ref_tvd_face_value.argtypes = [ctypes.c_int, dbl, dbl, dbl]
        R.ref_tvd_face_rhs.restype = dbl
        R.ref_tvd_face_rhs.argtypes = [ctypes.c_int, dbl, dbl, dbl, dbl]
        _ref = R
    return _ref


def set_threads(n):
    lib().orc_set_threads(int(n))


def terms(dt=0.0, diff=0.0, diff_cell=None, diff_comp=None, limiter=-1, vel=None, source=None,
          react=None, bc_type=None, bc_value=None, div_scheme=None):
    t = OTerms()
    if div_scheme is not None:          # Divergence time-scheme weights, e.g. (0.5, 0.5)
        t.div_levels = len(div_scheme)
        for i, w in enumerate(div_scheme):
            t.div_weights[i] = w
    t.dt, t.diff_const = dt, diff
    t.diff_cell = _p(diff_cell)
    if diff_comp is not None:
        t.use_diff_comp = 1
        for i, v in enumerate(diff_comp):
            t.diff_comp[i] = v
    t.conv_limiter = limiter
    if vel is not None:
        for d, v in enumerate(vel):
            t.vel[d] = _p(v)
    t.source = _p(source)
    t.react = _p(react)
    bt = list(bc_type or []) + [0] * 6
    bv = list(bc_value or []) + [0.0] * 6
    for f in range(6):
        t.bc_type[f] = bt[f]
        t.bc_value[f] = bv[f]
    # keep the arrays alive with the struct
    t._keep = (diff_cell, vel, source, react)
    return t


def faces_alloc(g):
    out = []
    for ax in range(g.dim):
        out.append(np.zeros(lib().orc_num_faces(ctypes.byref(g), ax) * g.ncomp))
    return out


def grad_faces(g, phi):
    f = faces_alloc(g) + [None] * (3 - g.dim)
    rc = lib().orc_grad_faces(ctypes.byref(g), _p(phi), _p(f[0]), _p(f[1]), _p(f[2]))
    assert rc == 0
    return f[:g.dim]


def div_faces(g, faces):
    out = np.zeros(lib().orc_num_cells_padded(ctypes.byref(g)) * g.ncomp)
    f = list(faces) + [None] * (3 - len(faces))
    rc = lib().orc_div_faces(ctypes.byref(g), _p(f[0]), _p(f[1]), _p(f[2]), _p(out))
    assert rc == 0
    return out


def interpolate_to_faces(g, ucell):
    """Cell-centred velocity fields (one padded array per axis) -> face arrays."""
    f = faces_alloc_1(g) + [None] * (3 - g.dim)
    u = list(ucell) + [None] * (3 - len(ucell))
    rc = lib().orc_interpolate_to_faces(ctypes.byref(g), _p(u[0]), _p(u[1]), _p(u[2]), _p(f[0]), _p(f[1]), _p(f[2]))
    assert rc == 0
    return f[:g.dim]


def faces_alloc_1(g):
    return [np.zeros(lib().orc_num_faces(ctypes.byref(g), ax)) for ax in range(g.dim)]


def tvd_face_flux(g, lim, phi, vel):
    f = faces_alloc(g) + [None] * (3 - g.dim)
    u = list(vel) + [None] * (3 - len(vel))
    rc = lib().orc_tvd_face_flux(ctypes.byref(g), lim, _p(phi), _p(u[0]), _p(u[1]), _p(u[2]),
                                 _p(f[0]), _p(f[1]), _p(f[2]))
    assert rc == 0
    return f[:g.dim]


def tvd_explicit_step(g, lim, phi, vel, dt):
    out = phi.copy()
    u = list(vel) + [None] * (3 - len(vel))
    rc = lib().orc_tvd_explicit_step(ctypes.byref(g), lim, _p(phi), _p(u[0]), _p(u[1]), _p(u[2]),
                                     dt, _p(out))
    assert rc == 0
    return out


def tvd_matrix_cell(g, lim, phi, vel, cell, comp=0):
    nb, c, rhs = np.zeros(6), np.zeros(6), np.zeros(6)
    u = list(vel) + [None] * (3 - len(vel))
    rc = lib().orc_tvd_matrix_cell(ctypes.byref(g), lim, _p(phi), _p(u[0]), _p(u[1]), _p(u[2]),
                                   cell, comp, _p(nb), _p(c), _p(rhs))
    assert rc == 0
    return nb, c, rhs


def ghost_fill(g, phi, modes, values=None):
    m = np.array(list(modes) + [0] * (6 - len(modes)), np.int32)
    v = np.array(values if values is not None else [0.0] * 6, np.float64)
    out = phi.copy()
    rc = lib().orc_ghost_fill(ctypes.byref(g), _p(m), _p(v), _p(out))
    assert rc == 0, rc
    return out


def assemble(g, t, phi_old):
    ncell = lib().orc_num_cells_internal(ctypes.byref(g))
    nrows = ncell * g.ncomp
    cap = nrows * (2 * g.dim + 2)
    rp = np.zeros(nrows + 1, np.int32)
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap, np.float64)
    rhs = np.zeros(nrows)
    x0 = np.zeros(nrows)
    nnz = lib().orc_assemble(ctypes.byref(g), ctypes.byref(t), _p(phi_old), _p(rp), _p(col),
                             _p(val), cap, _p(rhs), _p(x0))
    assert nnz >= 0, nnz
    return rp, col[:nnz].copy(), val[:nnz].copy(), rhs, x0


def spmv(rp, col, val, x):
    y = np.zeros(len(rp) - 1)
    lib().orc_spmv(len(rp) - 1, _p(rp), _p(col), _p(val), _p(np.ascontiguousarray(x)), _p(y))
    return y


def bicgstab2(rp, col, val, b, x0, tol, max_it=100, hist_cap=0):
    x = np.array(x0, dtype=np.float64, copy=True)
    it = ctypes.c_int32()
    res = ctypes.c_double()
    hist = np.zeros(max(hist_cap, 1))
    ok = lib().orc_bicgstab2(len(rp) - 1, _p(rp), _p(col), _p(val), _p(np.ascontiguousarray(b)),
                             _p(x), tol, max_it, ctypes.byref(it), ctypes.byref(res), _p(hist),
                             hist_cap)
    nh = min(hist_cap, it.value)
    return x, bool(ok), it.value, res.value, hist[:nh]


def copy_to_field(g, x, phi):
    out = phi.copy()
    lib().orc_copy_to_field(ctypes.byref(g), _p(x), _p(out))
    return out


def cell_centers(g):
    """Cell-centre coordinates of every padded cell, shape (padded..., dim)."""
    dx = [g.size[d] / g.n[d] for d in range(g.dim)]
    axes = [(np.arange(g.n[d] + 2 * g.ghost) - g.ghost + 0.5) * dx[d] for d in range(g.dim)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack(mesh, axis=-1)
