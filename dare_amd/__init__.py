"""dare_amd — MI355X-native finite-volume hot path of DaRe behind a C-ABI.

The compute lives in libdare_hip.so (hand-written HIP kernels for gfx950, see
include/dare_hip.h). This module loads it with ctypes and offers thin helpers that pass
torch tensors' device pointers across the boundary. torch provides device memory and the
stream only. There is no CPU fallback: if the library is missing every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdare_hip.so")

# limiter ids (include/dare_hip.h)
CDS, UPWIND, VANALBADA, MINMOD, VANLEER = 0, 1, 2, 3, 4
LIMITERS = {"CDS": CDS, "UPWIND": UPWIND, "VANALBADA": VANALBADA, "MINMOD": MINMOD, "VANLEER": VANLEER}
GHOST_LEAVE, GHOST_PERIODIC, GHOST_DIRICHLET, GHOST_NEUMANN, GHOST_CONSTANT = 0, 1, 2, 3, 4
BC_NONE, BC_DIRICHLET, BC_NEUMANN, BC_FIXED = 0, 1, 2, 3
PRECOND_NONE, PRECOND_JACOBI, PRECOND_MG, PRECOND_JACOBI_LEFT = 0, 1, 2, 3

DARE_OK = 0
DARE_ERR_BREAKDOWN = -5


class DareError(RuntimeError):
    pass


class Grid(ctypes.Structure):
    """dare_grid: one single-rank Cartesian subdomain (ghost-padded, last axis fastest)."""
    _fields_ = [("dim", ctypes.c_int32), ("ncomp", ctypes.c_int32), ("ghost", ctypes.c_int32),
                ("n", ctypes.c_int32 * 3), ("periodic", ctypes.c_int32 * 3),
                ("size", ctypes.c_double * 3)]

    @classmethod
    def make(cls, n, size=None, ghost=2, ncomp=1, periodic=None):
        n = list(n)
        dim = len(n)
        size = list(size) if size is not None else [1.0] * dim
        periodic = list(periodic) if periodic is not None else [0] * dim
        g = cls()
        g.dim, g.ncomp, g.ghost = dim, ncomp, ghost
        for d in range(3):
            g.n[d] = n[d] if d < dim else 1
            g.periodic[d] = int(periodic[d]) if d < dim else 0
            g.size[d] = float(size[d]) if d < dim else 1.0
        return g

    @property
    def res(self):
        return [self.n[d] for d in range(self.dim)]

    @property
    def padded(self):
        return [self.n[d] + 2 * self.ghost for d in range(self.dim)]

    @property
    def num_cells(self):
        return int(np.prod(self.res))

    @property
    def num_padded(self):
        return int(np.prod(self.padded))

    def face_shape(self, axis):
        return [self.n[d] + (1 if d == axis else 0) for d in range(self.dim)]

    def num_faces(self, axis):
        return int(np.prod(self.face_shape(axis)))

    # reference formulas (CartesianMesh.inl:75-83)
    def geometry(self):
        dx = [self.size[d] / float(self.n[d]) for d in range(self.dim)]
        V = 1.0
        for d in range(self.dim):
            V *= dx[d]
        A = [V / dx[d] for d in range(self.dim)]
        return dx, V, A


class Terms(ctypes.Structure):
    """dare_terms: the per-cell operator evaluated on the device by the assembly kernels."""
    _fields_ = [("dt", ctypes.c_double), ("diff_const", ctypes.c_double),
                ("diff_cell", ctypes.c_void_p), ("diff_comp", ctypes.c_double * 8),
                ("use_diff_comp", ctypes.c_int32), ("conv_limiter", ctypes.c_int32),
                ("vel", ctypes.c_void_p * 3), ("source", ctypes.c_void_p),
                ("react", ctypes.c_void_p), ("bc_type", ctypes.c_int32 * 6),
                ("bc_value", ctypes.c_double * 6), ("div_weights", ctypes.c_double * 3),
                ("div_levels", ctypes.c_int32)]

    @classmethod
    def make(cls, dt=0.0, diff=0.0, diff_cell=None, diff_comp=None, limiter=-1, vel=None,
             source=None, react=None, bc_type=None, bc_value=None, div_scheme=None):
        """div_scheme: the time-scheme weights of the divergence terms (EULER_BACKWARD,
        EULER_FORWARD, CRANK_NICHOLSON, ADAMS_BASHFORT, ADAMS_MOULTON below); None is
        EULER_BACKWARD."""
        t = cls()
        if div_scheme is not None:
            t.div_levels = len(div_scheme)
            for i, w in enumerate(div_scheme):
                t.div_weights[i] = w
        t.dt = dt
        t.diff_const = diff
        t.diff_cell = _ptr(diff_cell)
        if diff_comp is not None:
            t.use_diff_comp = 1
            for i, v in enumerate(diff_comp):
                t.diff_comp[i] = v
        t.conv_limiter = limiter
        if vel is not None:
            for d, v in enumerate(vel):
                t.vel[d] = _ptr(v)
        t.source = _ptr(source)
        t.react = _ptr(react)
        bt = list(bc_type or []) + [0] * 6
        bv = list(bc_value or []) + [0.0] * 6
        for f in range(6):
            t.bc_type[f] = bt[f]
            t.bc_value[f] = bv[f]
        t._keep = (diff_cell, vel, source, react)
        return t


# Divergence<Grid, TimeScheme> weights (Equations/TimeDiscretizationSchemes.h:29-76)
EULER_BACKWARD = (1.0,)
EULER_FORWARD = (0.0, 1.0)
CRANK_NICHOLSON = (0.5, 0.5)
ADAMS_BASHFORT = (0.0, 1.5, -0.5)
ADAMS_MOULTON = (5.0 / 12.0, 8.0 / 12.0, -1.0 / 12.0)


class SolverParams(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("tol_kind", ctypes.c_int32), ("max_it", ctypes.c_int32),
                ("inner_tol", ctypes.c_double), ("max_refine", ctypes.c_int32),
                ("inner_max_it", ctypes.c_int32), ("history", ctypes.POINTER(ctypes.c_double)),
                ("history_cap", ctypes.c_int32), ("precond", ctypes.c_int32)]


class SolverResult(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("status", ctypes.c_int32), ("refinements", ctypes.c_int32),
                ("residual", ctypes.c_double), ("bnorm", ctypes.c_double),
                ("history_len", ctypes.c_int32)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("tile_k", ctypes.c_int32), ("tile_i", ctypes.c_int32), ("grid", ctypes.c_int32),
                ("order", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _ptr(t):
    """Device pointer of a torch tensor (or an int address), None -> NULL."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


_lib = None


def lib():
    """The loaded libdare_hip.so. Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DareError(f"{LIB_PATH} not found: build it with `python -m dare_amd.build`")
    # torch's ROCm wheel ships its own HIP runtime with the same soname as the system one.
    # Loading torch first makes this library bind to that runtime: with two HIP runtimes in
    # one process, the one loaded second sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    G = ctypes.POINTER(Grid)
    T = ctypes.POINTER(Terms)
    P = ctypes.POINTER(SolverParams)
    R = ctypes.POINTER(SolverResult)
    sig = {
        "dare_ctx_create": (i32, [i32, ctypes.POINTER(vp)]),
        "dare_ctx_destroy": (i32, [vp]),
        "dare_ctx_set_stream": (i32, [vp, vp]),
        "dare_ctx_get_stream": (vp, [vp]),
        "dare_ctx_synchronize": (i32, [vp]),
        "dare_last_error": (ctypes.c_char_p, []),
        "dare_abi_version": (i32, []),
        "dare_malloc": (i32, [vp, ctypes.c_size_t, ctypes.POINTER(vp)]),
        "dare_free": (i32, [vp, vp]),
        "dare_memcpy_h2d": (i32, [vp, vp, vp, ctypes.c_size_t]),
        "dare_memcpy_d2h": (i32, [vp, vp, vp, ctypes.c_size_t]),
        "dare_memcpy_d2d": (i32, [vp, vp, vp, ctypes.c_size_t]),
        "dare_fill_f64": (i32, [vp, vp, i64, dbl]),
        "dare_stream_probe": (i32, [vp, i64, vp, vp, i32, i32, i32]),
        "dare_ghost_fill": (i32, [vp, G, vp, vp, vp]),
        "dare_grad_faces": (i32, [vp, G, vp, vp, vp, vp]),
        "dare_div_faces": (i32, [vp, G, vp, vp, vp, vp]),
        "dare_interpolate_to_faces": (i32, [vp, G, vp, vp, vp, vp, vp, vp]),
        "dare_tvd_face_flux": (i32, [vp, G, i32, vp, vp, vp, vp, vp, vp, vp]),
        "dare_tvd_explicit_step": (i32, [vp, G, i32, vp, vp, vp, vp, dbl, vp]),
        "dare_csr_create": (i32, [vp, ctypes.POINTER(vp)]),
        "dare_csr_destroy": (i32, [vp]),
        "dare_csr_build_structure": (i32, [vp, vp, G, T]),
        "dare_csr_fill_values": (i32, [vp, vp, G, T, vp, vp, vp]),
        "dare_csr_upload_host": (i32, [vp, vp, i64, i64, vp, vp, vp]),
        "dare_csr_info": (i32, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(vp),
                                ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "dare_csr_download": (i32, [vp, vp, vp, vp, vp]),
        "dare_copy_to_field": (i32, [vp, G, vp, vp]),
        "dare_spmv": (i32, [vp, vp, vp, vp]),
        "dare_dot": (i32, [vp, i64, vp, vp, ctypes.POINTER(dbl)]),
        "dare_spmv_dot": (i32, [vp, vp, vp, vp, vp, ctypes.POINTER(dbl)]),
        "dare_axpby": (i32, [vp, i64, dbl, vp, dbl, vp]),
        "dare_residual": (i32, [vp, vp, vp, vp, vp, ctypes.POINTER(dbl)]),
        "dare_csr_symmetric": (i32, [vp, vp, ctypes.POINTER(i32)]),
        "dare_csr_set_symmetric_traversal": (i32, [vp, i32]),
        "dare_bicgstab2_solve": (i32, [vp, vp, vp, vp, P, R]),
        "dare_bicgstab2_solve_noalloc": (i32, [vp, vp, vp, vp, P, R]),
        "dare_refine_solve_mixed": (i32, [vp, vp, vp, vp, P, R]),
        "dare_batch_create": (i32, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "dare_batch_destroy": (i32, [vp]),
        "dare_batch_vectors": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i64)]),
        "dare_batch_set_config": (i32, [vp, vp, ctypes.POINTER(BatchConfig)]),
        "dare_batch_get_config": (i32, [vp, ctypes.POINTER(BatchConfig)]),
        "dare_batch_autotune": (i32, [vp, vp, i32, ctypes.POINTER(BatchConfig), ctypes.POINTER(dbl)]),
        "dare_batch_solve": (i32, [vp, vp, P, i32, i32, R]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc, what=""):
    if rc != DARE_OK:
        msg = lib().dare_last_error().decode(errors="replace")
        raise DareError(f"{what} failed ({rc}): {msg}")
    return rc


class Context:
    """dare_ctx on one device; follows torch's current stream by default."""

    def __init__(self, device=0, stream=None):
        self.h = ctypes.c_void_p()
        check(lib().dare_ctx_create(device, ctypes.byref(self.h)), "dare_ctx_create")
        if stream is None:
            try:
                import torch
                stream = torch.cuda.current_stream(device).cuda_stream
            except Exception:
                stream = None
        check(lib().dare_ctx_set_stream(self.h, stream), "dare_ctx_set_stream")

    def synchronize(self):
        check(lib().dare_ctx_synchronize(self.h), "dare_ctx_synchronize")

    def close(self):
        if self.h:
            lib().dare_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Csr:
    """Device CSR handle (dare_csr)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        check(lib().dare_csr_create(ctx.h, ctypes.byref(self.h)), "dare_csr_create")

    def info(self):
        nr, nnz = ctypes.c_int64(), ctypes.c_int64()
        rp, col, val = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().dare_csr_info(self.h, ctypes.byref(nr), ctypes.byref(nnz), ctypes.byref(rp),
                                  ctypes.byref(col), ctypes.byref(val)), "dare_csr_info")
        return nr.value, nnz.value, rp.value, col.value, val.value

    @property
    def nrows(self):
        return self.info()[0]

    @property
    def nnz(self):
        return self.info()[1]

    def download(self):
        nr, nnz = self.info()[:2]
        rp = np.empty(nr + 1, np.int32)
        col = np.empty(max(nnz, 1), np.int32)
        val = np.empty(max(nnz, 1), np.float64)
        check(lib().dare_csr_download(self.ctx.h, self.h, rp.ctypes.data, col.ctypes.data,
                                      val.ctypes.data), "dare_csr_download")
        return rp, col[:nnz], val[:nnz]

    def close(self):
        if self.h:
            lib().dare_csr_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- thin wrappers (torch tensors)

def stream_probe(ctx, src, dst, ratio=1, grid=0, layout=0):
    """dst[j] = sum of `ratio` elements of src (src holds ratio * dst.numel()), through the
    library's streaming kernel: the attainable-bandwidth probe (ratio 1 = copy; layout 0 XCD
    slabs, 1 grid stride, 2 XCD slabs over `ratio` separate arrays, 3 the same read-only: dst[:512 grid]
    receives per-lane totals)."""
    assert src.numel() == ratio * dst.numel()
    check(lib().dare_stream_probe(ctx.h, dst.numel(), _ptr(src), _ptr(dst), ratio, grid, layout),
          "dare_stream_probe")


def ghost_fill(ctx, grid, phi, modes, values=None):
    m = (ctypes.c_int32 * 6)(*list(modes) + [0] * (6 - len(modes)))
    v = (ctypes.c_double * 6)(*(list(values) if values is not None else [0.0] * 6))
    check(lib().dare_ghost_fill(ctx.h, ctypes.byref(grid), m, v, _ptr(phi)), "dare_ghost_fill")


def grad_faces(ctx, grid, phi, faces):
    f = list(faces) + [None] * (3 - len(faces))
    check(lib().dare_grad_faces(ctx.h, ctypes.byref(grid), _ptr(phi), _ptr(f[0]), _ptr(f[1]),
                                _ptr(f[2])), "dare_grad_faces")


def div_faces(ctx, grid, faces, out):
    f = list(faces) + [None] * (3 - len(faces))
    check(lib().dare_div_faces(ctx.h, ctypes.byref(grid), _ptr(f[0]), _ptr(f[1]), _ptr(f[2]),
                               _ptr(out)), "dare_div_faces")


def interpolate_to_faces(ctx, grid, ucell, faces):
    """Cell-centred velocity fields (one padded tensor per axis) -> face arrays."""
    u = list(ucell) + [None] * (3 - len(ucell))
    f = list(faces) + [None] * (3 - len(faces))
    check(lib().dare_interpolate_to_faces(ctx.h, ctypes.byref(grid), _ptr(u[0]), _ptr(u[1]), _ptr(u[2]),
                                          _ptr(f[0]), _ptr(f[1]), _ptr(f[2])), "dare_interpolate_to_faces")


def tvd_face_flux(ctx, grid, limiter, phi, vel, faces):
    u = list(vel) + [None] * (3 - len(vel))
    f = list(faces) + [None] * (3 - len(faces))
    check(lib().dare_tvd_face_flux(ctx.h, ctypes.byref(grid), limiter, _ptr(phi), _ptr(u[0]),
                                   _ptr(u[1]), _ptr(u[2]), _ptr(f[0]), _ptr(f[1]), _ptr(f[2])),
          "dare_tvd_face_flux")


def tvd_explicit_step(ctx, grid, limiter, phi, vel, dt, phi_out):
    u = list(vel) + [None] * (3 - len(vel))
    check(lib().dare_tvd_explicit_step(ctx.h, ctypes.byref(grid), limiter, _ptr(phi), _ptr(u[0]),
                                       _ptr(u[1]), _ptr(u[2]), dt, _ptr(phi_out)),
          "dare_tvd_explicit_step")


def build_structure(ctx, csr, grid, terms):
    check(lib().dare_csr_build_structure(ctx.h, csr.h, ctypes.byref(grid), ctypes.byref(terms)),
          "dare_csr_build_structure")


def fill_values(ctx, csr, grid, terms, phi_old, rhs, x0=None):
    check(lib().dare_csr_fill_values(ctx.h, csr.h, ctypes.byref(grid), ctypes.byref(terms),
                                     _ptr(phi_old), _ptr(rhs), _ptr(x0)), "dare_csr_fill_values")


def upload_host_csr(ctx, csr, row_ptr, col, val):
    rp = np.ascontiguousarray(row_ptr, np.int32)
    c = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(val, np.float64)
    check(lib().dare_csr_upload_host(ctx.h, csr.h, len(rp) - 1, len(c), rp.ctypes.data,
                                     c.ctypes.data, v.ctypes.data), "dare_csr_upload_host")


def spmv(ctx, csr, x, y):
    check(lib().dare_spmv(ctx.h, csr.h, _ptr(x), _ptr(y)), "dare_spmv")


def spmv_dot(ctx, csr, x, y, w, readback=True):
    out = ctypes.c_double()
    check(lib().dare_spmv_dot(ctx.h, csr.h, _ptr(x), _ptr(y), _ptr(w),
                              ctypes.byref(out) if readback else None), "dare_spmv_dot")
    return out.value if readback else None


def dot(ctx, a, b):
    out = ctypes.c_double()
    check(lib().dare_dot(ctx.h, a.numel(), _ptr(a), _ptr(b), ctypes.byref(out)), "dare_dot")
    return out.value


def residual(ctx, csr, b, x, r, norm=False):
    """r = b - A x with the solver's operator; returns ||r|| when norm is True."""
    v = ctypes.c_double()
    check(lib().dare_residual(ctx.h, csr.h, _ptr(b), _ptr(x), _ptr(r), ctypes.byref(v) if norm else None),
          "dare_residual")
    return v.value if norm else None


def csr_symmetric(ctx, csr):
    """True when the FP64 operator applications and solves use the symmetric (four-slot)
    traversal of the matrix's diagonal view: the view is exactly symmetric."""
    v = ctypes.c_int32()
    check(lib().dare_csr_symmetric(ctx.h, csr.h, ctypes.byref(v)), "dare_csr_symmetric")
    return bool(v.value)


def set_symmetric_traversal(csr, enable):
    """enable=False: every traversal of the matrix reads its seven slots (same results)."""
    check(lib().dare_csr_set_symmetric_traversal(csr.h, 1 if enable else 0), "dare_csr_set_symmetric_traversal")


def copy_to_field(ctx, grid, x, phi):
    check(lib().dare_copy_to_field(ctx.h, ctypes.byref(grid), _ptr(x), _ptr(phi)), "dare_copy_to_field")


def _params(tol, tol_kind, max_it, inner_tol, max_refine, inner_max_it, hist_cap, precond=0):
    p = SolverParams()
    p.precond = precond
    p.tol, p.tol_kind, p.max_it = tol, tol_kind, max_it
    p.inner_tol, p.max_refine, p.inner_max_it = inner_tol, max_refine, inner_max_it
    hist = (ctypes.c_double * max(hist_cap, 1))()
    p.history = ctypes.cast(hist, ctypes.POINTER(ctypes.c_double))
    p.history_cap = hist_cap
    return p, hist


def bicgstab2(ctx, csr, b, x, tol=1e-8, tol_kind=0, max_it=100, history=0, precond=0):
    """FP64 BiCGStab(2) (BiCGStab2.inl:167-582). x holds the initial guess on entry.
    precond=1: Jacobi right preconditioning."""
    p, hist = _params(tol, tol_kind, max_it, 0.0, 0, 0, history, precond)
    r = SolverResult()
    check(lib().dare_bicgstab2_solve(ctx.h, csr.h, _ptr(b), _ptr(x), ctypes.byref(p), ctypes.byref(r)),
          "dare_bicgstab2_solve")
    out = {k: getattr(r, k) for k, _ in SolverResult._fields_}
    out["history"] = list(hist[:r.history_len])
    return out


def bicgstab2_noalloc(ctx, csr, b, x, tol=1e-8, tol_kind=0, max_it=100, history=0):
    """dare_bicgstab2_solve_noalloc: the FP64 solve with its workspace in the library's static
    device scratch (no device allocation; at most 262144 rows, unpreconditioned)."""
    p, hist = _params(tol, tol_kind, max_it, 0.0, 0, 0, history, 0)
    r = SolverResult()
    check(lib().dare_bicgstab2_solve_noalloc(ctx.h, csr.h, _ptr(b), _ptr(x), ctypes.byref(p), ctypes.byref(r)),
          "dare_bicgstab2_solve_noalloc")
    out = {k: getattr(r, k) for k, _ in SolverResult._fields_}
    out["history"] = list(hist[:r.history_len])
    return out


def refine_mixed(ctx, csr, b, x, tol=1e-8, tol_kind=1, max_it=1000, inner_tol=1e-4, max_refine=30,
                 inner_max_it=200, history=0, precond=0):
    """FP32 inner BiCGStab(2) inside FP64 iterative refinement (precond=1: Jacobi on the inner
    solves)."""
    p, hist = _params(tol, tol_kind, max_it, inner_tol, max_refine, inner_max_it, history, precond)
    r = SolverResult()
    check(lib().dare_refine_solve_mixed(ctx.h, csr.h, _ptr(b), _ptr(x), ctypes.byref(p), ctypes.byref(r)),
          "dare_refine_solve_mixed")
    out = {k: getattr(r, k) for k, _ in SolverResult._fields_}
    out["history"] = list(hist[:r.history_len])
    return out


class Batch:
    """Batched solves over many structured systems (dare_batch). `systems` are Csr handles
    built by build_structure on 3-D grids; the batch owns the concatenated b and x, and
    `rhs_ptr(s)` / `x_ptr(s)` are the device addresses of system s's rows in them."""

    def __init__(self, ctx, systems):
        self.ctx = ctx
        self.systems = list(systems)
        arr = (ctypes.c_void_p * len(self.systems))(*[c.h.value for c in self.systems])
        self.h = ctypes.c_void_p()
        check(lib().dare_batch_create(ctx.h, len(self.systems), arr, ctypes.byref(self.h)), "dare_batch_create")
        b, x = ctypes.c_void_p(), ctypes.c_void_p()
        offs = (ctypes.c_int64 * (len(self.systems) + 1))()
        check(lib().dare_batch_vectors(self.h, ctypes.byref(b), ctypes.byref(x), offs), "dare_batch_vectors")
        self.b_addr, self.x_addr = b.value, x.value
        self.offsets = np.array(offs[:], dtype=np.int64)

    @property
    def rows(self):
        return int(self.offsets[-1])

    def rhs_ptr(self, s):
        return self.b_addr + 8 * int(self.offsets[s])

    def x_ptr(self, s):
        return self.x_addr + 8 * int(self.offsets[s])

    def config(self):
        c = BatchConfig()
        check(lib().dare_batch_get_config(self.h, ctypes.byref(c)), "dare_batch_get_config")
        return c.as_dict()

    def set_config(self, tile_k=0, tile_i=16, grid=2048, order=1):
        c = BatchConfig(tile_k, tile_i, grid, order)
        check(lib().dare_batch_set_config(self.ctx.h, self.h, ctypes.byref(c)), "dare_batch_set_config")

    def autotune(self, reps=5):
        c, us = BatchConfig(), ctypes.c_double()
        check(lib().dare_batch_autotune(self.ctx.h, self.h, reps, ctypes.byref(c), ctypes.byref(us)),
              "dare_batch_autotune")
        return c.as_dict(), us.value

    def solve(self, tol=1e-8, tol_kind=1, max_it=1000, mixed=False, inner_tol=1e-4, max_refine=30,
              inner_max_it=1000, graph_passes=4):
        p, hist = _params(tol, tol_kind, max_it, inner_tol, max_refine, inner_max_it, 0)
        res = (SolverResult * len(self.systems))()
        check(lib().dare_batch_solve(self.ctx.h, self.h, ctypes.byref(p), 1 if mixed else 0, graph_passes, res),
              "dare_batch_solve")
        return [{k: getattr(r, k) for k, _ in SolverResult._fields_} for r in res]

    def close(self):
        if self.h:
            lib().dare_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
