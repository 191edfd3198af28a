"""Learned warm start for implicit steps (host side of dare_convnet).

WarmStartNet wraps the C-ABI network: a 3-level 3-D convolutional encoder-decoder
(16/32/64 channels, 3x3x3, BF16 weights and activations, FP32 accumulation, hand-written
matrix-core kernels in csrc/convnet.hip). It corrects a classical guess x_base — by default
the second-order extrapolation 2 phi^n - phi^{n-1}, computed in FP64. Inputs per cell:
phi^n / max|phi^n|, (phi^n - phi^{n-1}) / max|.|, r0 dt / V / S with r0 = b - A x_base the
residual of x_base and S = max|r0| dt / V, and optional static fields; the guess is
x_base + S * out. Training is online Adam on (inputs, solver solution) pairs that the plain
solver produces along its own trajectories; the solver then starts from the prediction and
still iterates to its FP64 tolerance.
"""
import ctypes
import json

import numpy as np

from . import DareError, _ptr, check, lib

LAYERS = [(8, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
          (64, 32, 1), (32, 32, 1), (32, 16, 0), (16, 16, 0), (16, 16, 0)]


def _sig():
    L = lib()
    if getattr(L, "_convnet_sig", False):
        return L
    vp, i32, i64, dbl, flt = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_float
    from . import Grid
    G = ctypes.POINTER(Grid)
    sig = {
        "dare_convnet_create": (i32, [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(vp)]),
        "dare_convnet_destroy": (i32, [vp]),
        "dare_convnet_num_params": (i32, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
        "dare_convnet_set_params": (i32, [vp, vp, vp]),
        "dare_convnet_get_params": (i32, [vp, vp, vp, vp]),
        "dare_convnet_pack": (i32, [vp, vp, G, vp, vp, vp, dbl, i32, ctypes.POINTER(vp), vp, vp, vp]),
        "dare_conv_net_fwd": (i32, [vp, vp]),
        "dare_convnet_set_inference": (i32, [vp, i32]),
        "dare_convnet_set_coarsening": (i32, [vp, i32]),
        "dare_convnet_predict": (i32, [vp, vp, G, vp, vp, vp, dbl, i32, ctypes.POINTER(vp), vp, vp, vp, vp]),
        "dare_convnet_unpack": (i32, [vp, vp, G, vp, vp, vp]),
        "dare_convnet_loss": (i32, [vp, vp, G, vp, vp, ctypes.POINTER(dbl), vp]),
        "dare_conv_net_bwd": (i32, [vp, vp]),
        "dare_conv_net_step": (i32, [vp, vp, flt, flt, flt, flt]),
        "dare_convnet_tensors": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(dbl)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    L._convnet_sig = True
    return L


def he_init(in_channels, seed=0, out_scale=1e-3):
    """He-normal weights for the ReLU layers (fan-in 27 x real input channels); the output
    layer starts small, so the first guesses are close to phi^n."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (cin, cout, _) in enumerate(LAYERS):
        fan_in = 27 * (in_channels if l == 0 else cin)
        w = rng.normal(0.0, np.sqrt(2.0 / fan_in), size=(27, cout, cin))
        if l == 0:
            w[:, :, in_channels:] = 0.0
        if l == len(LAYERS) - 1:
            w = rng.normal(0.0, out_scale / np.sqrt(27 * cin), size=(27, cout, cin))
            w[:, 1:, :] = 0.0          # only output channel 0 is read
        parts.append(w.ravel())
    return np.concatenate(parts).astype(np.float32)


class WarmStartNet:
    """max_dims: the largest grid box the network will see. coarsen = 2 runs the network at half
    resolution (2x2x2 means in, trilinear interpolation of the correction out): 8x fewer
    voxels for the smooth part of the step, which is what the solver is slow to remove."""

    def __init__(self, ctx, max_dims, in_channels=3, seed=0, params=None, coarsen=1):
        L = _sig()
        self.ctx = ctx
        self.in_channels = in_channels
        self.coarsen = int(coarsen)
        self.h = ctypes.c_void_p()
        self._inference = False
        dims = (ctypes.c_int32 * 3)(*[d // self.coarsen for d in max_dims])
        check(L.dare_convnet_create(ctx.h, in_channels, dims, ctypes.byref(self.h)), "dare_convnet_create")
        if self.coarsen != 1:
            check(L.dare_convnet_set_coarsening(self.h, self.coarsen), "dare_convnet_set_coarsening")
        n = ctypes.c_int64()
        offs = (ctypes.c_int64 * 12)()
        check(L.dare_convnet_num_params(self.h, ctypes.byref(n), offs), "dare_convnet_num_params")
        self.nparams = n.value
        self.offsets = list(offs)
        self.set_params(he_init(in_channels, seed) if params is None else params)
        self.loss_history = []

    # ------------------------------------------------------------------ parameters
    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.nparams:
            raise DareError(f"expected {self.nparams} parameters, got {w.size}")
        check(_sig().dare_convnet_set_params(self.ctx.h, self.h, w.ctypes.data), "dare_convnet_set_params")

    def params(self):
        w = np.empty(self.nparams, np.float32)
        check(_sig().dare_convnet_get_params(self.ctx.h, self.h, w.ctypes.data, None), "dare_convnet_get_params")
        return w

    def grads(self):
        g = np.empty(self.nparams, np.float32)
        check(_sig().dare_convnet_get_params(self.ctx.h, self.h, None, g.ctypes.data), "dare_convnet_get_params")
        return g

    def save(self, path, meta=None):
        info = {"in_channels": self.in_channels, "layers": LAYERS, "coarsen": self.coarsen, **(meta or {})}
        np.savez(path, params=self.params(), meta=np.array(json.dumps(info)))

    @staticmethod
    def load_params(path):
        """(parameters, meta); meta["coarsen"] is 1 for networks saved before coarsening existed."""
        with np.load(path, allow_pickle=False) as z:
            meta = json.loads(str(z["meta"]))
            meta.setdefault("coarsen", 1)
            return z["params"].astype(np.float32), meta

    @classmethod
    def load(cls, ctx, path, max_dims):
        w, meta = cls.load_params(path)
        return cls(ctx, max_dims, in_channels=meta["in_channels"], params=w, coarsen=meta["coarsen"])

    # ------------------------------------------------------------------ passes
    @staticmethod
    def _crop(crop):
        return None if crop is None else (ctypes.c_int32 * 6)(*crop)

    def pack(self, grid, phi, phi_prev, resid, resid_scale, extras=(), crop=None):
        ne = len(extras)
        fptr = (ctypes.c_void_p * max(ne, 1))(*[_ptr(f) for f, _, _ in extras])
        sc = np.array([s for _, s, _ in extras] or [0.0], np.float64)
        of = np.array([o for _, _, o in extras] or [0.0], np.float64)
        check(_sig().dare_convnet_pack(self.ctx.h, self.h, ctypes.byref(grid), _ptr(phi), _ptr(phi_prev), _ptr(resid),
                                       float(resid_scale), ne, fptr, sc.ctypes.data, of.ctypes.data,
                                       self._crop(crop)), "dare_convnet_pack")

    def forward(self, inference=False):
        """inference=True skips the decoder activations that only backward() reads."""
        if inference != self._inference:
            check(_sig().dare_convnet_set_inference(self.h, int(inference)), "dare_convnet_set_inference")
            self._inference = inference
        check(_sig().dare_conv_net_fwd(self.ctx.h, self.h), "dare_conv_net_fwd")

    def unpack(self, grid, x_base, x_out, crop=None):
        check(_sig().dare_convnet_unpack(self.ctx.h, self.h, ctypes.byref(grid), _ptr(x_base), _ptr(x_out),
                                         self._crop(crop)), "dare_convnet_unpack")

    def predict(self, grid, phi, phi_prev, x_base, resid, resid_scale, x_out, extras=()):
        """x_out (solver rows) = x_base + the network's correction; resid = b - A x_base. One
        graph launch (dare_convnet_predict), the same result as pack -> forward -> unpack."""
        ne = len(extras)
        fptr = (ctypes.c_void_p * max(ne, 1))(*[_ptr(f) for f, _, _ in extras])
        sc = np.array([s for _, s, _ in extras] or [0.0], np.float64)
        of = np.array([o for _, _, o in extras] or [0.0], np.float64)
        check(_sig().dare_convnet_predict(self.ctx.h, self.h, ctypes.byref(grid), _ptr(phi), _ptr(phi_prev),
                                          _ptr(resid), float(resid_scale), ne, fptr, sc.ctypes.data, of.ctypes.data,
                                          _ptr(x_base), _ptr(x_out)), "dare_convnet_predict")

    def predict_eager(self, grid, phi, phi_prev, x_base, resid, resid_scale, x_out, extras=()):
        """predict as three separate calls (no graph)."""
        self.pack(grid, phi, phi_prev, resid, resid_scale, extras)
        self.forward(inference=True)
        self.unpack(grid, x_base, x_out)

    def loss(self, grid, x_base, x_true, readback=True, crop=None):
        v = ctypes.c_double()
        check(_sig().dare_convnet_loss(self.ctx.h, self.h, ctypes.byref(grid), _ptr(x_base), _ptr(x_true),
                                       ctypes.byref(v) if readback else None, self._crop(crop)), "dare_convnet_loss")
        return v.value if readback else None

    def backward(self):
        check(_sig().dare_conv_net_bwd(self.ctx.h, self.h), "dare_conv_net_bwd")

    def step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        check(_sig().dare_conv_net_step(self.ctx.h, self.h, lr, beta1, beta2, eps), "dare_conv_net_step")

    def train_step(self, grid, phi, phi_prev, x_base, resid, resid_scale, x_true, lr=1e-3, extras=(),
                   readback=True, crop=None):
        self.pack(grid, phi, phi_prev, resid, resid_scale, extras, crop)
        self.forward()
        v = self.loss(grid, x_base, x_true, readback, crop)
        self.backward()
        self.step(lr)
        return v

    def tensors(self):
        p = (ctypes.c_void_p * 19)()
        s = (ctypes.c_double * 3)()
        check(_sig().dare_convnet_tensors(self.h, p, s), "dare_convnet_tensors")
        names = ["x0", "out", "dout", "a1", "e1", "p1", "a3", "e2", "p2", "a5", "e3", "u2", "a7", "s2", "d2",
                 "u1", "a9", "s1", "d1"]
        return dict(zip(names, [v or 0 for v in p])), list(s)

    def close(self):
        if self.h:
            _sig().dare_convnet_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
