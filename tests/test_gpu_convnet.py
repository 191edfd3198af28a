"""Warm-start network kernels (csrc/convnet.hip) against the CPU oracle (oracle/convnet.py).

Packing of the solver fields is bit-exact (same double arithmetic, same BF16 rounding).
Forward activations and weight gradients agree with the oracle run with BF16 rounding at the
same storage points; the remaining difference is the FP32 summation order of the matrix-core
accumulations (and BF16 rounding flips it causes), bounded here in relative L2 norm.
"""
import ctypes

import numpy as np
import pytest

import dare_amd as da
from dare_amd.learned import WarmStartNet, he_init
from oracle import convnet as cn

pytestmark = pytest.mark.gpu


def d2h_bf16(ctx, addr, shape):
    n = int(np.prod(shape))
    u = np.empty(n, np.uint16)
    da.check(da.lib().dare_memcpy_d2h(ctx.h, u.ctypes.data, addr, 2 * n), "d2h")
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64).reshape(shape)


def d2h_f32(ctx, addr, shape):
    n = int(np.prod(shape))
    a = np.empty(n, np.float32)
    da.check(da.lib().dare_memcpy_d2h(ctx.h, a.ctypes.data, addr, 4 * n), "d2h")
    return a.astype(np.float64).reshape(shape)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fields(rng, n, ghost=2):
    import torch
    pad = [m + 2 * ghost for m in n]
    phi = rng.normal(size=pad)
    phim = phi + 0.1 * rng.normal(size=pad)
    rhs = rng.normal(size=n)
    ext = rng.uniform(size=pad)
    to = lambda a: torch.from_numpy(np.ascontiguousarray(a).ravel()).cuda()
    sl = tuple(slice(ghost, ghost + m) for m in n)
    return (to(phi), to(phim), to(rhs), to(ext)), (phi[sl], phim[sl], rhs, ext[sl])


N = [8, 12, 16]


def setup(ctx, rng, in_ch=4):
    net = WarmStartNet(ctx, N, in_channels=in_ch, seed=3)
    w = he_init(in_ch, seed=3, out_scale=1.0)
    net.set_params(w)
    g = da.Grid.make(N, size=[1, 1, 1], ghost=2)
    dev, hostf = fields(rng, N)
    extras = [(dev[3], 2.0, 0.5)]
    net.pack(g, dev[0], dev[1], dev[2], 0.7, extras)
    return net, g, dev, hostf, w, extras


def test_pack_bit_exact(ctx, rng):
    net, g, dev, (phi, phim, rhs, ext), w, _ = setup(ctx, rng)
    ptr, scales = net.tensors()
    x0 = d2h_bf16(ctx, ptr["x0"], N + [8])
    exp, s = cn.pack(phi, phim, rhs.reshape(N), 0.7, extras=[(ext, 2.0, 0.5)])
    assert np.allclose(scales, s, rtol=0, atol=0)
    assert np.array_equal(x0, exp)
    net.close()


def test_pack_crop_matches_oracle(ctx, rng):
    """A crop packs the sub-box: same values as the oracle on the sliced fields."""
    net = WarmStartNet(ctx, N, in_channels=3, seed=3)
    big = [16, 16, 20]
    g = da.Grid.make(big, size=[1, 1, 1], ghost=2)
    dev, (phi, phim, rhs, ext) = fields(rng, big)
    o = [4, 2, 3]
    net.pack(g, dev[0], dev[1], dev[2], 0.7, crop=o + N)
    ptr, scales = net.tensors()
    x0 = d2h_bf16(ctx, ptr["x0"], N + [8])
    sl = tuple(slice(o[d], o[d] + N[d]) for d in range(3))
    exp, s = cn.pack(phi[sl], phim[sl], rhs.reshape(big)[sl], 0.7)
    assert np.allclose(scales, s, rtol=0, atol=0)
    assert np.array_equal(x0, exp)
    net.close()


def test_forward_matches_oracle(ctx, rng):
    net, g, dev, hostf, w, _ = setup(ctx, rng)
    net.forward()
    ctx.synchronize()
    ptr, _ = net.tensors()
    x0 = d2h_bf16(ctx, ptr["x0"], N + [8])
    out_ref, A = cn.forward(x0, cn.split(cn.bf16(w)), exact=False)
    shapes = {"a1": (0, 16), "e1": (0, 16), "p1": (1, 16), "a3": (1, 32), "e2": (1, 32), "p2": (2, 32),
              "a5": (2, 64), "e3": (2, 64), "a7": (1, 32), "s2": (1, 32), "d2": (1, 32), "a9": (0, 16),
              "s1": (0, 16), "d1": (0, 16)}
    for name, (lv, c) in shapes.items():
        got = d2h_bf16(ctx, ptr[name], [m >> lv for m in N] + [c])
        assert rel(got, A[name]) <= 1e-2, (name, rel(got, A[name]))
    out = d2h_f32(ctx, ptr["out"], N)
    assert rel(out, out_ref) <= 1e-2, rel(out, out_ref)
    net.close()


def test_forward_wide_box_matches_oracle(ctx, rng):
    """A box wide enough for every tiling of the conv kernels: 64- and 32-voxel x tiles, and
    z-marching blocks that carry their plane ring over several planes (16 x 64 x 256: 64
    tile columns at full resolution, so the resident grid splits z into blocks of two planes)."""
    import torch
    nb = [16, 64, 256]
    net = WarmStartNet(ctx, nb, in_channels=3, seed=5)
    w = he_init(3, seed=5, out_scale=1.0)
    net.set_params(w)
    g = da.Grid.make(nb, size=[1, 1, 1], ghost=2)
    dev, hostf = fields(rng, nb)
    net.pack(g, dev[0], dev[1], dev[2], 0.7)
    net.forward()
    ctx.synchronize()
    ptr, _ = net.tensors()
    x0 = d2h_bf16(ctx, ptr["x0"], nb + [8])
    out_ref, A = cn.forward(x0, cn.split(cn.bf16(w)), exact=False)
    shapes = {"e1": (0, 16), "e2": (1, 32), "e3": (2, 64), "s2": (1, 32), "s1": (0, 16), "d1": (0, 16)}
    for name, (lv, c) in shapes.items():
        got = d2h_bf16(ctx, ptr[name], [m >> lv for m in nb] + [c])
        assert rel(got, A[name]) <= 1e-2, (name, rel(got, A[name]))
    out = d2h_f32(ctx, ptr["out"], nb)
    assert rel(out, out_ref) <= 1e-2, rel(out, out_ref)
    net.close()


def test_inference_mode_skips_only_the_backward_masks(ctx, rng):
    """Inference mode: bit-identical output and skip sums, a7 / a9 left untouched (a sentinel
    survives), and the backward pass refused until a forward pass with the mode off."""
    import torch
    nb = [16, 64, 256]                      # both sub-pixel tilings of the decoder convs
    net = WarmStartNet(ctx, nb, in_channels=3, seed=5)
    g = da.Grid.make(nb, size=[1, 1, 1], ghost=2)
    dev, _ = fields(rng, nb)
    net.pack(g, dev[0], dev[1], dev[2], 0.7)
    net.forward()
    ctx.synchronize()
    ptr, _ = net.tensors()
    sizes = {"a7": int(np.prod(nb)) // 8 * 32, "a9": int(np.prod(nb)) * 16}

    def raw(name, n):
        u = np.empty(n, np.uint16)
        da.check(da.lib().dare_memcpy_d2h(ctx.h, u.ctypes.data, ptr[name], 2 * n), "d2h")
        return u
    ref = {k: raw(k, n) for k, n in [("s2", sizes["a7"]), ("s1", sizes["a9"]), ("a9", sizes["a9"])]}
    ref_out = raw("out", 2 * int(np.prod(nb)))
    for name, n in sizes.items():
        sentinel = np.full(n, 0x7fc1, np.uint16)
        da.check(da.lib().dare_memcpy_h2d(ctx.h, ptr[name], sentinel.ctypes.data, 2 * n), "h2d")
    net.forward(inference=True)
    ctx.synchronize()
    assert np.array_equal(raw("out", 2 * int(np.prod(nb))), ref_out)
    assert np.array_equal(raw("s2", sizes["a7"]), ref["s2"])
    assert np.array_equal(raw("s1", sizes["a9"]), ref["s1"])
    for name, n in sizes.items():
        assert np.all(raw(name, n) == 0x7fc1), name
    x_true = torch.from_numpy(rng.normal(size=int(np.prod(nb)))).cuda()
    base = dev[0].view(*[m + 4 for m in nb])[2:-2, 2:-2, 2:-2].contiguous().view(-1)
    net.loss(g, base, x_true)
    with pytest.raises(da.DareError) as e:
        net.backward()
    assert "outside inference mode" in str(e.value)
    net.forward()
    ctx.synchronize()
    assert np.array_equal(raw("a9", sizes["a9"]), ref["a9"])
    net.loss(g, base, x_true)
    net.backward()
    net.close()


def test_backward_matches_oracle(ctx, rng):
    import torch
    net, g, dev, (phi, phim, rhs, ext), w, _ = setup(ctx, rng)
    net.forward()
    x_true = torch.from_numpy(rng.normal(size=int(np.prod(N)))).cuda()
    base = dev[0].view(*[m + 4 for m in N])[2:-2, 2:-2, 2:-2].contiguous().view(-1)
    loss = net.loss(g, base, x_true)
    net.backward()
    grads = net.grads()
    ptr, scales = net.tensors()
    x0 = d2h_bf16(ctx, ptr["x0"], N + [8])
    Wb = cn.split(cn.bf16(w))
    out_ref, A = cn.forward(x0, Wb, exact=False)
    out_gpu = d2h_f32(ctx, ptr["out"], N)
    loss_ref, dout = cn.loss_and_dout(out_gpu, phi, x_true.cpu().numpy().reshape(N), scales[2])
    assert abs(loss - loss_ref) <= 1e-9 * abs(loss_ref)
    dW = cn.backward(dout, A, Wb, exact=False)
    G = cn.split(grads.astype(np.float64))
    for l in range(11):
        e = rel(G[l], dW[l])
        assert e <= 3e-2, (l, e)
    net.close()


def test_training_reduces_loss(ctx, rng):
    """Adam on one fixed sample: the loss falls to a fraction of its start."""
    import torch
    net = WarmStartNet(ctx, N, in_channels=3, seed=5)
    g = da.Grid.make(N, size=[1, 1, 1], ghost=2)
    dev, (phi, phim, rhs, ext) = fields(rng, N)
    # a target that is a smooth function of the inputs
    x_true = dev[0].view(*[m + 4 for m in N])[2:-2, 2:-2, 2:-2].contiguous().view(-1) * 1.05
    base = dev[0].view(*[m + 4 for m in N])[2:-2, 2:-2, 2:-2].contiguous().view(-1)
    losses = [net.train_step(g, dev[0], dev[1], base, dev[2], 0.7, x_true, lr=3e-3) for _ in range(120)]
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-10:]) < 0.25 * np.mean(losses[:10]), (losses[:10], losses[-10:])
    # the guess is phi^n + S * out, S = max|resid * resid_scale|
    xg = torch.empty_like(x_true)
    net.predict(g, dev[0], dev[1], base, dev[2], 0.7, xg)
    ptr, s = net.tensors()
    out = d2h_f32(ctx, ptr["out"], N)
    assert np.allclose(xg.cpu().numpy().reshape(N), phi + s[2] * out, rtol=1e-14, atol=1e-14)
    net.close()


def test_save_load_roundtrip(ctx, tmp_path):
    net = WarmStartNet(ctx, N, in_channels=3, seed=9)
    p = str(tmp_path / "w.npz")
    net.save(p, meta={"note": "test"})
    w, meta = WarmStartNet.load_params(p)
    assert np.array_equal(w, net.params()) and meta["in_channels"] == 3
    net2 = WarmStartNet(ctx, N, in_channels=3, params=w)
    assert np.array_equal(net2.params(), w)
    net.close()
    net2.close()


def test_adam_step_matches_oracle(ctx, rng):
    """dare_conv_net_step against the oracle's float32 Adam (oracle/convnet.py adam) on two
    consecutive steps with the gradients the backward pass produced: the FP32 master weights
    agree to one float32 rounding."""
    import torch
    net, g, dev, (phi, phim, rhs, ext), w, extras = setup(ctx, rng)
    base = dev[0].view(*[m + 4 for m in N])[2:-2, 2:-2, 2:-2].contiguous().view(-1)
    x_true = torch.from_numpy(rng.normal(size=int(np.prod(N)))).cuda()
    w0 = net.params()
    m = np.zeros_like(w0)
    v = np.zeros_like(w0)
    for t in (1, 2):
        net.pack(g, dev[0], dev[1], dev[2], 0.7, extras)
        net.forward()
        net.loss(g, base, x_true)
        net.backward()
        gr = net.grads()
        net.step(lr=2e-3, beta1=0.9, beta2=0.999, eps=1e-8)
        w1 = net.params()
        we, m, v = cn.adam(w0, gr, m, v, t, 2e-3, 0.9, 0.999, 1e-8)
        tol = 2 * np.finfo(np.float32).eps * (np.abs(we.astype(np.float64)) + 2e-3)
        assert np.all(np.abs(w1.astype(np.float64) - we.astype(np.float64)) <= tol), t
        assert np.any(w1 != w0)
        w0 = w1
    net.close()


NC = [16, 24, 32]       # a box whose halves are multiples of 4


def test_coarse_pack_unpack_loss_match_oracle(ctx, rng):
    """Half-resolution mode (coarsening 2): the packed input, the interpolated guess and the
    loss gradient are bit-exact against the oracle (oracle/convnet.py pack_coarse,
    unpack_coarse, loss_and_dout_coarse); the loss value agrees to rounding."""
    import torch
    net = WarmStartNet(ctx, NC, in_channels=4, seed=3, coarsen=2)
    net.set_params(he_init(4, seed=3, out_scale=1.0))
    g = da.Grid.make(NC, size=[1, 1, 1], ghost=2)
    dev, (phi, phim, rhs, ext) = fields(rng, NC)
    extras = [(dev[3], 2.0, 0.5)]
    net.pack(g, dev[0], dev[1], dev[2], 0.7, extras)
    ptr, scales = net.tensors()
    half = [m // 2 for m in NC]
    x0 = d2h_bf16(ctx, ptr["x0"], half + [8])
    exp, s = cn.pack_coarse(phi, phim, rhs.reshape(NC), 0.7, extras=[(ext, 2.0, 0.5)])
    assert np.allclose(scales, s, rtol=0, atol=0)
    assert np.array_equal(x0, exp)
    net.forward()
    base = dev[0].view(*[m + 4 for m in NC])[2:-2, 2:-2, 2:-2].contiguous().view(-1)
    xg = torch.empty_like(base)
    net.unpack(g, base, xg)
    out = d2h_f32(ctx, ptr["out"], half)
    assert np.array_equal(xg.cpu().numpy().reshape(NC), cn.unpack_coarse(phi, out, scales[2]))
    x_true = torch.from_numpy(rng.normal(size=int(np.prod(NC)))).cuda()
    loss = net.loss(g, base, x_true)
    loss_ref, dout = cn.loss_and_dout_coarse(out, phi, x_true.cpu().numpy().reshape(NC), scales[2])
    assert abs(loss - loss_ref) <= 1e-12 * abs(loss_ref)
    ptr, _ = net.tensors()
    got = d2h_bf16(ctx, ptr["dout"], half + [16])
    assert np.array_equal(got[..., 0], dout) and not np.any(got[..., 1:])
    net.close()


def test_coarse_mode_refuses_boxes_that_do_not_halve(ctx):
    net = WarmStartNet(ctx, [16, 16, 16], in_channels=3, coarsen=2)
    g = da.Grid.make([12, 16, 16], size=[1, 1, 1], ghost=2)
    import torch
    z = torch.zeros(16 * 20 * 20, dtype=torch.float64, device="cuda")
    r = torch.zeros(12 * 16 * 16, dtype=torch.float64, device="cuda")
    with pytest.raises(da.DareError, match="multiples of 8"):
        net.pack(g, z, z, r, 1.0)
    net.close()


@pytest.mark.parametrize("coarsen", [1, 2])
def test_graph_predict_equals_eager_on_every_replay(ctx, rng, coarsen):
    """dare_convnet_predict (one captured graph, replayed) against pack -> forward -> unpack
    called separately: new data in the same buffers before every replay, eager work and
    synchronisations between replays; every replay bit-identical to the eager result. A new
    pointer re-captures."""
    import torch
    net = WarmStartNet(ctx, NC, in_channels=4, seed=3, coarsen=coarsen)
    net.set_params(he_init(4, seed=3, out_scale=1.0))
    g = da.Grid.make(NC, size=[1, 1, 1], ghost=2)
    dev, _ = fields(rng, NC)
    extras = [(dev[3], 2.0, 0.5)]
    base = torch.empty(int(np.prod(NC)), dtype=torch.float64, device="cuda")
    xg = torch.empty_like(base)
    xe = torch.empty_like(base)
    for rep in range(4):
        for t in dev:
            t.copy_(torch.from_numpy(rng.normal(size=t.numel())))
        base.copy_(torch.from_numpy(rng.normal(size=base.numel())))
        net.predict(g, dev[0], dev[1], base, dev[2], 0.7, xg, extras=extras)
        ctx.synchronize()
        net.predict_eager(g, dev[0], dev[1], base, dev[2], 0.7, xe, extras=extras)
        ctx.synchronize()
        assert torch.equal(xg, xe), rep
    other = torch.empty_like(base)
    net.predict(g, dev[0], dev[1], base, dev[2], 0.7, other, extras=extras)
    ctx.synchronize()
    assert torch.equal(other, xe)
    # new weights after the capture (set_params, then an Adam step): the replayed graph uses
    # them, the sub-pixel weights of the decoder included (folded outside the graph)
    for update in ("set_params", "step"):
        if update == "set_params":
            net.set_params(he_init(4, seed=11, out_scale=1.0))
        else:
            net.train_step(g, dev[0], dev[1], base, dev[2], 0.7, base, lr=1e-2, extras=extras)
        net.predict(g, dev[0], dev[1], base, dev[2], 0.7, xg, extras=extras)
        ctx.synchronize()
        net.predict_eager(g, dev[0], dev[1], base, dev[2], 0.7, xe, extras=extras)
        ctx.synchronize()
        assert torch.equal(xg, xe), update
        assert not torch.equal(xe, other), update
        other.copy_(xe)
    net.close()
