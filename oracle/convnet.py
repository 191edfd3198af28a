"""CPU oracle of the warm-start network (test infrastructure only).

A numpy restatement of the 3-level encoder-decoder built in dare_amd/csrc/convnet.hip: same
layers, same tap order, same pooling/upsampling/skip structure, forward and backward. With
`exact=False` every tensor the GPU stores in BF16 is rounded to BF16 at the same point
(round to nearest even) -- including the folded weights of the two decoder convs that read an
upsampled tensor, which both compute in sub-pixel form -- so GPU and oracle differ only by the
FP32 summation order of the matrix-core accumulations; with `exact=True` nothing is rounded (float64 throughout), which is
the mode the finite-difference check of the backward pass uses.

The network has no counterpart in the reference; what pins it is the mathematics: the
analytic gradients of this oracle are checked against central finite differences, and the
GPU kernels are checked against this oracle.
"""
import numpy as np

# (cin, cout, level): level 0 full resolution, 1 half, 2 quarter
LAYERS = [(8, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2),
          (64, 32, 1), (32, 32, 1), (32, 16, 0), (16, 16, 0), (16, 16, 0)]


def offsets():
    off = [0]
    for cin, cout, _ in LAYERS:
        off.append(off[-1] + 27 * cin * cout)
    return off


def split(flat):
    """Flat parameter vector -> list of [27, cout, cin] arrays."""
    off = offsets()
    return [flat[off[l]:off[l + 1]].reshape(27, c[1], c[0]) for l, c in enumerate(LAYERS)]


def bf16(a):
    """Round to BF16 (through float32, round to nearest even), returned as float64."""
    x = np.asarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(a):
    x = np.asarray(a, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _taps():
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                yield (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1), dz, dy, dx


def _shift(xp, dz, dy, dx, shape):
    D, H, W = shape
    return xp[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]


def conv(x, w):
    """y[v][co] = sum_tap sum_ci x[v + off(tap)][ci] w[tap][co][ci], zero padding."""
    D, H, W, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((D, H, W, w.shape[1]))
    for t, dz, dy, dx in _taps():
        y += _shift(xp, dz, dy, dx, (D, H, W)) @ w[t].T
    return y


def conv_t(g, w):
    """Data gradient of conv: dx[v][ci] = sum_tap sum_co g[v - off(tap)][co] w[tap][co][ci]."""
    D, H, W, _ = g.shape
    gp = np.pad(g, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dx = np.zeros((D, H, W, w.shape[2]))
    for t, dz, dy, dx_ in _taps():
        dx += _shift(gp, -dz, -dy, -dx_, (D, H, W)) @ w[t]
    return dx


def wgrad(x, g):
    """dw[tap][co][ci] = sum_v g[v][co] x[v + off(tap)][ci]."""
    D, H, W, cin = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    G = g.reshape(-1, g.shape[3])
    dw = np.zeros((27, g.shape[3], cin))
    for t, dz, dy, dx in _taps():
        dw[t] = G.T @ _shift(xp, dz, dy, dx, (D, H, W)).reshape(-1, cin)
    return dw


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def up(x):
    return x.repeat(2, 0).repeat(2, 1).repeat(2, 2)


def up_bwd(g):
    D, H, W, C = g.shape
    return g.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def relu(x):
    return np.maximum(x, 0.0)


def fold(w, rnd=lambda a: a):
    """Sub-pixel weights of a conv that reads the 2x nearest upsampling: ws[r][o] ([8, 8, cout,
    cin], r = 4 rz + 2 ry + rx output parity, o = 4 oz + 2 oy + ox coarse tap) is the sum of the
    taps d whose coarse offset floor((r + d) / 2) equals o - 1 + r on every axis; `rnd` rounds
    the sums (the GPU stores them in BF16)."""
    ws = np.zeros((8, 8) + w.shape[1:])
    for r in range(8):
        rr = ((r >> 2) & 1, (r >> 1) & 1, r & 1)
        for o in range(8):
            oo = ((o >> 2) & 1, (o >> 1) & 1, o & 1)
            for t, dz, dy, dx in _taps():
                if all((ra + d) // 2 == oa - 1 + ra for ra, oa, d in zip(rr, oo, (dz, dy, dx))):
                    ws[r, o] += w[t]
    return rnd(ws)


def conv_up(x, w, rnd=lambda a: a):
    """conv(up(x), w) in sub-pixel form: y[2p + r] = sum_o x[p + o - 1 + r] ws[r][o]^T over
    the 8 coarse taps of each output parity r (zero padding). Equal to conv(up(x), w) up to the
    rounding `rnd` applies to the folded weights."""
    D, H, W, _ = x.shape
    ws = fold(w, rnd)
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((2 * D, 2 * H, 2 * W, w.shape[1]))
    for r in range(8):
        rz, ry, rx = (r >> 2) & 1, (r >> 1) & 1, r & 1
        acc = np.zeros((D, H, W, w.shape[1]))
        for o in range(8):
            oz, oy, ox = (o >> 2) & 1, (o >> 1) & 1, o & 1
            acc += _shift(xp, oz - 1 + rz, oy - 1 + ry, ox - 1 + rx, (D, H, W)) @ ws[r, o].T
        y[rz::2, ry::2, rx::2] = acc
    return y


def forward(x0, weights, exact=False):
    """x0: [D, H, W, 8]. Returns (out [D, H, W], activations dict)."""
    r = (lambda a: a) if exact else bf16
    W = weights
    A = {"x0": x0}
    A["a1"] = r(relu(conv(x0, W[0])))
    A["e1"] = r(relu(conv(A["a1"], W[1])))
    A["p1"] = r(pool(A["e1"]))
    A["a3"] = r(relu(conv(A["p1"], W[2])))
    A["e2"] = r(relu(conv(A["a3"], W[3])))
    A["p2"] = r(pool(A["e2"]))
    A["a5"] = r(relu(conv(A["p2"], W[4])))
    A["e3"] = r(relu(conv(A["a5"], W[5])))
    # decoder convs over the upsampled tensors: the GPU computes them in sub-pixel form with
    # BF16-folded weights; exactly (float64) that form equals conv(up(x), w)
    cu = (lambda x, w: conv(up(x), w)) if exact else (lambda x, w: conv_up(x, w, bf16))
    A["u2"] = up(A["e3"])
    A["a7"] = r(relu(cu(A["e3"], W[6])))
    A["s2"] = r(A["a7"] + A["e2"])
    A["d2"] = r(relu(conv(A["s2"], W[7])))
    A["u1"] = up(A["d2"])
    A["a9"] = r(relu(cu(A["d2"], W[8])))
    A["s1"] = r(A["a9"] + A["e1"])
    A["d1"] = r(relu(conv(A["s1"], W[9])))
    out = conv(A["d1"], W[10])[..., 0]
    if not exact:
        out = out.astype(np.float32).astype(np.float64)
    return out, A


def backward(dout, A, weights, exact=False):
    """dout: d loss / d out [D, H, W]. Returns the list of weight gradients [27, cout, cin]."""
    r = (lambda a: a) if exact else bf16
    W = weights
    msk = lambda g, y: g * (y > 0)
    G = np.zeros(dout.shape + (16,))
    G[..., 0] = r(dout)
    dW = [None] * 11
    dW[10] = wgrad(A["d1"], G)
    g10 = msk(r(conv_t(G, W[10])), A["d1"])
    dW[9] = wgrad(A["s1"], g10)
    ds1 = r(conv_t(g10, W[9]))
    g9 = msk(ds1, A["a9"])
    dW[8] = wgrad(A["u1"], g9)
    du1 = r(conv_t(g9, W[8]))
    g8 = msk(r(up_bwd(du1)), A["d2"])
    dW[7] = wgrad(A["s2"], g8)
    ds2 = r(conv_t(g8, W[7]))
    g7 = msk(ds2, A["a7"])
    dW[6] = wgrad(A["u2"], g7)
    du2 = r(conv_t(g7, W[6]))
    g6 = msk(r(up_bwd(du2)), A["e3"])
    dW[5] = wgrad(A["a5"], g6)
    g5 = msk(r(conv_t(g6, W[5])), A["a5"])
    dW[4] = wgrad(A["p2"], g5)
    dp2 = r(conv_t(g5, W[4]))
    g4 = msk(r(ds2 + 0.125 * up(dp2)), A["e2"])
    dW[3] = wgrad(A["a3"], g4)
    g3 = msk(r(conv_t(g4, W[3])), A["a3"])
    dW[2] = wgrad(A["p1"], g3)
    dp1 = r(conv_t(g3, W[2]))
    g2 = msk(r(ds1 + 0.125 * up(dp1)), A["e1"])
    dW[1] = wgrad(A["a1"], g2)
    g1 = msk(r(conv_t(g2, W[1])), A["a1"])
    dW[0] = wgrad(A["x0"], g1)
    return dW


# ---------------------------------------------------------------- solver-side packing
def scales(phi_int, phim_int, resid_int, resid_scale):
    s0 = max(np.max(np.abs(phi_int)), 1e-300)
    s1 = max(np.max(np.abs(phi_int - phim_int)), 1e-300)
    s2 = max(np.max(np.abs(resid_int * resid_scale)), 1e-300)
    return s0, s1, s2


def pack(phi_int, phim_int, resid_int, resid_scale, extras=(), in_ch=8):
    """Input tensor [D, H, W, 8] as BF16 values (arrays of shape [D, H, W] over the box);
    resid = b - A phi^n, resid_scale maps it to field units."""
    s0, s1, s2 = scales(phi_int, phim_int, resid_int, resid_scale)
    x = np.zeros(phi_int.shape + (in_ch,))
    x[..., 0] = bf16(phi_int * (1.0 / s0))
    x[..., 1] = bf16((phi_int - phim_int) * (1.0 / s1))
    x[..., 2] = bf16(resid_int * resid_scale * (1.0 / s2))
    for e, (f, scale, offset) in enumerate(extras):
        x[..., 3 + e] = bf16((f - offset) * scale)
    return x, (s0, s1, s2)


def loss_and_dout(out, phi_int, x_true_int, s2, exact=False):
    """MSE of the output against the step in units of s2 (the output scale)."""
    t = (x_true_int - phi_int) * (1.0 / s2)
    d = out - t
    n = d.size
    g = 2.0 * d / n
    return float(np.mean(d * d)), (g if exact else bf16(g))


def _block_sum(a, cf=2):
    """Sum over cf^3 blocks of a [D, H, W] array, added in the order (a, b, c) of the block
    offsets along (z, y, x), c fastest (the device's order, so the doubles agree bit for bit)."""
    acc = np.zeros(tuple(m // cf for m in a.shape))
    for i in range(cf):
        for j in range(cf):
            for k in range(cf):
                acc = acc + a[i::cf, j::cf, k::cf]
    return acc


def pack_coarse(phi_int, phim_int, resid_int, resid_scale, extras=(), in_ch=8, cf=2):
    """Input tensor at 1/cf resolution [D/cf, H/cf, W/cf, 8]: each voxel the mean of the
    channel values (as in pack) of its cf^3 cells; the scales are those of the cells."""
    s0, s1, s2 = scales(phi_int, phim_int, resid_int, resid_scale)
    avg = 1.0 / cf ** 3
    x = np.zeros(tuple(m // cf for m in phi_int.shape) + (in_ch,))
    x[..., 0] = bf16(_block_sum(phi_int * (1.0 / s0), cf) * avg)
    x[..., 1] = bf16(_block_sum((phi_int - phim_int) * (1.0 / s1), cf) * avg)
    x[..., 2] = bf16(_block_sum(resid_int * resid_scale * (1.0 / s2), cf) * avg)
    for e, (f, scale, offset) in enumerate(extras):
        x[..., 3 + e] = bf16(_block_sum((f - offset) * scale, cf) * avg)
    return x, (s0, s1, s2)


def prolong_trilinear(out):
    """Cell values of a half-resolution field [D, H, W] -> [2D, 2H, 2W]: per axis the cell
    centre lies a quarter voxel from its voxel's centre, so it takes 3/4 of that voxel and 1/4
    of the neighbour on its side (the voxel itself again at the box faces); the eight products
    are added with the weight (wz wy) wx in the order of (z, y, x) choices, nearest first."""
    nc = out.shape
    idx = []
    for d in range(3):
        p = np.arange(2 * nc[d])
        v0 = p >> 1
        v1 = np.clip(v0 + np.where(p & 1, 1, -1), 0, nc[d] - 1)
        idx.append((v0, v1))
    res = np.zeros(tuple(2 * m for m in nc))
    for a in range(2):
        for b in range(2):
            for c in range(2):
                w = (0.25 if a else 0.75) * (0.25 if b else 0.75) * (0.25 if c else 0.75)
                z = idx[0][a][:, None, None]
                y = idx[1][b][None, :, None]
                x = idx[2][c][None, None, :]
                res = res + w * out[z, y, x]
    return res


def unpack_coarse(base_int, out, s2):
    """The guess at the cells: base + s2 * (trilinear interpolation of the voxel output)."""
    return base_int + s2 * prolong_trilinear(np.asarray(out, np.float64))


def loss_and_dout_coarse(out, base_int, x_true_int, s2, cf=2, exact=False):
    """MSE of the voxel output against the cf^3 mean of the step in units of s2."""
    t = _block_sum((x_true_int - base_int) * (1.0 / s2), cf) * (1.0 / cf ** 3)
    d = out - t
    g = 2.0 * d / d.size
    return float(np.mean(d * d)), (g if exact else bf16(g))


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1) in float32, in the order the device
    kernel evaluates it: m = b1 m + (1 - b1) g; v = b2 v + (1 - b2) g g;
    w -= lr (m / c1) / (sqrt(v / c2) + eps) with c1 = 1 - b1^t, c2 = 1 - b2^t.
    Returns the new (w, m, v)."""
    f = np.float32
    w, g, m, v = (np.asarray(a, np.float32) for a in (w, g, m, v))
    b1, b2, lr, eps = f(b1), f(b2), f(lr), f(eps)
    c1 = f(1) - np.power(b1, f(t), dtype=np.float32)
    c2 = f(1) - np.power(b2, f(t), dtype=np.float32)
    m = b1 * m + (f(1) - b1) * g
    v = b2 * v + (f(1) - b2) * g * g
    w = w - lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return w, m, v
