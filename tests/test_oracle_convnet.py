"""The warm-start network's CPU oracle, pinned by the mathematics: its analytic weight
gradients agree with central finite differences of its loss (float64, no rounding), and its
building blocks satisfy their defining identities (the data gradient of a convolution is the
adjoint of the convolution; pooling and upsampling are adjoint up to the factor 8)."""
import numpy as np

from oracle import convnet as cn


def _weights(rng, in_ch=3):
    ws = []
    for l, (cin, cout, _) in enumerate(cn.LAYERS):
        fan = 27 * (in_ch if l == 0 else cin)
        w = rng.normal(0, np.sqrt(2.0 / fan), size=(27, cout, cin))
        if l == 0:
            w[:, :, in_ch:] = 0
        if l == 10:
            w[:, 1:, :] = 0
        ws.append(w)
    return ws


def test_conv_adjoint_and_resampling_identities():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(4, 6, 8, 16))
    g = rng.normal(size=(4, 6, 8, 32))
    w = rng.normal(size=(27, 32, 16))
    # <conv(x), g> = <x, conv_t(g)>
    assert np.isclose(np.sum(cn.conv(x, w) * g), np.sum(x * cn.conv_t(g, w)), rtol=1e-12)
    # <conv(x), g> = <w, wgrad(x, g)>
    assert np.isclose(np.sum(cn.conv(x, w) * g), np.sum(w * cn.wgrad(x, g)), rtol=1e-12)
    # centre tap only = a per-voxel matrix product
    wc = np.zeros_like(w)
    wc[13] = w[13]
    assert np.allclose(cn.conv(x, wc), x @ w[13].T)
    # pooling and upsampling: <pool(a), b> = <a, up(b)> / 8, up_bwd = 8 * adjoint of pool
    a = rng.normal(size=(4, 6, 8, 5))
    b = rng.normal(size=(2, 3, 4, 5))
    assert np.isclose(np.sum(cn.pool(a) * b), np.sum(a * cn.up(b)) / 8)
    assert np.isclose(np.sum(cn.up(b) * a), np.sum(b * cn.up_bwd(a)))


def test_subpixel_conv_equals_conv_of_upsampling(rng):
    """conv(up(x), w) = conv_up(x, w) in float64 (the sub-pixel form the GPU computes for the
    decoder convs): brute-force conv of the explicitly upsampled tensor on a ragged box, both
    channel shapes of the network's decoder layers. The fold puts every tap in exactly one
    coarse tap per parity, so the folded weights of each parity sum to the sum of all taps."""
    for cin, cout, shape in ((32, 16, (3, 2, 4)), (64, 32, (2, 3, 2))):
        x = rng.normal(size=shape + (cin,))
        w = rng.normal(size=(27, cout, cin))
        y0 = cn.conv(cn.up(x), w)
        y1 = cn.conv_up(x, w)
        assert np.max(np.abs(y1 - y0)) <= 1e-12 * np.max(np.abs(y0))
        ws = cn.fold(w)
        for r in range(8):
            assert np.allclose(ws[r].sum(axis=0), w.sum(axis=0), rtol=0, atol=1e-12)
    # one-hot input and weight: the single coarse voxel lands on exactly the fine voxels the
    # tap reaches through the upsampling
    x = np.zeros((2, 2, 2, 32))
    x[1, 0, 1, 3] = 1.0
    for t in (0, 4, 13, 22, 26):
        w = np.zeros((27, 16, 32))
        w[t, 5, 3] = 1.0
        np.testing.assert_array_equal(cn.conv_up(x, w), cn.conv(cn.up(x), w))


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    W = _weights(rng)
    x0 = np.zeros((4, 4, 8, 8))
    x0[..., :3] = rng.normal(size=(4, 4, 8, 3))
    t = rng.normal(size=(4, 4, 8))

    def loss(ws):
        out, _ = cn.forward(x0, ws, exact=True)
        return np.mean((out - t) ** 2)

    out, A = cn.forward(x0, W, exact=True)
    dout = 2 * (out - t) / out.size
    dW = cn.backward(dout, A, W, exact=True)
    checked = 0
    for l in range(11):
        cin, cout, _ = cn.LAYERS[l]
        for _ in range(3):
            tap = int(rng.integers(27))
            co = 0 if l == 10 else int(rng.integers(cout))
            ci = int(rng.integers(3 if l == 0 else cin))
            h = 1e-6
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][tap, co, ci] += h
            Wm[l][tap, co, ci] -= h
            fd = (loss(Wp) - loss(Wm)) / (2 * h)
            an = dW[l][tap, co, ci]
            # FD noise floor: eps * loss / h ~ 1e-10
            assert abs(fd - an) <= 1e-5 * abs(an) + 1e-9, (l, tap, co, ci, fd, an)
            checked += 1
    assert checked == 33


def test_bf16_rounding_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0])
    got = cn.bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -6 and got[3] == 1.0
    assert got[4] == -2.5 and got[5] == 0.0


def test_adam_oracle_first_step_and_moments():
    """The first Adam step moves every weight by lr g / (|g| + eps) (bias-corrected moments
    are g and g^2), and a constant gradient keeps every later step at that size; a zero
    gradient leaves the weight unchanged."""
    rng = np.random.default_rng(3)
    w0 = rng.normal(size=1000).astype(np.float32)
    g = rng.normal(size=1000).astype(np.float32)
    g[:10] = 0.0
    lr = 1e-3
    w1, m, v = cn.adam(w0, g, np.zeros_like(g), np.zeros_like(g), 1, lr)
    step = w0.astype(np.float64) - w1.astype(np.float64)
    exp = lr * g / (np.abs(g) + 1e-8)
    ulp = 2 * np.finfo(np.float32).eps * (np.abs(w0) + lr)     # rounding of the float32 update
    assert np.all(np.abs(step - exp) <= ulp)
    assert np.all(w1[:10] == w0[:10])
    w, mm, vv = w1, m, v
    for t in range(2, 6):
        w_new, mm, vv = cn.adam(w, g, mm, vv, t, lr)
        assert np.all(np.abs(w.astype(np.float64) - w_new.astype(np.float64) - exp) <= ulp + 1e-5 * lr)
        w = w_new


def _expand(c):
    """Each voxel repeated over its 2x2x2 cells."""
    return np.repeat(np.repeat(np.repeat(c, 2, 0), 2, 1), 2, 2)


def test_coarse_pack_of_block_constant_fields_is_the_pack_of_the_voxels(rng):
    """Fields constant over every 2x2x2 block (dyadic values, so every sum is exact): the
    half-resolution input is the full-resolution input of the voxel fields, bit for bit; and a
    non-constant block gives its mean (a swapped or dropped cell changes the mean)."""
    shape = (4, 6, 8)
    c = [rng.integers(-64, 64, size=shape) / 64.0 for _ in range(4)]
    f = [_expand(a) for a in c]
    xc, sc = cn.pack_coarse(f[0], f[0] - f[1], f[2], 0.5, extras=[(f[3], 2.0, 0.25)])
    xf, sf = cn.pack(c[0], c[0] - c[1], c[2], 0.5, extras=[(c[3], 2.0, 0.25)])
    assert sc == sf
    assert np.array_equal(xc, xf)
    phi = rng.integers(-64, 64, size=(8, 8, 8)) / 64.0
    x, (s0, _, _) = cn.pack_coarse(phi, phi, phi, 1.0)
    mean = phi.reshape(4, 2, 4, 2, 4, 2).mean(axis=(1, 3, 5))
    assert np.array_equal(x[..., 0], cn.bf16(mean / s0))


def test_trilinear_prolongation_reproduces_linear_fields():
    """Partition of unity, and exact reproduction of a linear field away from the box faces:
    the centre of cell z lies at voxel coordinate (z - 1/2) / 2. Swapping the 3/4 and 1/4
    weights, or taking the neighbour on the wrong side, breaks the linear case."""
    nc = (4, 5, 6)
    assert np.array_equal(cn.prolong_trilinear(np.full(nc, 3.5)), np.full(tuple(2 * m for m in nc), 3.5))
    Z, Y, X = np.meshgrid(*[np.arange(m, dtype=float) for m in nc], indexing="ij")
    out = 1.0 * Z + 2.0 * Y + 3.0 * X
    fine = cn.prolong_trilinear(out)
    z, y, x = np.meshgrid(*[(np.arange(2 * m) - 0.5) / 2 for m in nc], indexing="ij")
    exp = z + 2.0 * y + 3.0 * x
    inner = tuple(slice(1, 2 * m - 1) for m in nc)
    assert np.allclose(fine[inner], exp[inner], rtol=0, atol=1e-12)
    # at the faces the nearest voxel repeats: the first cell is 3/4 + 1/4 of voxel 0 along z
    assert np.allclose(fine[0, 1:-1, 1:-1] - exp[0, 1:-1, 1:-1], 0.25, atol=1e-12)
    base = np.ones(tuple(2 * m for m in nc))
    assert np.allclose(cn.unpack_coarse(base, out, 2.0), base + 2.0 * fine, rtol=0, atol=0)


def test_coarse_loss_is_the_loss_against_block_means(rng):
    """With a block-constant step the half-resolution loss is the full-resolution loss of the
    voxel fields; its gradient is the analytic derivative (central differences, float64)."""
    shape = (3, 4, 5)
    step = rng.integers(-64, 64, size=shape) / 64.0
    base = rng.integers(-64, 64, size=shape) / 64.0
    out = rng.normal(size=shape)
    lc, gc = cn.loss_and_dout_coarse(out, _expand(base), _expand(base + step), 0.5, exact=True)
    lf, gf = cn.loss_and_dout(out, base, base + step, 0.5, exact=True)
    assert np.isclose(lc, lf, rtol=1e-14) and np.allclose(gc, gf, rtol=1e-14)
    xb = rng.normal(size=(6, 8, 10))
    xt = rng.normal(size=(6, 8, 10))
    _, g = cn.loss_and_dout_coarse(out, xb, xt, 0.7, exact=True)
    h = 1e-6
    for idx in [(0, 0, 0), (2, 3, 4), (1, 2, 0)]:
        op, om = out.copy(), out.copy()
        op[idx] += h
        om[idx] -= h
        fd = (cn.loss_and_dout_coarse(op, xb, xt, 0.7, exact=True)[0]
              - cn.loss_and_dout_coarse(om, xb, xt, 0.7, exact=True)[0]) / (2 * h)
        assert np.isclose(fd, g[idx], rtol=1e-6, atol=1e-12)
