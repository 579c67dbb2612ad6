"""Pins the numpy restatement of the warm-start network (oracle/net_oracle.py):
its analytic gradients against central finite differences in FP64 (rounding
off), on an 8^3 grid, for entries of every layer; the convolution against a
direct triple loop; pooling / upsampling as adjoints of each other (their
backward passes are the transposes), and the BF16 rounding against the bit
pattern definition.  A planted mistake (a flipped tap in the backward data
pass) fails the finite-difference check."""
import numpy as np
import pytest

from oracle import net_oracle as no


def _problem(n=8, seed=0):
    rng = np.random.default_rng(seed)
    b = rng.normal(size=(n, n, n))
    flags = rng.integers(0, 3, size=(n, n, n)).astype(np.uint8)
    target = rng.normal(size=(n, n, n))
    return b, flags, target


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-30, 65504.0], np.float32)
    r = no.bf16(x)
    u = r.view(np.uint32)
    assert np.all(u & 0xFFFF == 0)
    assert r[0] == 1.0 and r[1] == 1.0  # 1 + 2^-8 is halfway: ties to even (mantissa 0)
    assert r[2] == np.float32(1.0078125)  # above halfway rounds up
    assert abs(r[3] - x[3]) <= 2 ** -8 * abs(x[3])


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(3, 4, 5, 2))
    W = rng.normal(size=(27, 2, 3))
    net = no.Net({}, exact=True)
    y = net.conv(x, W)
    want = np.zeros((3, 4, 5, 3))
    for z in range(3):
        for yy in range(4):
            for xx in range(5):
                for t in range(27):
                    dx, dy, dz = t % 3 - 1, t // 3 % 3 - 1, t // 9 - 1
                    zz, y2, x2 = z + dz, yy + dy, xx + dx
                    if 0 <= zz < 3 and 0 <= y2 < 4 and 0 <= x2 < 5:
                        want[z, yy, xx] += x[zz, y2, x2] @ W[t]
    assert np.allclose(y, want, atol=1e-12)
    # the backward data pass is the adjoint: <conv(x), g> = <x, conv_bwd_data(g)>
    g = rng.normal(size=y.shape)
    assert abs(np.sum(y * g) - np.sum(x * net.conv_bwd_data(g, W))) < 1e-9
    # and the weight gradient: <conv(x, W'), g> is linear in W' with gradient conv_wgrad(x, g)
    Wd = rng.normal(size=W.shape)
    assert abs(np.sum(net.conv(x, Wd) * g) - np.sum(Wd * net.conv_wgrad(x, g))) < 1e-9


def test_pool_up_adjoint():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4, 6, 8, 3))
    c = rng.normal(size=(2, 3, 4, 3))
    net = no.Net({}, exact=True)
    assert abs(np.sum(net.pool(x) * c) - np.sum(x * net.pool_bwd(c))) < 1e-12
    assert abs(np.sum(net.up(c) * x) - np.sum(c * net.up_bwd(x))) < 1e-12


def _loss(params, X0, target, fluid):
    return no.Net(params, exact=True).loss_and_grads(X0, target, fluid)[0]


@pytest.mark.parametrize("layer", ["enc0", "enc1", "enc2", "enc3", "dec2", "dec1", "dec0", "head"])
def test_gradients_finite_differences(layer):
    b, flags, target = _problem()
    params = {k: (np.asarray(w, np.float64), np.asarray(bb, np.float64)) for k, (w, bb) in
              no.random_params(3).items()}
    net = no.Net(params, exact=True)
    X0 = net.features(b, flags)
    fluid = flags == 1
    loss, g, _ = net.loss_and_grads(X0, target, fluid)
    rng = np.random.default_rng(4)
    W, bias = params[layer]
    gW, gb = g[layer]
    eps = 1e-6
    for _ in range(4):
        idx = tuple(rng.integers(0, s) for s in W.shape)
        Wp, Wm = W.copy(), W.copy()
        Wp[idx] += eps
        Wm[idx] -= eps
        num = (_loss({**params, layer: (Wp, bias)}, X0, target, fluid) -
               _loss({**params, layer: (Wm, bias)}, X0, target, fluid)) / (2 * eps)
        assert abs(num - gW[idx]) <= 1e-5 * max(1.0, abs(num)), (layer, idx, num, gW[idx])
    bp, bm = np.array(bias, copy=True), np.array(bias, copy=True)
    j = (0,) if np.ndim(bias) else ()
    if np.ndim(bias):
        bp[j] += eps
        bm[j] -= eps
    else:
        bp, bm = bias + eps, bias - eps
    num = (_loss({**params, layer: (W, bp)}, X0, target, fluid) -
           _loss({**params, layer: (W, bm)}, X0, target, fluid)) / (2 * eps)
    gbj = gb[j] if np.ndim(gb) else gb
    assert abs(num - gbj) <= 1e-5 * max(1.0, abs(num)), (layer, num, gbj)


def test_planted_flip_fails():
    b, flags, target = _problem()
    params = {k: (np.asarray(w, np.float64), np.asarray(bb, np.float64)) for k, (w, bb) in
              no.random_params(3).items()}

    class Wrong(no.Net):
        def conv_bwd_data(self, g, W):  # forgets to flip the taps
            acc = np.zeros(g.shape[:3] + (W.shape[1],), self.dt)
            for t, dx, dy, dz in self.taps():
                acc += self.shifted(g, dx, dy, dz) @ W[t].T.astype(self.dt)
            return acc

    net = Wrong(params, exact=True)
    X0 = net.features(b, flags)
    fluid = flags == 1
    _, g, _ = net.loss_and_grads(X0, target, fluid)
    W, bias = params["enc0"]
    idx = (13, 0, 3)
    eps = 1e-6
    Wp, Wm = W.copy(), W.copy()
    Wp[idx] += eps
    Wm[idx] -= eps
    num = (_loss({**params, "enc0": (Wp, bias)}, X0, target, fluid) -
           _loss({**params, "enc0": (Wm, bias)}, X0, target, fluid)) / (2 * eps)
    assert abs(num - g["enc0"][0][idx]) > 1e-5 * max(1.0, abs(num))


def _dense_A(flags, ih2):
    """Independent dense assembly of BuildSystem's matrix on FLUID rows (others 0)."""
    Z, Y, X = flags.shape
    N = flags.size
    A = np.zeros((N, N))
    idx = lambda i, j, k: i + X * (j + Y * k)
    for k in range(Z):
        for j in range(Y):
            for i in range(X):
                if flags[k, j, i] != 1:
                    continue
                r = idx(i, j, k)
                for (di, dj, dk, h2) in ((1, 0, 0, ih2[0]), (-1, 0, 0, ih2[0]), (0, 1, 0, ih2[1]),
                                         (0, -1, 0, ih2[1]), (0, 0, 1, ih2[2]), (0, 0, -1, ih2[2])):
                    ii, jj, kk = i + di, j + dj, k + dk
                    if not (0 <= ii < X and 0 <= jj < Y and 0 <= kk < Z) or flags[kk, jj, ii] == 2:
                        continue
                    A[r, r] += h2
                    if flags[kk, jj, ii] == 1:
                        A[r, idx(ii, jj, kk)] -= h2
    return A


def test_apply_A_matches_dense_assembly():
    rng = np.random.default_rng(6)
    flags = rng.choice([0, 1, 1, 1, 2], size=(4, 5, 6)).astype(np.uint8)
    ih2 = (36.0, 25.0, 16.0)
    p = rng.normal(size=flags.shape)
    A = _dense_A(flags, ih2)
    want = (A @ np.where(flags == 1, p, 0).reshape(-1)).reshape(flags.shape)
    assert np.allclose(no.apply_A(p, flags, ih2), want, atol=1e-12)


def test_energy_loss_gradients_and_meaning():
    """The energy loss is the relative A-norm error after the best scaling, so
    it is 0 for the exact solution (any scale) and its gradients match finite
    differences."""
    rng = np.random.default_rng(9)
    n = 8
    flags = rng.choice([0, 1, 1, 1, 2], size=(n, n, n)).astype(np.uint8)
    flags[0] = 0  # an AIR layer keeps the FLUID block nonsingular
    ih2 = (64.0, 64.0, 64.0)
    fluid = flags == 1
    x = np.where(fluid, rng.normal(size=flags.shape), 0)
    b = no.apply_A(x, flags, ih2)
    # the meaning: with the prediction replaced by 3x the solution the loss is 0
    p = 3.0 * x
    s = float((b * p)[fluid].sum())
    qq = float((p * no.apply_A(p, flags, ih2))[fluid].sum())
    assert abs(1 - s * s / (qq * float((b * x)[fluid].sum()))) < 1e-12
    params = {k: (np.asarray(w, np.float64), np.asarray(bb, np.float64)) for k, (w, bb) in
              no.random_params(5).items()}
    net = no.Net(params, exact=True)
    X0 = net.features(b, flags)
    kw = dict(loss_kind="energy", b=b, flags=flags, ih2=ih2)
    loss, g, _ = net.loss_and_grads(X0, x, fluid, **kw)
    assert 0.0 <= loss <= 1.0
    eps = 1e-6
    for layer, idx in (("dec0", (13, 5, 2)), ("enc2", (4, 7, 11)), ("head", (3,))):
        W, bias = params[layer]
        Wp, Wm = W.copy(), W.copy()
        Wp[idx] += eps
        Wm[idx] -= eps
        lp = no.Net({**params, layer: (Wp, bias)}, exact=True).loss_and_grads(X0, x, fluid, **kw)[0]
        lm = no.Net({**params, layer: (Wm, bias)}, exact=True).loss_and_grads(X0, x, fluid, **kw)[0]
        num = (lp - lm) / (2 * eps)
        assert abs(num - g[layer][0][idx]) <= 1e-5 * max(1e-3, abs(num)), (layer, num, g[layer][0][idx])
