"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

numpy restatement of the pressure warm-start network run by the HIP library
(no counterpart in the reference: its solvers start from zero or from the
previous pressure, src/common/array3.hpp:189-197).  It is the checker for the
GPU's forward and backward kernels: the same architecture, the same BF16
rounding points (weights, stored activations and the backward operands) with
FP32 accumulation, so GPU outputs and gradients agree to a few BF16 ulps.  With
``exact=True`` nothing is rounded and everything is FP64, which is what the
finite-difference checks of the restatement itself use.

Architecture (4 levels, 16-64 channels, 3x3x3 convolutions with zero padding,
channel-last arrays (z, y, x, c)):
  X0 (16 ch: b / rms(b), FLUID, AIR, 13 zeros)
  E0 = relu(conv(X0, 16->16))            level 0
  E1 = relu(conv(pool(E0), 16->32))      level 1  (pool: 2x2x2 mean)
  E2 = relu(conv(pool(E1), 32->48))      level 2
  E3 = relu(conv(pool(E2), 48->64))      level 3
  D2 = relu(conv([up(E3), E2], 112->48)) level 2  (up: nearest, 2x per axis)
  D1 = relu(conv([up(D2), E1], 80->32))  level 1
  D0 = relu(conv([up(D1), E0], 48->16))  level 0 (kept in FP32 for the head)
  p  = D0 . hw + hb                      (1x1, FP32)
Losses: 'mse', the mean over FLUID cells of (p - target)^2; 'energy', the
A-norm error left after the best scaling of p, relative to the solution's:
1 - (b.p)^2 / ((p.A.p) (b.x)) with x the solution of A x = b (p zero outside
FLUID), which is the quantity CG reduces.
Tap t of a 3x3x3 kernel has offsets (dx, dy, dz) = (t % 3 - 1, t // 3 % 3 - 1, t // 9 - 1);
weights are W[t, ci, co].
"""
import numpy as np

LAYERS = [("enc0", 16, 16), ("enc1", 16, 32), ("enc2", 32, 48), ("enc3", 48, 64),
          ("dec2", 112, 48), ("dec1", 80, 32), ("dec0", 48, 16)]
IN_CH = 16


def bf16(a):
    """Round FP32 to the nearest BF16 (ties to even), returned as FP32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


class Net:
    def __init__(self, params, exact=False):
        """params: dict name -> (W (27, ci, co), b (co,)) for the convolutions,
        and 'head' -> (hw (16,), hb scalar)."""
        self.p = params
        self.exact = exact
        self.dt = np.float64 if exact else np.float32

    def r(self, a):
        return a.astype(self.dt) if self.exact else bf16(a)

    # -- building blocks --------------------------------------------------
    @staticmethod
    def taps():
        for t in range(27):
            yield t, t % 3 - 1, t // 3 % 3 - 1, t // 9 - 1

    @staticmethod
    def shifted(x, dx, dy, dz):
        """y[z, y, x] = x[z + dz, y + dy, x + dx], zero outside."""
        out = np.zeros_like(x)
        Z, Y, X = x.shape[:3]
        zs, ys, xs = slice(max(0, -dz), Z - max(0, dz)), slice(max(0, -dy), Y - max(0, dy)), slice(
            max(0, -dx), X - max(0, dx))
        zd, yd, xd = slice(max(0, dz), Z - max(0, -dz)), slice(max(0, dy), Y - max(0, -dy)), slice(
            max(0, dx), X - max(0, -dx))
        out[zs, ys, xs] = x[zd, yd, xd]
        return out

    def conv(self, x, W):
        acc = np.zeros(x.shape[:3] + (W.shape[2],), self.dt)
        for t, dx, dy, dz in self.taps():
            acc += self.shifted(x, dx, dy, dz) @ W[t].astype(self.dt)
        return acc

    def conv_bwd_data(self, g, W):
        acc = np.zeros(g.shape[:3] + (W.shape[1],), self.dt)
        for t, dx, dy, dz in self.taps():
            acc += self.shifted(g, -dx, -dy, -dz) @ W[t].T.astype(self.dt)
        return acc

    def conv_wgrad(self, x, g):
        W = np.zeros((27, x.shape[3], g.shape[3]), self.dt)
        gf = g.reshape(-1, g.shape[3])
        for t, dx, dy, dz in self.taps():
            W[t] = self.shifted(x, dx, dy, dz).reshape(-1, x.shape[3]).T @ gf
        return W

    def pool(self, x):
        Z, Y, X, C = x.shape
        return self.r(x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).astype(self.dt).sum(axis=(1, 3, 5)) / 8)

    @staticmethod
    def up(x):
        return x.repeat(2, 0).repeat(2, 1).repeat(2, 2)

    @staticmethod
    def pool_bwd(g):
        return Net.up(g) / 8

    @staticmethod
    def up_bwd(g):
        Z, Y, X, C = g.shape
        return g.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))

    def layer(self, name, x, keep=False):
        W, b = self.p[name]
        pre = self.conv(x, self.r(W)) + b.astype(self.dt)
        y = np.maximum(pre, 0)
        return y.astype(self.dt) if keep else self.r(y)

    # -- network ----------------------------------------------------------
    def features(self, b, flags, air_value=0):
        """b (z, y, x) float, flags uint8 (0 AIR, 1 FLUID, 2 BOUNDRY)."""
        rms = np.sqrt(np.mean(np.asarray(b, np.float64) ** 2))
        x = np.zeros(b.shape + (IN_CH,), np.float32)
        x[..., 0] = (np.asarray(b, np.float64) / (rms if rms > 0 else 1.0)).astype(np.float32)
        x[..., 1] = flags == 1
        x[..., 2] = flags == 0
        return self.r(x)

    def forward(self, X0):
        a = {"X0": X0}
        a["E0"] = self.layer("enc0", X0)
        a["P1"] = self.pool(a["E0"])
        a["E1"] = self.layer("enc1", a["P1"])
        a["P2"] = self.pool(a["E1"])
        a["E2"] = self.layer("enc2", a["P2"])
        a["P3"] = self.pool(a["E2"])
        a["E3"] = self.layer("enc3", a["P3"])
        a["C2"] = np.concatenate([self.up(a["E3"]), a["E2"]], axis=3)
        a["D2"] = self.layer("dec2", a["C2"])
        a["C1"] = np.concatenate([self.up(a["D2"]), a["E1"]], axis=3)
        a["D1"] = self.layer("dec1", a["C1"])
        a["C0"] = np.concatenate([self.up(a["D1"]), a["E0"]], axis=3)
        a["D0"] = self.layer("dec0", a["C0"], keep=True)
        hw, hb = self.p["head"]
        a["out"] = a["D0"].astype(self.dt) @ np.asarray(hw, self.dt) + self.dt(hb)
        return a

    def loss_and_grads(self, X0, target, fluid, loss_kind="mse", b=None, flags=None, ih2=None):
        """The loss (see the module text; 'energy' takes the right-hand side b,
        the markers and 1/h^2 per axis, and target = the solution x) and the
        gradient of every parameter (dict like params)."""
        a = self.forward(X0)
        nf = max(int(fluid.sum()), 1)
        if loss_kind == "mse":
            diff = np.where(fluid, a["out"] - target.astype(self.dt), 0)
            loss = float((diff.astype(np.float64) ** 2).sum() / nf)
            dout = (2.0 * diff / nf).astype(self.dt)
        else:
            p = np.where(fluid, a["out"], 0).astype(np.float64)
            q = apply_A(p, flags, ih2)
            bb = np.asarray(b, np.float64)
            s = float((bb * p)[fluid].sum())
            qq = float((p * q)[fluid].sum())
            E = float((bb * np.asarray(target, np.float64))[fluid].sum())
            loss = 1.0 - s * s / (qq * E)
            coef = 2.0 * s / (qq * E)
            dout = np.where(fluid, coef * (s / qq * q - bb), 0).astype(self.dt)
        g = {}
        hw, hb = self.p["head"]
        g["head"] = ((a["D0"].astype(self.dt) * dout[..., None]).reshape(-1, 16).sum(0), dout.sum())
        dD0 = dout[..., None] * np.asarray(hw, self.dt)

        def back(name, x, y, dy):
            dpre = self.r(np.where(y > 0, dy, 0))
            W, _ = self.p[name]
            g[name] = (self.conv_wgrad(x, dpre), dpre.reshape(-1, dpre.shape[3]).sum(0))
            return self.conv_bwd_data(dpre, self.r(W))

        dC0 = back("dec0", a["C0"], a["D0"], dD0)
        dD1 = self.up_bwd(dC0[..., :32])
        dE0 = dC0[..., 32:].copy()
        dC1 = back("dec1", a["C1"], a["D1"], dD1)
        dD2 = self.up_bwd(dC1[..., :48])
        dE1 = dC1[..., 48:].copy()
        dC2 = back("dec2", a["C2"], a["D2"], dD2)
        dE3 = self.up_bwd(dC2[..., :64])
        dE2 = dC2[..., 64:].copy()
        dP3 = back("enc3", a["P3"], a["E3"], dE3)
        dE2 = dE2 + self.pool_bwd(dP3)
        dP2 = back("enc2", a["P2"], a["E2"], dE2)
        dE1 = dE1 + self.pool_bwd(dP2)
        dP1 = back("enc1", a["P1"], a["E1"], dE1)
        dE0 = dE0 + self.pool_bwd(dP1)
        back("enc0", a["X0"], a["E0"], dE0)
        return loss, g, a


def random_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co in LAYERS:
        p[name] = ((rng.normal(size=(27, ci, co)) * np.sqrt(2.0 / (27 * ci)) * scale).astype(np.float32),
                   (rng.normal(size=co) * 0.05).astype(np.float32))
    p["head"] = ((rng.normal(size=16) * 0.25).astype(np.float32), np.float32(0.01))
    return p


def apply_A(p, flags, ih2):
    """The pressure matrix of BuildSystem (src/fluid_solvers/single_phase_pressure_solver.cpp:99-164)
    applied to p on FLUID rows (0 elsewhere): the diagonal sums 1/h_a^2 over in-range
    non-BOUNDRY neighbours, FLUID neighbours couple with -1/h_a^2; p is taken as 0
    outside FLUID.  ih2 = (1/hx^2, 1/hy^2, 1/hz^2); arrays are (z, y, x)."""
    fluid = flags == 1
    p = np.where(fluid, np.asarray(p, np.float64), 0.0)
    out = np.zeros_like(p)
    for axis, h2 in ((2, ih2[0]), (1, ih2[1]), (0, ih2[2])):
        n = p.shape[axis]
        for step in (1, -1):
            nb_flags = np.full(p.shape, 255, np.uint8)  # 255: out of range
            nb_p = np.zeros_like(p)
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if step == 1:
                dst[axis], src[axis] = slice(0, n - 1), slice(1, n)
            else:
                dst[axis], src[axis] = slice(1, n), slice(0, n - 1)
            nb_flags[tuple(dst)] = flags[tuple(src)]
            nb_p[tuple(dst)] = p[tuple(src)]
            counted = (nb_flags != 255) & (nb_flags != 2)
            out += np.where(counted, h2 * p, 0.0)
            out -= np.where(nb_flags == 1, h2 * nb_p, 0.0)
    return np.where(fluid, out, 0.0)
