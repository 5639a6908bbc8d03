"""CPU checker of the learned surrogate action (test infrastructure only).

The surrogate has no counterpart in the reference; this numpy restatement
pins the GPU kernels, and its own derivatives are pinned by finite
differences in tests/test_surrogate_ref.py.

S_theta(U) = sum_x sum_o w3[o] h2[o](x)
  h2 = tanh(b2 + sum_t W2[t] h1(x + d_t)),  h1 = tanh(b1 + sum_t W1[t] phi(x + d_t))
over the 11 taps d_t = 0, +e_0, -e_0, ..., +e_4, -e_4 of the A4* lattice
(e_4 = (-1,-1,-1,-1)), and 65 gauge-invariant features per site:
  phi[2k], phi[2k+1]     Re, Im Tr P_ab(x) / N   for the 10 plaquettes a < b,
      P_ab(x) = U_a(x) U_b(x+a) U_a(x+b)^dag U_b(x)^dag
  phi[20+2k], [21+2k]    Re, Im Tr R_ab(x) / N   for the 20 rectangles a != b,
      R_ab(x) = U_a(x) U_a(x+a) U_b(x+2a) U_a(x+a+b)^dag U_a(x+b)^dag U_b(x)^dag
  phi[60+a]              Tr(U_a U_a^dag)(x) / N - 1.
The force follows the reference's convention (mom -= eps F, U += eps mom):
F = dS/dUbar, so dS = 2 Re sum conj(F) dU.
"""
import numpy as np

NUMLINK = 5
NTAP = 11
NCH = 16
NFEAT = 65
EV = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [-1, -1, -1, -1]])


def bf16(a):
    """round float32 values to the nearest bfloat16 (ties to even)"""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


class Geometry:
    """sites in the reference's even-first order; shift tables by vectors"""

    def __init__(self, dims):
        self.dims = tuple(dims)
        nx, ny, nz, nt = dims
        self.V = V = nx * ny * nz * nt
        c = np.zeros((V, 4), dtype=np.int64)
        for t in range(nt):
            for z in range(nz):
                for y in range(ny):
                    for x in range(nx):
                        lex = x + nx * (y + ny * (z + nz * t))
                        i = (lex + V) // 2 if (x + y + z + t) % 2 else lex // 2
                        c[i] = (x, y, z, t)
        self.coords = c
        self._cache = {}

    def index(self, c):
        nx, ny, nz, nt = self.dims
        c = np.mod(c, self.dims)
        lex = c[:, 0] + nx * (c[:, 1] + ny * (c[:, 2] + nz * c[:, 3]))
        odd = c.sum(axis=1) % 2 == 1
        return np.where(odd, (lex + self.V) // 2, lex // 2)

    def shift(self, vec):
        key = tuple(int(v) for v in vec)
        if key not in self._cache:
            self._cache[key] = self.index(self.coords + np.array(key))
        return self._cache[key]

    def tap(self, t):
        """site x + d_t for every x"""
        if t == 0:
            return np.arange(self.V)
        a, sg = (t - 1) // 2, (1 if t % 2 == 1 else -1)
        return self.shift(sg * EV[a])


def opposite_tap(t):
    return 0 if t == 0 else (t + 1 if t % 2 == 1 else t - 1)


def loops():
    """the 30 loops: lists of (link, offset vector, dagger)"""
    out = []
    for a in range(NUMLINK):
        for b in range(a + 1, NUMLINK):
            ea, eb = EV[a], EV[b]
            out.append([(a, 0 * ea, False), (b, ea, False), (a, eb, True), (b, 0 * ea, True)])
    for a in range(NUMLINK):
        for b in range(NUMLINK):
            if a == b:
                continue
            ea, eb = EV[a], EV[b]
            out.append([(a, 0 * ea, False), (a, ea, False), (b, 2 * ea, False),
                        (a, ea + eb, True), (a, eb, True), (b, 0 * ea, True)])
    return out


LOOPS = loops()


def dag(m):
    return np.conj(np.swapaxes(m, -1, -2))


def _factor(geo, U, f, x0=None):
    a, off, d = f
    idx = geo.shift(off)
    m = U[idx, a]
    return dag(m) if d else m


def _chain(ms):
    out = ms[0]
    for m in ms[1:]:
        out = out @ m
    return out


def features(geo, U):
    N = U.shape[-1]
    phi = np.zeros((geo.V, NFEAT))
    for k, lp in enumerate(LOOPS):
        tr = np.trace(_chain([_factor(geo, U, f) for f in lp]), axis1=-2, axis2=-1) / N
        phi[:, 2 * k] = tr.real
        phi[:, 2 * k + 1] = tr.imag
    for a in range(NUMLINK):
        phi[:, 60 + a] = np.sum(np.abs(U[:, a]) ** 2, axis=(-2, -1)) / N - 1.0
    return phi


def features_jvp(geo, U, R):
    """d/dt phi(U + t R) at t = 0"""
    N = U.shape[-1]
    v = np.zeros((geo.V, NFEAT))
    for k, lp in enumerate(LOOPS):
        ms = [_factor(geo, U, f) for f in lp]
        rs = [_factor(geo, R, f) for f in lp]
        tot = 0
        for m in range(len(lp)):
            tot = tot + np.trace(_chain(ms[:m] + [rs[m]] + ms[m + 1:]), axis1=-2, axis2=-1)
        tot = tot / N
        v[:, 2 * k] = tot.real
        v[:, 2 * k + 1] = tot.imag
    for a in range(NUMLINK):
        v[:, 60 + a] = 2 * np.sum((np.conj(U[:, a]) * R[:, a]).real, axis=(-2, -1)) / N
    return v


def init_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return {"W1": (rng.standard_normal((NTAP, NCH, NFEAT)) * scale / np.sqrt(NTAP * NFEAT)),
            "b1": np.zeros(NCH),
            "W2": (rng.standard_normal((NTAP, NCH, NCH)) * scale / np.sqrt(NTAP * NCH)),
            "b2": np.zeros(NCH),
            "w3": rng.standard_normal(NCH) * scale / np.sqrt(NCH)}


def weights_used(th, round_bf16=True):
    """the weights the kernels compute with: W1, W2, w3 rounded to bfloat16"""
    if not round_bf16:
        return {k: np.asarray(v, dtype=np.float64) for k, v in th.items()}
    out = dict(th)
    for k in ("W1", "W2", "w3"):
        out[k] = bf16(th[k]).astype(np.float64)
    out["b1"] = np.asarray(th["b1"], dtype=np.float32).astype(np.float64)
    out["b2"] = np.asarray(th["b2"], dtype=np.float32).astype(np.float64)
    return out


def conv(geo, W, x):
    """out[s, o] = sum_t sum_i W[t, o, i] x[s + d_t, i]"""
    out = np.zeros((geo.V, W.shape[1]))
    for t in range(NTAP):
        out += x[geo.tap(t)] @ W[t].T
    return out


def conv_t(geo, W, y):
    """adjoint of conv: out[s, i] = sum_t sum_o W[t, o, i] y[s - d_t, o]"""
    out = np.zeros((geo.V, W.shape[2]))
    for t in range(NTAP):
        out += y[geo.tap(opposite_tap(t))] @ W[t]
    return out


def forward(geo, th, phi):
    h1 = np.tanh(conv(geo, th["W1"], phi) + th["b1"])
    h2 = np.tanh(conv(geo, th["W2"], h1) + th["b2"])
    S = float(np.sum(h2 @ th["w3"]))
    return S, (phi, h1, h2)


def grad_phi(geo, th, cache):
    phi, h1, h2 = cache
    gz2 = (1 - h2 ** 2) * th["w3"]
    gz1 = (1 - h1 ** 2) * conv_t(geo, th["W2"], gz2)
    return conv_t(geo, th["W1"], gz1)


def force_from_gphi(geo, U, g):
    """F = dS/dUbar for feature gradients g[s, k] = dS/dphi_k(s)"""
    N = U.shape[-1]
    F = np.zeros_like(U)
    for k, lp in enumerate(LOOPS):
        G = (g[:, 2 * k] + 1j * g[:, 2 * k + 1]) / (2 * N)
        ms = [_factor(geo, U, f) for f in lp]
        n = len(lp)
        for m, (a, off, d) in enumerate(lp):
            rest = _chain([ms[(m + 1 + j) % n] for j in range(n - 1)])
            contrib = (np.conj(G)[:, None, None] * rest) if d else (G[:, None, None] * dag(rest))
            np.add.at(F, (geo.shift(off), a), contrib)
    for a in range(NUMLINK):
        F[:, a] += g[:, 60 + a][:, None, None] * U[:, a] / N
    return F


def action_and_force(geo, th, U):
    phi = features(geo, U)
    S, cache = forward(geo, th, phi)
    return S, force_from_gphi(geo, U, grad_phi(geo, th, cache))


def tangent_grad(geo, th, cache, v):
    """gradient over the parameters of the tangent S_dot = dS(phi + eps v)/deps"""
    phi, h1, h2 = cache
    a1, a2 = 1 - h1 ** 2, 1 - h2 ** 2
    z1d = conv(geo, th["W1"], v)
    h1d = a1 * z1d
    z2d = conv(geo, th["W2"], h1d)
    h2d = a2 * z2d
    w3 = th["w3"]
    g = {}
    g["w3"] = h2d.sum(axis=0)
    adj_z2d = a2 * w3
    adj_z2 = -2 * h2 * a2 * w3 * z2d
    g["W2"] = np.zeros_like(th["W2"])
    for t in range(NTAP):
        g["W2"][t] = adj_z2d.T @ h1d[geo.tap(t)] + adj_z2.T @ h1[geo.tap(t)]
    g["b2"] = adj_z2.sum(axis=0)
    adj_h1d = conv_t(geo, th["W2"], adj_z2d)
    adj_h1 = conv_t(geo, th["W2"], adj_z2) - 2 * h1 * adj_h1d * z1d
    adj_z1d = adj_h1d * a1
    adj_z1 = adj_h1 * a1
    g["W1"] = np.zeros_like(th["W1"])
    for t in range(NTAP):
        g["W1"][t] = adj_z1d.T @ v[geo.tap(t)] + adj_z1.T @ phi[geo.tap(t)]
    g["b1"] = adj_z1.sum(axis=0)
    return g


def force_loss(geo, th, U, Ftarget):
    """|F_theta - F*|^2 / |F*|^2"""
    _, F = action_and_force(geo, th, U)
    R = F - Ftarget
    return float(np.vdot(R, R).real / np.vdot(Ftarget, Ftarget).real)


def force_loss_grad(geo, th, U, Ftarget):
    """(loss, gradient over the parameters) of force_loss"""
    phi = features(geo, U)
    S, cache = forward(geo, th, phi)
    F = force_from_gphi(geo, U, grad_phi(geo, th, cache))
    R = F - Ftarget
    Z = np.vdot(Ftarget, Ftarget).real
    v = features_jvp(geo, U, R)
    g = tangent_grad(geo, th, cache, v)
    return float(np.vdot(R, R).real / Z), {k: val / Z for k, val in g.items()}


# parameter vector of the library: W1 [11][16][65], b1 [16], W2 [11][16][16],
# b2 [16], w3 [16] (float32)
PARAM_SHAPES = [("W1", (NTAP, NCH, NFEAT)), ("b1", (NCH,)), ("W2", (NTAP, NCH, NCH)),
                ("b2", (NCH,)), ("w3", (NCH,))]


def pack(th):
    return np.concatenate([np.asarray(th[k], dtype=np.float64).ravel()
                           for k, _ in PARAM_SHAPES]).astype(np.float32)


def unpack(vec):
    out, o = {}, 0
    for k, shp in PARAM_SHAPES:
        n = int(np.prod(shp))
        out[k] = np.asarray(vec[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out
