"""CPU oracle of the warm-start message-passing network (test infrastructure).

Only tests/ and the benchmark's CPU baseline may use this module; the product
path is the HIP implementation in samos_amd/csrc/mpnn.hip.

The network predicts, for every particle of a 2-D soft-disk configuration, the
displacement that FIRE relaxation will apply to it.  It has no counterpart in
the reference (SAMoS relaxes with IntegratorFIRE only,
src/integrators/integrator_fire.cpp:34-148); its training pairs come from that
relaxation.  Definition (H = 64):

  edges e = (i -> j) for every neighbour j of i within r_l = rcut + pad
  d_e = minimum image of x_j - x_i, rho_e = |d_e|, dhat_e = d_e / rho_e,
  ov_e = r_i + r_j - rho_e
  edge features   f_e = [rho_e / r_l, ov_e, r_i - 1, r_j - 1]
  node features   u_i = [r_i - 1, sum_e max(ov_e, 0), #{e : ov_e > 0} / 6]
  embedding       h_i = silu(u_i W0 + b0)
  3 layers        Z_e = [h_i, h_j, f_e] We + be,  A_i = (1/8) sum_{e in row i} silu(Z_e)
                  h_i <- h_i + silu([h_i, A_i] Wn + bn)
  head            U_e = silu([h_i, h_j, f_e] Wo + bo),  w_e = U_e . v + c
                  delta_i = sum_{e in row i} w_e dhat_e
  loss            (1 / 2N) sum_i |delta_i - delta*_i|^2

The output is a sum of edge unit vectors weighted by rotation- and
translation-invariant scalars, so delta rotates with the configuration.

GEMM operands are rounded to bfloat16 (round to nearest even) exactly where
the device kernels round them, with float32 accumulation; `exact=True` keeps
everything in float64 (used to pin the backward pass against finite
differences).
"""
import numpy as np

H = 64
K_NODE = 32       # node-feature operand width (3 used, zero padded)
K_EDGE = 160      # [h_i, h_j, f_e] operand width (132 used, zero padded)
K_UPD = 128       # [h_i, A_i]
N_LAYERS = 3
AGG = 0.125       # A_i = AGG * sum over row i of the messages

# parameter blocks in the flat order shared with the device (name, rows, cols)
LAYOUT = [("W0", K_NODE, H), ("b0", 1, H)]
for _l in range(N_LAYERS):
    LAYOUT += [(f"We{_l}", K_EDGE, H), (f"be{_l}", 1, H), (f"Wn{_l}", K_UPD, H), (f"bn{_l}", 1, H)]
LAYOUT += [("Wo", K_EDGE, H), ("bo", 1, H), ("v", 1, H), ("c", 1, 1)]
N_PARAMS = sum(r * c for _, r, c in LAYOUT)
USED_ROWS = {"W0": 3, "Wo": 2 * H + 4}
for _l in range(N_LAYERS):
    USED_ROWS[f"We{_l}"] = 2 * H + 4


def unflatten(flat):
    out, o = {}, 0
    for name, r, c in LAYOUT:
        out[name] = np.asarray(flat[o:o + r * c], dtype=np.float64).reshape(r, c) if r > 1 else \
            np.asarray(flat[o:o + r * c], dtype=np.float64).reshape(c)
        o += r * c
    return out


def flatten(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _, _ in LAYOUT])


def bf16(x):
    """Round float32 values to bfloat16 precision (nearest, ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def edges(x, y, L, cut):
    """All ordered pairs (i, j), i != j, with minimum-image distance^2 < cut^2,
    rows sorted by i (the device keeps each row in its list order)."""
    n = len(x)
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    m = (dx * dx + dy * dy < cut * cut) & ~np.eye(n, dtype=bool)
    src, dst = np.nonzero(m)
    return src, dst, dx[src, dst], dy[src, dst]


class Net:
    """Forward and backward of the network for one configuration."""

    def __init__(self, flat, exact=False):
        self.p = unflatten(flat)
        self.exact = exact

    def _q(self, a):
        return np.asarray(a, dtype=np.float64) if self.exact else bf16(a).astype(np.float64)

    def _acc(self, a):
        return np.asarray(a, dtype=np.float64) if self.exact else np.asarray(a, dtype=np.float32).astype(np.float64)

    def forward(self, x, y, r, L, cut):
        src, dst, dx, dy = edges(x, y, L, cut)
        n, E = len(x), len(src)
        rho = np.sqrt(dx * dx + dy * dy)
        ov = r[src] + r[dst] - rho
        dhx, dhy = dx / rho, dy / rho
        fe = np.stack([rho / cut, ov, r[src] - 1.0, r[dst] - 1.0], 1)
        S = np.bincount(src, weights=np.maximum(ov, 0.0), minlength=n)
        C = np.bincount(src, weights=(ov > 0).astype(float), minlength=n)
        u = np.zeros((n, K_NODE))
        u[:, 0], u[:, 1], u[:, 2] = r - 1.0, S, C / 6.0
        if not self.exact:
            fe = fe.astype(np.float32).astype(np.float64)
            u = u.astype(np.float32).astype(np.float64)
        p, q = self.p, self._q
        c = dict(src=src, dst=dst, dhx=dhx, dhy=dhy, fe=fe, n=n, E=E)
        X0 = q(u)
        Z0 = self._acc(X0 @ q(p["W0"])) + p["b0"]
        h = silu(Z0)
        c["X0"], c["Z0"], c["h"] = X0, Z0, [h]
        c["Xe"], c["Ze"], c["Xn"], c["Zn"] = [], [], [], []
        for l in range(N_LAYERS):
            Xe = np.zeros((E, K_EDGE))
            Xe[:, :H], Xe[:, H:2 * H], Xe[:, 2 * H:2 * H + 4] = q(h[src]), q(h[dst]), q(fe)
            Ze = self._acc(Xe @ q(p[f"We{l}"])) + p[f"be{l}"]
            Me = silu(Ze)
            A = np.zeros((n, H))
            np.add.at(A, src, Me)
            A *= AGG
            Xn = np.concatenate([q(h), q(A)], 1)
            Zn = self._acc(Xn @ q(p[f"Wn{l}"])) + p[f"bn{l}"]
            h = h + silu(Zn)
            c["Xe"].append(Xe); c["Ze"].append(Ze); c["Xn"].append(Xn); c["Zn"].append(Zn); c["h"].append(h)
        Xo = np.zeros((E, K_EDGE))
        Xo[:, :H], Xo[:, H:2 * H], Xo[:, 2 * H:2 * H + 4] = q(h[src]), q(h[dst]), q(fe)
        Zo = self._acc(Xo @ q(p["Wo"])) + p["bo"]
        U = silu(Zo)
        w = U @ p["v"] + p["c"][0]
        delta = np.zeros((n, 2))
        np.add.at(delta[:, 0], src, w * dhx)
        np.add.at(delta[:, 1], src, w * dhy)
        c.update(Xo=Xo, Zo=Zo, U=U, w=w)
        self.cache = c
        return delta

    def backward(self, delta, target):
        """Loss (1/2N) sum |delta - target|^2 and its gradient as a flat vector."""
        c, p, q = self.cache, self.p, self._q
        src, dst, n = c["src"], c["dst"], c["n"]
        diff = delta - target
        loss = 0.5 * float(np.sum(diff * diff)) / n
        gd = diff / n
        g = {name: np.zeros((r, cc)) if r > 1 else np.zeros(cc) for name, r, cc in LAYOUT}
        gw = gd[src, 0] * c["dhx"] + gd[src, 1] * c["dhy"]
        g["v"] = c["U"].T @ gw
        g["c"] = np.array([gw.sum()])
        gZo = q(np.outer(gw, p["v"]) * dsilu(c["Zo"]))
        g["Wo"] = self._acc(c["Xo"].T @ gZo)
        g["bo"] = gZo.sum(0)
        gXo = self._acc(gZo @ q(p["Wo"]).T)
        gh = np.zeros((n, H))
        np.add.at(gh, src, gXo[:, :H])
        np.add.at(gh, dst, gXo[:, H:2 * H])
        for l in reversed(range(N_LAYERS)):
            gZn = q(gh * dsilu(c["Zn"][l]))
            g[f"Wn{l}"] = self._acc(c["Xn"][l].T @ gZn)
            g[f"bn{l}"] = gZn.sum(0)
            gXn = self._acc(gZn @ q(p[f"Wn{l}"]).T)
            gA = gXn[:, H:]
            gh = gh + gXn[:, :H]
            gZe = q(AGG * gA[src] * dsilu(c["Ze"][l]))
            g[f"We{l}"] = self._acc(c["Xe"][l].T @ gZe)
            g[f"be{l}"] = gZe.sum(0)
            gXe = self._acc(gZe @ q(p[f"We{l}"]).T)
            np.add.at(gh, src, gXe[:, :H])
            np.add.at(gh, dst, gXe[:, H:2 * H])
        gZ0 = q(gh * dsilu(c["Z0"]))
        g["W0"] = self._acc(c["X0"].T @ gZ0)
        g["b0"] = gZ0.sum(0)
        return loss, flatten(g)


def init_params(seed=0):
    """Uniform(+-sqrt(6 / (fan_in + fan_out))) weights on the used rows, zero
    biases and zero padding rows (the device initialiser follows the same rule
    with its own generator)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, r, cc in LAYOUT:
        if r == 1:
            p[name] = np.zeros(cc)
            continue
        used = USED_ROWS.get(name, r)
        a = np.sqrt(6.0 / (used + cc))
        w = np.zeros((r, cc))
        w[:used] = rng.uniform(-a, a, size=(used, cc))
        p[name] = w
    p["v"] = rng.uniform(-0.1, 0.1, size=H)
    return flatten(p)
