"""CPU oracle of the delayed-acceptance surrogate network.  TEST
INFRASTRUCTURE ONLY: tests/ may import it; the product never does.

A numpy restatement of the network of smmp-svn_amd/csrc/device/mlp_common.hpp
and its training passes (mlp.hip): whole-chain features, forward pass,
mean-squared-error loss and its gradient by back-propagation, and the
sampler's screen (surrogate.hpp: first-layer sums on the 2^-40 integer grid).
With bf16=True the matrix operands are rounded to BF16 (round to nearest
even) at the points where the GPU kernels round them, so the GPU's results
can be compared closely; with bf16=False everything is FP64 (the form the
finite-difference check pins)."""
from __future__ import annotations

import numpy as np

HID, SLOTS, EMB, TYPES = 256, 10, 16, 32
RESCOLS = 2 * SLOTS + EMB   # inputs per residue
EMB_COL = 2 * SLOTS         # embedding columns of a residue start here

STD = ("ala", "arg", "asn", "asp", "cys", "gln", "glu", "gly", "his", "ile",
       "leu", "lys", "met", "phe", "pro", "ser", "thr", "trp", "tyr", "val")


class Layout:
    """Offsets of the parameters in the FP32 vector for nres residues."""

    def __init__(self, nres: int):
        self.nres = nres
        self.D = nres * RESCOLS
        self.W1 = 0
        self.B1 = self.W1 + self.D * HID
        self.W2 = self.B1 + HID
        self.B2 = self.W2 + HID * HID
        self.W3 = self.B2 + HID
        self.B3 = self.W3 + HID * HID
        self.W4 = self.B3 + HID
        self.B4 = self.W4 + HID
        self.EMBO = self.B4 + 16
        self.NPARAM = self.EMBO + TYPES * EMB


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def type_id(name: str) -> int:
    n = name.strip().lower()[:3]
    return STD.index(n) + 1 if n in STD else 21


class Residues:
    """The residues of a system (all chains in order): first variable
    (0-based), number of variables, embedding row."""

    def __init__(self, var0, nvar, types, chain):
        self.var0 = np.asarray(var0, dtype=np.int64)
        self.nvar = np.asarray(nvar, dtype=np.int64)
        self.type = np.asarray(types, dtype=np.int64)
        self.chain = np.asarray(chain, dtype=np.int64)
        self.nres = len(self.var0)
        self.lay = Layout(self.nres)

    @classmethod
    def from_context(cls, ctx) -> "Residues":
        v = ctx.view
        nml = int(v("mol_i.ntlml")[0])
        var0, nvar, types, chain = [], [], [], []
        seq = v("res_c.seq")
        for m in range(nml):
            for r in range(int(v("mol_i.irsml1")[m]), int(v("mol_i.irsml2")[m]) + 1):
                var0.append(int(v("res_i.ivrrs1")[r - 1]) - 1)
                nvar.append(int(v("res_i.nvrrs")[r - 1]))
                name = bytes(np.asarray(seq[:, r - 1]).tobytes()).decode() if seq.ndim == 2 else str(seq[r - 1])
                types.append(type_id(name))
                chain.append(m)
        return cls(var0, nvar, types, chain)

    def angle_columns(self) -> list[tuple[int, int]]:
        """(column, variable) of every angle input: sin at 36 r + 2 s, cos + 1."""
        out = []
        for r in range(self.nres):
            for s in range(self.nvar[r]):
                out.append((RESCOLS * r + 2 * s, self.var0[r] + s))
        return out

    def features(self, vlvr: np.ndarray, theta: np.ndarray, bf: bool) -> np.ndarray:
        """X (conformations x D) of conformations vlvr (nconf x nvt)."""
        vlvr = np.atleast_2d(np.asarray(vlvr, dtype=np.float64))
        X = np.zeros((len(vlvr), self.lay.D))
        for c, var in self.angle_columns():
            X[:, c] = np.sin(vlvr[:, var]).astype(np.float32)
            X[:, c + 1] = np.cos(vlvr[:, var]).astype(np.float32)
        emb = np.asarray(theta, dtype=np.float64)[self.lay.EMBO:self.lay.NPARAM].reshape(TYPES, EMB)
        for r in range(self.nres):
            X[:, RESCOLS * r + EMB_COL:RESCOLS * (r + 1)] = emb[self.type[r]]
        return bf16(X) if bf else X


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(lay: Layout, theta: np.ndarray, bf: bool):
    t = np.asarray(theta, dtype=np.float64)
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    return dict(W1=r(t[lay.W1:lay.B1].reshape(lay.D, HID)), b1=t[lay.B1:lay.W2],
                W2=r(t[lay.W2:lay.B2].reshape(HID, HID)), b2=t[lay.B2:lay.W3],
                W3=r(t[lay.W3:lay.B3].reshape(HID, HID)), b3=t[lay.B3:lay.W4], w4=t[lay.W4:lay.B4], b4=t[lay.B4])


def forward(res: Residues, theta, vlvr, bf: bool = True):
    """Per-conformation outputs and the intermediate values."""
    p = unpack(res.lay, theta, bf)
    X = res.features(vlvr, np.asarray(theta), bf)
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    Z1 = X @ p["W1"] + p["b1"]
    H1 = r(_silu(Z1))
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = r(_silu(Z2))
    Z3 = H2 @ p["W3"] + p["b3"]
    H3 = r(_silu(Z3))
    y = H3 @ p["w4"] + p["b4"]
    return y, dict(X=X, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Z3=Z3, H3=H3, y=y, p=p)


def loss_grad(res: Residues, theta, vlvr, target, bf: bool = True, pairs: bool = False):
    """Mean squared error of the outputs and its gradient (NPARAM).  With
    pairs the conformations come in pairs (x, x') and the targets are for
    the differences y(x') - y(x)."""
    lay = res.lay
    pred, c = forward(res, theta, vlvr, bf)
    target = np.asarray(target, dtype=np.float64)
    if pairs:
        n = len(pred) // 2
        e = pred[1::2] - pred[0::2] - target
        dy = np.empty(2 * n)
        dy[0::2] = -2.0 * e / n
        dy[1::2] = 2.0 * e / n
    else:
        n = len(pred)
        e = pred - target
        dy = 2.0 * e / n
    loss = float(np.mean(e * e))
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    p = c["p"]
    g = np.zeros(lay.NPARAM)
    g[lay.W4:lay.B4] = c["H3"].T @ dy
    g[lay.B4] = dy.sum()
    dZ3 = r(np.outer(dy, p["w4"]) * _silu_grad(c["Z3"]))
    g[lay.W3:lay.B3] = (c["H2"].T @ dZ3).ravel()
    g[lay.B3:lay.W4] = dZ3.sum(0)
    dZ2 = r((dZ3 @ p["W3"].T) * _silu_grad(c["Z2"]))
    g[lay.W2:lay.B2] = (c["H1"].T @ dZ2).ravel()
    g[lay.B2:lay.W3] = dZ2.sum(0)
    dZ1 = r((dZ2 @ p["W2"].T) * _silu_grad(c["Z1"]))
    g[lay.W1:lay.B1] = (c["X"].T @ dZ1).ravel()
    g[lay.B1:lay.W2] = dZ1.sum(0)
    dX = dZ1 @ p["W1"].T
    ge = np.zeros((TYPES, EMB))
    for q in range(res.nres):
        ge[res.type[q]] += dX[:, RESCOLS * q + EMB_COL:RESCOLS * (q + 1)].sum(0)
    g[lay.EMBO:] = ge.ravel()
    return loss, g, pred


def screen_output(res: Residues, theta, vlvr) -> np.ndarray:
    """The sampler's screen (surrogate.hpp) for conformations vlvr: layer-1
    sums as integers on the 2^-40 grid of bf16(W1) * bf16(x) products (each
    exact in FP64), then the FP32 layers on BF16 operands."""
    lay = res.lay
    th = np.asarray(theta, dtype=np.float32)
    X = res.features(vlvr, th, True)
    W1 = bf16(th[lay.W1:lay.B1]).reshape(lay.D, HID)
    f32 = np.float32
    out = []
    for x in np.atleast_2d(X):
        terms = np.rint(W1 * x[:, None] * 2.0 ** 40)  # exact products, rounded once to the grid
        z1 = terms.astype(np.int64).sum(0)
        z = (z1.astype(np.float64) * 2.0 ** -40).astype(f32) + th[lay.B1:lay.W2]
        h1 = bf16(z / (f32(1) + np.exp(-z)))
        z2 = (h1 @ bf16(th[lay.W2:lay.B2]).reshape(HID, HID)).astype(f32) + th[lay.B2:lay.W3]
        h2 = bf16(z2 / (f32(1) + np.exp(-z2)))
        z3 = (h2 @ bf16(th[lay.W3:lay.B3]).reshape(HID, HID)).astype(f32) + th[lay.B3:lay.W4]
        h3 = bf16(z3 / (f32(1) + np.exp(-z3)))
        out.append(float(h3 @ th[lay.W4:lay.B4].astype(np.float64)) + float(th[lay.B4]))
    return np.array(out)


def adam(theta, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (returns new theta, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    th = theta - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return th, m, v


def init_theta(lay: Layout, seed: int, scale: float = 1.0) -> np.ndarray:
    """Random parameters (fan-in scaled normal weights, zero biases)."""
    rng = np.random.default_rng(seed)
    t = np.zeros(lay.NPARAM)
    t[lay.W1:lay.B1] = rng.normal(0, scale / np.sqrt(lay.D), lay.D * HID)
    t[lay.W2:lay.B2] = rng.normal(0, scale / np.sqrt(HID), HID * HID)
    t[lay.W3:lay.B3] = rng.normal(0, scale / np.sqrt(HID), HID * HID)
    t[lay.W4:lay.B4] = rng.normal(0, scale / np.sqrt(HID), HID)
    t[lay.EMBO:] = rng.normal(0, 0.5, TYPES * EMB)
    return t.astype(np.float32)
