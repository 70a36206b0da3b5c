"""CPU oracle of the delayed-acceptance surrogate network.  TEST
INFRASTRUCTURE ONLY: tests/ may import it; the product never does.

A numpy restatement of the network of smmp-svn_amd/csrc/device/mlp_common.hpp
and its training passes (mlp.hip): features of residue windows, forward pass,
mean-squared-error loss and its gradient by back-propagation.  With
bf16=True the matrix operands are rounded to BF16 (round to nearest even) at
the points where the GPU kernels round them, so the GPU's results can be
compared closely; with bf16=False everything is FP64 (the form the
finite-difference check pins)."""
from __future__ import annotations

import numpy as np

IN, HID, SLOTS, EMB, TYPES = 128, 64, 10, 16, 32
ANGLE_COLS = 3 * 2 * SLOTS
EMB_COL = ANGLE_COLS
# offsets in the FP32 parameter vector
W1 = 0
B1 = W1 + IN * HID
W2 = B1 + HID
B2 = W2 + HID * HID
W3 = B2 + HID
B3 = W3 + HID * HID
W4 = B3 + HID
B4 = W4 + HID
EMBO = B4 + 16
NPARAM = EMBO + TYPES * EMB

STD = ("ala", "arg", "asn", "asp", "cys", "gln", "glu", "gly", "his", "ile",
       "leu", "lys", "met", "phe", "pro", "ser", "thr", "trp", "tyr", "val")


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
    """Residue windows of a system: per residue its first variable (0-based),
    number of variables, embedding row and chain."""

    def __init__(self, var0, nvar, types, chain):
        self.var0 = np.asarray(var0, dtype=np.int64)
        self.nvar = np.asarray(nvar, dtype=np.int64)
        self.type = np.asarray(types, dtype=np.int64)
        self.chain = np.asarray(chain, dtype=np.int64)
        self.nres = len(self.var0)

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

    def features(self, vlvr: np.ndarray, theta: np.ndarray, bf: bool) -> tuple[np.ndarray, np.ndarray]:
        """X (rows = conformations x residues, IN columns) and the window type
        ids (rows x 3) of conformations vlvr (nconf x nvt)."""
        vlvr = np.atleast_2d(np.asarray(vlvr, dtype=np.float64))
        nconf, L = len(vlvr), self.nres
        X = np.zeros((nconf * L, IN))
        T = np.zeros((nconf * L, 3), dtype=np.int64)
        emb = theta[EMBO:EMBO + TYPES * EMB].astype(np.float64).reshape(TYPES, EMB)
        for r in range(L):
            for d in range(3):
                q = r + d - 1
                if 0 <= q < L and self.chain[q] == self.chain[r]:
                    T[r::L, d] = self.type[q]
                    for s in range(self.nvar[q]):
                        a = vlvr[:, self.var0[q] + s]
                        X[r::L, 20 * d + 2 * s] = np.sin(a).astype(np.float32)
                        X[r::L, 20 * d + 2 * s + 1] = np.cos(a).astype(np.float32)
        for d in range(3):
            X[:, EMB_COL + EMB * d:EMB_COL + EMB * (d + 1)] = emb[T[:, d]]
        return (bf16(X) if bf else X), T


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(theta: np.ndarray, bf: bool):
    t = np.asarray(theta, dtype=np.float64)
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    return dict(W1=r(t[W1:B1].reshape(IN, HID)), b1=t[B1:W2], W2=r(t[W2:B2].reshape(HID, HID)), b2=t[B2:W3],
                W3=r(t[W3:B3].reshape(HID, HID)), b3=t[B3:W4], w4=t[W4:B4], b4=t[B4])


def forward(res: Residues, theta, vlvr, bf: bool = True):
    """Per-conformation predictions and the intermediate values."""
    p = unpack(theta, bf)
    X, T = res.features(vlvr, np.asarray(theta), bf)
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    Z1 = X @ p["W1"] + p["b1"]
    H1 = r(_silu(Z1))
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = r(_silu(Z2))
    Z3 = H2 @ p["W3"] + p["b3"]
    H3 = r(_silu(Z3))
    y = H3 @ p["w4"] + p["b4"]
    pred = y.reshape(-1, res.nres).sum(1)
    return pred, dict(X=X, T=T, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Z3=Z3, H3=H3, y=y, p=p)


def loss_grad(res: Residues, theta, vlvr, target, bf: bool = True, pairs: bool = False):
    """Mean squared error of the predictions and its gradient (NPARAM).
    With pairs the conformations come in pairs (x, x') and the targets are
    for the differences pred(x') - pred(x)."""
    pred, c = forward(res, theta, vlvr, bf)
    target = np.asarray(target, dtype=np.float64)
    if pairs:
        n = len(pred) // 2
        e = pred[1::2] - pred[0::2] - target
        gconf = np.empty(2 * n)
        gconf[0::2] = -2.0 * e / n
        gconf[1::2] = 2.0 * e / n
    else:
        n = len(pred)
        e = pred - target
        gconf = 2.0 * e / n
    loss = float(np.mean(e * e))
    r = (lambda a: bf16(a)) if bf else (lambda a: a)
    p = c["p"]
    dy = np.repeat(gconf, res.nres)
    g = np.zeros(NPARAM)
    g[W4:B4] = c["H3"].T @ dy
    g[B4] = dy.sum()
    dZ3 = r(np.outer(dy, p["w4"]) * _silu_grad(c["Z3"]))
    g[W3:B3] = (c["H2"].T @ dZ3).ravel()
    g[B3:W4] = dZ3.sum(0)
    dZ2 = r((dZ3 @ p["W3"].T) * _silu_grad(c["Z2"]))
    g[W2:B2] = (c["H1"].T @ dZ2).ravel()
    g[B2:W3] = dZ2.sum(0)
    dZ1 = r((dZ2 @ p["W2"].T) * _silu_grad(c["Z1"]))
    g[W1:B1] = (c["X"].T @ dZ1).ravel()
    g[B1:W2] = dZ1.sum(0)
    dX = dZ1 @ p["W1"].T
    ge = np.zeros((TYPES, EMB))
    for d in range(3):
        np.add.at(ge, c["T"][:, d], dX[:, EMB_COL + EMB * d:EMB_COL + EMB * (d + 1)])
    g[EMBO:] = ge.ravel()
    return loss, g, pred


def adam(theta, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (returns new theta, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    th = theta - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return th, m, v


def init_theta(seed: int, scale: float = 1.0) -> np.ndarray:
    """Random parameters (fan-in scaled normal weights, zero biases)."""
    rng = np.random.default_rng(seed)
    t = np.zeros(NPARAM)
    t[W1:B1] = rng.normal(0, scale / np.sqrt(IN), IN * HID)
    t[W2:B2] = rng.normal(0, scale / np.sqrt(HID), HID * HID)
    t[W3:B3] = rng.normal(0, scale / np.sqrt(HID), HID * HID)
    t[W4:B4] = rng.normal(0, scale / np.sqrt(HID), HID)
    t[EMBO:] = rng.normal(0, 0.5, TYPES * EMB)
    return t.astype(np.float32)
