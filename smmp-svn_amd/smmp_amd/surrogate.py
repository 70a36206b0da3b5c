"""The delayed-acceptance surrogate from Python.  The network lives with the
context on the device (smmp_mlp_set / smmp_mlp_fit train it there with the
library's BF16 MFMA kernels); this class keeps the scale and offset that
map network units to kcal/mol and drives training and delayed-acceptance
sampling.

The network predicts E~(x) = scale * y(f(x)) + offset from the whole
chain (include/smmp_mi355x.h: per residue sin/cos of its variables and a
16-d embedding of its type, three hidden layers of 256).  It is fitted to
exact energies the sampler produced, and in the sampler it only screens
trials: the exact energy judges every trial that passes, so the sampled
distribution is the exact one whatever the network's quality (delayed
acceptance, Christen and Fox, 2005)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi


def layout(ctx) -> dict:
    """The network's parameter layout for the context's system (smmp_mlp_layout)."""
    out = np.zeros(16, dtype=np.int32)
    ctx.check(_ffi.lib.smmp_mlp_layout(ctx.h, _ffi.iptr(out)))
    keys = ("n", "nres", "D", "hid", "w1", "b1", "w2", "b2", "w3", "b3", "w4", "b4", "emb", "emb_dim", "types",
            "slots")
    return {k: int(v) for k, v in zip(keys, out)}


def init_theta(ctx, seed: int = 0, init_scale: float = 1.0) -> np.ndarray:
    """Fan-in scaled normal weights, zero biases, embeddings of spread 0.5."""
    L = layout(ctx)
    rng = np.random.default_rng(seed)
    th = np.zeros(L["n"], dtype=np.float32)
    H, D = L["hid"], L["D"]
    th[L["w1"]:L["b1"]] = rng.normal(0, init_scale / np.sqrt(D), D * H)
    th[L["w2"]:L["b2"]] = rng.normal(0, init_scale / np.sqrt(H), H * H)
    th[L["w3"]:L["b3"]] = rng.normal(0, init_scale / np.sqrt(H), H * H)
    th[L["w4"]:L["b4"]] = rng.normal(0, init_scale / np.sqrt(H), H)
    th[L["emb"]:] = rng.normal(0, 0.5, L["types"] * L["emb_dim"])
    return th


class Surrogate:
    """The network of one context, with its energy scale."""

    def __init__(self, ctx, theta=None, seed: int = 0, init_scale: float = 1.0):
        self.ctx = ctx
        self.n = _ffi.lib.smmp_mlp_size(ctx.h)
        self.scale = 1.0   # kcal/mol per network unit
        self.offset = 0.0  # kcal/mol
        self.set_theta(init_theta(ctx, seed, init_scale) if theta is None else theta)

    @classmethod
    def zero(cls, ctx) -> "Surrogate":
        """A network whose output is zero everywhere (delayed acceptance with
        it is plain Metropolis, draw for draw)."""
        return cls(ctx, theta=np.zeros(_ffi.lib.smmp_mlp_size(ctx.h), dtype=np.float32))

    def set_theta(self, th) -> None:
        a = np.ascontiguousarray(th, dtype=np.float32)
        if a.size != self.n:
            raise ValueError(f"expected {self.n} parameters")
        self.ctx.check(_ffi.lib.smmp_mlp_set(self.ctx.h, a.ctypes.data_as(C.POINTER(C.c_float))))

    def theta(self) -> np.ndarray:
        a = np.zeros(self.n, dtype=np.float32)
        self.ctx.check(_ffi.lib.smmp_mlp_get(self.ctx.h, a.ctypes.data_as(C.POINTER(C.c_float))))
        return a

    def grad(self, vlvr, target=None, pairs: bool = False):
        """(loss, predictions in network units, gradient) for conformations
        vlvr (n x nvt) and targets in network units (or None); with pairs
        the conformations come in pairs (x, x') and the targets are for
        the differences of the outputs."""
        v = np.ascontiguousarray(vlvr, dtype=np.float64)
        n = len(v)
        pred = np.zeros(n, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        if target is None:
            self.ctx.check(_ffi.lib.smmp_mlp_grad(self.ctx.h, _ffi.dptr(v), n, None, 0, pred.ctypes.data_as(fp),
                                                  None, None))
            return None, pred, None
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.n, dtype=np.float32)
        loss = C.c_double(0.0)
        self.ctx.check(_ffi.lib.smmp_mlp_grad(self.ctx.h, _ffi.dptr(v), n, t.ctypes.data_as(fp), int(pairs),
                                              pred.ctypes.data_as(fp), g.ctypes.data_as(fp), C.byref(loss)))
        return loss.value, pred, g

    def predict(self, vlvr) -> np.ndarray:
        """E~ in kcal/mol."""
        _, pred, _ = self.grad(vlvr)
        return pred.astype(np.float64) * self.scale + self.offset

    def fit(self, vlvr, energies, epochs: int = 20, batch: int = 256, lr: float = 2e-3, seed: int = 0,
            clip: float | None = None, rescale: bool = True) -> np.ndarray:
        """Fits the network to exact energies (n x nvt conformations, n
        energies in kcal/mol); energies above clip are left out.  With
        rescale the targets are standardised first (offset = mean, scale =
        standard deviation).  Returns the loss of every epoch."""
        e = np.asarray(energies, dtype=np.float64)
        v = np.asarray(vlvr, dtype=np.float64)
        if clip is not None:
            keep = e <= clip
            e, v = e[keep], v[keep]
        if rescale:
            self.offset = float(e.mean())
            self.scale = float(max(e.std(), 1e-6))
        t = (e - self.offset) / self.scale
        return self._fit(v, t, 0, epochs, batch, lr, seed)

    def fit_moves(self, pairs_v, delta_e, epochs: int = 20, batch: int = 256, lr: float = 2e-3, seed: int = 0,
                  clip: float = 20.0, rescale: bool = True) -> np.ndarray:
        """Fits the network's energy change to exact changes of single moves:
        pairs_v (2n x nvt: x, x' alternating), delta_e (n, kcal/mol), the
        changes clipped to [-clip, clip] (beyond a few kT only the sign of a
        change matters to the first stage).  Returns the loss of every epoch."""
        d = np.clip(np.asarray(delta_e, dtype=np.float64), -clip, clip)
        if rescale:
            self.offset = 0.0
            self.scale = float(max(d.std(), 1e-6))
        return self._fit(np.asarray(pairs_v, dtype=np.float64), d / self.scale, 1, epochs, batch, lr, seed)

    def _fit(self, v, t, pairs, epochs, batch, lr, seed):
        v = np.ascontiguousarray(v, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.float32)
        hist = np.zeros(epochs)
        self.ctx.check(_ffi.lib.smmp_mlp_fit(self.ctx.h, _ffi.dptr(v), t.ctypes.data_as(C.POINTER(C.c_float)),
                                             int(pairs), len(t), epochs, batch, C.c_float(lr),
                                             C.c_uint32(seed & 0xFFFFFFFF), _ffi.dptr(hist)))
        return hist

    def sample(self, temps, seeds, nsweep, vlvr, **kw):
        """smmp_sample with delayed acceptance screened by this network."""
        kw.setdefault("da_scale", self.scale)
        return self.ctx.sample(temps, seeds, nsweep, vlvr, da=True, **kw)


def move_pairs(ctx, vlvr, seed: int = 0):
    """Training pairs of single moves from sampled conformations: for each
    conformation x, one free variable changed as the sampler proposes it
    (addang(v, axvr (u - 1/2)), metropolis.f:64-66) gives x'; both energies
    come from smmp_energy_batch (FP64).  Returns (pairs (2n x nvt, x and x'
    alternating), E(x') - E(x), E(x))."""
    v = np.asarray(vlvr, dtype=np.float64)
    n, nvt = v.shape
    rng = np.random.default_rng(seed)
    idvr = ctx.idvr - 1
    ax = np.asarray(ctx.view("var_r.axvr")[:nvt], dtype=np.float64)
    j = idvr[rng.integers(0, len(idvr), n)]
    x1 = v.copy()
    a = x1[np.arange(n), j] + ax[j] * (rng.random(n) - 0.5)
    x1[np.arange(n), j] = np.mod(a + np.pi, 2 * np.pi) - np.pi
    both = np.empty((2 * n, nvt))
    both[0::2] = v
    both[1::2] = x1
    e = _ffi.energy_batch([ctx], np.zeros(2 * n, dtype=np.int32), both.ravel(), precision=0)[:, 0]
    return both, e[1::2] - e[0::2], e[0::2]


def da_gain(plain: dict, da: dict, t_plain: float, t_da: float) -> dict:
    """Accepted moves per second with and without the network (every move a
    delayed-acceptance run accepts has passed the exact energy)."""
    a_plain = float(np.sum(plain["acc"]))
    a_da = float(np.sum(da["acc"]))
    st = da["da_stats"]
    return {"accepted_per_s_plain": a_plain / t_plain, "accepted_per_s_da": a_da / t_da,
            "gain": (a_da / t_da) / (a_plain / t_plain),
            "screened": float(st[:, 0].sum()), "rejected_by_network": float(st[:, 1].sum()),
            "exact_evaluations": float(st[:, 2].sum()), "accepted_da": float(st[:, 3].sum())}
