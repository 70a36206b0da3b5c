"""Metropolis weights for the GPU sweeps (the reference passes a Python
callback, smmp.pyf:3-10; the sweep kernels need it as data).

  Canonical(beta)                       w(x) = beta x            (canon.f:135-149)
  CanonicalCommon(ctx)                  w(x) = beta x, beta read from the
                                        context's /bet/ at each call (can_weight)
  Multicanonical(b, alpha, kmin, kmax, ebin, beta=None)
      beta None:  w(x) = b(k) x + alpha(k)              (muca_weight2, mulcan_par_mod.f90:398-412)
      beta given: w(x) = (beta + b(k)) x + alpha(k)     (muca_weight, mulcan_par_mod.f90:380-395)
      with k = int(clamp(x/ebin + sign(1/2, x), kmin, kmax)).
"""
from __future__ import annotations

import numpy as np

from . import _ffi


class Canonical:
    def __init__(self, beta: float):
        self.beta = float(beta)

    def descriptor(self) -> _ffi.Weight:
        return _ffi.Weight(kind=0, beta=self.beta)

    def __call__(self, x: float) -> float:
        return self.beta * x


class CanonicalCommon:
    """can_weight of canon.f:135-149: beta x with beta taken from the
    context's COMMON /bet/ when the sweep starts."""

    def __init__(self, ctx):
        self.ctx = ctx

    @property
    def beta(self) -> float:
        return float(self.ctx.view("bet.beta")[0])

    def descriptor(self) -> _ffi.Weight:
        return _ffi.Weight(kind=0, beta=self.beta)

    def __call__(self, x: float) -> float:
        return self.beta * x


class Multicanonical:
    def __init__(self, b, alpha, kmin: int, kmax: int, ebin: float, beta: float | None = None):
        n = kmax - kmin + 1
        self.b = np.ascontiguousarray(b, dtype=np.float64)
        self.alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        if self.b.shape != (n,) or self.alpha.shape != (n,):
            raise ValueError(f"Multicanonical: b and alpha need kmax-kmin+1 = {n} values")
        if not ebin > 0:
            raise ValueError("Multicanonical: ebin must be positive")
        self.kmin, self.kmax, self.ebin, self.beta = int(kmin), int(kmax), float(ebin), beta

    def bin(self, x: float) -> int:
        v = x / self.ebin + np.copysign(0.5, x)
        return int(min(max(v, self.kmin), self.kmax))

    def __call__(self, x: float) -> float:
        k = self.bin(x) - self.kmin
        slope = self.b[k] if self.beta is None else self.beta + self.b[k]
        return slope * x + self.alpha[k]

    def descriptor(self) -> _ffi.Weight:
        return _ffi.Weight(kind=2 if self.beta is None else 1, beta=0.0 if self.beta is None else self.beta,
                           kmin=self.kmin, kmax=self.kmax, ebin=self.ebin, b=_ffi.dptr(self.b),
                           alpha=_ffi.dptr(self.alpha))


def as_descriptor(w) -> _ffi.Weight | None:
    """The weight descriptor of w (a descriptor, or an object of this module),
    or None for any other callable: its values are known only by calling it,
    so the sweep must run where it can be called (the host-driven
    Metropolis loop).  Nothing is inferred from sample values."""
    if isinstance(w, _ffi.Weight):
        return w
    if hasattr(w, "descriptor"):
        return w.descriptor()
    if callable(w):
        return None
    raise TypeError(f"metropolis: unsupported weight {w!r}")
