"""Metropolis weights for the GPU sweeps (the reference passes a Python
callback, smmp.pyf:3-10; the sweep kernels need it as data).

  Canonical(beta)                       w(x) = beta x            (canon.f:135-149)
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


def as_descriptor(w) -> _ffi.Weight:
    """A weight descriptor for w: a Canonical/Multicanonical object, a
    descriptor, or a callable linear in the energy (read off as beta)."""
    if isinstance(w, _ffi.Weight):
        return w
    if hasattr(w, "descriptor"):
        return w.descriptor()
    if callable(w):
        k = float(w(1.0))
        ok = float(w(0.0)) == 0.0
        for x in (2.0, -3.5, 123.25):
            ok = ok and abs(float(w(x)) - k * x) <= 1e-12 * max(1.0, abs(k * x))
        if not ok:
            raise TypeError("metropolis: the weight callable is not linear in the energy; pass a "
                            "smmp_amd.weights.Multicanonical (the GPU sweep evaluates the weight itself)")
        return _ffi.Weight(kind=0, beta=k)
    raise TypeError(f"metropolis: unsupported weight {w!r}")
