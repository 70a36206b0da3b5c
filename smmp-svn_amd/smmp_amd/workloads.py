"""Synthetic workloads: random sequences and conformations of the benchmark configs.

Residue alphabet: the 20 neutral standard residues of the ECEPP/3 library
(SMMP/lib.sh3).  A chain may not start with proline (the reference renames a
leading ``pro`` to ``pron``, redseq.f:175-179, and its README asks for chains
that do not start with it), so a leading ``pro`` is redrawn.
"""
from __future__ import annotations

import math
import os

import numpy as np

RESIDUES = ("ala", "arg", "asn", "asp", "cys", "gln", "glu", "gly", "his",
            "ile", "leu", "lys", "met", "phe", "pro", "ser", "thr", "trp",
            "tyr", "val")

#: Gas constant in kcal/(mol K) as used by canon.f:57 and partem_s.f:36.
GASC = 1.98773e-3


def random_sequence(length: int, seed: int) -> list[str]:
    """Uniform random sequence of ``length`` residues; no leading proline."""
    rng = np.random.default_rng(seed)
    seq = [RESIDUES[i] for i in rng.integers(0, len(RESIDUES), size=length)]
    while seq[0] == "pro":
        seq[0] = RESIDUES[int(rng.integers(0, len(RESIDUES)))]
    return seq


def write_seq_file(path: str, seq: list[str], name: str = "") -> str:
    """Writes a sequence file in the reference's format (redseq.f:13-213)."""
    with open(path, "w") as f:
        f.write(f"# {name}\n" if name else "#\n")
        for i in range(0, len(seq), 10):
            f.write(" ".join(seq[i:i + 10]) + "\n")
    return path


def geometric_temperatures(n: int, tmin: float, tmax: float) -> np.ndarray:
    """T_k = tmin (tmax/tmin)^(k/(n-1)), k = 0..n-1."""
    if n == 1:
        return np.array([tmin])
    k = np.arange(n, dtype=np.float64)
    return tmin * (tmax / tmin) ** (k / (n - 1))


def write_temperatures(path: str, temps) -> str:
    """The ``temperatures`` file read by partem_s.f:89-96 (index, T)."""
    with open(path, "w") as f:
        for i, t in enumerate(temps, 1):
            f.write(f"{i} {float(t)!r}\n")
    return path


def c3_sequences(nseq: int = 256, lmin: int = 10, lmax: int = 60,
                 seed: int = 3) -> list[list[str]]:
    """Batched-energy workload: ``nseq`` sequences, lengths U{lmin..lmax}."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lmin, lmax + 1, size=nseq)
    return [random_sequence(int(L), seed * 100003 + i) for i, L in enumerate(lens)]


def uniform_dihedrals(nconf: int, nvr: int, seed: int) -> np.ndarray:
    """Dihedrals uniform in [-pi, pi) (as addang(pi, 2pi(u-1/2)))."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-math.pi, math.pi, size=(nconf, nvr))


def default_libdir() -> str:
    """Directory holding lib.sh3 and tes.dat: $SMMP_LIBDIR or the bundled copy."""
    env = os.environ.get("SMMP_LIBDIR")
    if env:
        return env.rstrip("/") + "/"
    here = os.path.dirname(os.path.abspath(__file__))
    return os.path.join(os.path.dirname(os.path.dirname(here)), "tests", "golden", "SMMP") + "/"
