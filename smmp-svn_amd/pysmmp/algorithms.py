"""CanonicalMonteCarlo (algorithms.py of the reference, same class and
methods, Python 3): sweeps run on the GPU through smmp.metropolis."""
from __future__ import annotations

import smmp


class CanonicalMonteCarlo:
    """Canonical Monte Carlo with Metropolis weights."""

    def __init__(self, myUniverse, nequi: int, sweeps: int, weight=smmp.can_weight):
        # the reference's default is lambda x: x * smmp.bet.beta (algorithms.py:19);
        # smmp.can_weight has that meaning and runs the sweep on the GPU, a
        # lambda passed here is honoured through the host-driven trial loop
        self._def_sweeps = sweeps
        self._def_equi = nequi
        self._univ = myUniverse
        self._acc = 0
        self._performed = 0
        self._min_energy = myUniverse.energy()
        self._min_counter = 0
        self._weight = weight

    def acceptanceRate(self) -> float:
        return float(self._acc) / (self._performed * int(smmp.mol_par.nvr))

    def minimumEnergy(self):
        return self._min_energy, self._min_counter

    def performedSweeps(self) -> int:
        return self._performed

    def run(self) -> None:
        self.sweeps(self._def_sweeps)

    def equilibrate(self) -> None:
        self.sweeps(self._def_equi)

    def sweeps(self, n: int) -> None:
        for _ in range(n):
            self.sweep()

    def sweep(self) -> None:
        """One Metropolis sweep over all free variables (metropolis.f)."""
        eol = self._univ.energy()
        eol, self._acc = smmp.metropolis(eol, self._acc, self._weight)
        self._performed += 1
        if eol < self._min_energy:
            self._min_energy = eol
            self._min_counter = self._performed
