"""Universe: force field, solvent model, temperature (universe.py of the
reference, same class and methods, Python 3), over the drop-in ``smmp``."""
from __future__ import annotations

import smmp

GASC = 1.98773e-3  # kcal/(mol K), canon.f:57

# force-field switches of epar_l per name (universe.py:18-32)
_FORCE_FIELDS = {"ecepp3": (0, 0), "ecepp2": (0, 1), "flex": (1, 0)}


class Universe:
    """The environment of a simulation: temperature, solvent, force field
    and the molecules in it."""

    def __init__(self, T: float = 300, ff: str = "ecepp3", st: int = 1, libdir: str = "SMMP/"):
        self.setTemperature(T)
        if ff not in _FORCE_FIELDS:
            raise ValueError(f"unknown force field {ff!r}")
        flex, sh2 = _FORCE_FIELDS[ff]
        smmp.epar_l.flex = flex
        smmp.epar_l.sh2 = sh2
        smmp.epar_l.epsd = 0
        smmp.epar_l.ientyp = 0
        smmp.epar_l.tesgrd = 0
        smmp.isolty.itysol = st
        smmp.init_energy(libdir)
        self._objects = []
        self._box_size = 10
        smmp.updchois.upchswitch = 0
        smmp.updchois.rndord = 0
        smmp.updchois.bgsprob = 0

    def add(self, m) -> None:
        self._objects.append(m)

    def energy(self) -> float:
        """Total energy of this universe."""
        return smmp.energy()

    def rgyr(self):
        return smmp.rgyr(0)[0]

    def endToEnd(self):
        return smmp.rgyr(0)[1]

    def hbond(self):
        return sum(o.hbond() for o in self._objects) + smmp.interhbond()

    def helix(self):
        return smmp.helix()

    def temperature(self) -> float:
        return 1.0 / (float(smmp.bet.beta) * GASC)

    def setTemperature(self, T: float) -> None:
        """Sets beta = 1/(k_B T) for T > 0 (ignored otherwise, as the reference)."""
        if T > 0:
            smmp.bet.beta = 1.0 / (T * GASC)

    def boxSize(self):
        return self._box_size

    def objects(self):
        return self._objects

    def save(self) -> None:
        pass
