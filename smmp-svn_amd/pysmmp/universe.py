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

    def sheet(self):
        """Candidate beta sheets between the proteins (universe.py:66-88):
        pairs (i, j, +1 parallel / -1 antiparallel) of proteins with at least
        two strand residues each, at least three H-bonds between them in
        h_bond.mmhb (as the last interhbond left it) and end-to-end
        directions with |cos| > 0.7."""
        min_strand, min_hb = 2, 3
        mmhb = smmp.h_bond.mmhb
        res = []
        objs = self._objects
        for i in range(len(objs)):
            if objs[i].strand() < min_strand:
                continue
            for j in range(i, len(objs)):
                if mmhb[objs[i].id()][objs[j].id()] >= min_hb and objs[j].strand() >= min_strand:
                    print(f"Found a candidate for a sheet between {i} and {j}.")
                    d1, d2 = objs[i].directionVector(), objs[j].directionVector()
                    c = sum(d1[k] * d2[k] for k in range(3))
                    if c > 0.7:
                        res.append([i, j, 1])
                    elif c < -0.7:
                        res.append([i, j, -1])
        return res

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
