"""Protein, Residue and Atom views of the COMMON state (protein.py of the
reference, same classes and methods, Python 3), over the drop-in ``smmp``.

Deviations: atom positions read the atom's own slot (the reference reads two
slots below it through the -1 lower bound of xat, INCL.H:237 vs
protein.py:42); atom lists include the last atom of the range (the
reference stops one short); residues() returns whole residue names (the
reference takes one character of each); saveVariables names fixed torsions
correctly (the reference reuses the previous variable's name); savePDB
writes this protein's chain (the reference passes its 0-based id to outpdb,
so the first protein writes every chain); directionVector normalises by the
vector's length (the reference's reduce leaves the first component
unsquared); strand counts the free variables of this chain only (the
reference's loop bound mixes variable and free-variable indices and runs
into the next chain).  PDB input,
contacts and rmsdfun are outside this engine."""
from __future__ import annotations

import sys

import numpy as np

import smmp

CRD = 57.2957795130823  # degrees per radian


def _chars(arr, i: int, n: int) -> str:
    return b"".join(arr[k][i] for k in range(n)).decode().strip()


def nursvr(ivr: int) -> int:
    """Residue (1-based) of variable index ivr (0-based), as nursvr.f:11-38
    (the reference's Python copy compares with '>' and misplaces the first
    variable of a residue)."""
    for i in range(int(smmp.mol_i.ntlml) - 1, -1, -1):
        ifirs = smmp.mol_i.irsml1[i] - 1
        if ivr + 1 >= smmp.res_i.ivrrs1[ifirs] and smmp.mol_i.nvrml[i] > 0:
            for j in range(smmp.mol_i.irsml2[i] - 1, ifirs - 1, -1):
                if ivr + 1 >= smmp.res_i.ivrrs1[j] and smmp.res_i.nvrrs[j] > 0:
                    return j + 1
    raise KeyError(f"nursvr: cannot find variable {ivr + 1}")


class Atom:
    """Read-only view of one atom (0-based index in SMMP's arrays)."""

    def __init__(self, id: int):
        self._id = id

    def position(self):
        # xat has lower bound -1 (INCL.H:237): Fortran xat(id+1) is element id+2 of the view
        return (float(smmp.atm_r.xat[self._id + 2]), float(smmp.atm_r.yat[self._id + 2]),
                float(smmp.atm_r.zat[self._id + 2]))

    def charge(self) -> float:
        return float(smmp.atm_r.cgat[self._id])

    def radius(self) -> float:
        return float(smmp.solvent.rvdw[self._id])

    def name(self) -> str:
        return _chars(smmp.atm_c.nmat, self._id, 4)

    def __str__(self) -> str:
        x, y, z = self.position()
        return f"{self.name()} at ({x}, {y}, {z})"


class Residue:
    def __init__(self, id: int):
        self._id = id

    def atoms(self):
        first = smmp.res_i.iatrs1[self._id]
        last = smmp.res_i.iatrs2[self._id]
        return [Atom(i) for i in range(first - 1, last)]


class Protein:
    """A polypeptide read from a sequence file and, optionally, a variable file."""

    def __init__(self, seq: str, var: str = " "):
        prev = int(smmp.mol_i.ntlml) - 1
        if seq[-3:].lower() == "pdb":
            raise NotImplementedError("Protein: PDB input is outside the energy-and-move engine")
        smmp.init_molecule(1, "nh2", "cooh", seq, var)
        self._id = int(smmp.mol_i.ntlml) - 1
        if self._id - prev > 1:
            raise NotImplementedError("more than one protein in one sequence file")

    def setOrigin(self, x, a) -> None:
        """Position of the first atom (x) and orientation angles (a) of the protein."""
        g = smmp.mol_r.gbpr
        for k in range(3):
            g[k][self._id] = x[k]
            g[3 + k][self._id] = a[k]

    def origin(self):
        g = smmp.mol_r.gbpr
        return [float(g[k][self._id]) for k in range(3)], [float(g[k][self._id]) for k in range(3, 6)]

    def _var_name(self, i: int) -> str:
        return _chars(smmp.var_c.nmvr, i, 3)

    def getDihedral(self, angle: str, residue) -> float:
        """Value in degrees of variable 'angle' of residue (index, or 'name n')."""
        try:
            idx = int(residue)
        except ValueError:
            name, n = residue.split()
            idx = self.findResidue(name.lower(), int(n))
        first = smmp.res_i.ivrrs1[idx]
        for i in range(first - 1, first - 1 + smmp.res_i.nvrrs[idx]):
            if self._var_name(i) == angle:
                return float(smmp.var_r.vlvr[i]) * CRD
        raise KeyError(f"{angle} not found for {residue}.")

    def findResidue(self, name: str, idx: int = 1) -> int:
        """Index of the idx-th residue named 'name'."""
        hits = [k for k, r in enumerate(self.residues()) if r.lower() == name.lower()]
        if len(hits) < idx:
            raise KeyError(f"{name}{idx} not found.")
        return hits[idx - 1] + int(smmp.mol_i.irsml1[self._id]) - 1

    def atoms(self):
        first = smmp.res_i.iatrs1[smmp.mol_i.irsml1[self._id] - 1]
        last = smmp.res_i.iatrs2[smmp.mol_i.irsml2[self._id] - 1]
        return [Atom(i) for i in range(first - 1, last)]

    def residues(self):
        first = smmp.mol_i.irsml1[self._id] - 1
        last = smmp.mol_i.irsml2[self._id]
        return [_chars(smmp.res_c.seq, i, 4) for i in range(first, last)]

    def id(self) -> int:
        return self._id

    def energy(self) -> float:
        return smmp.energy()

    def hbond(self):
        return smmp.hbond(self._id + 1, 0)

    def rgyr(self):
        return smmp.rgyr(self._id + 1)

    def contacts(self):
        return smmp.contacts()

    def savePDB(self, file: str = "") -> None:
        """The chain's coordinates in PDB format (outpdb.f), to file or stdout."""
        smmp.outpdb(self._id + 1, file)

    def strand(self) -> int:
        """Residues of this chain in a beta-strand-like configuration: free
        phi within 30 degrees of -150 with the next free variable (psi)
        within 30 of 150 (protein.py:184-205)."""
        max_dev, phi_sheet, psi_sheet = 30.0, -150.0, 150.0
        ifivr = int(smmp.mol_i.ivrml1[self._id])
        idvr = smmp.context().idvr
        nvr = len(idvr)
        count = 0
        i = 0
        while i < nvr and idvr[i] < ifivr:
            i += 1
        last = ifivr + int(smmp.mol_i.nvrml[self._id])  # first variable of the next chain
        for j in range(i, nvr):
            if idvr[j] >= last:
                break
            iv = int(idvr[j]) - 1
            if self._var_name(iv) == "phi" and j + 1 < nvr:
                phi = float(smmp.var_r.vlvr[iv]) * CRD
                psi = float(smmp.var_r.vlvr[int(idvr[j + 1]) - 1]) * CRD
                if abs(phi - phi_sheet) < max_dev and abs(psi - psi_sheet) < max_dev:
                    count += 1
        return count

    def directionVector(self):
        """Unit vector from the first to the last C-alpha atom of the chain
        (protein.py:207-221)."""
        beginning = end = None
        for a in self.atoms():
            if a.name().strip() == "ca":
                if beginning is None:
                    beginning = a.position()
                else:
                    end = a.position()
        if beginning is None or end is None:
            raise ValueError("directionVector: the chain needs two C-alpha atoms")
        v = np.array(end, dtype=np.float64) - np.array(beginning, dtype=np.float64)
        return list(v / np.linalg.norm(v))

    def zimmer(self) -> str:
        """Zimmerman codes of the chain's residues (zimmer.f, protein.py:265-270)."""
        first = int(smmp.mol_i.irsml1[self._id]) - 1
        last = int(smmp.mol_i.irsml2[self._id])
        smmp.zimmer(last)
        return smmp.zimme.zimm.ravel(order="F")[first:last].tobytes().decode()

    def saveVariables(self, file: str = "") -> None:
        """The conformation in the .var format (residue : name : degrees,
        '&' marks a fixed torsion), to file or standard output."""
        out = open(file, "w") if file else sys.stdout
        try:
            i0 = smmp.mol_i.ivrml1[self._id] - 1
            for i in range(i0, i0 + smmp.mol_i.nvrml[self._id]):
                name = self._var_name(i)
                value = float(smmp.var_r.vlvr[i]) * CRD
                if not smmp.var_l.fxvr[i]:
                    print(nursvr(i), ":", name, ":", value, file=out)
                elif smmp.var_i.ityvr[i] == 3:
                    print(nursvr(i), ":", name, ":", value, " &", file=out)
        finally:
            if file:
                out.close()

    def __len__(self) -> int:
        return int(smmp.mol_i.irsml2[self._id] - smmp.mol_i.irsml1[self._id] + 1)
