"""Drop-in replacement of the f2py extension module ``smmp`` (smmp.pyf).

The reference's Python layer (universe.py, protein.py, algorithms.py) calls
``smmp.<function>(...)`` and reads and writes COMMON-block variables as
``smmp.<block>.<field>``.  This module offers the same surface over the
MI355X engine's C-ABI (libsmmp_mi355x.so): one process-wide context holds
the COMMON state on the host; energies and Metropolis sweeps run on the GPU.

Differences from the f2py module (each stated where it applies):
  * errors raise ``smmp_amd._ffi.SmmpError`` instead of stopping the process;
  * ``metropolis``'s weight: ``smmp.can_weight`` and the objects of
    ``smmp_amd.weights`` run the whole sweep in the GPU kernel; any other
    Python callable is called twice per trial like the reference calls it,
    from a host-driven trial loop with each energy on the GPU (exact, one
    GPU energy per trial);
  * ``partem_s`` is exported (partem_s.f:13 has no f2py entry);
  * observables rgyr, helix, hbond, interhbond, zimmer, outpdb and the
    canon driver are provided, rmsdfun, and minim / gradient in vacuum;
    contacts, regul and anneal raise NotImplementedError.
"""
from __future__ import annotations

import os
import sys

import numpy as np

_here = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_here))
from smmp_amd import _ffi, weights  # noqa: E402

_ctx = _ffi.Context()


def context() -> _ffi.Context:
    """The process-wide engine context behind this module."""
    return _ctx


class _Block:
    """A COMMON block: attribute access yields numpy views of its fields
    (Fortran layout; scalars as 0-d arrays, as f2py gives them)."""

    def __init__(self, name: str):
        object.__setattr__(self, "_name", name)

    def _view(self, field: str) -> np.ndarray:
        v = _ctx.view(f"{self._name}.{field}")
        return v.reshape(()) if v.shape == (1,) and field in _SCALARS.get(self._name, ()) else v

    def __getattr__(self, field: str):
        if field.startswith("_"):
            raise AttributeError(field)
        return self._view(field)

    def __setattr__(self, field: str, value) -> None:
        self._view(field)[...] = value

    def __repr__(self) -> str:
        return f"<smmp COMMON block {self._name}>"


# scalar fields (rank 0 in the Fortran declarations)
_SCALARS = {
    "epar_l": ("flex", "tesgrd", "sh2", "epsd", "ientyp"),
    "isolty": ("itysol", "isolscl"),
    "updchois": ("rndord", "upchswitch", "bgsprob"),
    "bet": ("beta",),
    "bounds": ("boxsize",),
    "h_bond": ("ntyhb",),
    "updtparam": ("up2bmax", "up2bmin"),
    "mol_par": ("nvr",),
    "mol_i": ("ntlml",),
    "epar_r": ("conv", "plt", "slp"),
    "eny_r": ("eyel", "eyvw", "eyhb", "eyvr", "eysm", "eysl"),
    "pnt": ("npnt",),
    "v_sol": ("vdvol", "asa"),
    "myblock": ("mti",),
}


def _blocks() -> list[str]:
    n = _ffi.lib.smmp_state_count()
    names = {_ffi.lib.smmp_state_name(i).decode().split(".")[0] for i in range(n)}
    return sorted(names)


for _b in _blocks():
    globals()[_b] = _Block(_b)


# ---------------------------------------------------------------- functions
def init_energy(libdir: str) -> None:
    """smmp.pyf:4404, init_energy.f:14 (idempotent here)."""
    _ctx.init_energy(libdir)


def init_molecule(iabin: int, grpn: str, grpc: str, seqfile: str = " ", varfile: str = " ") -> None:
    """smmp.pyf:5011, init_molecule.f:13."""
    _ctx.init_molecule(int(iabin), grpn, grpc, seqfile, varfile)


def setvar(nml: int, vlvrx) -> None:
    """smmp.pyf:949, setvar.f:13 (host rebuild of molecule nml)."""
    _ctx.setvar(int(nml), np.asarray(vlvrx, dtype=np.float64))


def energy() -> float:
    """smmp.pyf:380, energy.f:12: total energy of the current variables (GPU)."""
    return _ctx.energy()


def enysol(nmol: int = 0) -> float:
    """smmp.pyf:664, enysol.f:13: solvent term of the current coordinates."""
    return _ctx.enysol(int(nmol))


def sgrnd(seed: int) -> None:
    """smmp.pyf:5620, twister.f:64."""
    _ctx.sgrnd(int(seed))


def grnd() -> float:
    """twister.f:104."""
    return _ctx.grnd()


def outvar(nml: int, fileout: str = "") -> None:
    """outvar.f:12: the variables of molecule nml (0 = all) in .var format."""
    _ctx.outvar(int(nml), fileout)


#: can_weight (canon.f:135-149): beta x with beta from /bet/; the sweep runs on the GPU
can_weight = weights.CanonicalCommon(_ctx)


def metropolis(eol1: float, acz: float, dummy) -> tuple[float, float]:
    """smmp.pyf:22, metropolis.f:13: one sweep over the free variables.
    ``dummy`` is the weight.  can_weight and the objects of smmp_amd.weights
    (canonical, multicanonical) carry their weight as data and the whole
    sweep runs in the GPU sweep kernel.  Any other callable is called as the
    reference calls it, twice per trial, from the host-driven loop
    (_metropolis_host): same draws, same decisions, one GPU energy per
    trial."""
    d = weights.as_descriptor(dummy)
    if d is None:
        return _metropolis_host(float(eol1), float(acz), dummy)
    return _ctx.metropolis(float(eol1), float(acz), d)


def _accanalyze(iupt: int, iupstate: int) -> None:
    """metropolis.f:186-193: counts of trials and accepted trials per type."""
    nc, na = updstats.ncalls, updstats.nacalls
    nc[4] += 1
    na[4] += iupstate
    nc[iupt - 1] += 1
    na[iupt - 1] += iupstate


def _accept(delta: float) -> bool:
    """metropolis.f:76-87: rd is drawn only when delta > 0."""
    if delta <= 0.0:
        return True
    rd = grnd()
    return float(np.exp(-min(delta, 100.0))) >= rd


def _metropolis_host(eol1: float, acz: float, dummy) -> tuple[float, float]:
    """metropolis.f:13-184 with a Python weight: the trial loop runs here in
    the reference's order (variable choice, update type, proposal, the
    callback on the new and the old energy, acceptance, restore), every
    energy() on the GPU, every random number from the context's twister."""
    nvr = int(mol_par.nvr)
    idv = _ctx.idvr
    ax = var_r.axvr
    vl = var_r.vlvr
    names = var_c.nmvr
    rndord, upch = int(updchois.rndord), int(updchois.upchswitch)
    eol = energy()
    for nsw in range(1, nvr + 1):
        ivar = nsw
        if rndord:
            ivar = min(1 + int(nvr * grnd()), nvr)  # u = 1 exactly would pass the end
        jv = int(idv[ivar - 1])
        if upch != 0 and names[0, jv - 1] != b"x":  # updtch1 / updtch2 (metropolis.f:196-220)
            b = float(bet.beta)
            prob = float(updchois.bgsprob) if upch == 1 else (b - float(updtparam.up2bmin)) / (
                float(updtparam.up2bmax) - float(updtparam.up2bmin))
            if grnd() < prob:  # BGS, inert in the reference (bgsnvar = 0)
                _accanalyze(2, 0)
                continue
        vrol = float(vl[jv - 1])
        dv = float(ax[jv - 1]) * (grnd() - 0.5)
        vl[jv - 1] = _addang(vrol, dv)
        enw = energy()
        if _accept(float(dummy(enw)) - float(dummy(eol))):
            eol, up = enw, 1
        else:
            vl[jv - 1] = vrol
            up = 0
        acz += up
        _accanalyze(1, up)
    gb = mol_r.gbpr
    box = float(bounds.boxsize)
    ntlml = int(mol_i.ntlml)
    for iml in range(ntlml if ntlml > 1 else 0):  # rigid moves of each chain (metropolis.f:94-168)
        for i in range(6):
            gbol = float(gb[i, iml])
            if i < 3:
                g = gbol + (grnd() - 0.5)
                if box > 0:
                    g = min(g, box) if g > box else max(g, -box)
            else:
                g = (grnd() - 0.5) * (np.pi if i == 4 else 2.0 * np.pi)
            gb[i, iml] = g
            enw = energy()
            if _accept(float(dummy(enw)) - float(dummy(eol))):
                eol, up = enw, 1
            else:
                gb[i, iml] = gbol
                up = 0
            acz += up
            _accanalyze(3 if i < 3 else 4, up)
    enw = energy()
    if abs(eol - enw) > 0.000001 and eol < 100000:
        raise _ffi.SmmpError(-4, "metropolis: eol and enw difference unacceptable")
    return eol, acz


def partem_s(num_rep: int, nequi: int, nswp: int, nmes: int = 1, nsave: int = 1, newsta: bool = True,
             switch: int = 1):
    """partem_s.f:13: parallel tempering on one random-number stream.
    Reads the 'temperatures' file of the current directory (partem_s.f:88-95)
    and writes 'time.d' every nmes sweeps in the reference's format (sweep,
    slot, temperature, energy, radius of gyration, C-alpha end-to-end
    distance: rgyr(0, ..) after the slot's sweep, partem_s.f:186-193).  At
    the end it writes the restart files 'par_R.in' and 'conf_NNNN.var'
    (partem_s.f:237-255); with newsta false it starts from them and runs
    sweeps nstart..nswp, nstart read from par_R.in (partem_s.f:97-111; the
    reference then leaves its replicas at zero torsions, this starts them
    from their conf files, and takes eol(k) from record k rather than the
    first value of every record).  Returns (eol, ipoi, energies per sweep and
    slot)."""
    temps = np.zeros(num_rep)
    with open("temperatures") as f:
        for _ in range(num_rep):
            j, t = f.readline().split()[:2]
            temps[int(j) - 1] = float(t)
    start, nstart = None, 1
    if not newsta:
        nstart, _ = read_par_r("par_R.in", num_rep)
        idvr = _ctx.idvr - 1
        start = np.stack([_ctx.redvar(conf_file_name(k + 1))[idvr] for k in range(num_rep)])
    nrun = max(int(nswp) - nstart + 1, 0)
    eol, ipoi, traj, obs, coor = _ctx.partem_s(temps, int(nequi), nrun, int(switch), with_obs=True, start=start,
                                               with_coor=True)
    with open("time.d", "w") as out:
        for sw in range(nrun):
            nsw = nstart + sw
            if nsw % max(int(nmes), 1) == 0:
                for k1 in range(num_rep):
                    out.write(f"{nsw:8d}{k1 + 1:3d}{temps[k1]:12.3f}{traj[sw, k1]:12.3f}"
                              f"{obs[sw, k1, 0]:12.3f}{obs[sw, k1, 1]:12.3f}\n")
    write_par_r("par_R.in", int(nswp), eol)
    keep = _ctx.vlvr()
    for k in range(num_rep):
        v = keep.copy()
        v[_ctx.idvr - 1] = coor[k]
        _ctx.set_vlvr(v)
        _ctx.outvar(0, conf_file_name(k + 1))
    _ctx.set_vlvr(keep)
    return eol, ipoi, traj


def conf_file_name(k: int) -> str:
    """fileNameMP('conf_0000.var', 6, 9, k) of utilities.f:141-170."""
    if k >= 10000:
        raise ValueError("too few characters available to label filenames")
    return f"conf_{k:04d}.var"


def write_par_r(path: str, nswp: int, eol) -> None:
    """par_R.in as partem_s.f:237-241 writes it: Fortran unformatted
    sequential records, per replica [nswp] and [eol(1..R)]."""
    e = np.ascontiguousarray(eol, dtype="<f8")
    with open(path, "wb") as f:
        for _ in range(len(e)):
            f.write(np.array([4, nswp, 4], dtype="<i4").tobytes())
            f.write(np.array([8 * len(e)], dtype="<i4").tobytes() + e.tobytes() + np.array([8 * len(e)], dtype="<i4").tobytes())


def read_par_r(path: str, num_rep: int):
    """(nstart, eol) from par_R.in: record 2k-1 holds nstart, record 2k the
    replicas' energies, of which replica k's is element k."""
    raw = open(path, "rb").read()
    pos, recs = 0, []
    while pos < len(raw):
        n = int(np.frombuffer(raw, dtype="<i4", count=1, offset=pos)[0])
        recs.append(raw[pos + 4:pos + 4 + n])
        pos += n + 8
    if len(recs) < 2 * num_rep:
        raise ValueError("par_R.in holds fewer replicas than num_rep")
    nstart = int(np.frombuffer(recs[0], dtype="<i4")[0])
    eol = np.array([np.frombuffer(recs[2 * k + 1], dtype="<f8")[k] if len(recs[2 * k + 1]) >= 8 * (k + 1)
                    else np.frombuffer(recs[2 * k + 1], dtype="<f8")[0] for k in range(num_rep)])
    return nstart, eol


def helix() -> tuple[int, int, int, int]:
    """smmp.pyf helix() -> (nhel, mhel, nbet, mbet), helix.f:13."""
    return _ctx.helix()


def rgyr(nml: int) -> tuple[float, float]:
    """smmp.pyf rgyr(nml) -> (rgy, ee), rgyr.f:13: radius of gyration and
    C-alpha end-to-end distance of chain nml (0: all) at the current
    coordinates."""
    return _ctx.rgyr(int(nml))


def _outside(name: str):
    def f(*args, **kwargs):
        raise NotImplementedError(f"smmp.{name}: outside the energy-and-move engine")
    f.__name__ = name
    return f


def hbond(nml: int, ipr: int = 0) -> int:
    """smmp.pyf:1806, hbond.f:14: H-bonds of molecule nml between residues
    more than four apart (mhb); the type table goes to h_bond.  ipr > 0
    (printing the bonds) is not provided."""
    return _ctx.hbond(int(nml))


def interhbond() -> int:
    """smmp.pyf:2092, hbond.f:228: H-bonds between chains (h_bond.mmhb)."""
    return _ctx.interhbond()


def zimmer(nresi: int) -> None:
    """smmp.pyf:3238, zimmer.f:15: Zimmerman codes into zimme.zimm."""
    _ctx.zimmer(int(nresi))


def outpdb(nml: int, filename: str = "") -> None:
    """smmp.pyf:3810, outpdb.f:12: PDB file of molecule nml (0: all)."""
    _ctx.outpdb(int(nml), filename)


def canon(nequi: int, nswp: int, nmes: int, temp: float, lrand: bool = True) -> None:
    """smmp.pyf:3522, canon.f:12-132: canonical Metropolis simulation at temp
    with can_weight.  Writes start.pdb, minen.pdb at each new minimum, every
    nmes sweeps a time.d line (sweep, energy, radius of gyration of chain 1,
    helix counts, H-bonds of the last chain, eyel, eyhb, eyvr, eysl in
    canon.f:110-112's format), then lastconf.var and final.pdb."""
    bet.beta = 1.0 / (temp * 1.98773e-3)
    if lrand:  # canon.f:62-67
        vl = _ctx.vlvr()
        ax = var_r.axvr
        for iv in _ctx.idvr:
            vl[iv - 1] = _addang(np.pi, float(ax[iv - 1]) * (grnd() - 0.5))
        _ctx.set_vlvr(vl)
    eol = energy()
    e_min = eol
    outpdb(0, "start.pdb")
    acz = 0.0
    for _ in range(int(nequi)):
        eol, acz = metropolis(eol, acz, can_weight)
    acz = 0.0
    ntlml = int(mol_i.ntlml)
    with open("time.d", "w") as td:
        for nsw in range(0, int(nswp) + 1):
            eol, acz = metropolis(eol, acz, can_weight)
            if eol < e_min:
                e_min = eol
                outpdb(0, "minen.pdb")
            if nsw % int(nmes) == 0:
                rgy, _ = rgyr(1)
                nhel, mhel, nbet, mbet = helix()
                mhb = 0
                for i in range(1, ntlml + 1):
                    mhb = hbond(i, 0)
                td.write(f"{nsw:7d}{eol:15.3f}{rgy:15.3f}{nhel:7d}{mhel:7d}{nbet:7d}{mbet:7d}{mhb:7d}"
                         f"{float(eny_r.eyel):15.3f}{float(eny_r.eyhb):15.3f}{float(eny_r.eyvr):15.3f}"
                         f"{float(eny_r.eysl):15.3f}\n")
    outvar(0, "lastconf.var")
    outpdb(0, "final.pdb")


def rmsdfun(nml: int, ir1: int, ir2: int, ixat, xrf, yrf, zrf, isl: int) -> float:
    """smmp.pyf:4095, rmsdfun.f:12: rmsd of residues ir1..ir2 of molecule nml
    from the reference xrf, yrf, zrf (ixat: reference index per atom) after
    the optimal superposition; isl 0 heavy atoms, 1 backbone, 2 C-alpha."""
    return _ctx.rmsdfun(nml, ir1, ir2, ixat, xrf, yrf, zrf, isl)


def minim(imin: int, maxit: int, eps: float) -> None:
    """smmp.pyf:1234, minim.f:13: local minimisation (imin 1 quasi-Newton,
    2 conjugate gradients) of the ECEPP/3 energy by the free variables, at
    most maxit function calls; the derivatives come from the GPU in vacuum
    (itysol 0; the reference stops for itysol > 0, and the device form of
    the analytic solvent's derivatives is switched off), without
    regularisation (ireg 0)."""
    _ctx.minim(int(imin), int(maxit), float(eps))


def gradient() -> None:
    """gradient.f:13 (vacuum): eny_r.eysm.. and eny_r.gdeyvr of the current
    variables."""
    _ctx.gradient()


for _n in ("contacts", "anneal", "regul", "esolan", "eninteract"):
    globals()[_n] = _outside(_n)


# ----------------------------------------------------------- multicanonical
class _Multicanonical:
    """The f2py module 'multicanonical' (mulcan_par_mod.f90): weight
    iteration and multicanonical simulation of the current molecule, one
    chain on the context's random-number stream, sweeps on the GPU."""

    def __init__(self):
        self.b = self.alpha = None
        self.kmin = self.kmax = 0
        self.ebin = 1.0

    @staticmethod
    def _random_start() -> None:
        # mulcan_par_mod.f90:93-98: vlvr(idvr(i)) = addang(pi, axvr (grnd() - 1/2))
        vl = _ctx.vlvr()
        ax = _ctx.view("var_r.axvr")
        for iv in _ctx.idvr:
            dv = float(ax[iv - 1]) * (grnd() - 0.5)
            vl[iv - 1] = _addang(np.pi, dv)
        _ctx.set_vlvr(vl)

    def mulcan_par(self, nsweep: int, nup: int, temp: float, kmin: int, kmax: int, binWidth: float,
                   l_iter: bool = False):
        """mulcan_par_mod.f90:23-190.  Files as the reference: mpar_full.d
        (i, xhist, g1, b, alpha, ent per bin after each update; read back when
        l_iter) and muca.d (i, b, alpha of the final weights).  The sweeps use
        muca_weight, (beta + b(k)) x + alpha(k), with beta of COMMON /bet/."""
        n = kmax - kmin + 1
        beta_t = 1.0 / (temp * 1.98773e-3)
        ihist = np.zeros(n, dtype=np.int32)
        if l_iter:
            rows = np.loadtxt("mpar_full.d", ndmin=2)
            xhist = np.zeros(n); g1 = np.zeros(n); b = np.zeros(n); alpha = np.zeros(n); ent = np.zeros(n)
            for r in rows[:n]:
                k = int(r[0]) - kmin
                xhist[k], g1[k], b[k], alpha[k], ent[k] = r[1:6]
        else:
            xhist, g1, b, alpha, ent = (np.zeros(n) for _ in range(5))
        self._random_start()
        eol = energy()
        acz = 0.0
        w = weights.Multicanonical(b, alpha, kmin, kmax, binWidth, beta=float(bet.beta))
        for nsw in range(1, nsweep + 1):
            w.beta = float(bet.beta)
            eol, acz = metropolis(eol, acz, w)
            ihist[w.bin(eol) - kmin] += 1
            if nsw % nup == 0:
                _ffi.lib.smmp_muca_update(kmin, kmax, binWidth, beta_t, _ffi.iptr(ihist), _ffi.dptr(xhist),
                                          _ffi.dptr(g1), _ffi.dptr(b), _ffi.dptr(alpha), _ffi.dptr(ent))
                with open("mpar_full.d", "w") as f:
                    for k in range(kmax, kmin - 1, -1):
                        i = k - kmin
                        f.write(f"{k:5d}{xhist[i]:15.4f}{g1[i]:15.4f}{b[i]:15.4f}{alpha[i]:15.4f}{ent[i]:15.4f}\n")
        _ffi.lib.smmp_muca_final(kmin, kmax, binWidth, beta_t, _ffi.dptr(b), _ffi.dptr(alpha))
        with open("muca.d", "w") as f:
            for k in range(kmin, kmax + 1):
                f.write(f"{k} {float(b[k - kmin])!r} {float(alpha[k - kmin])!r}\n")
        self.b, self.alpha, self.kmin, self.kmax, self.ebin = b, alpha, kmin, kmax, binWidth
        return eol, acz / (nsweep * int(mol_par.nvr)), xhist

    def mulcan_sim(self, nequi: int, nsweep: int, nmes: int, nsave: int, kmin: int, kmax: int, binWidth: float,
                   restart: bool = False):
        """mulcan_par_mod.f90:193-377 with the weights of muca.d (muca_weight2,
        b(k) x + alpha(k)).  start.d holds the restart state (sweep, energy,
        histogram, variables) every nsave sweeps; time.d gets 'nsw, eol' every
        nmes sweeps with the contact columns (outside this engine) as 0."""
        n = kmax - kmin + 1
        b = np.zeros(n); alpha = np.zeros(n)
        for r in np.loadtxt("muca.d", ndmin=2)[:n]:
            b[int(r[0]) - kmin], alpha[int(r[0]) - kmin] = r[1], r[2]
        w = weights.Multicanonical(b, alpha, kmin, kmax, binWidth)
        ihist = np.zeros(n, dtype=np.int64)
        xhist = np.zeros(n)
        nswm = 1
        if restart:
            with open("start.d") as f:
                nswm = int(float(f.readline().split()[0]))
                xhist[:] = [float(x) for x in f.readline().split()]
                vl = _ctx.vlvr()
                for _ in range(int(mol_par.nvr)):
                    j, x = f.readline().split()
                    vl[_ctx.idvr[int(j) - 1] - 1] = float(x)
                _ctx.set_vlvr(vl)
        else:
            self._random_start()
        eol = energy()
        acz = 0.0
        if not restart:
            for _ in range(nequi):
                eol, acz = metropolis(eol, acz, w)
        acz = 0.0
        with open("time.d", "w") as tf:
            for nsw in range(nswm, nsweep + 1):
                eol, acz = metropolis(eol, acz, w)
                ihist[w.bin(eol) - kmin] += 1
                if nsw % nsave == 0:
                    vl = _ctx.vlvr()
                    with open("start.d", "w") as f:
                        f.write(f"{nswm} {float(eol)!r}\n" + " ".join(repr(float(x)) for x in xhist) + "\n")
                        for i, iv in enumerate(_ctx.idvr, 1):
                            f.write(f"{i} {float(vl[iv - 1])!r}\n")
                if nsw % nmes == 0:
                    xhist += ihist
                    ihist[:] = 0
                    tf.write(f"{nsw:7d}{eol:12.2f}{0:8d}{0:8d}{0.0:12.4f}\n")
        return eol, acz / (max(nsweep - nswm + 1, 1) * int(mol_par.nvr)), xhist


multicanonical = _Multicanonical()


def _addang(a1: float, a2: float) -> float:
    """addang of difang.f:38-63: a1 + a2 wrapped into (-pi, pi]."""
    d = np.fmod(a1 + a2, 2.0 * np.pi)
    return float(d if abs(d) <= np.pi else d - np.copysign(2.0 * np.pi, d))
