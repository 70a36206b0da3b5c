"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: it checks the HIP engine, it is never
part of the product path.  Systems are built from the arrays of a product
context (whose topology the CPU tests pin against the Fortran fixtures) or
from a fixture file directly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "oracle", "liboracle.so")

I32P = C.POINTER(C.c_int32)
F64P = C.POINTER(C.c_double)


class Topo(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("nat", "nvt", "nms", "npnt")]
                + [(n, I32P) for n in ("ityat", "iowat", "nbdat", "ibdat")]
                + [(n, F64P) for n in ("cgat", "blat", "baat", "toat")]
                + [(n, I32P) for n in ("iatvr", "ityvr", "iclvr", "iorvr", "imsvr1", "latms1", "latms2",
                                       "ivwat1", "ivwat2", "lvwat1", "lvwat2", "i14at1", "i14at2", "l14at")]
                + [(n, F64P) for n in ("sigma", "rvdw", "cij", "aij", "a14", "chb", "ahb")]
                + [("ihbty", I32P)]
                + [(n, F64P) for n in ("e0to", "sgto", "rnto", "spoint")]
                + [(n, C.c_double) for n in ("conv", "plt", "slp")]
                + [(n, C.c_int) for n in ("epsd", "itysol", "i2", "i3")]
                + [("gbpr", C.c_double * 6), ("nml", C.c_int)]
                + [(n, C.c_int * 10) for n in ("mol_a1", "mol_a2", "mol_i2", "mol_i3")]
                + [("gbm", C.c_double * 60), ("boxsize", C.c_double)]
                + [(n, I32P) for n in ("iadvr1", "ladvr")] + [("nad", C.c_int)])


class Conf(C.Structure):
    _fields_ = [(n, F64P) for n in ("toat", "sntoat", "cstoat", "snbaat", "csbaat", "xat", "yat", "zat",
                                    "xtoat", "ytoat", "ztoat", "xbaat", "ybaat", "zbaat")] + [
        ("e", C.c_double * 9), ("ei", C.c_double * 4), ("gb", C.c_double * 60)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_int)]


class Weight(C.Structure):
    _fields_ = [("kind", C.c_int), ("beta", C.c_double), ("kmin", C.c_int), ("kmax", C.c_int),
                ("ebin", C.c_double), ("b", F64P), ("alpha", F64P)]


def _load():
    if not os.path.exists(LIB):
        raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
    lib = C.CDLL(LIB)
    P = C.POINTER
    lib.oracle_sgrnd.argtypes = [P(MT), C.c_int32]
    lib.oracle_grnd.argtypes = [P(MT)]
    lib.oracle_grnd.restype = C.c_double
    lib.oracle_grnd_u32.argtypes = [P(MT)]
    lib.oracle_grnd_u32.restype = C.c_uint32
    lib.oracle_conf_alloc.argtypes = [P(Conf), C.c_int]
    lib.oracle_conf_free.argtypes = [P(Conf)]
    lib.oracle_conf_init.argtypes = [P(Topo), P(Conf)]
    lib.oracle_setvar.argtypes = [P(Topo), P(Conf), F64P]
    lib.oracle_energy.argtypes = [P(Topo), P(Conf), F64P]
    lib.oracle_energy.restype = C.c_double
    lib.oracle_gradient.argtypes = [P(Topo), P(Conf), F64P, F64P]
    lib.oracle_gradient.restype = C.c_double
    lib.oracle_weight_eval.argtypes = [P(Weight), C.c_double]
    lib.oracle_weight_eval.restype = C.c_double
    lib.oracle_metropolis.argtypes = [P(Topo), P(Conf), P(MT), F64P, C.c_int, I32P, F64P, P(Weight), F64P, F64P]
    lib.oracle_partem.argtypes = [P(Topo), P(MT), F64P, C.c_int, I32P, F64P, C.c_int, F64P, C.c_int, C.c_int,
                                  C.c_int, F64P, I32P, F64P, F64P]
    lib.oracle_set_updchois.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(C.c_int8), C.c_double]
    lib.oracle_updstats.argtypes = [F64P]
    lib.oracle_set_trace.argtypes = [F64P, C.c_int, C.c_int]
    lib.oracle_trace_count.restype = C.c_int
    lib.oracle_muca_update.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, I32P, F64P, F64P, F64P, F64P]
    lib.oracle_esolan.argtypes = [P(Topo), P(Conf), P(MT), F64P]
    lib.oracle_esolan.restype = C.c_double
    lib.oracle_esolan_grad.argtypes = [P(Topo), P(Conf), P(MT), F64P, F64P]
    lib.oracle_esolan_grad.restype = C.c_double
    lib.oracle_set_solvent_mt.argtypes = [P(MT)]
    lib.oracle_esolan_force_rotation.argtypes = [C.c_double, C.c_double]
    lib.oracle_difang.argtypes = [C.c_double, C.c_double]
    lib.oracle_difang.restype = C.c_double
    lib.oracle_addang.argtypes = [C.c_double, C.c_double]
    lib.oracle_addang.restype = C.c_double
    return lib


lib = _load()


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OracleMT:
    def __init__(self, seed: int | None = None):
        self.s = MT()
        self.s.mti = 625
        if seed is not None:
            lib.oracle_sgrnd(C.byref(self.s), seed)

    def u32(self, n: int) -> np.ndarray:
        return np.array([lib.oracle_grnd_u32(C.byref(self.s)) for _ in range(n)], dtype=np.uint64)

    def grnd(self) -> float:
        return lib.oracle_grnd(C.byref(self.s))


class OracleSystem:
    """The chains of a product context (smmp_amd._ffi.Context) for the oracle."""

    def __init__(self, ctx, initial_torsions: bool = False):
        """With initial_torsions the oracle starts from the torsions as the
        molecule was set up (before any rebuild wraps them into [-pi, pi]),
        the state the reference is in when its first energy() applies new
        variables."""
        self.ctx = ctx  # the arrays below may be views of the context's memory
        v = ctx.view
        toat_setup = np.array(v("atm_r.toat"), dtype=np.float64)
        # a rebuild refreshes mol_i.ixrfpt (the first backbone atoms), which
        # init_molecule leaves from before the terminal groups were added
        vl0 = np.ascontiguousarray(v("var_r.vlvr"), dtype=np.float64)
        from smmp_amd import _ffi
        nml = int(v("mol_i.ntlml")[0])
        for m in range(1, nml + 1):
            ctx.check(_ffi.lib.smmp_setvar(ctx.h, m, vl0.ctypes.data_as(F64P)))
        last = nml - 1
        nat = int(v("res_i.iatrs2")[int(v("mol_i.irsml2")[last]) - 1])
        nvt = int(v("mol_i.ivrml1")[last] + v("mol_i.nvrml")[last] - 1)
        nms = int(v("mol_i.imsml1")[last] + v("mol_i.nmsml")[last] - 1)
        nvw = int(v("eny_i.ivwml1")[last] + v("eny_i.nvwml")[last] - 1)
        n14 = int(v("eny_i.i14ml1")[last] + v("eny_i.n14ml")[last] - 1)
        nad = int(v("mol_i.iadml1")[last] + v("mol_i.nadml")[last] - 1)
        self.nml = nml
        self.nat, self.nvt = nat, nvt
        npnt = int(v("pnt.npnt")[0])
        itysol = int(v("isolty.itysol")[0])
        self.keep = {
            "ityat": i32(v("atm_i.ityat")[:nat]), "iowat": i32(v("atm_i.iowat")[:nat]),
            "nbdat": i32(v("atm_i.nbdat")[:nat]), "ibdat": i32(v("atm_i.ibdat").ravel(order="F")[:3 * nat]),
            "cgat": f64(v("atm_r.cgat")[:nat]), "blat": f64(v("atm_r.blat")[:nat]),
            "baat": f64(v("atm_r.baat")[:nat]), "toat": f64(v("atm_r.toat")[:nat]),
            "iatvr": i32(v("var_i.iatvr")[:nvt]), "ityvr": i32(v("var_i.ityvr")[:nvt]),
            "iclvr": i32(v("var_i.iclvr")[:nvt]), "iorvr": i32(v("var_i.iorvr")[:nvt]),
            "imsvr1": i32(v("var_i.imsvr1")[:nvt]), "latms1": i32(v("var_i.latms1")[:nms]),
            "latms2": i32(v("var_i.latms2")[:nms]),
            "ivwat1": i32(v("eny_i.ivwat1")[:nat]), "ivwat2": i32(v("eny_i.ivwat2")[:nat]),
            "lvwat1": i32(v("eny_i.lvwat1")[:max(nvw, 1)]), "lvwat2": i32(v("eny_i.lvwat2")[:max(nvw, 1)]),
            "i14at1": i32(v("eny_i.i14at1")[:nat]), "i14at2": i32(v("eny_i.i14at2")[:nat]),
            "l14at": i32(v("eny_i.l14at")[:max(n14, 1)]),
            "sigma": f64(v("solvent.sigma")[:nat]), "rvdw": f64(v("solvent.rvdw")[:nat]),
            "cij": f64(v("epar_r.cij").ravel(order="F")), "aij": f64(v("epar_r.aij").ravel(order="F")),
            "a14": f64(v("epar_r.a14").ravel(order="F")), "chb": f64(v("epar_r.chb").ravel(order="F")),
            "ahb": f64(v("epar_r.ahb").ravel(order="F")), "ihbty": i32(v("epar_i.ihbty").ravel(order="F")),
            "e0to": f64(v("epar_r.e0to")), "sgto": f64(v("epar_r.sgto")), "rnto": f64(v("epar_r.rnto")),
            "spoint": f64(v("tes.spoint")[:max(npnt, 1), :].T.ravel()),
            "iadvr1": i32(v("var_i.iadvr1")[:nvt]), "ladvr": i32(v("var_i.ladvr")[:max(nad, 1)]),
        }
        if initial_torsions:
            self.keep["toat"] = f64(toat_setup[:nat])
        t = Topo()
        t.nat, t.nvt, t.nms, t.npnt = nat, nvt, nms, npnt
        t.nad = nad
        for k, a in self.keep.items():
            setattr(t, k, a.ctypes.data_as(I32P if a.dtype == np.int32 else F64P))
        t.conv = float(v("epar_r.conv")[0])
        t.plt = float(v("epar_r.plt")[0])
        t.slp = float(v("epar_r.slp")[0])
        t.epsd = int(v("epar_l.epsd")[0])
        t.itysol = itysol
        ix = v("mol_i.ixrfpt")
        t.i2, t.i3 = int(ix[1, 0]), int(ix[2, 0])
        for k in range(6):
            t.gbpr[k] = float(v("mol_r.gbpr")[k, 0])
        t.nml = nml
        for m in range(nml):
            t.mol_a1[m] = int(v("res_i.iatrs1")[int(v("mol_i.irsml1")[m]) - 1])
            t.mol_a2[m] = int(v("res_i.iatrs2")[int(v("mol_i.irsml2")[m]) - 1])
            t.mol_i2[m], t.mol_i3[m] = int(ix[1, m]), int(ix[2, m])
            for k in range(6):
                t.gbm[6 * m + k] = float(v("mol_r.gbpr")[k, m])
        t.boxsize = float(v("bounds.boxsize")[0])
        self.t = t
        self.nvr = int(v("mol_par.nvr")[0])
        self.idvr = i32(v("var_i.idvr")[:self.nvr])
        self.axvr = f64(v("var_r.axvr")[:nvt])
        self.vlvr0 = f64(v("var_r.vlvr")[:nvt])
        self.c = Conf()
        if lib.oracle_conf_alloc(C.byref(self.c), nat) != 0:
            raise MemoryError
        lib.oracle_conf_init(C.byref(self.t), C.byref(self.c))
        lib.oracle_set_updchois(0, 0, 0.0, 4.0, 0.5, None, -1.0)  # trials in order, no update-type draws

    def __del__(self):
        if getattr(self, "c", None) is not None:
            lib.oracle_conf_free(C.byref(self.c))
            self.c = None

    def reset(self):
        lib.oracle_conf_init(C.byref(self.t), C.byref(self.c))

    def energy(self, vlvr_all, mt: "OracleMT | None" = None) -> np.ndarray:
        """[total, eysm, eyel, eyvw, eyhb, eyvr, eysl, asa, vdvol]; mt gives
        the draws of the analytic solvent's random rotations (itysol < 0)."""
        vl = f64(vlvr_all)
        lib.oracle_set_solvent_mt(C.byref(mt.s) if mt is not None else None)
        try:
            lib.oracle_energy(C.byref(self.t), C.byref(self.c), vl.ctypes.data_as(F64P))
        finally:
            lib.oracle_set_solvent_mt(None)
        return np.array(self.c.e[:], dtype=np.float64)

    def gradient(self, vlvr_all, mt: "OracleMT | None" = None) -> tuple[np.ndarray, np.ndarray]:
        """gradient.f: ([eysm, eyel, eyvw, eyhb, eyvr], gdeyvr[nvt]) in vacuum;
        with the analytic solvent (itysol < 0) ([eysm, eyel, eyvw, eyhb, eyvr,
        eysl], gdeyvr) with esolan's rotations drawn from mt."""
        vl = f64(vlvr_all)
        gd = np.zeros(self.nvt)
        lib.oracle_set_solvent_mt(C.byref(mt.s) if mt is not None else None)
        try:
            e = lib.oracle_gradient(C.byref(self.t), C.byref(self.c), vl.ctypes.data_as(F64P),
                                    gd.ctypes.data_as(F64P))
        finally:
            lib.oracle_set_solvent_mt(None)
        if not np.isfinite(e):
            raise ValueError("oracle gradient: vacuum or the analytic solvent of one chain only")
        n = 6 if self.t.itysol < 0 else 5
        return np.array(self.c.e[1:1 + n], dtype=np.float64), gd

    def setvar(self, vlvr_all) -> None:
        """Applies the variables and rebuilds (setvar.f), without an energy."""
        vl = f64(vlvr_all)
        lib.oracle_setvar(C.byref(self.t), C.byref(self.c), vl.ctypes.data_as(F64P))

    def esolan_gradient(self, mt: "OracleMT | None" = None) -> tuple[float, np.ndarray]:
        """Analytic solvent term and its Cartesian gradient (nat x 3,
        esolan.f's gradan) of the current coordinates."""
        g = np.zeros(3 * self.nat)
        e = lib.oracle_esolan_grad(C.byref(self.t), C.byref(self.c), C.byref(mt.s) if mt is not None else None,
                                   None, g.ctypes.data_as(F64P))
        return e, g.reshape(self.nat, 3)

    def esolan_areas(self, mt: "OracleMT | None" = None) -> tuple[float, np.ndarray]:
        """Analytic solvent term and per-atom areas of the current coordinates."""
        a = np.zeros(self.nat)
        e = lib.oracle_esolan(C.byref(self.t), C.byref(self.c), C.byref(mt.s) if mt is not None else None,
                              a.ctypes.data_as(F64P))
        return e, a

    def inter(self) -> np.ndarray:
        """[eysmi, eyeli, eyvwi, eyhbi] of the last energy"""
        return np.array(self.c.ei[:], dtype=np.float64)

    def gbpr(self) -> np.ndarray:
        return np.array(self.c.gb[:6 * self.nml], dtype=np.float64)

    def set_gbpr(self, gb) -> None:
        for k, x in enumerate(np.asarray(gb, dtype=np.float64).ravel()):
            self.c.gb[k] = float(x)

    def coords(self) -> np.ndarray:
        n = self.nat
        return np.stack([np.ctypeslib.as_array(self.c.xat, (n,)).copy(),
                         np.ctypeslib.as_array(self.c.yat, (n,)).copy(),
                         np.ctypeslib.as_array(self.c.zat, (n,)).copy()], 1)

    def set_updchois(self, rndord: int = 0, upchswitch: int = 0, bgsprob: float = 0.0, ubeta: float = -1.0,
                     up2bmax: float = 4.0, up2bmin: float = 0.5) -> None:
        """Trial order and update-type draws of the following metropolis
        calls (metropolis.f:42-57); clears the accanalyze counts."""
        names = self.ctx.view("var_c.nmvr")  # (3, mxvr) characters
        self._xname = np.array([names[0, k] == b"x" for k in range(self.nvt)], dtype=np.int8)
        lib.oracle_set_updchois(rndord, upchswitch, bgsprob, up2bmax, up2bmin,
                                self._xname.ctypes.data_as(C.POINTER(C.c_int8)), ubeta)

    @staticmethod
    def updstats() -> np.ndarray:
        """ncalls(1..5), nacalls(1..5) since set_updchois."""
        out = np.zeros(10)
        lib.oracle_updstats(out.ctypes.data_as(F64P))
        return out

    def metropolis(self, mt: OracleMT, vlvr, beta: float, nsweep: int, weight: Weight | None = None):
        vl = f64(vlvr).copy()
        w = weight or Weight(kind=0, beta=beta)
        eol = C.c_double(0.0)
        acz = C.c_double(0.0)
        out = []
        for _ in range(nsweep):
            bad = lib.oracle_metropolis(C.byref(self.t), C.byref(self.c), C.byref(mt.s), vl.ctypes.data_as(F64P),
                                        self.nvr, self.idvr.ctypes.data_as(I32P), self.axvr.ctypes.data_as(F64P),
                                        C.byref(w), C.byref(eol), C.byref(acz))
            out.append((eol.value, acz.value, bad))
        return vl, out

    def partem(self, mt: OracleMT, temps, nequi: int, nswp: int, switch: int = 1):
        R = len(temps)
        temps = f64(temps)
        eol = np.zeros(R)
        ipoi = np.zeros(R, dtype=np.int32)
        coor = np.zeros(R * self.nvr)
        traj = np.zeros(max(nswp, 1) * R)
        rc = lib.oracle_partem(C.byref(self.t), C.byref(mt.s), self.vlvr0.ctypes.data_as(F64P), self.nvr,
                               self.idvr.ctypes.data_as(I32P), self.axvr.ctypes.data_as(F64P), R,
                               temps.ctypes.data_as(F64P), nequi, nswp, switch, eol.ctypes.data_as(F64P),
                               ipoi.ctypes.data_as(I32P), coor.ctypes.data_as(F64P), traj.ctypes.data_as(F64P))
        return rc, eol, ipoi, coor.reshape(R, self.nvr), traj[:nswp * R].reshape(nswp, R)


class Spheres:
    """A bare set of spheres (centres, radii, solvation parameters) for the
    oracle's analytic surface, without a molecule around it."""

    def __init__(self, xyz, r, sigma=None):
        xyz = np.asarray(xyz, dtype=np.float64)
        self.n = len(xyz)
        self.keep = [f64(xyz[:, 0]), f64(xyz[:, 1]), f64(xyz[:, 2]), f64(r),
                     f64(np.ones(self.n) if sigma is None else sigma)]
        self.t = Topo()
        self.t.nat = self.n
        self.t.rvdw = self.keep[3].ctypes.data_as(F64P)
        self.t.sigma = self.keep[4].ctypes.data_as(F64P)
        self.c = Conf()
        self.c.xat, self.c.yat, self.c.zat = (a.ctypes.data_as(F64P) for a in self.keep[:3])

    def esolan(self, mt: OracleMT | None = None) -> tuple[float, np.ndarray]:
        a = np.zeros(self.n)
        e = lib.oracle_esolan(C.byref(self.t), C.byref(self.c), C.byref(mt.s) if mt is not None else None,
                              a.ctypes.data_as(F64P))
        return e, a

    def esolan_grad(self, mt: OracleMT | None = None) -> tuple[float, np.ndarray]:
        """(sum sigma_i As_i, gradan n x 3): the reference's Cartesian gradient."""
        g = np.zeros(3 * self.n)
        e = lib.oracle_esolan_grad(C.byref(self.t), C.byref(self.c), C.byref(mt.s) if mt is not None else None,
                                   None, g.ctypes.data_as(F64P))
        return e, g.reshape(self.n, 3)

    def move(self, j: int, k: int, h: float) -> None:
        """Shifts coordinate k of sphere j by h (in place)."""
        self.keep[k][j] += h


def force_rotation(uga: float, ufi: float) -> None:
    """Makes the oracle's esolan rotate every atom by these numbers (u_gamma,
    u_phi in [0, 1)) instead of drawing; negative values switch it off."""
    lib.oracle_esolan_force_rotation(uga, ufi)
