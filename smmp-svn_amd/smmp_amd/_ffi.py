"""ctypes binding of the C-ABI in include/smmp_mi355x.h.

Loads ``libsmmp_mi355x.so`` built in-tree (``smmp-svn_amd/csrc/build``).
There is no fallback: a missing or unloadable library raises ImportError.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIBPATH = os.environ.get("SMMP_LIB") or os.path.join(PKG, "csrc", "build", "libsmmp_mi355x.so")

SMMP_I32, SMMP_F64, SMMP_CHR = 1, 2, 3
ERRORS = {-1: "invalid argument", -2: "input error", -3: "device error", -4: "energy drift",
          -5: "state error"}


class SmmpError(RuntimeError):
    """An error status of the C-ABI (the reference stops the process instead)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Weight(C.Structure):
    _fields_ = [("kind", C.c_int32), ("beta", C.c_double), ("kmin", C.c_int32), ("kmax", C.c_int32),
                ("ebin", C.c_double), ("b", C.POINTER(C.c_double)), ("alpha", C.POINTER(C.c_double))]


class SamplerCfg(C.Structure):
    _fields_ = [("nrep", C.c_int32), ("nsweep", C.c_int32), ("mode", C.c_int32), ("precision", C.c_int32),
                ("temps", C.POINTER(C.c_double)), ("seeds", C.POINTER(C.c_uint32)),
                ("w", C.POINTER(Weight)), ("resync_every", C.c_int32), ("verify", C.c_int32),
                ("groups", C.c_int32), ("graph_sweeps", C.c_int32), ("gbpr", C.POINTER(C.c_double)),
                ("da", C.c_int32), ("da_scale", C.c_double), ("da_stats", C.POINTER(C.c_double)),
                ("snap_vlvr", C.POINTER(C.c_double)), ("snap_e", C.POINTER(C.c_double)), ("snap_every", C.c_int32),
                ("obs", C.POINTER(C.c_double))]


#: fg(user, x, &f, g) of smmp_minqsn / smmp_mincjg
FG_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError(f"libsmmp_mi355x.so not built ({LIBPATH}); run __graft_entry__.build()")
    lib = C.CDLL(LIBPATH)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    fp = C.POINTER(C.c_float)
    sig = {
        "smmp_ctx_create": (C.c_int, [C.POINTER(P)]),
        "smmp_ctx_destroy": (None, [P]),
        "smmp_last_error": (C.c_char_p, [P]),
        "smmp_state_view": (C.c_int, [P, C.c_char_p, C.POINTER(P), ip, ip, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "smmp_state_count": (C.c_int, []),
        "smmp_state_name": (C.c_char_p, [C.c_int]),
        "smmp_init_energy": (C.c_int, [P, C.c_char_p]),
        "smmp_init_molecule": (C.c_int, [P, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]),
        "smmp_setvar": (C.c_int, [P, C.c_int, dp]),
        "smmp_energy": (C.c_int, [P, dp]),
        "smmp_enysol": (C.c_int, [P, C.c_int, dp]),
        "smmp_sgrnd": (C.c_int, [P, C.c_int32]),
        "smmp_grnd": (C.c_int, [P, dp]),
        "smmp_outvar": (C.c_int, [P, C.c_int, C.c_char_p]),
        "smmp_metropolis": (C.c_int, [P, dp, dp, C.POINTER(Weight)]),
        "smmp_partem_s": (C.c_int, [P, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, ip, dp, dp, dp, dp]),
        "smmp_redvar": (C.c_int, [P, C.c_char_p, dp]),
        "smmp_rgyr": (C.c_int, [P, C.c_int, dp, dp]),
        "smmp_helix": (C.c_int, [P, ip]),
        "smmp_hbond": (C.c_int, [P, C.c_int, ip]),
        "smmp_interhbond": (C.c_int, [P, ip]),
        "smmp_zimmer": (C.c_int, [P, C.c_int]),
        "smmp_outpdb": (C.c_int, [P, C.c_int, C.c_char_p]),
        "smmp_rmsdfun": (C.c_int, [P, C.c_int, C.c_int, C.c_int, ip, dp, dp, dp, C.c_int, dp]),
        "smmp_gradient": (C.c_int, [P]),
        "smmp_gradient_batch": (C.c_int, [P, C.c_int, dp, dp, dp]),
        "smmp_minim": (C.c_int, [P, C.c_int, C.c_int, C.c_double, ip]),
        "smmp_minim_batch": (C.c_int, [P, C.c_int, C.c_int, dp, C.c_int, C.c_double, dp, ip]),
        "smmp_minqsn": (C.c_int, [C.c_int, dp, dp, dp, dp, C.c_double, C.c_int, FG_FN, C.c_void_p, ip]),
        "smmp_mincjg": (C.c_int, [C.c_int, dp, dp, dp, C.c_double, C.c_int, FG_FN, C.c_void_p, ip, ip]),
        "smmp_energy_batch": (C.c_int, [C.POINTER(P), C.c_int, ip, C.c_int, dp, C.c_int, dp]),
        "smmp_energy_batch_noalloc": (C.c_int, [C.POINTER(P), C.c_int, ip, C.c_int, dp, C.c_int, dp]),
        "smmp_batch_create": (C.c_int, [C.POINTER(P), C.c_int, ip, C.c_int, C.POINTER(P), C.POINTER(C.c_int64)]),
        "smmp_batch_load": (C.c_int, [P, dp, C.c_int64]),
        "smmp_batch_run": (C.c_int, [P, C.c_int]),
        "smmp_batch_tune": (C.c_int, [P, C.c_int, ip, ip]),
        "smmp_batch_read": (C.c_int, [P, dp]),
        "smmp_batch_destroy": (None, [P]),
        "smmp_sample": (C.c_int, [P, C.POINTER(SamplerCfg), dp, dp, dp, ip, ip, dp]),
        "smmp_build_info": (C.c_char_p, []),
        "smmp_profile": (C.c_int, [P, C.c_int]),
        "smmp_profile_read": (C.c_int, [P, C.c_int, dp, C.POINTER(C.c_int64), dp]),
        "smmp_profile_phases": (C.c_int, [P, dp]),
        "smmp_sampler_info": (C.c_int, [P, ip]),
        "smmp_mlp_size": (C.c_int, [P]),
        "smmp_mlp_layout": (C.c_int, [P, ip]),
        "smmp_mlp_set": (C.c_int, [P, fp]),
        "smmp_mlp_get": (C.c_int, [P, fp]),
        "smmp_mlp_grad": (C.c_int, [P, dp, C.c_int, fp, C.c_int, fp, fp, dp]),
        "smmp_mlp_fit": (C.c_int, [P, dp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_uint32, dp]),
        "smmp_muca_update": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, ip, dp, dp, dp, dp, dp]),
        "smmp_muca_final": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, dp, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

#: every symbol the header declares (checked by the CPU tests)
EXPORTS = ("smmp_ctx_create", "smmp_ctx_destroy", "smmp_last_error", "smmp_state_view", "smmp_state_count",
           "smmp_state_name", "smmp_init_energy", "smmp_init_molecule", "smmp_setvar", "smmp_energy",
           "smmp_enysol", "smmp_sgrnd", "smmp_grnd", "smmp_outvar", "smmp_metropolis", "smmp_partem_s", "smmp_rgyr", "smmp_redvar", "smmp_helix",
           "smmp_hbond", "smmp_interhbond", "smmp_zimmer", "smmp_outpdb", "smmp_rmsdfun",
           "smmp_gradient", "smmp_gradient_batch", "smmp_minim", "smmp_minim_batch", "smmp_minqsn", "smmp_mincjg",
           "smmp_energy_batch", "smmp_energy_batch_noalloc", "smmp_batch_create", "smmp_batch_load", "smmp_batch_run", "smmp_batch_tune", "smmp_batch_read",
           "smmp_batch_destroy", "smmp_sample", "smmp_build_info", "smmp_profile", "smmp_profile_read",
           "smmp_profile_phases", "smmp_muca_update", "smmp_muca_final", "smmp_sampler_info",
           "smmp_mlp_size", "smmp_mlp_layout", "smmp_mlp_set", "smmp_mlp_get", "smmp_mlp_grad", "smmp_mlp_fit")

KERNEL_SWEEP, KERNEL_EXCHANGE, KERNEL_ENERGY, KERNEL_ESOLAN, KERNEL_GRADIENT = 0, 1, 2, 3, 4


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


# Live handles, torn down in order by close_all() at interpreter exit.
_live_batches: "weakref.WeakSet" = weakref.WeakSet()
_live_contexts: "weakref.WeakSet" = weakref.WeakSet()


def close_all() -> None:
    """Destroys every live resident batch, then every live context (graph
    executables, events, streams, device memory).  Registered with atexit,
    so this runs before the interpreter finalises objects in arbitrary order
    and before the HIP runtime's and a profiler's own exit handlers."""
    for b in list(_live_batches):
        b.close()
    for c in list(_live_contexts):
        c.close()


atexit.register(close_all)


class Context:
    """Owns one smmp_ctx (one system: the reference's COMMON blocks)."""

    def __init__(self):
        h = C.c_void_p()
        self.check(lib.smmp_ctx_create(C.byref(h)), None)
        self.h = h
        _live_contexts.add(self)

    def close(self) -> None:
        """Releases the context and its device state now (idempotent)."""
        h = getattr(self, "h", None)
        if h:
            self.h = None
            lib.smmp_ctx_destroy(h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    def check(self, rc: int, h=None) -> None:
        if rc != 0:
            hh = h if h is not None else getattr(self, "h", None)
            msg = lib.smmp_last_error(hh).decode() if hh else ""
            raise SmmpError(rc, msg)

    def view(self, name: str) -> np.ndarray:
        """numpy view (Fortran order) of a COMMON variable, e.g. 'atm_r.xat'."""
        ptr = C.c_void_p()
        dt = C.c_int32()
        nd = C.c_int32()
        shape = (C.c_int64 * 2)()
        lb = (C.c_int64 * 2)()
        self.check(lib.smmp_state_view(self.h, name.encode(), C.byref(ptr), C.byref(dt), C.byref(nd), shape, lb))
        if dt.value == SMMP_I32:
            ct, npdt = C.c_int32, np.int32
        elif dt.value == SMMP_F64:
            ct, npdt = C.c_double, np.float64
        else:
            ct, npdt = C.c_char, "S1"
        if nd.value == 0:
            n = 1
            shp = (1,)
        elif nd.value == 1:
            n = shape[0]
            shp = (shape[0],)
        else:
            n = shape[0] * shape[1]
            shp = (shape[0], shape[1])
        buf = (ct * n).from_address(ptr.value)
        arr = np.ctypeslib.as_array(buf) if npdt != "S1" else np.frombuffer(buf, dtype="S1", count=n)
        arr = arr.view(npdt) if npdt != "S1" else arr
        return arr.reshape(shp, order="F")

    def init_energy(self, libdir: str) -> None:
        self.check(lib.smmp_init_energy(self.h, libdir.encode()))

    def init_molecule(self, iabin: int, grpn: str, grpc: str, seqfile: str, varfile: str = " ") -> None:
        self.check(lib.smmp_init_molecule(self.h, iabin, grpn.encode(), grpc.encode(), seqfile.encode(),
                                          varfile.encode()))

    def energy(self) -> float:
        e = C.c_double()
        self.check(lib.smmp_energy(self.h, C.byref(e)))
        return e.value

    # ---- convenience accessors of the chain
    @property
    def nvt(self) -> int:
        v = self.view
        nt = int(v("mol_i.ntlml")[0])
        return int(v("mol_i.ivrml1")[nt - 1] + v("mol_i.nvrml")[nt - 1] - 1) if nt else 0

    @property
    def nvr(self) -> int:
        return int(self.view("mol_par.nvr")[0])

    @property
    def idvr(self) -> np.ndarray:
        return self.view("var_i.idvr")[:self.nvr].copy()

    def vlvr(self) -> np.ndarray:
        return self.view("var_r.vlvr")[:self.nvt].copy()

    def set_vlvr(self, vl) -> None:
        """Writes the variables (the next energy()/metropolis() builds from them)."""
        self.view("var_r.vlvr")[:self.nvt] = np.asarray(vl, dtype=np.float64)[:self.nvt]

    def enysol(self, nmol: int = 0) -> float:
        e = C.c_double()
        self.check(lib.smmp_enysol(self.h, int(nmol), C.byref(e)))
        return e.value

    def outvar(self, nml: int, path: str = "") -> None:
        self.check(lib.smmp_outvar(self.h, int(nml), path.encode()))

    def setvar(self, nml: int, vl) -> None:
        full = np.array(self.view("var_r.vlvr"), dtype=np.float64)
        full[:len(vl)] = vl
        self.check(lib.smmp_setvar(self.h, nml, dptr(full)))

    PHASES = ("propose", "rotate", "pairs", "solvent", "reduce", "accept", "full", "sweep_end")

    def profile(self, enable=True) -> None:
        """enable: False/0 off, True/1 kernel timing and work, 2 also sweep phases."""
        self.check(lib.smmp_profile(self.h, int(enable)))

    def sampler_info(self) -> dict:
        """How the last sample() ran (smmp_sampler_info)."""
        out = np.zeros(4, dtype=np.int32)
        self.check(lib.smmp_sampler_info(self.h, iptr(out)))
        return {"groups": int(out[0]), "persistent": bool(out[1]), "graph_replays": int(out[2]),
                "graph_reused": bool(out[3])}

    def profile_phases(self) -> dict:
        out = np.zeros(22)
        self.check(lib.smmp_profile_phases(self.h, dptr(out)))
        d = {k: out[i] for i, k in enumerate(self.PHASES)}
        d["full_trials"] = out[8]
        d["core_mhz"] = out[9] / (out[10] * 1e3) if out[10] > 0 else 0.0
        for i, k in enumerate(("sol_max_wave", "sol_gather", "sol_points", "sol_atoms", "sol_contacts",
                               "pair_max_wave", "pair_sum", "mark_sum", "torsion", "sync_wait", "sync_latency")):
            d[k] = out[11 + i]
        return d

    def profile_read(self, kernel: int) -> dict:
        ms = C.c_double()
        n = C.c_int64()
        work = np.zeros(4)
        self.check(lib.smmp_profile_read(self.h, kernel, C.byref(ms), C.byref(n), dptr(work)))
        return {"ms": ms.value, "launches": n.value, "pairs": work[0], "contacts": work[1],
                "builds": work[2], "moved": work[3], "work": work.copy()}

    def sgrnd(self, seed: int) -> None:
        self.check(lib.smmp_sgrnd(self.h, int(np.int32(np.uint32(seed & 0xffffffff)))))

    def grnd(self) -> float:
        u = C.c_double()
        self.check(lib.smmp_grnd(self.h, C.byref(u)))
        return u.value

    def metropolis(self, eol: float, acz: float, weight: Weight) -> tuple[float, float]:
        e = C.c_double(eol)
        a = C.c_double(acz)
        self.check(lib.smmp_metropolis(self.h, C.byref(e), C.byref(a), C.byref(weight)))
        return e.value, a.value

    def partem_s(self, temps, nequi: int, nswp: int, switch: int = 1, with_obs: bool = False, start=None,
                 with_coor: bool = False):
        """(eol, ipoi, energies per sweep and slot[, radius of gyration and
        end-to-end distance per sweep and slot][, final free variables per
        replica]).  start (R x nvr free variables): restart from them (no
        equilibration, no random start)."""
        temps = np.ascontiguousarray(temps, dtype=np.float64)
        R = len(temps)
        eol = np.zeros(R)
        ipoi = np.zeros(R, dtype=np.int32)
        traj = np.zeros(max(nswp, 1) * R)
        obs = np.zeros(max(nswp, 1) * R * 2) if with_obs else None
        st = np.ascontiguousarray(start, dtype=np.float64).reshape(R, self.nvr) if start is not None else None
        coor = np.zeros((R, self.nvr)) if with_coor else None
        rc = lib.smmp_partem_s(self.h, R, nequi, nswp, switch, dptr(temps), dptr(eol), iptr(ipoi), dptr(traj),
                               dptr(obs) if obs is not None else None, dptr(st) if st is not None else None,
                               dptr(coor) if coor is not None else None)
        self.check(rc)
        out = [eol, ipoi, traj[:nswp * R].reshape(nswp, R)]
        if with_obs:
            out.append(obs[:nswp * R * 2].reshape(nswp, R, 2))
        if with_coor:
            out.append(coor)
        return tuple(out)

    def helix(self) -> tuple[int, int, int, int]:
        """(helical residues, helical segments, sheet residues, sheet segments)."""
        out = np.zeros(4, dtype=np.int32)
        self.check(lib.smmp_helix(self.h, iptr(out)))
        return tuple(int(x) for x in out)

    def hbond(self, nml: int) -> int:
        """hbond.f: H-bonds of molecule nml (>= 1) between residues more than
        four apart; the type table is left in h_bond.ntyhb/ixtyhb/nutyhb."""
        m = C.c_int32()
        self.check(lib.smmp_hbond(self.h, int(nml), C.byref(m)))
        return int(m.value)

    def interhbond(self) -> int:
        """H-bonds between chains (h_bond.mmhb per pair of chains)."""
        m = C.c_int32()
        self.check(lib.smmp_interhbond(self.h, C.byref(m)))
        return int(m.value)

    def zimmer(self, nresi: int) -> str:
        """Zimmerman code of residues 1..nresi (zimme.zimm)."""
        self.check(lib.smmp_zimmer(self.h, int(nresi)))
        return self.view("zimme.zimm").ravel(order="F")[:nresi].tobytes().decode()

    def gradient(self) -> tuple[float, np.ndarray]:
        """gradient.f (vacuum): eysm and the derivatives by all variables
        (also left in eny_r)."""
        self.check(lib.smmp_gradient(self.h))
        return float(self.view("eny_r.eysm")[0]), np.array(self.view("eny_r.gdeyvr")[:self.nvt])

    def gradient_batch(self, vlvr) -> tuple[np.ndarray, np.ndarray]:
        """Energies (eysm, eyel, eyvw, eyhb, eyvr) and derivatives of many
        full variable vectors (rows of vlvr) in one launch."""
        vl = np.ascontiguousarray(vlvr, dtype=np.float64).reshape(-1, self.nvt)
        n = len(vl)
        e = np.zeros((n, 5))
        g = np.zeros((n, self.nvt))
        self.check(lib.smmp_gradient_batch(self.h, n, dptr(vl), dptr(e), dptr(g)))
        return e, g

    def minim(self, imin: int, maxit: int, eps: float) -> int:
        """minim.f: minimises from the current variables; returns the calls used."""
        k = C.c_int32(0)
        self.check(lib.smmp_minim(self.h, int(imin), int(maxit), float(eps), C.byref(k)))
        return int(k.value)

    def minim_batch(self, imin: int, vlvr, maxit: int, eps: float):
        """minim.f for many starts side by side: (best variables, their
        energies, calls used)."""
        vl = np.ascontiguousarray(vlvr, dtype=np.float64).reshape(-1, self.nvt).copy()
        n = len(vl)
        f = np.zeros(n)
        k = np.zeros(n, dtype=np.int32)
        self.check(lib.smmp_minim_batch(self.h, int(imin), n, dptr(vl), int(maxit), float(eps), dptr(f), iptr(k)))
        return vl, f, k

    def rmsdfun(self, nml: int, ir1: int, ir2: int, ixat, xrf, yrf, zrf, isl: int) -> float:
        """rmsdfun.f: rmsd after superposition on the reference coordinates."""
        ix = np.ascontiguousarray(ixat, dtype=np.int32)
        xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xrf, yrf, zrf)]
        out = C.c_double(0.0)
        self.check(lib.smmp_rmsdfun(self.h, int(nml), int(ir1), int(ir2), iptr(ix), dptr(xs[0]), dptr(xs[1]),
                                    dptr(xs[2]), int(isl), C.byref(out)))
        return float(out.value)

    def outpdb(self, nml: int, path: str = "") -> None:
        self.check(lib.smmp_outpdb(self.h, int(nml), path.encode()))

    def redvar(self, path: str) -> np.ndarray:
        """All variables after reading a .var file over the current ones
        (the context is not changed)."""
        v = np.zeros(self.nvt)
        self.check(lib.smmp_redvar(self.h, path.encode(), dptr(v)))
        return v

    def rgyr(self, nml: int = 0) -> tuple[float, float]:
        """(radius of gyration, C-alpha end-to-end distance) of the current coordinates."""
        rgy, ee = C.c_double(), C.c_double()
        self.check(lib.smmp_rgyr(self.h, int(nml), C.byref(rgy), C.byref(ee)))
        return rgy.value, ee.value

    def sample(self, temps, seeds, nsweep: int, vlvr, mode: int = 0, weight: Weight | None = None,
               precision: int = 0, with_traj: bool = True, allow_drift: bool = False, verify: bool = False,
               groups: int = 0, graph_sweeps: int = 0, gbpr=None, da: bool = False,
               da_scale: float = 1.0, snap_every: int = 0, obs: bool = False):
        """Independent-stream replica sampler (smmp_sample).  groups: workgroups
        per replica (0 automatic, -1 measured); graph_sweeps: sweeps per
        replay of a captured graph (0: direct launches); gbpr: several chains,
        R x ntlml x 6 start values of the chains' global parameters (returned
        as "gbpr"; None: every replica starts from the context's); da: delayed
        acceptance with the context's network (per-replica counts returned as
        "da_stats"); snap_every > 0: each
        replica's variables and energy every snap_every sweeps ("snap_vlvr",
        "snap_e"); obs: radius of gyration and end-to-end distance after every
        sweep ("obs", nsweep x R x 2, slot order)."""
        temps = np.ascontiguousarray(temps, dtype=np.float64)
        R = len(temps)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        need = R + (1 if mode == 1 else 0)
        if len(seeds) < need:
            raise ValueError(f"need {need} seeds")
        vl = np.ascontiguousarray(vlvr, dtype=np.float64).reshape(R, self.nvt).copy()
        eol = np.zeros(R)
        acc = np.zeros(R)
        ipoi = np.zeros(R, dtype=np.int32)
        nb = (weight.kmax - weight.kmin + 1) if weight is not None else 1
        hist = np.zeros(nb, dtype=np.int32)
        traj = np.zeros(max(nsweep, 1) * R) if with_traj else None
        cfg = SamplerCfg(nrep=R, nsweep=nsweep, mode=mode, precision=precision,
                         temps=dptr(temps), seeds=seeds.ctypes.data_as(C.POINTER(C.c_uint32)),
                         w=C.pointer(weight) if weight is not None else None, resync_every=1,
                         verify=1 if verify else 0, groups=groups, graph_sweeps=graph_sweeps)
        da_stats = None
        if da:
            da_stats = np.zeros((R, 4))
            cfg.da = 1
            cfg.da_scale = float(da_scale)
            cfg.da_stats = dptr(da_stats)
        snap_v = snap_e = None
        if snap_every > 0:
            ns = nsweep // snap_every
            snap_v = np.zeros((max(ns, 1), R, self.nvt))
            snap_e = np.zeros((max(ns, 1), R))
            cfg.snap_vlvr, cfg.snap_e, cfg.snap_every = dptr(snap_v), dptr(snap_e), snap_every
        ob = None
        if obs:
            ob = np.zeros((max(nsweep, 1), R, 2))
            cfg.obs = dptr(ob)
        nml = int(self.view("mol_i.ntlml")[0])
        gb = None
        if gbpr is not None:
            gb = np.ascontiguousarray(gbpr, dtype=np.float64).reshape(R, nml, 6).copy()
            cfg.gbpr = dptr(gb)
        rc = lib.smmp_sample(self.h, C.byref(cfg), dptr(vl), dptr(eol), dptr(acc), iptr(ipoi),
                             iptr(hist) if weight is not None else None,
                             dptr(traj) if traj is not None else None)
        if rc != 0 and not (allow_drift and rc == -4):
            self.check(rc)
        if snap_v is not None:
            snap_v, snap_e = snap_v[:nsweep // snap_every], snap_e[:nsweep // snap_every]
        return {"vlvr": vl, "eol": eol, "acc": acc, "ipoi": ipoi, "hist": hist, "rc": rc, "gbpr": gb,
                "da_stats": da_stats, "snap_vlvr": snap_v, "snap_e": snap_e,
                "obs": ob[:nsweep] if ob is not None else None,
                "traj": traj[:nsweep * R].reshape(nsweep, R) if traj is not None else None}


def energy_batch(ctxs, conf_sys, vlvr_flat, precision: int = 0) -> np.ndarray:
    """Energies of many conformations (rows: total, eysm, eyel, eyvw, eyhb, eyvr, eysl, asa)."""
    arr = (C.c_void_p * len(ctxs))(*[c.h for c in ctxs])
    cs = np.ascontiguousarray(conf_sys, dtype=np.int32)
    vl = np.ascontiguousarray(vlvr_flat, dtype=np.float64)
    out = np.zeros((len(cs), 8))
    rc = lib.smmp_energy_batch(arr, len(ctxs), iptr(cs), len(cs), dptr(vl), precision, dptr(out))
    ctxs[0].check(rc)
    return out


def energy_batch_noalloc(ctxs, conf_sys, vlvr_flat, precision: int = 0) -> np.ndarray:
    """energy_batch through the entry point that allocates no device memory
    (smmp_energy_batch_noalloc)."""
    arr = (C.c_void_p * len(ctxs))(*[c.h for c in ctxs])
    cs = np.ascontiguousarray(conf_sys, dtype=np.int32)
    vl = np.ascontiguousarray(vlvr_flat, dtype=np.float64)
    out = np.zeros((len(cs), 8))
    rc = lib.smmp_energy_batch_noalloc(arr, len(ctxs), iptr(cs), len(cs), dptr(vl), precision, dptr(out))
    ctxs[0].check(rc)
    return out


class EnergyBatch:
    """Resident batch (smmp_batch_*): the systems of ctxs and the conformations
    stay on the device; load() fills them, run() evaluates, read() returns
    rows total, eysm, eyel, eyvw, eyhb, eyvr, eysl, asa.  The contexts must
    outlive the batch."""

    def __init__(self, ctxs, conf_sys):
        self._ctxs = list(ctxs)
        arr = (C.c_void_p * len(self._ctxs))(*[c.h for c in self._ctxs])
        self.conf_sys = np.ascontiguousarray(conf_sys, dtype=np.int32)
        h = C.c_void_p()
        nv = C.c_int64()
        self._ctxs[0].check(lib.smmp_batch_create(arr, len(self._ctxs), iptr(self.conf_sys), len(self.conf_sys),
                                                  C.byref(h), C.byref(nv)))
        self.h, self.nvals = h, int(nv.value)
        _live_batches.add(self)

    def load(self, vlvr_flat) -> None:
        vl = np.ascontiguousarray(vlvr_flat, dtype=np.float64).ravel()
        self._vl = vl  # kept alive while the copy may still read it
        self._ctxs[0].check(lib.smmp_batch_load(self.h, dptr(vl), vl.size))

    def run(self, precision: int = 0) -> None:
        self._ctxs[0].check(lib.smmp_batch_run(self.h, precision))

    def tune(self, precision: int = 0) -> list[tuple[int, int]]:
        """Measured launch shapes (smmp_batch_tune): (threads, bits in LDS) per class."""
        shapes = np.zeros(4, dtype=np.int32)
        n = C.c_int32(0)
        self._ctxs[0].check(lib.smmp_batch_tune(self.h, precision, iptr(shapes), C.byref(n)))
        return [(int(shapes[2 * k]), int(shapes[2 * k + 1])) for k in range(n.value)]

    def read(self) -> np.ndarray:
        out = np.zeros((len(self.conf_sys), 8))
        self._ctxs[0].check(lib.smmp_batch_read(self.h, dptr(out)))
        return out

    def close(self) -> None:
        h = getattr(self, "h", None)
        if h:
            self.h = None
            lib.smmp_batch_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _minimise(kind: int, fg, x, scal, acur: float, maxfun: int):
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    n = len(x)
    f0, g0 = fg(x.copy())
    g = np.ascontiguousarray(g0, dtype=np.float64).copy()
    f = C.c_double(float(f0))
    err = []

    def cb(_user, xp, fp, gp):
        try:
            fv, gv = fg(np.ctypeslib.as_array(xp, (n,)).copy())
            fp[0] = float(fv)
            gv = np.asarray(gv, dtype=np.float64)
            for i in range(n):
                gp[i] = float(gv[i])
        except Exception as e:  # noqa: BLE001  (reported after the call)
            err.append(e)
            fp[0] = float("nan")

    cfun = FG_FN(cb)
    k = C.c_int32(0)
    ier = C.c_int32(0)
    if kind == 1:
        s = np.ascontiguousarray(np.ones(n) if scal is None else scal, dtype=np.float64)
        rc = lib.smmp_minqsn(n, dptr(x), C.byref(f), dptr(g), dptr(s), float(acur), int(maxfun), cfun, None,
                             C.byref(k))
    else:
        rc = lib.smmp_mincjg(n, dptr(x), C.byref(f), dptr(g), float(acur), int(maxfun), cfun, None, C.byref(k),
                             C.byref(ier))
    if err:
        raise err[0]
    if rc != 0:
        raise SmmpError(rc, "minimiser arguments")
    return x, float(f.value), g, int(k.value), int(ier.value)


def minqsn(fg, x, scal=None, acur: float = 1e-9, maxfun: int = 1000):
    """minqsn.f on fg(x) -> (f, gradient): (x, f, g, calls) at the best point."""
    x, f, g, k, _ = _minimise(1, fg, x, scal, acur, maxfun)
    return x, f, g, k


def mincjg(fg, x, acur: float = 1e-9, maxfun: int = 1000):
    """mincjg.f on fg(x) -> (f, gradient): (x, f, g, calls, ier)."""
    return _minimise(2, fg, x, None, acur, maxfun)
