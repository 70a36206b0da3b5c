#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the SMMP Fortran reference.

Runs ``oracle/_ref/ref_driver`` (the reference's own objects compiled with
amdflang by ``make -C oracle ref``, linked with our driver) in a scratch
directory and stores its outputs as small ``.npz`` files (loadable with
``allow_pickle=False``).  Needs /root/reference; run it in the build container,
never on a GPU box.

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py multi      # only the named groups
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "smmp-svn_amd"))
from smmp_amd import workloads  # noqa: E402

DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")
LIBDIR = os.path.join(HERE, "SMMP") + "/"


def parse(text: str) -> dict[str, list]:
    """Parses '@name type count' blocks; repeated names become lists of arrays."""
    out: dict[str, list] = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i]
        if not line.startswith("@"):
            i += 1
            continue
        name, typ, n = line[1:].split()
        n = int(n)
        vals = lines[i + 1:i + 1 + n]
        if typ == "int":
            arr = np.array([int(v) for v in vals], dtype=np.int64)
        elif typ == "double":
            arr = np.array([float(v.replace("D", "E")) for v in vals], dtype=np.float64)
        else:
            arr = np.array([v.rstrip() for v in vals], dtype="<U8")
        out.setdefault(name, []).append(arr)
        i += n + 1
    return out


def _big_stack():
    # esolan.f keeps arrays of mxat*(mxat+1) elements as locals
    import resource
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run(args: list[str], cwd: str) -> dict[str, list]:
    res = subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, preexec_fn=_big_stack)
    if res.returncode != 0:
        raise RuntimeError(f"ref_driver {' '.join(args)} failed:\n{res.stdout[-2000:]}\n{res.stderr[-2000:]}")
    return parse(res.stdout)


def first(d: dict[str, list]) -> dict[str, np.ndarray]:
    return {k: v[0] for k, v in d.items()}


def vars_file(path: str, topo: dict[str, np.ndarray], varfile: str, fix_omega: bool) -> str:
    """Maps a reference .var file (residue : name : degrees, and 'om*:..&')
    onto variable indices for the driver's 'index value fixed' format."""
    nm = [s.strip() for s in topo["nmvr"]]
    ivrrs1 = topo["ivrrs1"]
    nvrrs = topo["nvrrs"]
    vals: dict[int, tuple[float, int]] = {}
    for line in open(varfile):
        line = line.strip()
        if not line or "#" in line:
            continue
        parts = [p.strip() for p in line.split(":")]
        if len(parts) != 3:
            continue
        r, name, val = int(parts[0]), parts[1], float(parts[2])
        for iv in range(ivrrs1[r - 1], ivrrs1[r - 1] + nvrrs[r - 1]):
            if nm[iv - 1] == name:
                vals[iv] = (val, 0)
    if fix_omega:
        for iv, name in enumerate(nm, 1):
            if name.startswith("om"):
                vals[iv] = (180.0, 1)
    with open(path, "w") as f:
        for iv in sorted(vals):
            f.write(f"{iv} {vals[iv][0]!r} {vals[iv][1]}\n")
    return path


def save(name: str, **arrays) -> None:
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def topo_fixture(tag: str, seq: list[str], tmp: str, nconf: int, seed: int,
                 varf: str = "-") -> dict[str, np.ndarray]:
    seqf = workloads.write_seq_file(os.path.join(tmp, f"{tag}.seq"), seq, tag)
    topo = first(run(["topo", LIBDIR, seqf, varf, "1"], tmp))
    ene = run(["energy", LIBDIR, seqf, varf, "1", str(nconf), str(seed)], tmp)
    arrays = dict(topo)
    arrays["seq_in"] = np.array(seq, dtype="<U8")
    arrays["conf_vlvr"] = np.stack(ene["vlvr"])
    arrays["conf_energy"] = np.stack(ene["energy"])
    arrays["conf_asa"] = np.stack(ene["asa"])
    arrays["conf_xyz"] = np.stack(ene["xyz"])
    arrays["conf_seed"] = np.array([seed])
    save(f"topo_{tag}.npz", **arrays)
    return topo


MULTI_CASES = {
    # two Met-enkephalins (EXAMPLES/enkefatimes2.seq), OONS, a box that the
    # second chain's position reaches (metropolis.f:100-106)
    "enk2": dict(chains=[["tyr", "gly", "gly", "phe", "met"]] * 2, itysol=1, boxsize=9.2,
                 gbpr={2: (9.0, -3.0, 2.5, 0.7, -0.4, 1.9)}, seed=4357, temp=300.0, nequi=30, nsweep=8),
    # three chains of different lengths, vacuum, no box
    "three": dict(chains=[["tyr", "gly", "gly", "phe", "met"], ["ala", "gly", "ser"], ["val", "lys", "trp", "glu"]],
                  itysol=0, boxsize=0.0,
                  gbpr={2: (7.5, 1.0, -2.0, -1.2, 0.3, 2.6), 3: (-2.0, 8.0, 3.5, 2.2, -1.1, -0.5)},
                  seed=81236, temp=400.0, nequi=30, nsweep=8),
}


def write_multi_seq(path: str, chains: list[list[str]]) -> str:
    """A sequence file with one '#'-headed block per chain (redseq.f)."""
    with open(path, "w") as f:
        for k, ch in enumerate(chains):
            f.write(f"#  chain {k + 1}\n")
            f.write("   " + " ".join(r.capitalize() for r in ch) + "\n")
    return path


def multi_fixture(tag: str, case: dict, tmp: str) -> None:
    seqf = write_multi_seq(os.path.join(tmp, f"multi_{tag}.seq"), case["chains"])
    gbf = os.path.join(tmp, f"multi_{tag}.gb")
    with open(gbf, "w") as f:
        for iml, vals in case["gbpr"].items():
            for k, v in enumerate(vals):
                f.write(f"{iml} {k + 1} {v!r}\n")
    out = run(["multi", LIBDIR, seqf, str(case["itysol"]), gbf, repr(case["boxsize"]), str(case["seed"]), "4",
               repr(case["temp"]), str(case["nequi"]), str(case["nsweep"])], tmp)
    arrays = {k: v[0] for k, v in out.items() if len(v) == 1}
    # energies: the start state, then the random conformations
    arrays["energies"] = np.stack(out["energy"])
    arrays["inter"] = np.stack(out["inter"])
    arrays["asa_all"] = np.stack(out["asa"])
    arrays["xyz_all"] = np.stack(out["xyz"])
    arrays["conf_vlvr"] = np.stack(out["vlvr"])
    arrays["sweeps"] = np.stack(out["sweep"])
    arrays["chains"] = np.array(["-".join(c) for c in case["chains"]], dtype="<U64")
    arrays["itysol_in"] = np.array([case["itysol"]])
    arrays["boxsize_in"] = np.array([case["boxsize"]])
    arrays["temp_in"] = np.array([case["temp"]])
    arrays["gbpr_in"] = np.array([[iml] + list(v) for iml, v in case["gbpr"].items()])
    save(f"multi_{tag}.npz", **arrays)


CDR = 1.745329251994330e-2  # degrees to radians as the reference's INCL.H has it


def esolan_rotation_fixture(tmp: str) -> None:
    """A Met-enkephalin conformation where one neighbour's surface passes
    through an atom's north point (esolan.f:188-216 then draws a random
    rotation): one torsion is bisected (on the CPU oracle's coordinates,
    which equal the reference's) onto the root of |N_i - c_k|^2 - R_k^2;
    the reference then evaluates the conformation (its variables in
    radians, applied by the start energy's own rebuild) and two random ones
    after it."""
    sys.path.insert(0, os.path.dirname(HERE))
    from conftest import make_ctx
    from _oracle import OracleSystem

    c = make_ctx(tag="enkefa", itysol=-1)
    o = OracleSystem(c)
    R = np.asarray(c.view("solvent.rvdw")[:o.nat], dtype=np.float64)
    rng = np.random.default_rng(1)

    def fmat(v):
        o.energy(v)
        x = o.coords()
        north = x.copy()
        north[:, 2] += R
        d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
        nb = (d < R[:, None] + R[None, :]) & (d > np.abs(R[:, None] - R[None, :]))
        np.fill_diagonal(nb, False)
        return ((north[:, None, :] - x[None, :, :]) ** 2).sum(-1) - R[None, :] ** 2, nb

    while True:
        deg = o.vlvr0 / CDR
        deg[o.idvr - 1] = rng.uniform(-180.0, 180.0, o.nvr)
        j = o.idvr[rng.integers(o.nvr)] - 1

        def at(dj):
            dd = deg.copy()
            dd[j] = dj
            v = o.vlvr0.copy()
            v[o.idvr - 1] = dd[o.idvr - 1] * CDR
            return fmat(v)
        fa, nba = at(deg[j])
        fb, nbb = at(deg[j] + 10.0)
        cand = np.argwhere(nba & nbb & (np.sign(fa) != np.sign(fb)))
        if len(cand):
            break
    i, k = cand[0]
    lo, hi, flo = deg[j], deg[j] + 10.0, fa[i, k]
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        fm = at(mid)[0][i, k]
        if np.sign(fm) == np.sign(flo):
            lo, flo = mid, fm
        else:
            hi = mid
    deg[j] = lo
    assert abs(at(lo)[0][i, k]) < 1e-9
    rad = deg[o.idvr - 1] * CDR
    vf = os.path.join(tmp, "esolan_rot.rad")
    with open(vf, "w") as f:
        for iv, x in zip(o.idvr, rad):
            f.write(f"{iv} {float(x)!r}\n")
    out = run(["esolan", LIBDIR, os.path.join(HERE, "enkefa.seq"), vf, "4357", "2"], tmp)
    save("esolan_rotation_enkefa.npz", var_idx=o.idvr.astype(np.int64), var_rad=rad,
         atoms=np.array([i + 1, k + 1]), seed=np.array([4357]), energy=np.stack(out["energy"]),
         mti=np.concatenate(out["mti"]), xyz0=out["xyz"][0], conf_vlvr=np.stack(out["vlvr"]))


def esolan_mc_fixtures(tmp: str) -> None:
    """Metropolis (canon) and serial parallel tempering (partem_s) with the
    analytic solvent, itysol -1: every trial's energy() runs esolan, whose
    random rotations share the trials' twister."""
    enk = [s.lower() for s in open(os.path.join(HERE, "enkefa.seq")).read().split()[2:]]
    seqf = os.path.join(HERE, "enkefa.seq")
    t_enk = first(run(["topo", LIBDIR, seqf, "-", "1"], tmp))
    vf = vars_file(os.path.join(tmp, "enk_es.vars"), t_enk, os.path.join(HERE, "enkefa.var"), False)
    out = run(["canon", LIBDIR, seqf, vf, "-1", "4357", "300.0", "8"], tmp)
    vars_in = np.loadtxt(vf, ndmin=2)
    save("canon_esolan_enkefa.npz", e0=out["canon_e0"][0], sweeps=np.stack(out["sweep"]), vlvr=out["vlvr"][0],
         mti=out["mti"][0], seed=np.array([4357]), temp=np.array([300.0]),
         var_idx=vars_in[:, 0].astype(np.int64), var_deg=vars_in[:, 1])
    from smmp_amd import workloads as wl
    seq = wl.random_sequence(10, 10)
    seqf = wl.write_seq_file(os.path.join(tmp, "pt_es.seq"), seq, "pt")
    nrep, nequi, nswp = 3, 2, 6
    temps = wl.geometric_temperatures(nrep, 250.0, 1000.0)
    wl.write_temperatures(os.path.join(tmp, "temperatures"), temps)
    run(["partem", LIBDIR, seqf, "-1", "4242", str(nrep), str(nequi), str(nswp), "-1"], tmp)
    td = np.loadtxt(os.path.join(tmp, "time.d"))
    raw = open(os.path.join(tmp, "par_R.in"), "rb").read()
    rec = 4 + 4 + 4 + 4 + 8 * nrep + 4
    assert len(raw) == nrep * rec, len(raw)
    eol_final = np.frombuffer(raw[16:16 + 8 * nrep], dtype="<f8").copy()
    save("partem_esolan_len10.npz", seq_in=np.array(seq, dtype="<U8"), temps=temps, time_d=td, eol_final=eol_final,
         seed=np.array([4242]), nequi=np.array([nequi]), nswp=np.array([nswp]), switch=np.array([-1]))


def restart_fixture(tmp: str, nrep: int, nswp: int, eol_final) -> None:
    """The restart files (par_R.in, conf_NNNN.var) of the partem_s run just
    made in tmp, as the reference wrote them."""
    confs = {f"conf_{k:04d}": np.frombuffer(open(os.path.join(tmp, f"conf_{k:04d}.var"), "rb").read(), dtype=np.uint8)
             for k in range(1, nrep + 1)}
    raw = open(os.path.join(tmp, "par_R.in"), "rb").read()
    save("partem_restart_len10.npz", par_r=np.frombuffer(raw, dtype=np.uint8), eol_final=np.asarray(eol_final),
         nswp=np.array([nswp]), **confs)


def muca_table(kmin: int, kmax: int, ebin: float, t_low: float, t_high: float) -> tuple[np.ndarray, np.ndarray]:
    """A continuous piecewise-linear multicanonical weight: slope b(k) =
    1/(k_B T_k) with T_k geometric from t_low (kmin) to t_high (kmax), alpha
    by the continuity rule of mulcan_par_mod.f90:147-150 (alpha(kmax) = 0)."""
    k = np.arange(kmin, kmax + 1)
    temps = t_low * (t_high / t_low) ** ((k - kmin) / (kmax - kmin))
    b = 1.0 / (temps * 1.98773e-3)
    al = np.zeros_like(b)
    for i in range(len(k) - 2, -1, -1):
        al[i] = al[i + 1] + (b[i + 1] - b[i]) * (k[i] + 0.5) * ebin
    return b, al


def metro_fixtures(tmp: str) -> None:
    """metropolis.f with the update-type draws (upchswitch 1 and 2, rndord)
    and with a multicanonical weight b(k)x + alpha(k), on Met-enkephalin with
    the 19 free torsions of enkefa.var (omegas fixed)."""
    enk = [s.lower() for s in open(os.path.join(HERE, "enkefa.seq")).read().split()[2:]]
    t_enk = topo_fixture("enkefa", enk, tmp, nconf=8, seed=11)
    vf_fix = vars_file(os.path.join(tmp, "enk_fix.vars"), t_enk, os.path.join(HERE, "enkefa.var"), True)
    seqf = os.path.join(tmp, "enkefa.seq")
    kmin, kmax, ebin = -15, 120, 1.0
    b, al = muca_table(kmin, kmax, ebin, 300.0, 1000.0)
    wf = os.path.join(tmp, "muca_w.txt")
    with open(wf, "w") as f:
        f.write(f"{kmin} {kmax} {ebin!r}\n")
        for x, y in zip(b, al):
            f.write(f"{float(x):.17g} {float(y):.17g}\n")
    runs = {"upch1": ("4357", "300.0", "20", "1", "0.3", "0", "-"),
            "upch2": ("81236", "300.0", "20", "2", "0.0", "1", "-"),
            "muca": ("81236", "300.0", "30", "0", "0.0", "0", wf),
            "muca_upch1": ("2024", "300.0", "15", "1", "0.2", "1", wf)}
    out = {"muca_kmin": np.array([kmin]), "muca_kmax": np.array([kmax]), "muca_ebin": np.array([ebin]),
           "muca_b": b, "muca_alpha": al}
    for tag, (seed, temp, nsw, upch, bgs, rnd, w) in runs.items():
        r = run(["metro", LIBDIR, seqf, vf_fix, "0", seed, temp, nsw, upch, bgs, rnd, w], tmp)
        out.update({f"{tag}_args": np.array([int(seed), float(temp), int(nsw), int(upch), float(bgs), int(rnd)]),
                    f"{tag}_e0": r["canon_e0"][0], f"{tag}_sweeps": np.stack(r["sweep"]),
                    f"{tag}_vlvr": r["vlvr"][0], f"{tag}_mti": r["mti"][0], f"{tag}_ncalls": r["ncalls"][0]})
    save("metro_enkefa.npz", **out)


OBS_CASES = {
    # (chains, seed, conformations, gbpr of chains 2.. or None)
    "c2_20mer": ([workloads.random_sequence(20, 2)], 77, 30, None),
    "len10": ([workloads.random_sequence(10, 10)], 78, 30, None),
    "pair": ([["gly", "ser", "asn", "gln", "ala"], ["thr", "gln", "asn", "ser"]], 79, 40,
             {2: (2.5, 1.5, 0.5, 0.4, -0.3, 1.1)}),
}


def obs_fixtures(tmp: str) -> None:
    """hbond, interhbond, zimmer and outpdb (hbond.f, zimmer.f, outpdb.f) of
    conformations drawn near helix and sheet angles."""
    for tag, (chains, seed, nconf, gb) in OBS_CASES.items():
        seqf = write_multi_seq(os.path.join(tmp, f"obs_{tag}.seq"), chains)
        prefix = os.path.join(tmp, f"obs_{tag}")
        args = ["obs", LIBDIR, seqf, "1", str(seed), str(nconf), prefix]
        if gb:
            gbf = os.path.join(tmp, f"obs_{tag}.gb")
            with open(gbf, "w") as f:
                for iml, vals in gb.items():
                    for k, v in enumerate(vals):
                        f.write(f"{iml} {k + 1} {v!r}\n")
            args += [gbf, "0.0"]
        out = run(args, tmp)
        nch = len(chains)
        hb = out["hbond"]  # nconf * nch blocks: mhb, ntyhb, ixtyhb.., nutyhb..
        mhb = np.array([h[0] for h in hb]).reshape(nconf, nch)
        ntyhb = np.array([h[1] for h in hb]).reshape(nconf, nch)
        types = np.zeros((nconf, nch, 2, 100), dtype=np.int64)
        for k, h in enumerate(hb):
            n = int(h[1])
            types[k // nch, k % nch, 0, :n] = h[2:2 + n]
            types[k // nch, k % nch, 1, :n] = h[2 + n:2 + 2 * n]
        inter = np.stack(out["interhb"])
        pdbs = [open(f"{prefix}_{k}.pdb").read() for k in (1, 2)]
        save(f"obs_{tag}.npz", chains=np.array(["-".join(c) for c in chains], dtype="<U128"),
             seed=np.array([seed]), gbpr_in=np.array([[iml] + list(v) for iml, v in (gb or {}).items()]).reshape(-1, 7),
             vlvr=np.stack(out["vlvr"]), mhb=mhb, ntyhb=ntyhb, hbtypes=types, inter_mhb=inter[:, 0],
             mmhb=inter[:, 1:], zimm=np.stack(out["zimm"]), pdb1=np.array(pdbs[0]), pdb2=np.array(pdbs[1]))


def grad_fixtures(tmp: str) -> None:
    """gradient.f / opeshe.f derivatives and minim.f runs (quasi-Newton
    minqsn.f and conjugate gradients mincjg.f) in vacuum: Met-enkephalin
    from enkefa.var (the README's minimisation) and random conformations of
    the 10-residue sequence."""
    enk = [s.lower() for s in open(os.path.join(HERE, "enkefa.seq")).read().split()[2:]]
    t_enk = topo_fixture("enkefa", enk, tmp, nconf=8, seed=11)
    vf = vars_file(os.path.join(tmp, "enk.vars"), t_enk, os.path.join(HERE, "enkefa.var"), False)
    seqf = os.path.join(tmp, "enkefa.seq")
    seq10 = list(np.load(os.path.join(HERE, "topo_len10.npz"))["seq_in"])
    seqf10 = workloads.write_seq_file(os.path.join(tmp, "grad_len10.seq"), seq10, "len10")
    out = {}
    for tag, (sf, varf, nconf, seed, epsd) in {"enkefa": (seqf, vf, 6, 515, 0), "enkefa_epsd": (seqf, vf, 4, 516, 1),
                                              "len10": (seqf10, "-", 6, 517, 0)}.items():
        r = run(["grad", LIBDIR, sf, varf, "0", str(nconf), str(seed), str(epsd)], tmp)
        out[f"{tag}_vlvr"] = np.stack(r["vlvr"])
        out[f"{tag}_energy"] = np.stack(r["grad_energy"])
        out[f"{tag}_grad"] = np.stack(r["gdeyvr"])
    save("grad_vacuum.npz", **out)
    out = {}
    for tag, (sf, varf, imin, maxit, eps, nconf, seed) in {
            "enkefa_qn": (seqf, vf, 1, 10000, 1e-9, 0, 1), "enkefa_cg": (seqf, vf, 2, 10000, 1e-9, 0, 1),
            "enkefa_cg60": (seqf, vf, 2, 60, 1e-9, 0, 1), "enkefa_qn_rand": (seqf, vf, 1, 25, 1e-9, 3, 520),
            "len10_qn": (seqf10, "-", 1, 30, 1e-9, 2, 518), "len10_cg": (seqf10, "-", 2, 30, 1e-9, 2, 519)}.items():
        r = run(["minim", LIBDIR, sf, varf, "0", str(imin), str(maxit), repr(eps), str(nconf), str(seed)], tmp)
        out[f"{tag}_args"] = np.array([imin, maxit, eps])
        out[f"{tag}_start"] = np.stack(r["vlvr"])
        out[f"{tag}_vlvr"] = np.stack(r["min_vlvr"])
        out[f"{tag}_energy"] = np.stack(r["min_energy"])
        out[f"{tag}_nfun"] = np.concatenate(r["min_nfun"])
    save("minim_vacuum.npz", **out)
    # the README's printed minimisation of Met-enkephalin (minim.f's move
    # lines): step, energy, sum of squared derivatives
    readme = "/root/reference/trunk/README"
    rows = re.findall(r"Step\s+(\d+): energy\s+(\S+)\s+\(\s*(\S+)\s*\)", open(readme).read())
    path = os.path.join(HERE, "readme_minim_steps.txt")
    with open(path, "w") as f:
        f.write("# the reference README's minimisation of Met-enkephalin (enkefa.var, vacuum, minim(1, 10000, 1e-9)):\n"
                "# step, energy, sum of squared derivatives as printed\n")
        for k, e, g in rows:
            f.write(f"{k} {float(e):.6E} {float(g):.6E}\n")
    print(f"wrote {path} ({len(rows)} steps)")


def grad_esolan_fixtures(tmp: str) -> None:
    """gradient.f with the analytic solvent (itysol -1: opeshe.f plus
    opesol.f over esolan.f's gradient) and minim.f runs on it: Met-enkephalin
    from enkefa.var and random conformations of it and of the 10-residue
    sequence (the random conformations and esolan's rotations share the
    driver's twister, so the vlvr sequence also pins the draws)."""
    enk = [s.lower() for s in open(os.path.join(HERE, "enkefa.seq")).read().split()[2:]]
    t_enk = topo_fixture("enkefa", enk, tmp, nconf=8, seed=11)
    vf = vars_file(os.path.join(tmp, "enk.vars"), t_enk, os.path.join(HERE, "enkefa.var"), False)
    seqf = os.path.join(tmp, "enkefa.seq")
    seq10 = list(np.load(os.path.join(HERE, "topo_len10.npz"))["seq_in"])
    seqf10 = workloads.write_seq_file(os.path.join(tmp, "grad_len10.seq"), seq10, "len10")
    out = {}
    for tag, (sf, varf, nconf, seed) in {"enkefa": (seqf, vf, 5, 531), "len10": (seqf10, "-", 5, 532)}.items():
        r = run(["grad", LIBDIR, sf, varf, "-1", str(nconf), str(seed), "0"], tmp)
        out[f"{tag}_seed"] = np.array([seed])
        out[f"{tag}_vlvr"] = np.stack(r["vlvr"])
        out[f"{tag}_energy"] = np.stack(r["grad_energy"])
        out[f"{tag}_grad"] = np.stack(r["gdeyvr"])
    for tag, (sf, varf, imin, maxit, eps, nconf, seed) in {
            "enkefa_qn": (seqf, vf, 1, 40, 1e-9, 0, 1), "enkefa_cg": (seqf, vf, 2, 40, 1e-9, 0, 1),
            "len10_qn": (seqf10, "-", 1, 20, 1e-9, 1, 533)}.items():
        r = run(["minim", LIBDIR, sf, varf, "-1", str(imin), str(maxit), repr(eps), str(nconf), str(seed)], tmp)
        out[f"min_{tag}_args"] = np.array([imin, maxit, eps, seed])
        out[f"min_{tag}_start"] = np.stack(r["vlvr"])
        out[f"min_{tag}_vlvr"] = np.stack(r["min_vlvr"])
        out[f"min_{tag}_energy"] = np.stack(r["min_energy"])
        out[f"min_{tag}_nfun"] = np.concatenate(r["min_nfun"])
    save("grad_esolan.npz", **out)


def rmsd_fixtures(tmp: str) -> None:
    """rmsdfun.f (quaternion superposition, rmsdfun/rmsdopt/fitmol/jacobi)
    of a start conformation against a turned, shifted and perturbed copy."""
    enk = [s.lower() for s in open(os.path.join(HERE, "enkefa.seq")).read().split()[2:]]
    t_enk = topo_fixture("enkefa", enk, tmp, nconf=8, seed=11)
    vf = vars_file(os.path.join(tmp, "enk.vars"), t_enk, os.path.join(HERE, "enkefa.var"), False)
    seq10 = list(np.load(os.path.join(HERE, "topo_len10.npz"))["seq_in"])
    seqf10 = workloads.write_seq_file(os.path.join(tmp, "rmsd_len10.seq"), seq10, "len10")
    out = {}
    for tag, sf, varf, seed in (("enkefa", os.path.join(tmp, "enkefa.seq"), vf, 31),
                                ("len10", seqf10, "-", 32)):
        r = run(["rmsd", LIBDIR, sf, varf, "0", str(seed)], tmp)
        out[f"{tag}_ref"] = np.asarray(r["rmsd_ref"][0]).reshape(-1, 3)
        out[f"{tag}_ixat"] = np.asarray(r["rmsd_ixat"][0], dtype=np.int32)
        out[f"{tag}_rmsd"] = np.asarray(r["rmsd"][0])
        out[f"{tag}_xyz"] = np.asarray(r["xyz"][0]).reshape(-1, 3)
    save("rmsd.npz", **out)


def main(only: list[str]) -> None:
    if not os.path.exists(DRIVER):
        raise SystemExit("build the reference first: make -C oracle ref")
    tmp = tempfile.mkdtemp(prefix="smmp_golden_")
    if not only or "multi" in only:
        for tag, case in MULTI_CASES.items():
            multi_fixture(tag, case, tmp)
    if not only or "esolan" in only:
        # the analytic solvent (itysol -1, esolan.f): start and random conformations
        cases = {"enkefa": ([s.lower() for s in open(os.path.join(HERE, "enkefa.seq")).read().split()[2:]], 8, 4357),
                 "c2_20mer": (workloads.random_sequence(20, 2), 6, 1234),
                 "len10": (workloads.random_sequence(10, 10), 6, 99)}
        for tag, (seq, nconf, seed) in cases.items():
            seqf = workloads.write_seq_file(os.path.join(tmp, f"esolan_{tag}.seq"), seq, tag)
            out = run(["esolan", LIBDIR, seqf, "-", str(seed), str(nconf)], tmp)
            save(f"esolan_{tag}.npz", seq_in=np.array(seq, dtype="<U8"), seed=np.array([seed]),
                 energy=np.stack(out["energy"]), mti=np.concatenate(out["mti"]), xyz0=out["xyz"][0],
                 conf_vlvr=np.stack(out["vlvr"]))
    if not only or "esolan" in only:
        esolan_rotation_fixture(tmp)
    if not only or "esolan_mc" in only:
        esolan_mc_fixtures(tmp)
    if not only or "helix" in only:  # helix.f on the 20-mer and the 10-mer
        for tag in ("c2_20mer", "len10"):
            seq = list(np.load(os.path.join(HERE, f"topo_{tag}.npz"))["seq_in"])
            seqf = workloads.write_seq_file(os.path.join(tmp, f"hx_{tag}.seq"), seq, tag)
            out = run(["helix", LIBDIR, seqf, "1", "31", "40"], tmp)
            save(f"helix_{tag}.npz", seq_in=np.array(seq, dtype="<U8"), vlvr=np.stack(out["vlvr"]),
                 helix=np.stack(out["helix"]))
    if not only or "obs" in only:
        obs_fixtures(tmp)
    if not only or "metro" in only:
        metro_fixtures(tmp)
    if not only or "grad" in only:
        grad_fixtures(tmp)
    if not only or "grad_esolan" in only:
        grad_esolan_fixtures(tmp)
    if not only or "rmsd" in only:
        rmsd_fixtures(tmp)
    if only and "restart" in only:  # the len10 run of partem_ext_len10.npz again, for its files
        seq = list(np.load(os.path.join(HERE, "topo_len10.npz"))["seq_in"])
        seqf = workloads.write_seq_file(os.path.join(tmp, "pt_len10.seq"), seq, "len10")
        temps = workloads.geometric_temperatures(4, 250.0, 1000.0)
        workloads.write_temperatures(os.path.join(tmp, "temperatures"), temps)
        run(["partem", LIBDIR, seqf, "1", "4242", "4", "3", "25", "-1"], tmp)
        raw = open(os.path.join(tmp, "par_R.in"), "rb").read()
        eol_final = np.frombuffer(raw[16:16 + 8 * 4], dtype="<f8").copy()
        assert np.array_equal(eol_final, np.load(os.path.join(HERE, "partem_ext_len10.npz"))["eol_final"])
        restart_fixture(tmp, 4, 25, eol_final)
    if only:
        return

    # MT19937 streams (twister.f): the default seed and a few others.
    mt = {}
    for s in (4357, 81236, 1, 2147483647, -5):
        u = run(["mt", str(s), "2000"], tmp)["mt"][0]
        mt[f"seed_{s}".replace("-", "m")] = np.rint(u * 4294967295.0).astype(np.uint64)
    save("mt19937.npz", **mt)

    # Met-enkephalin (EXAMPLES/enkefa.seq): topology, energies, the known answer.
    enk = [s.lower() for s in open(os.path.join(HERE, "enkefa.seq")).read().split()[2:]]
    t_enk = topo_fixture("enkefa", enk, tmp, nconf=8, seed=11)
    vf = vars_file(os.path.join(tmp, "enk.vars"), t_enk, os.path.join(HERE, "enkefa.var"), False)
    seqf = os.path.join(tmp, "enkefa.seq")
    kn = first(run(["topo", LIBDIR, seqf, vf, "0"], tmp))
    vf_fix = vars_file(os.path.join(tmp, "enk_fix.vars"), t_enk, os.path.join(HERE, "enkefa.var"), True)
    kn_fix = first(run(["topo", LIBDIR, seqf, vf_fix, "0"], tmp))
    canon = run(["canon", LIBDIR, seqf, vf_fix, "0", "81236", "300.0", "40"], tmp)
    canon_sol = run(["canon", LIBDIR, seqf, vf, "1", "4357", "500.0", "10"], tmp)
    save("enkefa_runs.npz",
         known_energy=kn["energy"], known_vlvr=kn["vlvr_all"], known_xyz=kn["xyz"],
         fixed_nvr=kn_fix["nvr"], fixed_idvr=kn_fix["idvr"], fixed_vlvr=kn_fix["vlvr_all"],
         fixed_axvr=kn_fix["axvr"], fixed_energy=kn_fix["energy"],
         canon_e0=canon["canon_e0"][0], canon_sweeps=np.stack(canon["sweep"]),
         canon_vlvr=canon["vlvr"][0], canon_mti=canon["mti"][0],
         canon_sol_e0=canon_sol["canon_e0"][0], canon_sol_sweeps=np.stack(canon_sol["sweep"]),
         canon_sol_vlvr=canon_sol["vlvr"][0])

    # Sequences of the benchmark shapes and edge cases of the topology builder.
    cases = {
        "c2_20mer": workloads.random_sequence(20, 2),
        "c5_36mer": workloads.random_sequence(36, 5),
        "len10": workloads.random_sequence(10, 10),
        "len60": workloads.random_sequence(60, 60),
        "glyends": ["gly", "trp", "his", "arg", "pro", "cys", "gly"],
        "pro_first": ["pro", "lys", "gln", "glu", "asn", "asp", "ile", "thr", "ser", "gly"],
        "allres": list(workloads.RESIDUES),
    }
    for tag, seq in cases.items():
        topo_fixture(tag, seq, tmp, nconf=6, seed=len(seq) + 7)

    # partem_s on a 10-mer: 4 replicas, time.d at every sweep, par_R.in at the end.
    seqf = workloads.write_seq_file(os.path.join(tmp, "pt.seq"), cases["len10"], "pt")
    temps = workloads.geometric_temperatures(4, 250.0, 1000.0)
    workloads.write_temperatures(os.path.join(tmp, "temperatures"), temps)
    run(["partem", LIBDIR, seqf, "1", "777", "4", "2", "12"], tmp)
    td = np.loadtxt(os.path.join(tmp, "time.d"))
    raw = open(os.path.join(tmp, "par_R.in"), "rb").read()
    # unformatted sequential: [4-byte len][int nsw][4-byte len][4-byte len][R doubles][4-byte len]
    nrep = 4
    rec = 4 + 4 + 4 + 4 + 8 * nrep + 4
    eol_final = np.frombuffer(raw[12 + 4:12 + 4 + 8 * nrep], dtype="<f8").copy()
    assert len(raw) == nrep * rec, len(raw)
    save("partem_len10.npz", temps=temps, time_d=td, eol_final=eol_final,
         seed=np.array([777]), nequi=np.array([2]), nswp=np.array([12]))

    # partem_s from the extended chain (switch < 0: every free torsion at 180
    # degrees), so the run stays in the well-conditioned energy range.
    for tag, nrep, nequi, nswp in (("len10", 4, 3, 25), ("c2_20mer", 4, 2, 10)):
        seqf = workloads.write_seq_file(os.path.join(tmp, f"pt_{tag}.seq"), cases[tag], tag)
        temps = workloads.geometric_temperatures(nrep, 250.0, 1000.0)
        workloads.write_temperatures(os.path.join(tmp, "temperatures"), temps)
        run(["partem", LIBDIR, seqf, "1", "4242", str(nrep), str(nequi), str(nswp), "-1"], tmp)
        td = np.loadtxt(os.path.join(tmp, "time.d"))
        raw = open(os.path.join(tmp, "par_R.in"), "rb").read()
        rec = 4 + 4 + 4 + 4 + 8 * nrep + 4
        assert len(raw) == nrep * rec, len(raw)
        eol_final = np.frombuffer(raw[16:16 + 8 * nrep], dtype="<f8").copy()
        save(f"partem_ext_{tag}.npz", temps=temps, time_d=td, eol_final=eol_final, seed=np.array([4242]),
             nequi=np.array([nequi]), nswp=np.array([nswp]), switch=np.array([-1]))
        if tag == "len10":
            restart_fixture(tmp, nrep, nswp, eol_final)


if __name__ == "__main__":
    main(sys.argv[1:])
