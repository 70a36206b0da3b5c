"""Scenarios run through the drop-in Python API (smmp-svn_amd/pysmmp:
smmp, universe, protein, algorithms) in a fresh process each, since the
smmp module holds one global state like the reference's COMMON blocks.
Usage: python dropin_scenarios.py <scenario> <json args>; prints one JSON line."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "smmp-svn_amd", "pysmmp"))
GOLDEN = os.path.join(HERE, "golden")
LIBDIR = os.path.join(GOLDEN, "SMMP") + "/"


def setup_cpu(args):
    import smmp
    import protein
    import universe
    u = universe.Universe(T=300, st=1, libdir=LIBDIR)
    p = protein.Protein(os.path.join(GOLDEN, "enkefa.seq"), os.path.join(GOLDEN, "enkefa.var"))
    u.add(p)
    out = {"T": u.temperature(), "beta": float(smmp.bet.beta), "len": len(p), "residues": p.residues(),
           "phi1": p.getDihedral("phi", 0), "psi2": p.getDihedral("psi", "gly 1"),
           "nvr": int(smmp.mol_par.nvr), "itysol": int(smmp.isolty.itysol),
           "natoms": len(p.atoms()), "first_atom": p.atoms()[0].name()}
    p.setOrigin((1.0, 2.0, 3.0), (0.1, 0.2, 0.3))
    out["origin"] = p.origin()
    p.saveVariables(args["varout"])
    rgy, ee = smmp.rgyr(0)
    out["rgyr"] = [rgy, ee, u.rgyr(), u.endToEnd() if hasattr(u, "endToEnd") else ee]
    # the same from the coordinates: all atoms, and the C-alpha atoms of the
    # first and the last residue
    xyz = np.array([a.position() for a in p.atoms()])
    names = [a.name().strip().lower() for a in p.atoms()]
    ca = [i for i, nm in enumerate(names) if nm.startswith("ca")]
    r0, r1 = int(smmp.res_i.iatrs2[0]), int(smmp.res_i.iatrs1[len(p) - 1]) - 1
    c_first = max(i for i in ca if i < r0)
    c_last = max(i for i in ca if i >= r1)
    out["rgyr_np"] = [float(np.sqrt(((xyz - xyz.mean(0)) ** 2).sum(1).mean())),
                      float(np.linalg.norm(xyz[c_last] - xyz[c_first]))]
    from smmp_amd import weights
    out["beta_from_can_weight"] = weights.as_descriptor(smmp.can_weight).beta
    out["lambda_descriptor"] = repr(weights.as_descriptor(lambda x: x * smmp.bet.beta))  # None: host loop
    return out


def sheet_cpu(args):
    """Two strands (every free phi -150, psi 150) side by side: strand
    counts, direction vectors, Zimmerman codes, hbond through Protein and
    Universe, and sheet() with the chain-pair H-bond table set by hand."""
    import smmp
    import protein
    import universe
    u = universe.Universe(T=300, st=0, libdir=LIBDIR)
    seq = os.path.join(args["dir"], "s.seq")
    with open(seq, "w") as f:
        f.write("#  strand\nGly Ala Ser Val Thr Ala Gly\n")  # '#': a new molecule per file (redseq.f:58-74)
    p1 = protein.Protein(seq)
    p2 = protein.Protein(seq)
    u.add(p1)
    u.add(p2)
    v = smmp.var_r.vlvr
    for i in smmp.context().idvr:
        nm = b"".join(smmp.var_c.nmvr[:, i - 1]).decode()
        if nm == "phi":
            v[i - 1] = np.radians(-150.0)
        elif nm == "psi":
            v[i - 1] = np.radians(150.0)
    smmp.mol_r.gbpr[0, 1] = 4.8   # the second chain 4.8 A along x
    for m in (1, 2):              # coordinates of both chains (host rebuild)
        smmp.setvar(m, v)
    nphi = sum(1 for i in smmp.context().idvr if b"".join(smmp.var_c.nmvr[:, i - 1]) == b"phi")
    out = {"strand": [p1.strand(), p2.strand()], "nphi": nphi,
           "dir": [p1.directionVector(), p2.directionVector()], "zimmer": [p1.zimmer(), p2.zimmer()],
           "hbond": [p1.hbond(), p2.hbond()], "u_hbond": u.hbond()}
    out["sheet_none"] = u.sheet() if int(smmp.h_bond.mmhb[0, 1]) < 3 else "skip"
    smmp.h_bond.mmhb[0, 1] = 3
    out["sheet"] = u.sheet()
    return out


def canonical_gpu(args):
    import smmp
    import algorithms
    import protein
    import universe
    u = universe.Universe(T=300, st=0, libdir=LIBDIR)
    p = protein.Protein(os.path.join(GOLDEN, "enkefa.seq"), os.path.join(GOLDEN, "enkefa_c1.var"))
    u.add(p)
    smmp.sgrnd(args["seed"])
    if args.get("weight") == "lambda":  # the reference's default weight, through the host-driven loop
        mc = algorithms.CanonicalMonteCarlo(u, 0, args["sweeps"], weight=lambda x: x * smmp.bet.beta)
    else:
        mc = algorithms.CanonicalMonteCarlo(u, 0, args["sweeps"])
    energies, accs = [], []
    for _ in range(args["sweeps"]):
        mc.sweep()
        energies.append(u.energy())
        accs.append(mc._acc)
    return {"energies": energies, "acc": accs, "rate": mc.acceptanceRate(), "e0": mc.minimumEnergy()[0]}


def callback_gpu(args):
    """smmp.metropolis with a Python weight callable (smmp.pyf:3-10) over the
    table of tests/golden/metro_enkefa.npz, the update-type switches set."""
    import smmp
    import protein
    import universe
    g = np.load(os.path.join(GOLDEN, "metro_enkefa.npz"), allow_pickle=False)
    tag = args["tag"]
    seed, temp, nsw, upch, bgs, rnd = g[f"{tag}_args"]
    u = universe.Universe(T=float(temp), st=0, libdir=LIBDIR)
    u.add(protein.Protein(os.path.join(GOLDEN, "enkefa.seq"), os.path.join(GOLDEN, "enkefa_c1.var")))
    smmp.updchois.upchswitch = int(upch)
    smmp.updchois.bgsprob = float(bgs)
    smmp.updchois.rndord = int(rnd)
    kmin, ebin = int(g["muca_kmin"][0]), float(g["muca_ebin"][0])
    b, al = g["muca_b"], g["muca_alpha"]
    kmax = kmin + len(b) - 1

    def muca(x):  # muca_weight2 as a plain Python function
        k = int(min(kmax, max(kmin, x / ebin + np.copysign(0.5, x))))
        return b[k - kmin] * x + al[k - kmin]

    w = muca if tag.startswith("muca") else (lambda x: x * smmp.bet.beta)
    smmp.sgrnd(int(seed))
    eol, acz, sweeps = 0.0, 0.0, []
    for _ in range(int(nsw)):
        eol, acz = smmp.metropolis(eol, acz, w)
        sweeps.append([eol, acz])
    v = smmp.var_r.vlvr
    return {"sweeps": sweeps, "vlvr": [float(v[i - 1]) for i in smmp.context().idvr],
            "mti": int(smmp.myblock.mti), "ncalls": [int(x) for x in smmp.updstats.ncalls] +
            [int(x) for x in smmp.updstats.nacalls]}


def canon_gpu(args):
    """smmp.canon (canon.f) from enkefa_c1.var without a random start."""
    import smmp
    import protein
    import universe
    os.chdir(args["dir"])
    u = universe.Universe(T=300, st=0, libdir=LIBDIR)
    u.add(protein.Protein(os.path.join(GOLDEN, "enkefa.seq"), os.path.join(GOLDEN, "enkefa_c1.var")))
    smmp.sgrnd(args["seed"])
    smmp.canon(0, args["nswp"], args["nmes"], 300.0, False)
    return {"files": sorted(os.listdir("."))}


def minimization_gpu(args):
    """EXAMPLES/Python/minimization.py: Universe(st=0), Met-enkephalin from
    enkefa.var, smmp.minim(1, 10000, 1e-9); the energies and variables after."""
    import smmp
    import protein
    import universe
    u = universe.Universe(T=300, ff="ecepp3", st=0, libdir=LIBDIR)
    p = protein.Protein(os.path.join(GOLDEN, "enkefa.seq"), os.path.join(GOLDEN, "enkefa.var"))
    u.add(p)
    smmp.minim(args.get("imin", 1), args.get("maxit", 10000), args.get("eps", 1e-9))
    nvt = int(smmp.mol_i.ivrml1[0] + smmp.mol_i.nvrml[0] - 1)
    return {"eysm": float(smmp.eny_r.eysm), "eyel": float(smmp.eny_r.eyel), "eyvw": float(smmp.eny_r.eyvw),
            "eyhb": float(smmp.eny_r.eyhb), "eyvr": float(smmp.eny_r.eyvr), "energy": u.energy(),
            "vlvr": [float(x) for x in smmp.var_r.vlvr[:nvt]]}


def energy_gpu(args):
    import universe
    import protein
    u = universe.Universe(T=300, st=args["st"], libdir=LIBDIR)
    protein.Protein(os.path.join(GOLDEN, "enkefa.seq"), os.path.join(GOLDEN, "enkefa.var"))
    import smmp
    e = u.energy()
    return {"e": e, "eyel": float(smmp.eny_r.eyel), "eyvw": float(smmp.eny_r.eyvw),
            "eyhb": float(smmp.eny_r.eyhb), "eyvr": float(smmp.eny_r.eyvr), "eysl": float(smmp.eny_r.eysl)}


def partem_gpu(args):
    import smmp
    import universe
    import protein
    os.chdir(args["dir"])
    universe.Universe(T=300, st=1, libdir=LIBDIR)
    protein.Protein(args["seq"])
    smmp.sgrnd(args["seed"])
    eol, ipoi, traj = smmp.partem_s(args["nrep"], args["nequi"], args["nswp"], 1, 1, True, args["switch"])
    return {"eol": [float(x) for x in eol], "time_d": open("time.d").read()}


def partem_restart_gpu(args):
    """smmp.partem_s for nswp sweeps, then again with newsta false and a larger
    nswp: time.d of the second run starts at the stored sweep, the restart
    files are rewritten."""
    import smmp
    import universe
    import protein
    os.chdir(args["dir"])
    universe.Universe(T=300, st=1, libdir=LIBDIR)
    protein.Protein(args["seq"])
    smmp.sgrnd(args["seed"])
    smmp.partem_s(args["nrep"], args["nequi"], args["nswp1"], 1, 1, True, args["switch"])
    first = open("time.d").read()
    nstart, eol1 = smmp.read_par_r("par_R.in", args["nrep"])
    confs1 = [open(smmp.conf_file_name(k + 1)).read() for k in range(args["nrep"])]
    eol, ipoi, traj = smmp.partem_s(args["nrep"], args["nequi"], args["nswp2"], 1, 1, False, args["switch"])
    nstart2, eol2 = smmp.read_par_r("par_R.in", args["nrep"])
    return {"first": first, "second": open("time.d").read(), "nstart": nstart, "nstart2": nstart2,
            "eol1": [float(x) for x in eol1], "eol2": [float(x) for x in eol2], "eol": [float(x) for x in eol],
            "confs1": confs1}


def muca_gpu(args):
    import smmp
    import universe
    import protein
    os.chdir(args["dir"])
    universe.Universe(T=300, st=0, libdir=LIBDIR)
    protein.Protein(os.path.join(GOLDEN, "enkefa.seq"))
    smmp.sgrnd(args["seed"])
    eol, acz, xhist = smmp.multicanonical.mulcan_par(args["nsweep"], args["nup"], args["temp"], args["kmin"],
                                                      args["kmax"], args["ebin"], False)
    b = smmp.multicanonical.b
    alpha = smmp.multicanonical.alpha
    out = {"eol": eol, "acz": acz, "xhist": [float(x) for x in xhist], "b": [float(x) for x in b],
           "alpha": [float(x) for x in alpha], "muca_d": open("muca.d").read()}
    eol2, acz2, xh2 = smmp.multicanonical.mulcan_sim(2, args["nsim"], 2, 3, args["kmin"], args["kmax"],
                                                      args["ebin"], False)
    out.update({"sim_eol": eol2, "sim_hist": float(np.sum(xh2)), "time_d": open("time.d").read()})
    return out


if __name__ == "__main__":
    name, a = sys.argv[1], json.loads(sys.argv[2]) if len(sys.argv) > 2 else {}
    print(json.dumps(globals()[name](a)), flush=True)
