#!/usr/bin/env python3
"""Benchmark: Monte Carlo sweeps/s over all replicas, parallel tempering.

Workload (one step = one sweep of every replica + one exchange attempt):
64 replicas of one random 20-residue sequence (uniform over the 20 amino
acids, seed 2), ECEPP/3 + OONS solvent, temperatures geometric 250-1000 K,
exchange every sweep, FP64 energies; replicas start from random dihedrals
and are warmed up for W sweeps before K timed sweeps.

Prints one JSON line with the throughput, the roofline of the sweep kernel
(algorithmic FP64 flops from the kernel's own work counters over its
HIP-event time, against the MI355X FP64 vector peak) and a CPU baseline:
the SMMP Fortran reference (oracle/_ref, when built) or the C oracle, timed
on a bounded sample of the same workload on one host core.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "smmp-svn_amd")
sys.path.insert(0, PKG)

NREP, LENGTH, SEQ_SEED = 64, 20, 2
TMIN, TMAX = 250.0, 1000.0
PEAK_FP64 = 78.6  # TFLOP/s, MI355X FP64 vector
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 vector
# algorithmic flops per unit of work (divide and sqrt count 1):
FLOP_PAIR = 22      # ECEPP pair term: 3 sub + 5 r^2 + 1 rsqrt + 3 powers + 2 Coulomb + 4 LJ/HB + 3 sums
FLOP_POINT = 9      # one tessellation point against one solvent neighbour
FLOP_BUILD = 60     # one atom placed by the Eyring step
FLOP_ROTATE = 21    # one atom rotated rigidly about a bond (Rodrigues)
FLOP_PAIR_GRAD = 36  # one pair with its force: the 22 of the energy, dE/dr / r (5), c r_ij (3), both atoms' sums (6)
# the analytic solvent (esolan.f), per unit of its kernel's work counters:
FLOP_ES_SCAN = 9    # one atom pair of the neighbour search: 3 sub, 3 mul, 2 add, 1 compare
FLOP_ES_PAIR = 48   # one pair of circles classified: 6 div, 21 discriminant, 3 sqrt with their 12, 6 tests
FLOP_ES_ARC = 30    # one arc: midpoint (cos, sin, 6) and, when exposed, its term (2 cos, 2 sin, atan, 15)
FLOP_ES_COVER = 7   # one circle tested against an arc midpoint: 3 mul, 3 add, 1 compare
# HBM-side bytes per sweep of this workload in FP64: FETCH_SIZE x 2.0 plus
# WRITE_SIZE x 1.0 (the factors make a 256 MiB device copy read 256 MiB and
# write 256 MiB) of the timed launch, rocprofv3 --pmc, 200 sweeps:
# 577,014 KiB x 2.0 + 2,632,251 KiB.
HBM_BYTES_PER_SWEEP = 19.386e6
# The same for one batch of the batched-energy workload (4,096 conformations,
# the two dispatches): (18,224 + 6,910) KiB FETCH_SIZE x 2.0 + 427 KiB
# WRITE_SIZE x 1.0 (each XCD takes runs of one system's conformations, so a
# system's pair bits are fetched into one L2 rather than all eight).
C3_HBM_BYTES_PER_BATCH = 51.9e6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_system(tmp, length=LENGTH, seed=SEQ_SEED, tag="c2", itysol=1):
    from smmp_amd import _ffi, workloads
    seq = workloads.random_sequence(length, seed)
    seqf = workloads.write_seq_file(os.path.join(tmp, f"{tag}.seq"), seq, tag)
    c = _ffi.Context()
    for k, v in (("epar_l.flex", 0), ("epar_l.sh2", 0), ("epar_l.epsd", 0), ("epar_l.ientyp", 0),
                 ("isolty.itysol", itysol)):
        c.view(k)[0] = v
    c.init_energy(workloads.default_libdir())
    c.init_molecule(1, "nh2", "cooh", seqf, " ")
    return c, seq, seqf


def random_starts(c, nrep, seed):
    rng = np.random.default_rng(seed)
    v = np.tile(c.vlvr(), (nrep, 1))
    idv = c.idvr - 1
    u = rng.random((nrep, len(idv)))
    # addang(pi, 2 pi (u - 1/2)) as partem_s.f:119-120, i.e. uniform in [-pi, pi)
    v[:, idv] = np.mod(np.pi + 2 * np.pi * (u - 0.5) + np.pi, 2 * np.pi) - np.pi
    return v


def cpu_baseline(seqf, budget_s: float):
    """Fortran reference (oracle/_ref/ref_driver partem) on a bounded sample."""
    from smmp_amd import workloads
    drv = os.path.join(REPO, "oracle", "_ref", "ref_driver")
    nrep, nsw = 2, 40
    temps = workloads.geometric_temperatures(nrep, TMIN, TMAX)
    if os.path.exists(drv):
        with tempfile.TemporaryDirectory() as d:
            workloads.write_temperatures(os.path.join(d, "temperatures"), temps)
            t0 = time.perf_counter()
            subprocess.run([drv, "partem", workloads.default_libdir(), seqf, "1", "777", str(nrep), "0", str(nsw)],
                           cwd=d, check=True, stdout=subprocess.DEVNULL, timeout=max(60.0, 10 * budget_s))
            dt = time.perf_counter() - t0
        kind = "reference"
    else:  # the C restatement of the same algorithm
        sys.path.insert(0, os.path.join(REPO, "tests"))
        from smmp_amd import _ffi
        from _oracle import OracleMT, OracleSystem
        c = _ffi.Context()
        c.view("isolty.itysol")[0] = 1
        c.init_energy(workloads.default_libdir())
        c.init_molecule(1, "nh2", "cooh", seqf, " ")
        o = OracleSystem(c)
        t0 = time.perf_counter()
        o.partem(OracleMT(777), temps, 0, nsw)
        dt = time.perf_counter() - t0
        kind = "port"
    out = {"value": round(nrep * nsw / dt, 4), "unit": "sweeps/s", "cores": 1, "kind": kind,
           "sample": f"PT {nrep} replicas x {nsw} sweeps of the {LENGTH}-residue system "
                     f"(ECEPP/3+OONS, {TMIN:g}-{TMAX:g} K), one host core, {dt:.1f} s"}
    if kind == "reference":  # the same sample in one process per host core (the reference is serial)
        ncore = max(1, min(16, len(os.sched_getaffinity(0))))
        with tempfile.TemporaryDirectory() as d:
            procs = []
            t0 = time.perf_counter()
            for k in range(ncore):
                wd = os.path.join(d, str(k))
                os.mkdir(wd)
                workloads.write_temperatures(os.path.join(wd, "temperatures"), temps)
                procs.append(subprocess.Popen([drv, "partem", workloads.default_libdir(), seqf, "1", str(777 + k),
                                               str(nrep), "0", str(nsw)], cwd=wd, stdout=subprocess.DEVNULL))
            for pr in procs:
                pr.wait(timeout=max(120.0, 20 * budget_s))
            dtn = time.perf_counter() - t0
        out["all_cores"] = {"value": round(ncore * nrep * nsw / dtn, 3), "unit": "sweeps/s", "cores": ncore,
                            "sample": f"{ncore} independent processes of the same sample, {dtn:.1f} s"}
    return out


def autocorr_time(x: np.ndarray) -> float:
    """Integrated autocorrelation time of a series (Sokal's window: the
    smallest W with W >= 5 tau(W))."""
    x = np.asarray(x, dtype=np.float64) - np.mean(x)
    n = len(x)
    if n < 8 or not np.any(x):
        return 1.0
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf /= acf[0]
    tau = 1.0
    for w in range(1, n):
        tau = 1.0 + 2.0 * acf[1:w + 1].sum()
        if w >= 5.0 * tau:
            break
    return max(tau, 1.0)


def thinned(traj: np.ndarray, burn: int, taus) -> np.ndarray:
    """Pooled samples of every slot's energy series after burn sweeps, each
    slot thinned by twice its autocorrelation time."""
    out = []
    for k in range(traj.shape[1]):
        step = max(1, int(np.ceil(2.0 * taus[k])))
        out.append(traj[burn::step, k])
    return np.concatenate(out)


def da_experiment(c, R, temps, warmup, rounds, collect, moves_per_state, epochs, softs, timed, ks_sweeps, ks_burn,
                  log_tag, keep_traj=False):
    """Delayed acceptance on parallel tempering against the plain sampler.
    Online training: 'rounds' times, 'collect' sweeps of the exact sampler
    with a snapshot of every replica each sweep, 'moves_per_state' random
    single moves per snapshot with their exact energy changes (FP64,
    smmp_energy_batch: two energies per move), and a cosine-scheduled fit of
    the network on all moves so far.  Then, from one state, 'timed' sweeps
    of the plain sampler and of delayed acceptance at each softening factor
    (timed by the sweep kernel's HIP events), and for the KS test on the
    pooled energies 'ks_burn' sweeps of the plain sampler, then from that
    state three independent runs of 'ks_sweeps' sweeps (plain, delayed
    acceptance at the best factor, a second plain run as the control), each
    slot thinned by twice the largest autocorrelation time of its three
    runs."""
    from scipy.stats import ks_2samp
    from smmp_amd import _ffi
    from smmp_amd.surrogate import Surrogate, move_pairs
    seeds = lambda k: np.arange(k, k + R + 1, dtype=np.uint32)  # noqa: E731
    t0 = time.perf_counter()
    state = c.sample(temps, seeds(1), warmup, random_starts(c, R, 12345), mode=1, with_traj=False)["vlvr"]
    log(f"{log_tag}: warm-up {warmup} sweeps {time.perf_counter() - t0:.1f} s")
    net = Surrogate(c, seed=3)
    Ps, Ds = [], []
    t_collect = t_pairs = t_train = 0.0
    hist = np.zeros(1)
    for rnd in range(rounds):
        t0 = time.perf_counter()
        d = c.sample(temps, seeds(100 + 7 * rnd), collect, state, mode=1, with_traj=False, snap_every=1)
        t_collect += time.perf_counter() - t0
        state = d["vlvr"]
        t0 = time.perf_counter()
        V = np.repeat(d["snap_vlvr"].reshape(-1, c.nvt), moves_per_state, axis=0)
        pv, de, _ = move_pairs(c, V, seed=rnd)
        t_pairs += time.perf_counter() - t0
        Ps.append(pv)
        Ds.append(de)
        P, D = np.concatenate(Ps), np.concatenate(Ds)
        t0 = time.perf_counter()
        hs = []
        for ep in range(epochs):  # cosine schedule, Adam moments kept across epochs
            lr = 1e-4 + 0.5 * (2e-3 - 1e-4) * (1.0 + math.cos(math.pi * ep / epochs))
            hs.append(net.fit_moves(P, D, epochs=1, batch=1024, lr=lr, seed=1000 * rnd + ep, clip=10.0,
                                    rescale=(rnd == 0 and ep == 0))[0])
        hist = np.array(hs)
        t_train += time.perf_counter() - t0
        log(f"{log_tag}: round {rnd}: {len(D)} moves ({2 * len(D)} exact energies), loss {hist[0]:.4f} -> "
            f"{hist[-1]:.4f}, train {t_train:.1f} s")
    runs = {}
    cases = [("plain", lambda: c.sample(temps, seeds(900), timed, state, mode=1, with_traj=False))]
    for f in softs:  # the network's change scaled by f in both stages (Christen-Fox stays exact)
        cases.append((f"da{f:g}", lambda f=f: net.sample(temps, seeds(900), timed, state, mode=1, with_traj=False,
                                                         da_scale=net.scale * f)))
    for name, fn in cases:
        c.profile(1)
        out = fn()
        ks = c.profile_read(_ffi.KERNEL_SWEEP)
        c.profile(False)
        runs[name] = (out, ks["ms"] / 1e3)
        log(f"{log_tag}: {name}: {timed} sweeps in {ks['ms']:.1f} ms of sweep kernel, accepted {out['acc'].sum():.0f}")
    (plain, tp) = runs["plain"]
    rate = lambda o, t: float(o["acc"].sum()) / t  # noqa: E731
    gains = {f"{f:g}": round(rate(*runs[f"da{f:g}"]) / rate(plain, tp), 3) for f in softs}
    best = max(softs, key=lambda f: gains[f"{f:g}"])
    da, td = runs[f"da{best:g}"]
    ntrial = float(R * timed * c.nvr)
    st = da["da_stats"]
    res = {"gain": gains[f"{best:g}"], "softening": best, "gain_by_softening": gains,
           "accepted_per_s_plain": round(rate(plain, tp), 1), "accepted_per_s_da": round(rate(da, td), 1),
           "sweeps_per_s_plain": round(R * timed / tp, 1), "sweeps_per_s_da": round(R * timed / td, 1),
           "acceptance_plain": round(float(plain["acc"].sum()) / ntrial, 4),
           "acceptance_da": round(float(da["acc"].sum()) / ntrial, 4),
           "rejected_by_network": round(float(st[:, 1].sum() / st[:, 0].sum()), 4),
           "exact_energies_per_accept_plain": round(ntrial / float(plain["acc"].sum()), 3),
           "exact_energies_per_accept_da": round(float(st[:, 2].sum() / max(st[:, 3].sum(), 1.0)), 3),
           "training": {"moves": int(sum(len(x) for x in Ds)), "exact_energies": int(2 * sum(len(x) for x in Ds)),
                        "rounds": rounds, "sweeps_per_round": collect, "moves_per_snapshot": moves_per_state,
                        "epochs_per_round": epochs, "final_loss": round(float(hist[-1]), 5),
                        "scale_kcal_per_unit": round(net.scale, 3), "parameters": int(net.n),
                        "collect_s": round(t_collect, 2), "pairs_s": round(t_pairs, 2), "train_s": round(t_train, 2)}}
    if ks_sweeps > 0:
        # an exact sampler started from the stationary distribution stays in
        # it: equilibrate one state with the plain sampler, then run all
        # three samplers from it and compare their energies at every sweep
        t0 = time.perf_counter()
        eq = c.sample(temps, seeds(1500), ks_burn, state, mode=1, with_traj=False)["vlvr"]
        log(f"{log_tag}: KS equilibration {ks_burn} sweeps done, {time.perf_counter() - t0:.0f} s")
        a = c.sample(temps, seeds(2000), ks_sweeps, eq, mode=1)["traj"]
        log(f"{log_tag}: KS plain run done, {time.perf_counter() - t0:.0f} s")
        b = net.sample(temps, seeds(3000), ks_sweeps, eq, mode=1, da_scale=net.scale * best)["traj"]
        log(f"{log_tag}: KS delayed-acceptance run done, {time.perf_counter() - t0:.0f} s")
        z = c.sample(temps, seeds(4000), ks_sweeps, eq, mode=1)["traj"]
        log(f"{log_tag}: KS control run done, {time.perf_counter() - t0:.0f} s")
        burn = 0
        # each slot thinned by twice the largest of its three runs'
        # autocorrelation times (the slowest slots hold few independent
        # samples; no run's estimate is preferred)
        taus = [max(autocorr_time(a[:, k]), autocorr_time(b[:, k]), autocorr_time(z[:, k])) for k in range(R)]
        sa, sb, sz = thinned(a, burn, taus), thinned(b, burn, taus), thinned(z, burn, taus)
        res["ks"] = {"sweeps": ks_sweeps, "equilibration_sweeps": ks_burn, "samples_per_run": int(len(sa)),
                     "autocorr_sweeps_median": round(float(np.median(taus)), 2),
                     "autocorr_sweeps_max": round(float(np.max(taus)), 2),
                     "pooled_p_da_vs_plain": round(float(ks_2samp(sa, sb).pvalue), 4),
                     "pooled_p_control_plain_vs_plain": round(float(ks_2samp(sa, sz).pvalue), 4),
                     "run_s": round(time.perf_counter() - t0, 1)}
        log(f"{log_tag}: KS {res['ks']}")
        if keep_traj:
            res["traj"] = {"plain": a, "da": b, "control": z}
    return res


def run_da(args):
    """Delayed acceptance against the plain sampler on parallel tempering
    with 128 replicas of a random 36-residue sequence (seed 5), ECEPP/3 +
    OONS, 250-1000 K, the network trained online on the sampler's own
    moves."""
    from smmp_amd import workloads
    R, L = 128, 36
    tmp = tempfile.mkdtemp(prefix="smmp_da_")
    c, seq, _ = make_system(tmp, length=L, seed=5, tag="c5")
    temps = workloads.geometric_temperatures(R, TMIN, TMAX)
    log(f"delayed acceptance: {R} replicas x {L} residues, {c.nvr} free torsions")
    softs = [float(x) for x in str(args.da_soften).split(",")]
    r = da_experiment(c, R, temps, args.da_warmup, args.da_rounds, args.da_collect, args.da_moves, args.da_epochs,
                      softs, args.steps, args.da_ks, args.da_ks_burn, "da")
    line = {
        "metric": "DA-surrogate vs plain GPU: exact-energy-equivalent accepted moves/s",
        "value": r["gain"], "unit": "x", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
        "dtype": "f64 (energies), bf16 (network)",
        "data": "synthetic: random 36-residue sequence (seed 5), random initial dihedrals; network trained on the "
                "sampler's own moves",
        "config": {"workload": f"parallel tempering, {R} replicas x {L}-residue random sequence, ECEPP/3+OONS, "
                               f"T geometric {TMIN:g}-{TMAX:g} K, exchange every sweep",
                   "replicas": R, "residues": L, "free_torsions": c.nvr},
    }
    line.update({k: v for k, v in r.items() if k != "gain"})
    print(json.dumps(line), flush=True)


def run_esolan(args):
    """Energies with the analytic solvent (itysol -1, esolan.f) of conformations
    sampled at the parallel-tempering workload (64 replicas of the random
    20-residue sequence, seed 2, 250-1000 K): --es-warmup sweeps from random
    dihedrals with OONS, then one snapshot of every replica per sweep until
    --es-confs conformations; each timed step is one smmp_energy_batch call
    over all of them (ECEPP/3 + analytic solvent, FP64)."""
    import torch
    from smmp_amd import _ffi, workloads
    tmp = tempfile.mkdtemp(prefix="smmp_es_")
    c, seq, seqf = make_system(tmp)
    ca, _, _ = make_system(tmp, itysol=-1, tag="c2a")
    temps = workloads.geometric_temperatures(NREP, TMIN, TMAX)
    nsnap = max(1, args.es_confs // NREP)
    log(f"analytic solvent: sampling {NREP} replicas, {args.es_warmup} + {nsnap} sweeps")
    w = c.sample(temps, np.arange(1, NREP + 2, dtype=np.uint32), args.es_warmup, random_starts(c, NREP, 12345),
                 mode=1, with_traj=False)
    d = c.sample(temps, np.arange(500, 500 + NREP + 1, dtype=np.uint32), nsnap, w["vlvr"], mode=1, with_traj=False,
                 snap_every=1)
    V = np.ascontiguousarray(d["snap_vlvr"].reshape(-1, c.nvt))
    n = len(V)
    cs = np.zeros(n, dtype=np.int32)
    ca.sgrnd(4357)
    _ffi.energy_batch([ca], cs, V.ravel())  # untimed: device buffers and the intersection capacity
    ca.profile(True)
    torch.cuda.synchronize()
    K = args.steps
    log(f"timed {K} batches of {n} conformations")
    t0 = time.perf_counter()
    for _ in range(K):
        out = _ffi.energy_batch([ca], cs, V.ravel())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ke = ca.profile_read(_ffi.KERNEL_ENERGY)
    ks = ca.profile_read(_ffi.KERNEL_ESOLAN)
    wk = list(ks["work"])  # atom pairs scanned, circle pairs, arcs, coverage tests
    flops = FLOP_ES_SCAN * wk[0] + FLOP_ES_PAIR * wk[1] + FLOP_ES_ARC * wk[2] + FLOP_ES_COVER * wk[3]
    achieved = flops / (ks["ms"] * 1e-3) / 1e12
    log(f"esolan kernels {ks['ms'] / K:.2f} ms per batch ({ks['launches'] / K:.1f} launches), k_energy "
        f"{ke['ms'] / K:.2f} ms; per conformation: scan {wk[0] / K / n:.0f}, circle pairs {wk[1] / K / n:.0f}, "
        f"arcs {wk[2] / K / n:.0f}, coverage tests {wk[3] / K / n:.0f}")
    log(f"eysl mean {out[:, 6].mean():.3f}, total mean {out[:, 0].mean():.3f}")
    # the replica sampler with the analytic solvent: the replicas' trials in
    # lockstep, one device energy pass per trial step (parallel tempering)
    nsw2 = 2
    t0 = time.perf_counter()
    ca.sample(temps, np.arange(900, 900 + NREP + 1, dtype=np.uint32), nsw2, V[:NREP], mode=1, with_traj=False)
    torch.cuda.synchronize()
    t_smp = time.perf_counter() - t0
    log(f"analytic-solvent parallel tempering: {NREP} replicas x {nsw2} sweeps in {t_smp:.2f} s")
    cpu = None
    smp_cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        from _oracle import OracleMT, OracleSystem
        o = OracleSystem(ca)
        mt = OracleMT(4357)
        m, t1 = 0, time.perf_counter()
        while m < n and time.perf_counter() - t1 < 10.0:
            o.energy(V[m], mt)
            m += 1
        tc = time.perf_counter() - t1
        cpu = {"value": round(m / tc, 3), "unit": "energy evaluations/s", "cores": 1, "kind": "port",
               "sample": f"the first {m} of the same conformations through the C oracle (ECEPP/3 + esolan "
                         f"restatement), one host core, {tc:.1f} s"}
        t1 = time.perf_counter()
        o.metropolis(OracleMT(5), V[0], 1.0 / (TMIN * 1.98773e-3), 1)
        smp_cpu = round(1.0 / (time.perf_counter() - t1), 4)
    line = {
        "metric": "energy evaluations/s (ECEPP/3 + analytic solvent)",
        "value": round(K * n / dt, 2),
        "unit": "energy evaluations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": 1,
        "ms_per_step": round(1e3 * dt / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: random 20-residue sequence (seed 2); conformations sampled by parallel tempering",
        "config": {"workload": f"smmp_energy_batch of {n} conformations of the {LENGTH}-residue sequence sampled at "
                               f"{NREP} temperatures {TMIN:g}-{TMAX:g} K, ECEPP/3 + analytic solvent (itysol -1)",
                   "conformations": n, "atoms": int(ca.view("res_i.iatrs2")[LENGTH - 1])},
        "roofline": {"bound": "valu-fp64", "achieved": round(achieved, 4), "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64, 5), "traffic": None,
                     "kernel": "k_esolan (both passes) + k_esolan_sum"},
        "cpu_baseline": cpu,
        "sampler": {"replica_sweeps_per_s": round(NREP * nsw2 / t_smp, 2), "replicas": NREP, "sweeps": nsw2,
                    "mode": "parallel tempering (smmp_sample mode 1), analytic solvent",
                    "cpu_oracle_replica_sweeps_per_s": smp_cpu},
    }
    print(json.dumps(line), flush=True)


def c3_headline(K: int) -> dict:
    """Energy evaluations/s of the C3 batch (4,096 conformations over 256
    random sequences of 10-60 residues, ECEPP/3 + OONS, resident on the
    device) in the mixed and FP64 modes, with the energy kernel's roofline
    fractions (mixed against the FP32 vector peak, FP64 against FP64)."""
    from smmp_amd import _ffi, workloads
    nseq, per = 256, 16
    tmp = tempfile.mkdtemp(prefix="smmp_c3h_")
    seqs = workloads.c3_sequences(nseq, 10, 60, seed=3)
    ctxs, rows, cs = [], [], []
    for i, sq in enumerate(seqs):
        seqf = workloads.write_seq_file(os.path.join(tmp, f"s{i}.seq"), sq, f"s{i}")
        c = _ffi.Context()
        c.view("isolty.itysol")[0] = 1
        c.init_energy(workloads.default_libdir())
        c.init_molecule(1, "nh2", "cooh", seqf, " ")
        ctxs.append(c)
        v = np.tile(c.vlvr(), (per, 1))
        v[:, c.idvr - 1] = workloads.uniform_dihedrals(per, c.nvr, 1000 + i)
        rows.append(v.ravel())
        cs += [i] * per
    cs = np.asarray(cs, dtype=np.int32)
    n = len(cs)
    B = _ffi.EnergyBatch(ctxs, cs)
    B.load(np.concatenate(rows))
    out = {}
    for prec, name, peak in ((1, "mixed", PEAK_FP32), (0, "fp64", PEAK_FP64)):
        B.run(prec)
        B.read()
        shapes = B.tune(prec)  # the autotuner's launch shapes
        B.run(prec)
        B.read()
        ctxs[0].profile(True)
        t0 = time.perf_counter()
        for _ in range(K):
            B.run(prec)
        B.read()
        dt = time.perf_counter() - t0
        ke = ctxs[0].profile_read(_ffi.KERNEL_ENERGY)
        ctxs[0].profile(False)
        wk = ke["work"]
        flops = FLOP_PAIR * wk[0] + FLOP_POINT * 122 * wk[1] + FLOP_BUILD * wk[2]
        ach = flops / (ke["ms"] * 1e-3) / 1e12
        out[name] = {"evals_per_s": round(K * n / dt, 1), "k_energy_ms_per_batch": round(ke["ms"] / K, 3),
                     "launch_shapes": shapes,
                     "roofline_frac": round(ach / peak, 5), "achieved_tflops": round(ach, 3), "peak_tflops": peak,
                     "traffic_bytes_per_batch": C3_HBM_BYTES_PER_BATCH if (nseq, per) == (256, 16) else None}
    B.close()
    for c in ctxs:
        c.close()
    log(f"C3 headline: {out}")
    return {"workload": f"4,096 conformations over {nseq} random sequences of 10-60 residues, ECEPP/3+OONS, "
                        f"resident batch, {K} timed batches", **out}


def minim_headline(K: int, nconf: int = 4096, nmin: int = 1024, maxit: int = 300) -> dict:
    """Derivatives and local minimisation in vacuum (gradient.f, minim.f)
    on the C2 sequence (random 20-mer, seed 2): K timed smmp_gradient_batch
    calls over nconf random conformations (k_gradient's time from its HIP
    events, with its FP64 roofline: FLOP_PAIR_GRAD per pair counted once
    and FLOP_BUILD per atom placed), then nmin quasi-Newton minimisations
    from random starts side by side (smmp_minim_batch, maxit calls each)."""
    from smmp_amd import _ffi, workloads
    tmp = tempfile.mkdtemp(prefix="smmp_min_")
    c, _, _ = make_system(tmp, itysol=0)
    V = random_starts(c, nconf, 4242)
    c.gradient_batch(V)
    c.profile(True)
    t0 = time.perf_counter()
    for _ in range(K):
        c.gradient_batch(V)
    dt = time.perf_counter() - t0
    kg = c.profile_read(_ffi.KERNEL_GRADIENT)
    c.profile(False)
    wk = kg["work"]
    flops = FLOP_PAIR_GRAD * wk[0] + FLOP_BUILD * wk[2]
    ach = flops / (kg["ms"] * 1e-3) / 1e12
    out = {"gradient_evals_per_s": round(K * nconf / (kg["ms"] * 1e-3), 1),
           "gradient_evals_per_s_with_transfers": round(K * nconf / dt, 1),
           "k_gradient_ms_per_launch": round(kg["ms"] / max(kg["launches"], 1), 4),
           "roofline": {"bound": "valu", "achieved": round(ach, 4), "peak": PEAK_FP64, "unit": "TFLOP/s",
                        "frac": round(ach / PEAK_FP64, 5), "flop_per_launch": round(flops / max(kg["launches"], 1), 1)}}
    S = random_starts(c, nmin, 777)
    e0, _ = c.gradient_batch(S)
    t0 = time.perf_counter()
    Vb, fb, nb = c.minim_batch(1, S, maxit, 1e-9)
    dtm = time.perf_counter() - t0
    out["minimisation"] = {"starts": nmin, "maxit": maxit, "minimisations_per_s": round(nmin / dtm, 1),
                           "function_calls_per_s": round(float(nb.sum()) / dtm, 1),
                           "median_calls": float(np.median(nb)), "converged": int((nb < maxit).sum()),
                           "median_energy_start": float(np.median(e0[:, 0])), "median_energy_end": float(np.median(fb)),
                           "wall_s": round(dtm, 2)}
    c.close()
    log(f"minimisation headline: {out}")
    return {"workload": f"{LENGTH}-residue random sequence (seed {SEQ_SEED}), ECEPP/3 in vacuum: {nconf} "
                        f"conformations per derivative launch, {K} timed launches; {nmin} quasi-Newton "
                        f"minimisations from random starts", **out}


def run_minim(args):
    line = {"metric": "derivative evaluations/s (gradient.f) and minimisations/s (minim.f)", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f64", "data": "synthetic: random sequence and dihedrals"}
    r = minim_headline(max(args.steps, 1))
    line["value"] = r["gradient_evals_per_s"]
    line["unit"] = "gradient evals/s"
    line.update(r)
    print(json.dumps(line), flush=True)


def run_c3(args):
    """Batched energy-only evaluation: 4,096 conformations over 256 random
    sequences of 10-60 residues (16 conformations each, dihedrals uniform in
    [-pi, pi)), ECEPP/3 + OONS.  Each timed step is one smmp_batch_run over
    all of them with the conformations resident on the device (the same
    through smmp_energy_batch from host buffers is reported beside it); the
    headline is the mixed-precision path (FP32 pair terms,
    FP64 accumulation), with every structure's error against the FP64 path
    reported, and FP64 throughput beside it."""
    from smmp_amd import _ffi, workloads
    nseq, per = args.c3_seqs, args.c3_per_seq
    tmp = tempfile.mkdtemp(prefix="smmp_c3_")
    seqs = workloads.c3_sequences(nseq, 10, 60, seed=3)
    t0 = time.perf_counter()
    ctxs = []
    for i, s in enumerate(seqs):
        seqf = workloads.write_seq_file(os.path.join(tmp, f"s{i}.seq"), s, f"s{i}")
        c = _ffi.Context()
        for k, v in (("epar_l.flex", 0), ("epar_l.sh2", 0), ("epar_l.epsd", 0), ("epar_l.ientyp", 0),
                     ("isolty.itysol", 1)):
            c.view(k)[0] = v
        c.init_energy(workloads.default_libdir())
        c.init_molecule(1, "nh2", "cooh", seqf, " ")
        ctxs.append(c)
    nat = np.array([int(c.view("res_i.iatrs2")[len(s) - 1]) for c, s in zip(ctxs, seqs)])
    log(f"C3: {nseq} sequences ({min(map(len, seqs))}-{max(map(len, seqs))} residues, {nat.min()}-{nat.max()} atoms, "
        f"mean {nat.mean():.0f}) set up in {time.perf_counter() - t0:.1f} s")
    rows, cs = [], []
    for i, c in enumerate(ctxs):
        v = np.tile(c.vlvr(), (per, 1))
        v[:, c.idvr - 1] = workloads.uniform_dihedrals(per, c.nvr, 1000 + i)
        rows.append(v.ravel())
        cs += [i] * per
    V = np.concatenate(rows)
    cs = np.asarray(cs, dtype=np.int32)
    n = len(cs)
    K = args.steps
    B = _ffi.EnergyBatch(ctxs, cs)  # systems uploaded once; conformations resident
    B.load(V)
    res = {}
    shapes = {}
    for prec in (1, 0):
        B.run(prec)
        B.read()  # untimed first run
        t0 = time.perf_counter()
        for _ in range(K):
            B.run(prec)
        B.read()
        dt_fixed = time.perf_counter() - t0
        shapes[prec] = B.tune(prec)  # the autotuner's launch shapes for this batch
        B.run(prec)
        B.read()
        ctxs[0].profile(True)
        t0 = time.perf_counter()
        for _ in range(K):
            B.run(prec)
        out = B.read()
        dt = time.perf_counter() - t0
        ke = ctxs[0].profile_read(_ffi.KERNEL_ENERGY)
        ctxs[0].profile(False)
        log(f"precision {prec}: tuned shapes (threads, pair bits in LDS) per class {shapes[prec]}: "
            f"{1e3 * dt / K:.2f} ms per run, fixed rule {1e3 * dt_fixed / K:.2f} ms")
        # the host-buffer entry point (conformations in, energies out over PCIe each call)
        _ffi.energy_batch(ctxs, cs, V, prec)
        t1 = time.perf_counter()
        for _ in range(K):
            _ffi.energy_batch(ctxs, cs, V, prec)
        dth = time.perf_counter() - t1
        res[prec] = (out, dt, ke, dth)
        log(f"precision {prec}: {K} resident runs {1e3 * dt / K:.2f} ms each, k_energy {ke['ms'] / K:.2f} ms per batch; "
            f"smmp_energy_batch from host buffers {1e3 * dth / K:.2f} ms per call")
    (om, dtm, kem, dthm), (od, dtd, ked, dthd) = res[1], res[0]
    err = np.abs(om[:, 0] - od[:, 0])
    rel = err / np.maximum(np.abs(od[:, 0]), 1.0)
    # the mixed mode's stated bound: 2e-5 x (|E_el| + |E_vw| + |E_hb|) + 1e-3
    bound = 2e-5 * (np.abs(od[:, 2]) + np.abs(od[:, 3]) + np.abs(od[:, 4])) + 1e-3
    log(f"mixed vs FP64 per structure: max |dE| {err.max():.3e}, max rel {rel.max():.3e}, median rel "
        f"{np.median(rel):.3e}; within bound {np.mean(err <= bound):.4f}")

    def roof(ke, peak):
        wk = ke["work"]
        flops = FLOP_PAIR * wk[0] + FLOP_POINT * 122 * wk[1] + FLOP_BUILD * wk[2]
        a = flops / (ke["ms"] * 1e-3) / 1e12
        return flops / K, a, {"achieved": round(a, 4), "peak": peak, "frac": round(a / peak, 5)}

    wk = kem["work"] / (K * n)
    log(f"work per conformation: pairs {wk[0]:.0f}, solvent contacts {wk[1]:.0f} (x122 point tests), "
        f"atoms placed {wk[2]:.0f}")
    fl, am, rm = roof(kem, PEAK_FP32)
    _, ad, rd = roof(ked, PEAK_FP64)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        from _oracle import OracleSystem
        oracles = {}
        m, worst, t1 = 0, 0.0, time.perf_counter()
        step = max(1, n // 512)  # spread the sample over the sequences
        for k in range(0, n, step):
            if time.perf_counter() - t1 >= 15.0:
                break
            s = int(cs[k])
            o = oracles.get(s) or oracles.setdefault(s, OracleSystem(ctxs[s]))
            off = int(sum(ctxs[int(j)].nvt for j in cs[:k]))
            e = o.energy(V[off:off + ctxs[s].nvt])
            worst = max(worst, abs(od[k, 0] - e[0]) / max(abs(e[0]), 1.0))
            m += 1
        tc = time.perf_counter() - t1
        cpu = {"value": round(m / tc, 3), "unit": "energy evaluations/s", "cores": 1, "kind": "port",
               "sample": f"{m} of the conformations (every {step}th) through the C oracle (FP64), one host core, "
                         f"{tc:.1f} s (incl. system set-up); GPU FP64 vs oracle max rel error {worst:.2e}"}
    line = {
        "metric": "energy evals/s",
        "value": round(K * n / dtm, 1),
        "unit": "energy evaluations/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": 1,
        "ms_per_step": round(1e3 * dtm / K, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair terms + f64 accumulation",
        "data": f"synthetic: {nseq} random sequences of 10-60 residues (seed 3), {per} conformations each with "
                "dihedrals uniform in [-pi, pi)",
        "config": {"workload": f"energies of {n} resident conformations (smmp_batch_run) over {nseq} sequences of "
                               "uneven length, ECEPP/3 + OONS", "conformations": n, "sequences": nseq,
                   "atoms_min_mean_max": [int(nat.min()), round(float(nat.mean()), 1), int(nat.max())]},
        "fp64": {"value": round(K * n / dtd, 1), "ms_per_step": round(1e3 * dtd / K, 3),
                 "kernel_ms": round(ked["ms"] / K, 3), "roofline": rd,
                 "host_buffer_calls_per_s": round(K * n / dthd, 1)},
        "host_buffer_calls_evals_per_s": round(K * n / dthm, 1),
        "error_mixed_vs_fp64": {"max_abs": float(f"{err.max():.4e}"), "max_rel": float(f"{rel.max():.4e}"),
                                "median_rel": float(f"{np.median(rel):.4e}"),
                                "frac_within_bound": round(float(np.mean(err <= bound)), 5)},
        "roofline": {"bound": "valu-fp32", "unit": "TFLOP/s",
                     "traffic": C3_HBM_BYTES_PER_BATCH if (nseq, per) == (256, 16) else None,
                     "kernel": "k_energy<mixed> (both LDS classes, one batch)",
                     "avg_launch_ms": round(kem["ms"] / max(kem["launches"], 1), 4),
                     "flop_per_launch": round(fl / max(kem["launches"] / K, 1), 1), **rm},
        "kernel_evals_per_s": round(K * n / (kem["ms"] * 1e-3), 1),
        "launch_shapes": {"mixed": shapes[1], "fp64": shapes[0],
                          "note": "(threads per workgroup, pair bits in LDS) per LDS class, by smmp_batch_tune"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def run_muca(args):
    """Multicanonical sampling: 128 independent replicas of a random
    36-residue sequence (seed 5), ECEPP/3 in vacuo, 200 energy bins.  The
    weights come from mulcan_par-style iterations on the device sampler
    (every replica's sweep-end histogram pooled per update, Berg's recursion
    in smmp_muca_update, muca.d tables from smmp_muca_final); the timed run
    samples with those weights (muca_weight2) in mixed precision (FP32 pair
    terms in full recomputations and trial changes, FP64 sums, chain and
    decisions) with the sweeps replayed from a captured HIP graph, then the
    same run in FP64 and the rate of decisions the FP32 terms flip."""
    from smmp_amd import _ffi, workloads
    R, L, NB = 128, 36, 200
    tmp = tempfile.mkdtemp(prefix="smmp_muca_")
    c, seq, _ = make_system(tmp, length=L, seed=5, tag="c4", itysol=0)
    seeds = lambda k: np.arange(k, k + R + 1, dtype=np.uint32)  # noqa: E731
    log(f"multicanonical: {R} replicas x {L} residues, {c.nvr} free torsions, ECEPP/3 in vacuo")
    temps = workloads.geometric_temperatures(R, TMIN, TMAX)
    t0 = time.perf_counter()
    d = c.sample(temps, seeds(1), args.mu_warmup, random_starts(c, R, 12345), mode=1, precision=1)
    e = d["traj"][args.mu_warmup // 2:]
    lo, hi = float(np.quantile(e[:, 0], 0.01)), float(np.quantile(e[:, -1], 0.99))
    ebin = (hi - lo) / (NB - 2)
    kmin = int(math.floor(lo / ebin)) - 1
    kmax = kmin + NB - 1
    log(f"warm-up {args.mu_warmup} PT sweeps in {time.perf_counter() - t0:.1f} s: E {lo:.1f} .. {hi:.1f}, "
        f"ebin {ebin:.3f}, bins {kmin}..{kmax}")
    state = d["vlvr"]
    # weight iteration (mulcan_par_mod.f90:23-190), sweeps at (beta + b(k)) x + alpha(k)
    beta_t = 1.0 / (TMAX * 1.98773e-3)
    b, alpha, xhist, g1, ent = (np.zeros(NB) for _ in range(5))
    ih = np.zeros(NB, dtype=np.int32)
    t0 = time.perf_counter()
    for it in range(args.mu_iters):
        w = _ffi.Weight(kind=1, beta=beta_t, kmin=kmin, kmax=kmax, ebin=ebin, b=_ffi.dptr(b), alpha=_ffi.dptr(alpha))
        d = c.sample(np.full(R, TMAX), seeds(100 + 3 * it), args.mu_nup, state, mode=2, weight=w, precision=1,
                     with_traj=False)
        state = d["vlvr"]
        ih[:] = d["hist"]
        _ffi.lib.smmp_muca_update(kmin, kmax, ebin, beta_t, _ffi.iptr(ih), _ffi.dptr(xhist), _ffi.dptr(g1),
                                  _ffi.dptr(b), _ffi.dptr(alpha), _ffi.dptr(ent))
    _ffi.lib.smmp_muca_final(kmin, kmax, ebin, beta_t, _ffi.dptr(b), _ffi.dptr(alpha))
    # The recursion leaves the slope near the top of the range (and, clamped,
    # above it) at d ln g/dE, below beta(TMAX): walkers could drift to ever
    # higher energies.  Slopes are kept at least canonical at TMAX, and alpha
    # follows from them as in mulcan_par_mod.f90:170-174.
    b[:] = np.maximum(b, beta_t)
    alpha[-1] = 0.0
    for i in range(NB - 2, -1, -1):
        alpha[i] = alpha[i + 1] + (b[i + 1] - b[i]) * (kmin + i + 0.5) * ebin
    log(f"{args.mu_iters} weight updates of {args.mu_nup} sweeps x {R} replicas in {time.perf_counter() - t0:.1f} s; "
        f"bins visited {np.mean(xhist > 0):.2f}")
    w2 = _ffi.Weight(kind=2, beta=0.0, kmin=kmin, kmax=kmax, ebin=ebin, b=_ffi.dptr(b), alpha=_ffi.dptr(alpha))
    K, W = args.steps, args.warmup
    g = args.graph_sweeps or math.gcd(K, W) or K
    kw = dict(mode=2, weight=w2, precision=1, with_traj=False, graph_sweeps=g, groups=args.groups)
    state = c.sample(np.full(R, TMAX), seeds(500), W, state, **kw)["vlvr"]
    c.profile(1)
    t0 = time.perf_counter()
    out = c.sample(np.full(R, TMAX), seeds(600), K, state, **kw)
    dt = time.perf_counter() - t0
    ks = c.profile_read(_ffi.KERNEL_SWEEP)
    info = c.sampler_info()
    c.profile(False)
    hist = out["hist"].astype(np.float64)
    vis = hist > 0
    span = np.flatnonzero(vis)
    flat = float(hist[vis].min() / hist[vis].mean()) if vis.any() else 0.0
    flops = FLOP_PAIR * ks["pairs"] + FLOP_POINT * 122 * ks["contacts"] + FLOP_BUILD * ks["builds"] \
        + FLOP_ROTATE * ks["moved"]
    achieved = flops / (ks["ms"] * 1e-3) / 1e12
    log(f"timed {K} sweeps: {R * K / dt:.1f} sweeps/s; kernel {ks['launches']} launches, {ks['ms']:.1f} ms; "
        f"graph replays {info['graph_replays']}; bins visited {vis.mean():.2f} ({span[0] if vis.any() else -1}.."
        f"{span[-1] if vis.any() else -1}), min/mean {flat:.3f}")
    # the same timed run in FP64, and the acceptance decisions that the FP32
    # pair terms flip: both precisions from the same states and streams, the
    # first sweep at which a replica's variables differ holds its first flip
    kw64 = dict(kw, precision=0)
    c.sample(np.full(R, TMAX), seeds(500), W, state, **kw64)
    t0 = time.perf_counter()
    c.sample(np.full(R, TMAX), seeds(600), K, state, **kw64)
    dt64 = time.perf_counter() - t0
    S = 40
    runs = [c.sample(np.full(R, TMAX), seeds(700), S, state, mode=2, weight=w2, precision=p, with_traj=False,
                     snap_every=1, groups=args.groups)["snap_vlvr"] for p in (0, 1)]
    diff = np.abs(runs[0] - runs[1]).max(axis=2) > 0.0          # [sweep][replica]
    first = np.where(diff.any(axis=0), diff.argmax(axis=0), S)  # sweep of the first flip, or S
    nflip = int((first < S).sum())
    trials = int(np.minimum(first + 1, S).sum()) * c.nvr
    flips = {"trials": trials, "flipped_replicas": nflip, "rate_per_trial": nflip / trials,
             "sweeps": S, "replicas": R,
             "estimate": "first differing sweep per replica; its sweep counted whole, so the rate is a lower bound"}
    log(f"FP64 timed run: {R * K / dt64:.1f} sweeps/s (mixed {R * K / dt:.1f}); decision flips: {nflip} of {R} "
        f"replicas within {S} sweeps, {nflip / trials:.2e} per trial")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "tests"))
        from _oracle import OracleMT, OracleSystem
        o = OracleSystem(c)
        m, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < 15.0 and m < R:
            o.metropolis(OracleMT(7 + m), state[m], 1.0 / (TMAX * 1.98773e-3), 1)
            m += 1
        tc = time.perf_counter() - t1
        cpu = {"value": round(m / tc, 4), "unit": "sweeps/s", "cores": 1, "kind": "port",
               "sample": f"{m} single sweeps of the C oracle's Metropolis (FP64, canonical weight at {TMAX:g} K: the "
                         f"same trial work as a multicanonical sweep) from the replicas' start states, one host "
                         f"core, {tc:.1f} s"}
    line = {
        "metric": "MC sweeps/s over all replicas",
        "value": round(R * K / dt, 2),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * dt / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair terms (full recomputations and trial changes) + f64 sums, chain and decisions",
        "data": "synthetic: random 36-residue sequence (seed 5), random initial dihedrals, weights iterated on the "
                "device sampler",
        "config": {"workload": f"multicanonical sampling, {R} independent replicas x {L}-residue random sequence, "
                               f"ECEPP/3 in vacuo, {NB} energy bins, hipGraph replay of {g} sweeps",
                   "replicas": R, "residues": L, "free_torsions": c.nvr, "bins": NB, "ebin": round(ebin, 4),
                   "kmin": kmin, "weight_updates": args.mu_iters, "sweeps_per_update": args.mu_nup,
                   "workgroups_per_replica": info["groups"], "graph_replays": info["graph_replays"]},
        "histogram": {"bins_visited": round(float(vis.mean()), 3), "min_over_mean_visited": round(flat, 4)},
        "fp64_same_run": {"value": round(R * K / dt64, 2), "unit": "sweeps/s", "mixed_over_fp64": round(dt64 / dt, 4)},
        "mixed_decision_flips": flips,
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64, 5), "traffic": None, "kernel": "k_sweep<mixed>",
                     "avg_launch_ms": round(ks["ms"] / max(ks["launches"], 1), 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed sweeps")
    ap.add_argument("--warmup", type=int, default=30, help="untimed sweeps from the random start")
    ap.add_argument("--precision", type=int, default=0, help="0 FP64, 1 FP32 pair terms")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="leave out the energy-evaluation (C3) and delayed-acceptance (C5) figures of the default line")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (calibrates HBM byte counters) and exit")
    ap.add_argument("--groups", type=int, default=0,
                    help="workgroups per replica: 0 automatic, -1 measured (autotuner), >0 upper bound")
    ap.add_argument("--graph-sweeps", type=int, default=0,
                    help="replay a captured HIP graph of this many sweeps (must divide --steps and --warmup)")
    ap.add_argument("--da", action="store_true",
                    help="delayed acceptance with the trained surrogate against the plain sampler "
                         "(128 replicas x 36 residues); --steps sets the compared sweeps")
    ap.add_argument("--da-warmup", type=int, default=300)
    ap.add_argument("--da-soften", default="1,0.7,0.5,0.35",
                    help="factors on the network's energy change in both stages (comma list; the best one is "
                         "reported and used for the KS runs)")
    ap.add_argument("--da-rounds", type=int, default=2, help="collect-and-train rounds")
    ap.add_argument("--da-collect", type=int, default=40, help="sweeps per round (snapshots every sweep)")
    ap.add_argument("--da-moves", type=int, default=50, help="random single moves per snapshot")
    ap.add_argument("--da-epochs", type=int, default=8, help="epochs per round")
    ap.add_argument("--da-ks", type=int, default=8000, help="sweeps of each KS run (0: none)")
    ap.add_argument("--da-ks-burn", type=int, default=4000, help="plain sweeps that equilibrate the KS runs' start")
    ap.add_argument("--esolan", action="store_true",
                    help="energies with the analytic solvent (itysol -1) of sampled conformations; "
                         "--steps sets the timed batches")
    ap.add_argument("--es-warmup", type=int, default=200, help="sampling sweeps before the snapshots")
    ap.add_argument("--es-confs", type=int, default=1024, help="conformations per batch")
    ap.add_argument("--c3", action="store_true",
                    help="batched energies of 4,096 conformations over 256 sequences of 10-60 residues; "
                         "--steps sets the timed batches")
    ap.add_argument("--minim", action="store_true",
                    help="derivatives (gradient.f) of 4,096 conformations per launch and 1,024 batched minimisations "
                         "(minim.f) of the 20-residue system in vacuum; --steps sets the timed launches")
    ap.add_argument("--muca", action="store_true",
                    help="multicanonical sampling, 128 replicas x 36 residues, 200 bins, mixed precision, graph replay")
    ap.add_argument("--mu-warmup", type=int, default=200, help="parallel-tempering sweeps that set the energy range")
    ap.add_argument("--mu-iters", type=int, default=20, help="weight updates")
    ap.add_argument("--mu-nup", type=int, default=50, help="sweeps per weight update")
    ap.add_argument("--c3-seqs", type=int, default=256)
    ap.add_argument("--c3-per-seq", type=int, default=16)
    ap.add_argument("--phases", action="store_true",
                    help="also time the sweep kernel's phases on the device clock (adds barriers)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from smmp_amd import _ffi, workloads

    torch.cuda.set_device(0)
    if args.da:
        run_da(args)
        return
    if args.esolan:
        run_esolan(args)
        return
    if args.c3:
        run_c3(args)
        return
    if args.muca:
        run_muca(args)
        return
    if args.minim:
        run_minim(args)
        return
    if args.copy_probe:
        x = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {args.copy_probe} MiB x 5")
        return
    tmp = tempfile.mkdtemp(prefix="smmp_bench_")
    c, seq, seqf = make_system(tmp)
    temps = workloads.geometric_temperatures(NREP, TMIN, TMAX)
    nvr = c.nvr
    log(f"system: {LENGTH} residues, {int(c.view('res_i.iatrs2')[LENGTH - 1])} atoms, {nvr} free torsions")
    v0 = random_starts(c, NREP, 12345)
    seeds_w = np.arange(1, NREP + 2, dtype=np.uint32)
    seeds_t = np.arange(1001, 1001 + NREP + 1, dtype=np.uint32)
    log(f"warmup {args.warmup} sweeps")
    kw = dict(mode=1, precision=args.precision, with_traj=False, groups=args.groups, graph_sweeps=args.graph_sweeps)
    out = c.sample(temps, seeds_w, args.warmup, v0, **kw)
    v1 = out["vlvr"]
    c.profile(2 if args.phases else 1)
    torch.cuda.synchronize()
    log(f"timed {args.steps} sweeps")
    t0 = time.perf_counter()
    out = c.sample(temps, seeds_t, args.steps, v1, **kw)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    info = c.sampler_info()
    log(f"launch: {info['groups']} workgroups per replica, persistent {info['persistent']}, "
        f"graph replays {info['graph_replays']} (kept graph {info['graph_reused']})")
    ks = c.profile_read(_ffi.KERNEL_SWEEP)
    kx = c.profile_read(_ffi.KERNEL_EXCHANGE)
    ntrial = NREP * args.steps * nvr
    log("work per trial: " + ", ".join(f"{k} {ks[k] / ntrial:.1f}" for k in ("pairs", "contacts", "builds", "moved")))
    if args.phases:
        ph = c.profile_phases()
        log("phase us per trial: " + ", ".join(f"{k} {ph[k] * 1e3 / ntrial:.2f}" for k in _ffi.Context.PHASES)
            + f"; full-path trials {ph['full_trials']:.0f}; core clock {ph['core_mhz']:.0f} MHz")
        na = max(ph["sol_atoms"], 1.0)
        log(f"solvent per trial (block 0 of each replica): slowest wave {ph['sol_max_wave'] * 1e3 / ntrial:.2f} us; "
            f"per atom: gather {ph['sol_gather'] * 1e3 / na:.3f} us, point tests {ph['sol_points'] * 1e3 / na:.3f} us; "
            f"atoms per trial {na / ntrial:.1f}")
        log(f"pair change per trial: slowest wave {ph['pair_max_wave'] * 1e3 / ntrial:.2f} us, "
            f"wave average {ph['pair_sum'] * 1e3 / ntrial / 8:.2f} us; contact marking {ph['mark_sum'] * 1e3 / ntrial / 8:.2f} us; "
            f"torsion term (thread 0) {ph['torsion'] * 1e3 / ntrial:.2f} us")
        log(f"reduction per trial (workgroup 0 of each replica): waiting for the last arrival "
            f"{ph['sync_wait'] * 1e3 / ntrial:.2f} us, from the last arrival to release {ph['sync_latency'] * 1e3 / ntrial:.2f} us")
    c.profile(False)
    value = NREP * args.steps / dt
    npnt = int(c.view("pnt.npnt")[0])
    flops = (FLOP_PAIR * ks["pairs"] + FLOP_POINT * npnt * ks["contacts"] + FLOP_BUILD * ks["builds"]
             + FLOP_ROTATE * ks["moved"])
    kern_s = ks["ms"] / 1e3
    achieved = flops / kern_s / 1e12 if kern_s > 0 else 0.0
    log(f"sweep kernel: {ks['launches']} launches, {ks['ms'] / max(ks['launches'], 1):.3f} ms avg; "
        f"exchange {kx['ms'] / max(kx['launches'], 1):.4f} ms avg; accept {out['acc'].sum() / (NREP * args.steps * nvr):.3f}")
    cpu = None if args.no_cpu else cpu_baseline(seqf, 20.0)
    line = {
        "metric": "MC sweeps/s over all replicas",
        "value": round(value, 2),
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32+f64",
        "data": "synthetic: random 20-residue sequence (seed 2), random initial dihedrals",
        "config": {"workload": f"parallel tempering, {NREP} replicas x {LENGTH}-residue random sequence, "
                               f"ECEPP/3+OONS, T geometric {TMIN:g}-{TMAX:g} K, exchange every sweep",
                   "replicas": NREP, "residues": LENGTH, "free_torsions": nvr, "solvent": "OONS",
                   "precision": "fp64" if args.precision == 0 else "mixed",
                   "workgroups_per_replica": info["groups"]},
        "energy_evals_per_s_equiv": round(value * (nvr + 2), 1),
        "roofline": {"bound": "valu", "achieved": round(achieved, 4), "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64, 5),
                     "traffic": (round(HBM_BYTES_PER_SWEEP * args.steps / max(ks["launches"], 1))
                                 if args.precision == 0 else None),
                     "kernel": "k_sweep", "avg_launch_ms": round(ks["ms"] / max(ks["launches"], 1), 4),
                     "flop_per_launch": round(flops / max(ks["launches"], 1), 1)},
        "cpu_baseline": cpu,
    }
    if not args.no_extras:
        # the other two parts of the metric: energy evaluations/s (C3) and
        # delayed acceptance against the plain sampler (C5)
        line["energy_evals_per_s"] = c3_headline(5)
        line["minimisation"] = minim_headline(5)
        R5 = 128
        c5, _, _ = make_system(tmp, length=36, seed=5, tag="c5")
        t0 = time.perf_counter()
        da = da_experiment(c5, R5, workloads.geometric_temperatures(R5, TMIN, TMAX), 300, 2, 40, 50, 8,
                           [1.0, 0.5, 0.35], 200, 0, 0, "da")
        da["workload"] = ("parallel tempering, 128 replicas x 36-residue random sequence (seed 5), ECEPP/3+OONS, "
                          "250-1000 K; network trained online on the sampler's own moves (1,024,000 exact "
                          "energies); 200 timed sweeps per sampler")
        da["total_s"] = round(time.perf_counter() - t0, 1)
        line["da_gain"] = da
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
