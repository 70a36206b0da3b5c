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
# algorithmic flops per unit of work (divide and sqrt count 1):
FLOP_PAIR = 22      # ECEPP pair term: 3 sub + 5 r^2 + 1 rsqrt + 3 powers + 2 Coulomb + 4 LJ/HB + 3 sums
FLOP_POINT = 9      # one tessellation point against one solvent neighbour
FLOP_BUILD = 60     # one atom placed by the Eyring step
FLOP_ROTATE = 21    # one atom rotated rigidly about a bond (Rodrigues)
# HBM-side bytes per sweep of this workload in FP64: FETCH_SIZE x 2.0 plus
# WRITE_SIZE x 1.0 (the factors make a 256 MiB device copy read 256 MiB and
# write 256 MiB) of the timed launch, rocprofv3 --pmc, 1000 sweeps:
# 4,587,491 KiB + 9,686,131 KiB.
HBM_BYTES_PER_SWEEP = 14.616e6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_system(tmp):
    from smmp_amd import _ffi, workloads
    seq = workloads.random_sequence(LENGTH, SEQ_SEED)
    seqf = workloads.write_seq_file(os.path.join(tmp, "c2.seq"), seq, "c2")
    c = _ffi.Context()
    for k, v in (("epar_l.flex", 0), ("epar_l.sh2", 0), ("epar_l.epsd", 0), ("epar_l.ientyp", 0),
                 ("isolty.itysol", 1)):
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
    return {"value": round(nrep * nsw / dt, 4), "unit": "sweeps/s", "cores": 1, "kind": kind,
            "sample": f"PT {nrep} replicas x {nsw} sweeps of the {LENGTH}-residue system "
                      f"(ECEPP/3+OONS, {TMIN:g}-{TMAX:g} K), one host core, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed sweeps")
    ap.add_argument("--warmup", type=int, default=30, help="untimed sweeps from the random start")
    ap.add_argument("--precision", type=int, default=0, help="0 FP64, 1 FP32 pair terms")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (calibrates HBM byte counters) and exit")
    ap.add_argument("--groups", type=int, default=0,
                    help="workgroups per replica: 0 automatic, -1 measured (autotuner), >0 upper bound")
    ap.add_argument("--graph-sweeps", type=int, default=0,
                    help="replay a captured HIP graph of this many sweeps (must divide --steps and --warmup)")
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
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
