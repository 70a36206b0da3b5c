"""The drop-in Python API on the GPU: Universe.energy(), the canonical Monte
Carlo of algorithms.py and smmp.partem_s reproduce the reference's numbers."""
import os

import numpy as np
import pytest

from conftest import golden
from test_dropin import run

pytestmark = pytest.mark.gpu


def test_universe_energy_readme_known_answer():
    """README:389-393 (Met-enkephalin at enkefa.var, ECEPP/3, vacuum)."""
    out = run("energy_gpu", {"st": 0})
    assert round(out["e"], 1) == 1035.4
    assert round(out["eyel"], 2) == 19.91 and round(out["eyvw"], 1) == 114.2 and round(out["eyhb"], 1) == 900.8
    assert round(out["eyvr"], 4) == 0.4511


def test_minimization_example_readme_known_answer():
    """EXAMPLES/Python/minimization.py through the drop-in API: README's
    final energies (Total -0.12429E+02, Coulomb 0.2143E+02, Lennard-Jones
    -0.2923E+02, HB -0.6706E+01, Variables 0.2084E+01) and the Fortran's
    minimum (tests/golden/minim_vacuum.npz)."""
    m = golden("minim_vacuum.npz")
    out = run("minimization_gpu", {})
    assert float(f"{out['eysm']:.4E}") == -12.429
    assert float(f"{out['eyel']:.3E}") == 21.43 and float(f"{out['eyvw']:.3E}") == -29.23
    assert float(f"{out['eyhb']:.3E}") == -6.706 and float(f"{out['eyvr']:.3E}") == 2.084
    assert abs(out["energy"] - m["enkefa_qn_energy"][0][0]) < 1e-8
    assert np.abs(np.array(out["vlvr"]) - m["enkefa_qn_vlvr"][0]).max() < 1e-7


def test_canonical_monte_carlo_trajectory_vs_reference():
    """algorithms.CanonicalMonteCarlo sweeps = the reference's metropolis
    trajectory (C1: 19 free torsions, 300 K, seed 81236)."""
    r = golden("enkefa_runs.npz")
    out = run("canonical_gpu", {"seed": 81236, "sweeps": 40})
    assert np.array_equal(np.array(out["acc"]), r["canon_sweeps"][:, 1])
    e = np.array(out["energies"])
    assert np.abs(e - r["canon_sweeps"][:, 0]).max() <= 1e-8 * np.abs(e).max()


def test_canonical_monte_carlo_with_a_lambda_weight_vs_reference():
    """The reference's default weight, lambda x: x * smmp.bet.beta, goes
    through the host-driven trial loop and gives the same trajectory."""
    r = golden("enkefa_runs.npz")
    out = run("canonical_gpu", {"seed": 81236, "sweeps": 12, "weight": "lambda"})
    assert np.array_equal(np.array(out["acc"]), r["canon_sweeps"][:12, 1])
    e = np.array(out["energies"])
    assert np.abs(e - r["canon_sweeps"][:12, 0]).max() <= 1e-8 * np.abs(e).max()


@pytest.mark.parametrize("tag", ["muca", "upch1", "muca_upch1"])
def test_python_weight_callback_vs_reference(tag):
    """A Python weight callable (muca_weight2 written as a plain function, or
    a lambda) with update-type draws: move for move the Fortran's run."""
    g = golden("metro_enkefa.npz")
    out = run("callback_gpu", {"tag": tag})
    sw = np.array(out["sweeps"])
    assert np.array_equal(sw[:, 1], g[f"{tag}_sweeps"][:, 1])
    assert np.abs(sw[:, 0] - g[f"{tag}_sweeps"][:, 0]).max() <= 1e-8 * max(1.0, np.abs(sw[:, 0]).max())
    assert np.abs(np.array(out["vlvr"]) - g[f"{tag}_vlvr"]).max() < 1e-12
    assert out["mti"] == int(g[f"{tag}_mti"][0])
    assert out["ncalls"] == [int(x) for x in g[f"{tag}_ncalls"]]


def test_canon_time_d_vs_reference(tmp_path):
    """canon.f's driver: time.d carries the energies of the reference's
    Metropolis run (sweeps 0..nswp, canon.f:81-112, format i7,2f15.3,5i7,
    4f15.3), and the files canon.f writes exist."""
    r = golden("enkefa_runs.npz")
    out = run("canon_gpu", {"dir": str(tmp_path), "seed": 81236, "nswp": 19, "nmes": 1})
    for f in ("start.pdb", "final.pdb", "lastconf.var", "time.d"):
        assert f in out["files"], f
    td = [line for line in open(tmp_path / "time.d").read().splitlines()]
    assert len(td) == 20 and all(len(line) == 7 + 30 + 35 + 60 for line in td)
    e = np.array([float(line[7:22]) for line in td])
    assert np.abs(e - r["canon_sweeps"][:20, 0]).max() <= 6e-4
    assert [int(line[:7]) for line in td] == list(range(20))


def test_partem_s_files_vs_reference(tmp_path):
    p = golden("partem_ext_len10.npz")
    from smmp_amd import workloads
    seq = list(golden("topo_len10.npz")["seq_in"])
    seqf = workloads.write_seq_file(str(tmp_path / "s.seq"), seq, "t")
    R, nswp = len(p["temps"]), int(p["nswp"][0])
    workloads.write_temperatures(str(tmp_path / "temperatures"), p["temps"])
    out = run("partem_gpu", {"dir": str(tmp_path), "seq": seqf, "seed": int(p["seed"][0]), "nrep": R,
                             "nequi": int(p["nequi"][0]), "nswp": nswp, "switch": int(p["switch"][0])})
    rows = np.array([[float(v) for v in line.split()] for line in out["time_d"].splitlines()])
    assert np.array_equal(rows[:, :3], p["time_d"][:, :3])
    # energy, radius of gyration, C-alpha end-to-end distance (3 decimals)
    assert np.abs(rows[:, 3:6] - p["time_d"][:, 3:6]).max() <= 6e-4


def test_partem_s_restart_files(tmp_path):
    """partem_s writes par_R.in and conf_NNNN.var; with newsta false it runs
    sweeps nstart..nswp from those conformations (partem_s.f:97-111, 179-255)."""
    from smmp_amd import workloads
    seq = list(golden("topo_len10.npz")["seq_in"])
    seqf = workloads.write_seq_file(str(tmp_path / "s.seq"), seq, "t")
    temps = workloads.geometric_temperatures(3, 250.0, 1000.0)
    workloads.write_temperatures(str(tmp_path / "temperatures"), temps)
    out = run("partem_restart_gpu", {"dir": str(tmp_path), "seq": seqf, "seed": 4242, "nrep": 3, "nequi": 2,
                                     "nswp1": 4, "nswp2": 7, "switch": -1})
    assert out["nstart"] == 4 and out["nstart2"] == 7
    first = [line.split() for line in out["first"].splitlines()]
    second = [line.split() for line in out["second"].splitlines()]
    assert [int(r[0]) for r in first] == [s for s in range(1, 5) for _ in range(3)]
    assert [int(r[0]) for r in second] == [s for s in range(4, 8) for _ in range(3)]
    assert np.allclose(out["eol2"], out["eol"])
    assert all(len(text.splitlines()) > 0 for text in out["confs1"])


def test_mulcan_par_matches_oracle_iteration(tmp_path):
    """smmp.multicanonical.mulcan_par (GPU sweeps + host recursion) =
    the same iteration driven by the oracle's Metropolis and recursion:
    random start, muca_weight sweeps, histogram updates every nup sweeps."""
    from _oracle import F64P, I32P, OracleMT, OracleSystem, Weight, lib as olib
    from conftest import GOLDEN, make_ctx
    seed, nsweep, nup, temp, kmin, kmax, ebin = 4321, 12, 4, 300.0, -10, 120, 5.0
    out = run("muca_gpu", {"dir": str(tmp_path), "seed": seed, "nsweep": nsweep, "nup": nup, "temp": temp,
                           "kmin": kmin, "kmax": kmax, "ebin": ebin, "nsim": 6})
    c = make_ctx(tag="enkefa", itysol=0)
    o = OracleSystem(c)
    mt = OracleMT(seed)
    vl = c.vlvr()
    ax = c.view("var_r.axvr")
    for iv in o.idvr:
        dv = ax[iv - 1] * (mt.grnd() - 0.5)
        d = np.fmod(np.pi + dv, 2 * np.pi)
        vl[iv - 1] = d if abs(d) <= np.pi else d - np.copysign(2 * np.pi, d)
    n = kmax - kmin + 1
    b, alpha, xh, g1 = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    ih = np.zeros(n, dtype=np.int32)
    beta = 1.0 / (300.0 * 1.98773e-3)  # COMMON /bet/ set by Universe(T=300)
    for nsw in range(1, nsweep + 1):
        w = Weight(kind=1, beta=beta, kmin=kmin, kmax=kmax, ebin=ebin, b=b.ctypes.data_as(F64P),
                   alpha=alpha.ctypes.data_as(F64P))
        o.reset()
        vl, tr = o.metropolis(mt, vl, beta, 1, weight=w)
        e = tr[-1][0]
        v = e / ebin + np.copysign(0.5, e)
        ih[int(min(max(v, kmin), kmax)) - kmin] += 1
        if nsw % nup == 0:
            olib.oracle_muca_update(kmin, kmax, ebin, beta, ih.ctypes.data_as(I32P), xh.ctypes.data_as(F64P),
                                    g1.ctypes.data_as(F64P), b.ctypes.data_as(F64P), alpha.ctypes.data_as(F64P))
            ih[:] = 0
    b = b + beta
    for i in range(n - 2, -1, -1):
        alpha[i] = alpha[i + 1] + (b[i + 1] - b[i]) * (kmin + i + 0.5) * ebin
    alpha[-1] = 0.0
    assert np.array_equal(np.array(out["xhist"]), xh)
    assert np.allclose(np.array(out["b"]), b, rtol=1e-12, atol=1e-12)
    assert np.allclose(np.array(out["alpha"]), alpha, rtol=1e-10, atol=1e-9)
    assert abs(out["eol"] - e) <= 1e-8 * max(1.0, abs(e))
    assert out["sim_hist"] == 6.0  # mulcan_sim: one bin count per production sweep (nmes = 2 divides 6)
    assert len(out["time_d"].splitlines()) == 3
