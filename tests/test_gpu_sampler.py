"""Metropolis, parallel tempering and multicanonical sweeps on the GPU.

Trajectories are compared move by move: the GPU draws from the same
twister.f stream as the oracle / the Fortran reference, so acceptance
counts must be identical and energies agree to 1e-8 kcal/mol (the GPU tracks
the energy incrementally between the per-sweep full recomputations).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden, make_ctx

pytestmark = pytest.mark.gpu
GASC = 1.98773e-3


def say(*a):
    print("[gpu-sampler]", *a, flush=True)


def test_metropolis_c1_trajectory_vs_reference():
    """C1: Met-enkephalin, omegas fixed (19 free), ECEPP/3 vacuum, 300 K."""
    from smmp_amd import _ffi
    r = golden("enkefa_runs.npz")
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa_c1.var"))
    assert c.nvr == 19
    c.sgrnd(81236)
    beta = 1.0 / (300.0 * GASC)
    c.view("bet.beta")[0] = beta
    w = _ffi.Weight(kind=0, beta=beta)
    eol, acz = 0.0, 0.0
    say("smmp_metropolis C1 40 sweeps")
    for k in range(40):
        eol, acz = c.metropolis(eol, acz, w)
        assert acz == r["canon_sweeps"][k, 1], k
        assert abs(eol - r["canon_sweeps"][k, 0]) <= 1e-8 * max(1.0, abs(eol)), k
    assert np.abs(c.vlvr()[c.idvr - 1] - r["canon_vlvr"]).max() < 1e-12
    assert int(c.view("myblock.mti")[0]) == int(r["canon_mti"][0])


@pytest.mark.parametrize("tag", ["upch1", "upch2", "muca", "muca_upch1"])
def test_metropolis_variants_vs_reference(tag):
    """metropolis.f with the update-type draws (upchswitch 1 and 2, rndord)
    and a multicanonical weight b(k)x + alpha(k), move for move against the
    Fortran's runs: acceptances and energies per sweep, the final torsions,
    the twister position and accanalyze's counts (/updstats/)."""
    from smmp_amd import _ffi
    g = golden("metro_enkefa.npz")
    seed, temp, nsw, upch, bgs, rnd = g[f"{tag}_args"]
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa_c1.var"))
    beta = 1.0 / (temp * GASC)
    c.view("bet.beta")[0] = beta
    c.view("updchois.upchswitch")[0] = int(upch)
    c.view("updchois.bgsprob")[0] = float(bgs)
    c.view("updchois.rndord")[0] = int(rnd)
    c.sgrnd(int(seed))
    if tag.startswith("muca"):
        b = np.ascontiguousarray(g["muca_b"])
        al = np.ascontiguousarray(g["muca_alpha"])
        w = _ffi.Weight(kind=2, beta=beta, kmin=int(g["muca_kmin"][0]), kmax=int(g["muca_kmax"][0]),
                        ebin=float(g["muca_ebin"][0]), b=_ffi.dptr(b), alpha=_ffi.dptr(al))
    else:
        w = _ffi.Weight(kind=0, beta=beta)
    sw = g[f"{tag}_sweeps"]
    eol, acz = 0.0, 0.0
    say(f"smmp_metropolis enkefa {tag} {int(nsw)} sweeps")
    for k in range(int(nsw)):
        eol, acz = c.metropolis(eol, acz, w)
        assert acz == sw[k, 1], (tag, k)
        assert abs(eol - sw[k, 0]) <= 1e-8 * max(1.0, abs(eol)), (tag, k)
    assert np.abs(c.vlvr()[c.idvr - 1] - g[f"{tag}_vlvr"]).max() < 1e-12
    assert int(c.view("myblock.mti")[0]) == int(g[f"{tag}_mti"][0])
    counts = np.concatenate([c.view("updstats.ncalls"), c.view("updstats.nacalls")])
    assert np.array_equal(counts.astype(np.int64), g[f"{tag}_ncalls"])


def test_metropolis_with_solvent_vs_reference():
    from smmp_amd import _ffi
    r = golden("enkefa_runs.npz")
    c = make_ctx(tag="enkefa", itysol=1, varfile=os.path.join(GOLDEN, "enkefa.var"))
    c.sgrnd(4357)
    beta = 1.0 / (500.0 * GASC)
    w = _ffi.Weight(kind=0, beta=beta)
    eol, acz = 0.0, 0.0
    say("smmp_metropolis enkefa OONS 10 sweeps")
    for k in range(10):
        eol, acz = c.metropolis(eol, acz, w)
        assert acz == r["canon_sol_sweeps"][k, 1], k
        assert abs(eol - r["canon_sol_sweeps"][k, 0]) <= 1e-8 * max(1.0, abs(eol)), k


@pytest.mark.parametrize("tag", ["len10", "c2_20mer"])
def test_partem_s_serial_compatible_vs_reference(tag):
    """partem_s.f on one stream from the extended chain (switch < 0): the
    energies of time.d (3 decimals), its radius of gyration and C-alpha
    end-to-end distance (rgyr after each slot's sweep), and the energies of
    par_R.in.  (From random starts the
    first sweeps sit at clash energies ~1e13 kcal/mol where a move's energy
    change is below the rounding of the total, so the reference's own
    decisions there depend on its summation order.)"""
    p = golden(f"partem_ext_{tag}.npz")
    nswp, R = int(p["nswp"][0]), len(p["temps"])
    c = make_ctx(tag=tag, itysol=1)
    c.sgrnd(int(p["seed"][0]))
    say(f"smmp_partem_s {tag} {R} replicas {nswp} sweeps")
    eol, ipoi, traj, obs = c.partem_s(p["temps"], int(p["nequi"][0]), nswp, int(p["switch"][0]), with_obs=True)
    ref = p["time_d"][:, 3].reshape(nswp, R)
    assert np.abs(traj - ref).max() <= 6e-4
    assert np.abs(obs[:, :, 0] - p["time_d"][:, 4].reshape(nswp, R)).max() <= 6e-4
    assert np.abs(obs[:, :, 1] - p["time_d"][:, 5].reshape(nswp, R)).max() <= 6e-4
    assert np.abs(eol - p["eol_final"]).max() <= 1e-8 * max(1.0, np.abs(p["eol_final"]).max())


def test_batched_canonical_replicas_match_oracle_streams():
    """Each replica r of the batched sampler = oracle metropolis on stream seed_r."""
    from _oracle import OracleMT, OracleSystem
    c = make_ctx(tag="len10", itysol=1)
    osys = OracleSystem(c)
    R, nsw = 6, 4
    temps = np.array([250.0, 300.0, 400.0, 550.0, 750.0, 1000.0])
    seeds = np.array([11, 22, 33, 44, 55, 66], dtype=np.uint32)
    rng = np.random.default_rng(1)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, osys.idvr - 1] = rng.uniform(-np.pi, np.pi, size=(R, osys.nvr))
    say("smmp_sample canonical len10 R=6 4 sweeps")
    out = c.sample(temps, seeds, nsw, v0, mode=0)
    for r in range(R):
        mt = OracleMT(int(seeds[r]))
        osys.reset()
        vl, tr = osys.metropolis(mt, v0[r], 1.0 / (temps[r] * GASC), nsw)
        assert out["acc"][r] == tr[-1][1], r
        for k in range(nsw):
            assert abs(out["traj"][k, r] - tr[k][0]) <= 1e-8 * max(1.0, abs(tr[k][0])), (r, k)
        assert np.abs(out["vlvr"][r] - vl).max() < 1e-12


def test_batched_parallel_tempering_matches_python_restatement():
    """PT with one stream per replica plus an exchange stream (partem_s.f:209-232 rule)."""
    from _oracle import OracleMT, OracleSystem
    c = make_ctx(tag="len10", itysol=1)
    osys = OracleSystem(c)
    R, nsw = 4, 5
    temps = np.array([250.0, 400.0, 630.0, 1000.0])
    seeds = np.array([7, 8, 9, 10, 99], dtype=np.uint32)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, osys.idvr - 1] = np.pi  # extended chain: well-conditioned energies from the start
    say("smmp_sample PT len10 R=4 5 sweeps")
    out = c.sample(temps, seeds, nsw, v0, mode=1)
    # restatement: replica k sweeps at the temperature of its slot, then exchanges
    mts = [OracleMT(int(s)) for s in seeds[:R]]
    mtx = OracleMT(int(seeds[R]))
    vl = [v0[k].copy() for k in range(R)]
    eol = np.zeros(R)
    ipoi = list(range(1, R + 1))
    pbe = 1.0 / (temps * GASC)
    iswitch = 1
    for sw in range(nsw):
        slot_of = {ipoi[s] - 1: s for s in range(R)}
        for k in range(R):
            osys.reset()
            vl[k], tr = osys.metropolis(mts[k], vl[k], pbe[slot_of[k]], 1)
            eol[k] = tr[-1][0]
        if iswitch == 1:
            nu, last, iswitch = 1, R - 1, 2
        else:
            nu, last, iswitch = 2, R, 1
        for i in range(nu, last + 1, 2):
            j = 1 if i == R else i + 1
            a, b = ipoi[i - 1], ipoi[j - 1]
            d = -pbe[i - 1] * eol[b - 1] - pbe[j - 1] * eol[a - 1] + pbe[i - 1] * eol[a - 1] + pbe[j - 1] * eol[b - 1]
            if mtx.grnd() <= np.exp(d):
                ipoi[i - 1], ipoi[j - 1] = b, a
    assert list(out["ipoi"]) == ipoi
    assert np.abs(out["eol"] - eol).max() <= 1e-8 * max(1.0, np.abs(eol).max())


def test_multicanonical_histogram_and_consistency():
    from smmp_amd import _ffi
    from _oracle import OracleSystem
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa.var"))
    osys = OracleSystem(c)
    kmin, kmax, ebin = -20, 179, 1.0
    nb = kmax - kmin + 1
    b = np.full(nb, 1.0 / (600.0 * GASC))
    alpha = np.zeros(nb)
    w = _ffi.Weight(kind=2, beta=0.0, kmin=kmin, kmax=kmax, ebin=ebin,
                    b=_ffi.dptr(b), alpha=_ffi.dptr(alpha))
    R, nsw = 4, 6
    v0 = np.tile(c.vlvr(), (R, 1))
    say("smmp_sample multicanonical enkefa R=4")
    out = c.sample(np.full(R, 600.0), np.arange(1, R + 1, dtype=np.uint32), nsw, v0, mode=2, weight=w)
    assert out["hist"].sum() == R * nsw
    for r in range(R):
        ref = osys.energy(out["vlvr"][r])
        assert abs(out["eol"][r] - ref[0]) <= 1e-9 * max(1.0, abs(ref[0]))


def test_verify_mode_incremental_equals_full_recomputation():
    """Every trial's incremental energy change (moved/fixed pairs, torsion,
    solvent of the affected atoms) equals a full recomputation, C2 system."""
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 4
    v0 = np.tile(c.vlvr(), (R, 1))
    rng = np.random.default_rng(4)
    idv = c.idvr - 1
    v0[:2, idv] = rng.uniform(-np.pi, np.pi, size=(2, len(idv)))  # clash starts: full-path trials
    v0[2:, idv] = np.pi + rng.normal(0.0, 0.3, size=(2, len(idv)))  # near-extended starts
    say("smmp_sample verify c2_20mer R=4 2 sweeps")
    out = c.sample(np.array([250.0, 400.0, 630.0, 1000.0]), np.arange(31, 35, dtype=np.uint32), 2, v0,
                   mode=0, verify=True)
    assert out["rc"] == 0


def test_c2_system_sweeps_consistent_with_oracle():
    from _oracle import OracleSystem
    c = make_ctx(tag="c2_20mer", itysol=1)
    o = OracleSystem(c)
    R = 16
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, o.idvr - 1] = np.random.default_rng(8).uniform(-np.pi, np.pi, size=(R, o.nvr))
    say("smmp_sample PT c2_20mer R=16 15 sweeps")
    out = c.sample(np.geomspace(250.0, 1000.0, R), np.arange(1, R + 2, dtype=np.uint32), 15, v0, mode=1)
    for r in range(R):
        ref = o.energy(out["vlvr"][r])
        assert abs(out["eol"][r] - ref[0]) <= 1e-9 * max(1.0, abs(ref[0]))
    assert sorted(out["ipoi"]) == list(range(1, R + 1))


def test_energy_histogram_matches_oracle_sampler_ks():
    """Sampled energies of the batched GPU sampler and of the CPU oracle's
    Metropolis chains (different streams) come from one distribution."""
    from scipy.stats import ks_2samp
    from _oracle import OracleMT, OracleSystem
    c = make_ctx(tag="enkefa", itysol=1)
    o = OracleSystem(c)
    T, nsw, burn, every = 600.0, 240, 80, 20
    R = 64
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, o.idvr - 1] = np.pi
    say("smmp_sample canonical enkefa R=64 240 sweeps")
    out = c.sample(np.full(R, T), np.arange(100, 100 + R, dtype=np.uint32), nsw, v0, mode=0)
    gpu = out["traj"][burn - 1::every, :].ravel()
    cpu = []
    for k in range(8):
        o.reset()
        _, tr = o.metropolis(OracleMT(9000 + k), v0[0], 1.0 / (T * GASC), nsw)
        e = np.array([t[0] for t in tr])
        cpu.append(e[burn - 1::every])
    cpu = np.concatenate(cpu)
    p = ks_2samp(gpu, cpu).pvalue
    say(f"KS p = {p:.3g}, gpu mean {gpu.mean():.3f}, cpu mean {cpu.mean():.3f}")
    assert p > 1e-3


@pytest.mark.parametrize("groups", [2, 3, 4, 8])
def test_group_split_matches_single_workgroup(groups):
    """G workgroups per replica (device reductions every trial) take the same
    decisions as one workgroup; energies agree to rounding."""
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 8
    temps = np.geomspace(250.0, 1000.0, R)
    seeds = np.arange(1, R + 2, dtype=np.uint32)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, c.idvr - 1] = np.pi + np.random.default_rng(5).normal(0.0, 0.5, size=(R, c.nvr))
    say(f"smmp_sample PT c2_20mer R=8 6 sweeps groups=1 vs {groups}")
    a = c.sample(temps, seeds, 6, v0, mode=1, groups=1)
    b = c.sample(temps, seeds, 6, v0, mode=1, groups=groups)
    assert np.array_equal(a["acc"], b["acc"])
    assert np.array_equal(a["ipoi"], b["ipoi"])
    assert np.abs(a["traj"] - b["traj"]).max() <= 1e-8 * max(1.0, np.abs(a["traj"]).max())
    assert np.abs(a["vlvr"] - b["vlvr"]).max() < 1e-12


@pytest.mark.parametrize("groups", [2, 4])
def test_speculative_batches_match_sequential(groups, monkeypatch):
    """SMMP_SPEC=1: the G workgroups of a replica evaluate the next G trials
    at once, each assuming the earlier ones rejected, and take the decisions
    in order: the same acceptances, exchanges and twister positions as one
    trial at a time (energies to rounding), with the solvent term and the
    update-type draws."""
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 8
    temps = np.geomspace(250.0, 1000.0, R)
    seeds = np.arange(11, R + 12, dtype=np.uint32)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, c.idvr - 1] = np.pi + np.random.default_rng(9).normal(0.0, 0.5, size=(R, c.nvr))
    c.view("updchois.upchswitch")[0] = 1
    c.view("updchois.bgsprob")[0] = 0.2
    say(f"smmp_sample PT c2_20mer R=8 5 sweeps sequential vs speculative G={groups}")
    a = c.sample(temps, seeds, 5, v0, mode=1, groups=1)
    monkeypatch.setenv("SMMP_SPEC", "1")
    b = c.sample(temps, seeds, 5, v0, mode=1, groups=groups)
    monkeypatch.delenv("SMMP_SPEC")
    assert c.sampler_info()["groups"] == groups
    assert np.array_equal(a["acc"], b["acc"])
    assert np.array_equal(a["ipoi"], b["ipoi"])
    assert np.abs(a["traj"] - b["traj"]).max() <= 1e-8 * max(1.0, np.abs(a["traj"]).max())
    assert np.abs(a["vlvr"] - b["vlvr"]).max() < 1e-12


def test_mixed_precision_sampler_runs_and_tracks_fp64():
    """precision=1: FP32 pair terms in the full recomputations, FP64 trial
    changes; the drift check uses the mixed bound.  Final energies against
    the FP64 oracle at the sampled conformations."""
    from _oracle import OracleSystem
    c = make_ctx(tag="c2_20mer", itysol=1)
    o = OracleSystem(c)
    R = 8
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, o.idvr - 1] = np.pi + np.random.default_rng(12).normal(0.0, 0.6, size=(R, o.nvr))
    say("smmp_sample PT c2_20mer R=8 12 sweeps mixed precision")
    out = c.sample(np.geomspace(250.0, 1000.0, R), np.arange(1, R + 2, dtype=np.uint32), 12, v0, mode=1,
                   precision=1)
    worst = 0.0
    for r in range(R):
        ref = o.energy(out["vlvr"][r])
        scale = abs(ref[2]) + abs(ref[3]) + abs(ref[4])
        err = abs(out["eol"][r] - ref[0])
        worst = max(worst, err / (scale + 1.0))
        assert err <= 2e-5 * scale + 1e-3, (r, out["eol"][r], ref[0])
    say(f"mixed: worst |E - E_fp64| / (|el|+|vw|+|hb|+1) = {worst:.2e}")


def _pt_inputs(c, R, seed):
    temps = np.geomspace(250.0, 1000.0, R)
    seeds = np.arange(7, 7 + R + 1, dtype=np.uint32)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, c.idvr - 1] = np.random.default_rng(seed).uniform(-np.pi, np.pi, size=(R, c.nvr))
    return temps, seeds, v0


def _same_run(a, b):
    for k in ("acc", "ipoi", "eol", "vlvr", "traj"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("launch", ["persistent", "per-sweep", "canonical"])
def test_graph_replay_matches_direct_launches(launch, monkeypatch):
    """Sweeps replayed from a captured HIP graph (3 sweeps per replay) give
    the run of direct launches bit for bit; a later call with the same
    configuration replays the kept graph, and its trajectory rows start
    again at sweep 0."""
    if launch == "per-sweep":
        monkeypatch.setenv("SMMP_PT_LAUNCH", "per-sweep")
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 8
    mode = 0 if launch == "canonical" else 1
    temps, seeds, v0 = _pt_inputs(c, R, 31)
    say(f"smmp_sample c2_20mer R=8 12 sweeps {launch}: direct vs graph of 3 sweeps")
    a = c.sample(temps, seeds, 12, v0, mode=mode, groups=4)
    assert c.sampler_info()["graph_replays"] == 0
    b = c.sample(temps, seeds, 12, v0, mode=mode, groups=4, graph_sweeps=3)
    info = c.sampler_info()
    assert info["graph_replays"] == 4 and not info["graph_reused"]
    assert info["persistent"] == (launch == "persistent")
    _same_run(a, b)
    say("second graph call (kept graph)")
    b2 = c.sample(temps, seeds, 12, v0, mode=mode, groups=4, graph_sweeps=3)
    assert c.sampler_info()["graph_reused"]
    _same_run(a, b2)
    say("6 sweeps: direct vs graph")
    a6 = c.sample(temps, seeds, 6, v0, mode=mode, groups=4)
    b6 = c.sample(temps, seeds, 6, v0, mode=mode, groups=4, graph_sweeps=3)
    assert c.sampler_info()["graph_replays"] == 2
    _same_run(a6, b6)
    assert np.array_equal(a6["traj"], a["traj"][:6])


def test_graph_sweeps_must_divide_nsweep():
    c = make_ctx(tag="len10", itysol=1)
    temps, seeds, v0 = _pt_inputs(c, 2, 3)
    with pytest.raises(RuntimeError, match="graph_sweeps"):
        c.sample(temps, seeds, 7, v0, mode=1, graph_sweeps=3)


def test_measured_groups_keep_the_start_state():
    """groups=-1 times 1, 2, 4, 8 workgroups per replica on the call's own
    replicas and puts their start state back: the run equals an explicit
    run at the chosen count, and the choice is kept for the next call."""
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 16
    temps, seeds, v0 = _pt_inputs(c, R, 44)
    say("smmp_sample c2_20mer R=16 5 sweeps groups=-1")
    t = c.sample(temps, seeds, 5, v0, mode=1, groups=-1)
    G = c.sampler_info()["groups"]
    say(f"measured groups: {G}")
    assert G in (1, 2, 4, 8)
    e = c.sample(temps, seeds, 5, v0, mode=1, groups=G)
    _same_run(t, e)
    t2 = c.sample(temps, seeds, 5, v0, mode=1, groups=-1)
    assert c.sampler_info()["groups"] == G
    _same_run(t2, e)


def test_sample_observables_match_the_final_coordinates():
    """smmp_sample's per-sweep radius of gyration and end-to-end distance:
    the last sweep's values equal rgyr of each replica's final variables."""
    c = make_ctx(tag="len10", itysol=1)
    R, nsw = 4, 3
    temps = np.array([300.0, 450.0, 650.0, 900.0])
    rng = np.random.default_rng(4)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, c.idvr - 1] = np.pi + rng.normal(0, 0.3, size=(R, c.nvr))
    say("smmp_sample canonical len10 R=4 3 sweeps with observables")
    out = c.sample(temps, np.arange(7, 7 + R, dtype=np.uint32), nsw, v0, mode=0, obs=True)
    assert out["obs"].shape == (nsw, R, 2)
    for r in range(R):
        c.set_vlvr(out["vlvr"][r])
        c.energy()  # rebuilds the coordinates of these variables
        rgy, ee = c.rgyr(0)
        assert abs(out["obs"][-1, r, 0] - rgy) <= 1e-9 * max(1.0, rgy), r
        assert abs(out["obs"][-1, r, 1] - ee) <= 1e-9 * max(1.0, ee), r


def test_partem_s_restart_continues_the_run():
    """One replica (no exchanges): 3 sweeps, then a restart from the returned
    variables for 2 more on the same twister, equal the 5 sweeps of one run
    (the sweep-end energy and coordinates are full recomputations, so a
    restart from the variables loses nothing)."""
    c = make_ctx(tag="len10", itysol=1)
    T = np.array([400.0])
    c.sgrnd(99)
    say("smmp_partem_s len10 one replica: 3 sweeps, restart for 2, and 5 in one run")
    _, _, tra, ca = c.partem_s(T, 2, 3, -1, with_coor=True)
    eb, _, trb, cb = c.partem_s(T, 0, 2, -1, start=ca, with_coor=True)
    c.sgrnd(99)
    ec, _, trc, cc = c.partem_s(T, 2, 5, -1, with_coor=True)
    assert np.array_equal(tra, trc[:3])
    assert np.array_equal(trb, trc[3:])
    assert np.array_equal(cb, cc) and np.array_equal(eb, ec)
