"""The analytic solvent term (itysol < 0, esolan.f) on the GPU through the
C-ABI: smmp_energy against the Fortran reference's own outputs (start and
random conformations of three sequences, with the twister's position after
each energy, and a conformation whose degenerate north point draws a random
rotation from the context's twister), smmp_energy_batch against the CPU
oracle on many conformations of one and of two systems, and smmp_metropolis
and smmp_partem_s against the Fortran's canon and partem_s runs."""
import os
import sys

import numpy as np
import pytest

from conftest import golden, make_ctx

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__))))
from _oracle import OracleMT, OracleSystem, lib as olib  # noqa: E402

pytestmark = pytest.mark.gpu
RTOL = 1e-9  # FP64: areas from the device's atan2/sin/cos, sums in wave order


def say(*a):
    print("[gpu-esolan]", *a, flush=True)


def _ctx(tag, d):
    if tag == "enkefa":
        return make_ctx(tag="enkefa", itysol=-1)
    return make_ctx(seq=list(d["seq_in"]), itysol=-1)


def _components(c):
    v = c.view
    return np.array([float(v(f"eny_r.{k}")[0]) for k in ("eysm", "eyel", "eyvw", "eyhb", "eyvr", "eysl")])


def _close(a, b, rtol=RTOL):
    return abs(a - b) <= rtol * max(1.0, abs(b))


@pytest.mark.parametrize("tag", ["enkefa", "c2_20mer", "len10"])
def test_energy_matches_fortran(tag):
    d = golden(f"esolan_{tag}.npz")
    c = _ctx(tag, d)
    c.sgrnd(int(d["seed"][0]))
    say(f"smmp_energy {tag} start")
    e = c.energy()
    comp = _components(c)
    ref = d["energy"][0]
    assert _close(comp[5], ref[6]), (comp[5], ref[6])
    assert _close(e, ref[0]), (e, ref[0])
    assert int(c.view("myblock.mti")[0]) == d["mti"][0]
    idvr = c.idvr - 1
    ax = np.asarray(c.view("var_r.axvr")[:c.nvt], dtype=np.float64)
    for k, want in enumerate(d["conf_vlvr"]):
        v = c.vlvr()
        for j in idvr:
            v[j] = olib.oracle_addang(np.pi, ax[j] * (c.grnd() - 0.5))
        assert np.array_equal(v[idvr], want)
        c.set_vlvr(v)
        say(f"smmp_energy {tag} conformation {k}")
        e = c.energy()
        comp = _components(c)
        ref = d["energy"][k + 1]
        assert _close(comp[5], ref[6]), (k, comp[5], ref[6])
        assert _close(e, ref[0]), (k, e, ref[0])
        assert int(c.view("myblock.mti")[0]) == d["mti"][k + 1]


def test_degenerate_north_point_draws_like_fortran():
    d = golden("esolan_rotation_enkefa.npz")
    c = make_ctx(tag="enkefa", itysol=-1)
    v = c.vlvr()
    v[d["var_idx"] - 1] = d["var_rad"]
    c.set_vlvr(v)
    c.sgrnd(int(d["seed"][0]))
    say("smmp_energy enkefa, a neighbour's surface through an atom's north point")
    e = c.energy()
    assert int(c.view("myblock.mti")[0]) == 2  # one rotation: u_gamma, u_phi
    comp = _components(c)
    assert _close(comp[5], d["energy"][0][6]), (comp[5], d["energy"][0][6])
    assert _close(e, d["energy"][0][0])


def _random_confs(c, n, seed):
    rng = np.random.default_rng(seed)
    v = np.tile(c.vlvr(), (n, 1))
    v[:, c.idvr - 1] = rng.uniform(-np.pi, np.pi, size=(n, c.nvr))
    return v


def test_energy_batch_matches_oracle():
    from smmp_amd import _ffi
    d = golden("esolan_c2_20mer.npz")
    c = make_ctx(seq=list(d["seq_in"]), itysol=-1)
    o = OracleSystem(c)
    V = _random_confs(c, 48, 7)
    c.sgrnd(11)
    c.profile(True)
    say("smmp_energy_batch c2_20mer 48 conformations")
    out = _ffi.energy_batch([c], np.zeros(len(V), dtype=np.int32), V.ravel(), precision=0)
    st = c.profile_read(3)
    say(f"esolan kernels {st['ms']:.2f} ms, launches {st['launches']}, circle pairs {st['work'][1]:.0f}")
    mt = OracleMT(11)
    for k in range(len(V)):
        ref = o.energy(V[k], mt)
        assert _close(out[k, 6], ref[6]), (k, out[k, 6], ref[6])
        assert _close(out[k, 0], ref[0]), (k, out[k, 0], ref[0])
    assert int(c.view("myblock.mti")[0]) == mt.s.mti


@pytest.mark.parametrize("forms", [1, 2, 3])
def test_general_forms_match_oracle(forms, monkeypatch):
    """The forms the kernel takes for large neighbour sets (the pair loop
    on the circle list instead of a slot mask, bit 0; bucket entries and
    angles in the wave's scratch, bit 1), forced on every atom: the same
    energies as the oracle on sampled-size conformations."""
    from smmp_amd import _ffi
    d = golden("esolan_c2_20mer.npz")
    c = make_ctx(seq=list(d["seq_in"]), itysol=-1)
    o = OracleSystem(c)
    V = _random_confs(c, 16, 9)
    c.sgrnd(13)
    monkeypatch.setenv("SMMP_ESOLAN_FORMS", str(forms))
    say(f"smmp_energy_batch c2_20mer 16 conformations, general forms {forms}")
    out = _ffi.energy_batch([c], np.zeros(len(V), dtype=np.int32), V.ravel(), precision=0)
    mt = OracleMT(13)
    for k in range(len(V)):
        ref = o.energy(V[k], mt)
        assert _close(out[k, 6], ref[6]), (k, out[k, 6], ref[6])
        assert _close(out[k, 0], ref[0]), (k, out[k, 0], ref[0])
    assert int(c.view("myblock.mti")[0]) == mt.s.mti


def test_energy_batch_two_systems():
    from smmp_amd import _ffi
    a = make_ctx(tag="enkefa", itysol=-1)
    d = golden("esolan_len10.npz")
    b = make_ctx(seq=list(d["seq_in"]), itysol=-1)
    Va, Vb = _random_confs(a, 5, 1), _random_confs(b, 7, 2)
    cs = np.array([0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 1], dtype=np.int32)
    flat, ia, ib = [], 0, 0
    for s in cs:
        if s == 0:
            flat.append(Va[ia]); ia += 1
        else:
            flat.append(Vb[ib]); ib += 1
    say("smmp_energy_batch enkefa + len10, 12 conformations")
    out = _ffi.energy_batch([a, b], cs, np.concatenate(flat), precision=0)
    oa, ob = OracleSystem(a), OracleSystem(b)
    for k, s in enumerate(cs):
        ref = (oa if s == 0 else ob).energy(flat[k], OracleMT(1))
        assert _close(out[k, 6], ref[6]), (k, out[k, 6], ref[6])
        assert _close(out[k, 0], ref[0]), (k, out[k, 0], ref[0])


GASC = 1.98773e-3
CDR = 1.745329251994330e-2


def test_metropolis_matches_fortran():
    """smmp_metropolis with itysol -1 (trial loop on the host, every energy
    on the device, rotations on the trials' twister) against the Fortran's
    canon run: sweep energies, acceptances, final variables, twister."""
    from smmp_amd import _ffi
    d = golden("canon_esolan_enkefa.npz")
    c = make_ctx(tag="enkefa", itysol=-1)
    v = c.vlvr()
    v[d["var_idx"] - 1] = d["var_deg"] * CDR
    c.setvar(1, v)  # the reference applies the variable file with a rebuild,
    iat = np.asarray(c.view("var_i.iatvr")[:c.nvt]) - 1
    c.set_vlvr(np.asarray(c.view("atm_r.toat"))[iat])  # then reads the variables back from the torsions
    c.sgrnd(int(d["seed"][0]))
    say("smmp_energy + 8 x smmp_metropolis enkefa itysol -1")
    e0 = c.energy()
    assert _close(e0, d["e0"][0])
    w = _ffi.Weight()
    w.kind, w.beta = 0, 1.0 / (float(d["temp"][0]) * GASC)
    eol, acz = e0, 0.0
    for k, (we, wa) in enumerate(d["sweeps"]):
        eol, acz = c.metropolis(eol, acz, w)
        assert _close(eol, we) and acz == wa, (k, eol, we, acz, wa)
    assert np.abs(c.vlvr()[c.idvr - 1] - d["vlvr"]).max() < 1e-9
    assert int(c.view("myblock.mti")[0]) == d["mti"][0]


def test_partem_s_matches_fortran():
    d = golden("partem_esolan_len10.npz")
    c = make_ctx(seq=list(d["seq_in"]), itysol=-1)
    nswp, R = int(d["nswp"][0]), len(d["temps"])
    c.sgrnd(int(d["seed"][0]))
    say(f"smmp_partem_s len10 itysol -1, {R} replicas, {nswp} sweeps")
    eol, ipoi, traj, obs = c.partem_s(d["temps"], int(d["nequi"][0]), nswp, int(d["switch"][0]), with_obs=True)
    assert np.abs(traj - d["time_d"][:, 3].reshape(nswp, R)).max() <= 6e-4
    assert np.abs(obs[:, :, 0] - d["time_d"][:, 4].reshape(nswp, R)).max() <= 6e-4
    assert np.abs(obs[:, :, 1] - d["time_d"][:, 5].reshape(nswp, R)).max() <= 6e-4
    assert np.abs(eol - d["eol_final"]).max() <= 1e-8 * max(1.0, np.abs(d["eol_final"]).max())


def test_mixed_solvent_batch_and_samplers_are_rejected():
    from smmp_amd import _ffi
    a = make_ctx(tag="enkefa", itysol=-1)
    b = make_ctx(tag="enkefa", itysol=1)
    V = np.tile(a.vlvr(), (2, 1))
    with pytest.raises(_ffi.SmmpError, match="analytic solvent"):
        _ffi.energy_batch([a, b], np.array([0, 1], dtype=np.int32), V.ravel())
    with pytest.raises(_ffi.SmmpError, match="analytic solvent"):
        a.enysol()


def _near_extended(c, R, seed):
    rng = np.random.default_rng(seed)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, c.idvr - 1] = np.pi + rng.normal(0, 0.25, size=(R, c.nvr))
    return v0


def test_batched_canonical_replicas_match_oracle_streams():
    """smmp_sample mode 0 with itysol -1 (replicas in lockstep, one device
    pass per trial step): replica r is the oracle's Metropolis with esolan on
    stream seeds[r]."""
    c = _ctx("len10", golden("esolan_len10.npz"))
    o = OracleSystem(c)
    R, nsw = 4, 2
    temps = np.array([250.0, 400.0, 650.0, 1000.0])
    seeds = np.array([11, 22, 33, 44], dtype=np.uint32)
    v0 = _near_extended(c, R, 5)
    say(f"smmp_sample canonical len10 itysol -1 R={R} {nsw} sweeps")
    out = c.sample(temps, seeds, nsw, v0, mode=0, obs=True)
    for r in range(R):
        mt = OracleMT(int(seeds[r]))
        o.reset()
        vl, tr = o.metropolis(mt, v0[r], 1.0 / (temps[r] * GASC), nsw)
        assert out["acc"][r] == tr[-1][1], r
        for k in range(nsw):
            assert _close(out["traj"][k, r], tr[k][0]), (r, k, out["traj"][k, r], tr[k][0])
        assert np.abs(out["vlvr"][r] - vl).max() < 1e-9
    assert out["obs"].shape == (nsw, R, 2) and np.all(out["obs"][:, :, 0] > 0)


def test_batched_parallel_tempering_matches_oracle():
    """smmp_sample mode 1 with itysol -1: per-replica streams for the sweeps,
    the exchange stream seeds[R] for partem_s.f's exchange rule."""
    c = _ctx("len10", golden("esolan_len10.npz"))
    o = OracleSystem(c)
    R, nsw = 3, 3
    temps = np.array([300.0, 550.0, 1000.0])
    seeds = np.array([5, 6, 7, 8], dtype=np.uint32)
    v0 = _near_extended(c, R, 9)
    say(f"smmp_sample parallel tempering len10 itysol -1 R={R} {nsw} sweeps")
    out = c.sample(temps, seeds, nsw, v0, mode=1)
    beta = 1.0 / (temps * GASC)
    mts = [OracleMT(int(s)) for s in seeds[:R]]
    mx = OracleMT(int(seeds[R]))
    cur, eol, acc = v0.copy(), np.zeros(R), np.zeros(R)
    ipoi, slot = list(range(1, R + 1)), list(range(R))
    traj = np.zeros((nsw, R))
    for sw in range(nsw):
        for r in range(R):
            vl, tr = o.metropolis(mts[r], cur[r], beta[slot[r]], 1)
            cur[r], eol[r] = vl, tr[-1][0]
            acc[r] += tr[-1][1]
            traj[sw, slot[r]] = eol[r]
        nu, top = (1, R - 1) if sw % 2 == 0 else (2, R)
        for i in range(nu, top + 1, 2):
            j = 1 if i == R else i + 1
            a, b = ipoi[i - 1], ipoi[j - 1]
            delta = -beta[i - 1] * eol[b - 1] - beta[j - 1] * eol[a - 1] + beta[i - 1] * eol[a - 1] + beta[j - 1] * eol[b - 1]
            if mx.grnd() <= np.exp(delta):
                ipoi[i - 1], ipoi[j - 1] = b, a
                slot[b - 1], slot[a - 1] = i - 1, j - 1
    assert list(out["ipoi"]) == ipoi
    assert np.array_equal(out["acc"], acc)
    assert np.all(np.abs(out["traj"] - traj) <= RTOL * np.maximum(1.0, np.abs(traj)))


def test_several_chains_energy_and_metropolis_match_oracle():
    """Two chains with the analytic solvent: esolan(0) over the atoms of all
    chains (energy.f:79-82), and smmp_metropolis with the chains' rigid moves
    (metropolis.f:94-168) on the context's twister, against the oracle."""
    from conftest import make_multi_ctx
    from smmp_amd import _ffi
    c, g = make_multi_ctx("enk2", itysol=-1)
    o = OracleSystem(c)
    rng = np.random.default_rng(2)
    v = c.vlvr()
    v[c.idvr - 1] = np.pi + rng.normal(0, 0.3, c.nvr)
    c.set_vlvr(v)
    say("smmp_energy two chains itysol -1")
    e = c.energy()
    ref = o.energy(v, OracleMT(1))
    assert _close(e, ref[0]) and _close(_components(c)[5], ref[6])
    c.sgrnd(77)
    w = _ffi.Weight()
    w.kind, w.beta = 0, 1.0 / (400.0 * GASC)
    say("2 x smmp_metropolis two chains itysol -1")
    eol, acz = e, 0.0
    mt = OracleMT(77)
    o.reset()
    vl, tr = o.metropolis(mt, v, w.beta, 2)
    for k in range(2):
        eol, acz = c.metropolis(eol, acz, w)
        assert _close(eol, tr[k][0]) and acz == tr[k][1], (k, eol, tr[k])
    assert np.abs(c.vlvr() - vl).max() < 1e-9
    assert np.abs(np.asarray(c.view("mol_r.gbpr"))[:, :2].T.ravel() - o.gbpr()[:12]).max() < 1e-9
    assert int(c.view("myblock.mti")[0]) == mt.s.mti
