"""Systems of several chains on the GPU against the Fortran reference's
outputs (tests/golden/multi_*.npz) and the oracle: energies with the terms
between chains (energy.f:114, eninteract.f), the Metropolis run with the
rigid trials of metropolis.f:94-168 move for move, and the batched sampler
(verify mode, workgroup split, per-replica global parameters)."""
import numpy as np
import pytest

from conftest import make_multi_ctx

pytestmark = pytest.mark.gpu
GASC = 1.98773e-3
MULTI_TAGS = ["enk2", "three"]


def say(*a):
    print("[gpu-multichain]", *a, flush=True)


def _confs(g):
    vl = np.array(g["vlvr_all"], dtype=np.float64)
    idvr = np.asarray(g["idvr"]) - 1
    out = [vl.copy()]
    for row in g["conf_vlvr"]:
        v = vl.copy()
        v[idvr] = row
        out.append(v)
    return out


@pytest.mark.parametrize("tag", MULTI_TAGS)
def test_energy_components_and_interaction_match_reference(tag):
    c, g = make_multi_ctx(tag)
    nvt = len(g["vlvr_all"])
    for k, v in enumerate(_confs(g)):
        c.view("var_r.vlvr")[:nvt] = v
        say(f"smmp_energy {tag} conformation {k}")
        e = c.energy()
        ref = g["energies"][k]
        scale = np.abs(ref).max() + 1.0
        assert abs(e - ref[0]) <= 1e-10 * scale, (tag, k, e, ref[0])
        comps = [c.view(f"eny_r.{n}")[0] for n in ("eysm", "eyel", "eyvw", "eyhb", "eyvr", "eysl")]
        assert np.abs(np.array(comps) - ref[1:7]).max() <= 1e-10 * scale, (tag, k)
        inter = [c.view(f"eny_r.{n}")[0] for n in ("eysmi", "eyeli", "eyvwi", "eyhbi")]
        assert np.abs(np.array(inter) - g["inter"][k]).max() <= 1e-10 * scale, (tag, k)
        if k < len(g["xyz_all"]):
            nat = len(g["xyz_all"][k]) // 3
            xyz = np.stack([c.view("atm_r.xat")[2:nat + 2], c.view("atm_r.yat")[2:nat + 2],
                            c.view("atm_r.zat")[2:nat + 2]], 1).ravel()
            assert np.abs(xyz - g["xyz_all"][k]).max() < 1e-10, (tag, k)


@pytest.mark.parametrize("tag", MULTI_TAGS)
def test_energy_batch_of_several_chains(tag):
    from smmp_amd import _ffi
    c, g = make_multi_ctx(tag)
    confs = _confs(g)
    say(f"smmp_energy_batch {tag} {len(confs)} conformations, FP64 and mixed")
    out = _ffi.energy_batch([c], np.zeros(len(confs), dtype=np.int32), np.concatenate(confs), precision=0)
    for k in range(len(confs)):
        ref = g["energies"][k]
        assert abs(out[k, 0] - ref[0]) <= 1e-10 * (np.abs(ref).max() + 1.0), (tag, k)
    mixed = _ffi.energy_batch([c], np.zeros(len(confs), dtype=np.int32), np.concatenate(confs), precision=1)
    for k in range(len(confs)):
        ref = g["energies"][k]
        if abs(ref[0]) < 1e4:
            assert abs(mixed[k, 0] - ref[0]) <= 2e-5 * (abs(ref[2]) + abs(ref[3]) + abs(ref[4]) +
                                                        np.abs(g["inter"][k]).sum()) + 1e-3, (tag, k)


@pytest.mark.parametrize("tag", MULTI_TAGS)
def test_metropolis_with_rigid_moves_matches_reference(tag):
    """smmp_metropolis from the reference's state after its equilibration:
    the same draws, the same decisions, sweep by sweep."""
    from smmp_amd import _ffi
    c, g = make_multi_ctx(tag)
    nml = len(g["chains"])
    nvt = len(g["vlvr_start"])
    c.view("var_r.vlvr")[:nvt] = g["vlvr_start"]
    c.view("mol_r.gbpr")[:, :nml] = np.asarray(g["gbpr_start"]).reshape(nml, 6).T
    st = np.asarray(g["mt_state"], dtype=np.int64)
    c.view("myblock.mt")[:624] = st[:624].astype(np.int32)
    c.view("myblock.mti")[0] = int(st[624])
    beta = 1.0 / (float(g["temp_in"][0]) * GASC)
    c.view("bet.beta")[0] = beta
    w = _ffi.Weight(kind=0, beta=beta)
    eol, acz = 0.0, 0.0
    say(f"smmp_metropolis {tag} {len(g['sweeps'])} sweeps with rigid moves")
    for k in range(len(g["sweeps"])):
        eol, acz = c.metropolis(eol, acz, w)
        assert acz == g["sweeps"][k, 1], (tag, k, acz, g["sweeps"][k])
        assert abs(eol - g["sweeps"][k, 0]) <= 1e-8 * max(1.0, abs(eol)), (tag, k)
    assert np.abs(c.view("var_r.vlvr")[:nvt] - g["vlvr_end"]).max() < 1e-10
    gb = np.asarray(c.view("mol_r.gbpr")[:, :nml]).T.ravel()
    assert np.abs(gb - g["gbpr_end"]).max() < 1e-10
    assert int(c.view("myblock.mti")[0]) == int(g["mti"][0])


def _replicas(c, g, R, seed):
    nml = len(g["chains"])
    v0 = np.tile(np.asarray(g["vlvr_start"], dtype=np.float64), (R, 1))
    rng = np.random.default_rng(seed)
    v0[:, np.asarray(g["idvr"]) - 1] += rng.normal(0.0, 0.15, size=(R, len(g["idvr"])))
    gb = np.tile(np.asarray(g["gbpr_start"]).reshape(nml, 6), (R, 1, 1))
    gb[:, :, :3] += rng.normal(0.0, 0.3, size=(R, nml, 3))
    return v0, gb


@pytest.mark.parametrize("tag", MULTI_TAGS)
def test_sampler_verify_mode_with_rigid_moves(tag):
    """Every incremental trial, torsion or rigid, against a full recomputation."""
    c, g = make_multi_ctx(tag)
    R = 4
    v0, gb = _replicas(c, g, R, 3)
    say(f"smmp_sample {tag} R=4 verify 3 sweeps")
    out = c.sample(np.geomspace(300.0, 900.0, R), np.arange(11, 11 + R + 1, dtype=np.uint32), 3, v0, mode=1,
                   verify=True, gbpr=gb)
    assert out["rc"] == 0
    assert not np.array_equal(out["gbpr"], gb)  # some rigid moves were taken


@pytest.mark.parametrize("tag", MULTI_TAGS)
def test_sampler_several_chains_against_oracle(tag):
    """Parallel tempering with per-replica global parameters: the final
    energies equal the oracle's at the final variables and positions, and
    G workgroups per replica take the decisions of one."""
    from _oracle import OracleSystem
    c, g = make_multi_ctx(tag)
    o = OracleSystem(c)
    R = 8
    v0, gb = _replicas(c, g, R, 9)
    temps = np.geomspace(250.0, 1000.0, R)
    seeds = np.arange(5, 5 + R + 1, dtype=np.uint32)
    say(f"smmp_sample {tag} R=8 8 sweeps, groups 1 and 4")
    a = c.sample(temps, seeds, 8, v0, mode=1, groups=1, gbpr=gb)
    b = c.sample(temps, seeds, 8, v0, mode=1, groups=4, gbpr=gb)
    assert np.array_equal(a["acc"], b["acc"])
    assert np.array_equal(a["ipoi"], b["ipoi"])
    assert np.abs(a["gbpr"] - b["gbpr"]).max() < 1e-12
    for r in range(R):
        o.reset()
        o.set_gbpr(a["gbpr"][r].ravel())
        ref = o.energy(a["vlvr"][r])[0]
        assert abs(a["eol"][r] - ref) <= 1e-9 * max(1.0, abs(ref)), (tag, r, a["eol"][r], ref)
