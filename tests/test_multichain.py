"""Systems of several chains (energy.f:40-114 with eninteract.f; the rigid
trials of metropolis.f:94-168): host topology, oracle energies and the
oracle's Metropolis run against the Fortran reference's outputs for the same
chains (tests/golden/multi_*.npz, written by make_golden.py from the
reference compiled here).

Cases: two Met-enkephalins with OONS in a box the second chain's position
reaches, and three chains of 5, 3 and 4 residues in vacuum without a box."""
import numpy as np
import pytest

from conftest import make_multi_ctx
from test_topology import INT_ARRAYS, REAL_ARRAYS

MULTI_TAGS = ["enk2", "three"]
GASC = 1.98773e-3


@pytest.fixture(scope="module", params=MULTI_TAGS)
def multi(request):
    c, g = make_multi_ctx(request.param)
    return request.param, c, g


def test_chain_ranges_and_topology_match_reference(multi):
    tag, c, g = multi
    nml = len(g["chains"])
    assert int(c.view("mol_i.ntlml")[0]) == nml
    for name in ("irsml1", "irsml2", "ivrml1", "nvrml"):
        assert np.array_equal(c.view(f"mol_i.{name}")[:nml], g[name]), (tag, name)
    for name, (view, n) in INT_ARRAYS.items():
        k = int(g[n][0])
        assert np.array_equal(c.view(view)[:k], g[name]), (tag, name)
    for name, (view, n) in REAL_ARRAYS.items():
        k = int(g[n][0])
        np.testing.assert_allclose(c.view(view)[:k], g[name], rtol=0, atol=1e-12, err_msg=f"{tag} {name}")


def _oracle(c):
    from _oracle import OracleSystem
    return OracleSystem(c)


def test_oracle_energies_and_interaction_terms_match_reference(multi):
    """energies[0]: the library conformation at the fixture's chain
    positions; then random conformations (all variables free)."""
    tag, c, g = multi
    o = _oracle(c)
    vl = np.array(g["vlvr_all"], dtype=np.float64)
    idvr = np.asarray(g["idvr"]) - 1
    confs = [vl.copy()]
    for row in g["conf_vlvr"]:
        v = vl.copy()
        v[idvr] = row
        confs.append(v)
    for k, v in enumerate(confs):
        o.reset()
        e = o.energy(v)
        ref = g["energies"][k]
        scale = np.abs(ref).max() + 1.0
        assert np.abs(e[:7] - ref).max() <= 1e-11 * scale, (tag, k, e[:7], ref)
        ei = o.inter()
        np.testing.assert_allclose(ei, g["inter"][k], rtol=1e-11, atol=1e-11 * scale, err_msg=f"{tag} {k}")
        np.testing.assert_allclose(e[7], g["asa_all"][k][0], rtol=1e-11, err_msg=f"{tag} {k} asa")
        if k < len(g["xyz_all"]):
            xyz = o.coords().ravel()
            assert np.abs(xyz - g["xyz_all"][k]).max() < 1e-10, (tag, k)
    # the interaction terms are in the total, not in the components
    e = g["energies"][1]
    i = g["inter"][1]
    assert abs(e[0] - (e[1] + e[6] + i[0])) <= 1e-9 * (abs(e[0]) + 1.0)


def test_oracle_metropolis_with_rigid_moves_matches_reference(multi):
    """Torsion trials of all chains, then per chain three shifts (clamped to
    the box) and three new Euler angles, move for move."""
    from _oracle import OracleMT
    tag, c, g = multi
    o = _oracle(c)
    o.reset()
    o.set_gbpr(g["gbpr_start"])
    mt = OracleMT()
    st = np.asarray(g["mt_state"])
    for k in range(624):
        mt.s.mt[k] = int(st[k]) & 0xFFFFFFFF
    mt.s.mti = int(st[624])
    beta = 1.0 / (float(g["temp_in"][0]) * GASC)
    vl, out = o.metropolis(mt, g["vlvr_start"], beta, len(g["sweeps"]))
    for k, (eol, acz, bad) in enumerate(out):
        assert not bad
        assert acz == g["sweeps"][k, 1], (tag, k, acz, g["sweeps"][k])
        assert abs(eol - g["sweeps"][k, 0]) <= 1e-9 * max(1.0, abs(eol)), (tag, k)
    assert np.abs(vl - g["vlvr_end"]).max() < 1e-11
    assert np.abs(o.gbpr() - g["gbpr_end"]).max() < 1e-11
    assert mt.s.mti == int(g["mti"][0])
    # the box of enk2 was reached: some x positions sit on its edge
    if float(g["boxsize_in"][0]) > 0:
        b = float(g["boxsize_in"][0])
        assert np.all(np.abs(g["gbpr_end"].reshape(-1, 6)[:, :3]) <= b)
