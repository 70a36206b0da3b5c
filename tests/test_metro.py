"""metropolis.f with the update-type draws (updtch1/updtch2, rndord) and
with a multicanonical weight b(k)x + alpha(k): the oracle against the
Fortran's own runs (tests/golden/metro_enkefa.npz, written by
tests/golden/make_golden.py 'metro' from the reference compiled here).

Met-enkephalin from enkefa.var with the omegas fixed (19 free torsions),
ECEPP/3 in vacuo, T = 300 K."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden, make_ctx

GASC = 1.98773e-3
RUNS = ("upch1", "upch2", "muca", "muca_upch1")


def muca_weight(g):
    """The fixture's weight table as the oracle's descriptor (kind 2 =
    muca_weight2, mulcan_par_mod.f90:398-412)."""
    from _oracle import F64P, Weight
    b = np.ascontiguousarray(g["muca_b"])
    al = np.ascontiguousarray(g["muca_alpha"])
    w = Weight(kind=2, beta=0.0, kmin=int(g["muca_kmin"][0]), kmax=int(g["muca_kmax"][0]),
               ebin=float(g["muca_ebin"][0]), b=b.ctypes.data_as(F64P), alpha=al.ctypes.data_as(F64P))
    w._keep = (b, al)
    return w


@pytest.mark.parametrize("tag", RUNS)
def test_oracle_metropolis_variants_match_reference(tag):
    from _oracle import OracleMT, OracleSystem
    g = golden("metro_enkefa.npz")
    seed, temp, nsw, upch, bgs, rnd = g[f"{tag}_args"]
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa_c1.var"))
    o = OracleSystem(c)
    beta = 1.0 / (temp * GASC)
    o.set_updchois(rndord=int(rnd), upchswitch=int(upch), bgsprob=float(bgs), ubeta=beta)
    w = muca_weight(g) if tag.startswith("muca") else None
    e0 = o.energy(o.vlvr0)[0]
    assert abs(e0 - g[f"{tag}_e0"][0]) < 1e-10
    mt = OracleMT(int(seed))
    vl, tr = o.metropolis(mt, o.vlvr0, beta, int(nsw), weight=w)
    sw = g[f"{tag}_sweeps"]
    assert np.array_equal([t[1] for t in tr], sw[:, 1]), tag            # acceptances, sweep by sweep
    assert np.abs(np.array([t[0] for t in tr]) - sw[:, 0]).max() < 1e-10
    assert np.array_equal(vl[o.idvr - 1], g[f"{tag}_vlvr"])              # final torsions, bit for bit
    assert mt.s.mti == int(g[f"{tag}_mti"][0])                            # twister position
    assert np.array_equal(o.updstats().astype(np.int64), g[f"{tag}_ncalls"])


def test_update_draws_change_the_stream():
    """A planted mistake fails: the run with the draws switched off does not
    reproduce the upchswitch = 1 fixture (the pinning has power)."""
    from _oracle import OracleMT, OracleSystem
    g = golden("metro_enkefa.npz")
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa_c1.var"))
    o = OracleSystem(c)
    o.set_updchois(rndord=0, upchswitch=0)
    mt = OracleMT(int(g["upch1_args"][0]))
    vl, tr = o.metropolis(mt, o.vlvr0, 1.0 / (300.0 * GASC), int(g["upch1_args"][2]))
    assert mt.s.mti != int(g["upch1_mti"][0])
    assert not np.array_equal(vl[o.idvr - 1], g["upch1_vlvr"])
