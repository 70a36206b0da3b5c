"""hbond, interhbond, zimmer and outpdb (hbond.f:14-256, zimmer.f:15-93,
outpdb.f:12-185) against the Fortran's own outputs on conformations drawn
near helix and sheet angles (tests/golden/obs_*.npz): the H-bond counts and
type tables, the chain-pair counts, the Zimmerman codes, and the PDB files
byte for byte."""
import os
import tempfile

import numpy as np
import pytest

from conftest import golden, make_ctx, write_multi_seq

TAGS = ("c2_20mer", "len10", "pair")


def _ctx(g):
    chains = [str(c).split("-") for c in g["chains"]]
    d = tempfile.mkdtemp(prefix="smmp_obs_")
    c = make_ctx(itysol=1, seqfile=write_multi_seq(os.path.join(d, "o.seq"), chains))
    gb = c.view("mol_r.gbpr")
    for row in g["gbpr_in"]:
        gb[:, int(row[0]) - 1] = row[1:]
    return c, len(chains), d


def _apply(c, nch, free):
    v = c.vlvr()
    v[c.idvr - 1] = free
    c.set_vlvr(v)
    for m in range(1, nch + 1):
        c.setvar(m, v)


@pytest.mark.parametrize("tag", TAGS)
def test_hbond_interhbond_zimmer_match_reference(tag):
    g = golden(f"obs_{tag}.npz")
    c, nch, _ = _ctx(g)
    nres = int(c.view("mol_i.irsml2")[nch - 1])
    for k in range(len(g["vlvr"])):
        _apply(c, nch, g["vlvr"][k])
        for m in range(nch):
            assert c.hbond(m + 1) == g["mhb"][k, m], (tag, k, m)
            n = int(c.view("h_bond.ntyhb")[0])
            assert n == g["ntyhb"][k, m]
            assert np.array_equal(c.view("h_bond.ixtyhb")[:n], g["hbtypes"][k, m, 0, :n])
            assert np.array_equal(c.view("h_bond.nutyhb")[:n], g["hbtypes"][k, m, 1, :n])
        assert c.interhbond() == g["inter_mhb"][k]
        mm = c.view("h_bond.mmhb")[:nch, :nch]
        assert np.array_equal(mm.ravel(order="F"), g["mmhb"][k])
        assert [ord(ch) for ch in c.zimmer(nres)] == list(g["zimm"][k]), (tag, k)


@pytest.mark.parametrize("tag", TAGS)
def test_outpdb_matches_reference_byte_for_byte(tag):
    g = golden(f"obs_{tag}.npz")
    c, nch, d = _ctx(g)
    for k, key in ((0, "pdb1"), (1, "pdb2")):
        _apply(c, nch, g["vlvr"][k])
        p = os.path.join(d, f"{key}.pdb")
        c.outpdb(0, p)
        assert open(p).read() == str(g[key]), (tag, key)


def test_hbond_counts_a_planted_helix_bond():
    """A planted change fails: a conformation with bonds found, its hydrogen
    moved 1 A further from the acceptor, loses that bond."""
    g = golden("obs_c2_20mer.npz")
    c, nch, _ = _ctx(g)
    k = int(np.nonzero(g["ntyhb"][:, 0])[0][0])
    _apply(c, nch, g["vlvr"][k])
    n0 = c.hbond(1)
    n_types = int(c.view("h_bond.ntyhb")[0])
    ix = int(c.view("h_bond.ixtyhb")[0])
    ih, ia = ix // 2600, ix % 2600
    x, y, z = c.view("atm_r.xat"), c.view("atm_r.yat"), c.view("atm_r.zat")
    off = 1  # xat has lower bound -1: Fortran index i is element i + 1
    d = np.array([x[ih + off] - x[ia + off], y[ih + off] - y[ia + off], z[ih + off] - z[ia + off]])
    d /= np.linalg.norm(d)
    x[ih + off] += d[0]
    y[ih + off] += d[1]
    z[ih + off] += d[2]
    c.hbond(1)
    assert int(c.view("h_bond.ntyhb")[0]) == n_types - 1
    assert n0 >= 0
