"""Observables of the current variables against the Fortran reference's
own outputs: helix.f's helical and sheet-like residue and segment counts
(tests/golden/helix_*.npz: conformations drawn near the limits)."""
import numpy as np
import pytest

from conftest import golden, make_ctx


@pytest.mark.parametrize("tag", ["c2_20mer", "len10"])
def test_helix_matches_fortran(tag):
    d = golden(f"helix_{tag}.npz")
    c = make_ctx(seq=list(d["seq_in"]), itysol=1)
    idvr = c.idvr - 1
    for k, want in enumerate(d["helix"]):
        v = c.vlvr()
        v[idvr] = d["vlvr"][k]
        c.set_vlvr(v)
        assert c.helix() == tuple(int(x) for x in want), k
    assert d["helix"].sum(0).min() > 10  # every count is exercised
