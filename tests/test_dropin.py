"""The drop-in Python API (smmp module, Universe, Protein) on the host:
state set-up, views of the COMMON blocks, variable-file output."""
import json
import os
import subprocess
import sys

import numpy as np

from conftest import GOLDEN

HERE = os.path.dirname(os.path.abspath(__file__))


def run(name, args, timeout=300):
    r = subprocess.run([sys.executable, os.path.join(HERE, "dropin_scenarios.py"), name, json.dumps(args)],
                       capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def read_var(path):
    vals = []
    for line in open(path):
        parts = [x.strip() for x in line.split(":")]
        if len(parts) >= 3 and parts[0]:
            vals.append((int(parts[0]), parts[1], float(parts[2].split("&")[0])))
    return vals


def test_universe_and_protein_state(tmp_path):
    out = run("setup_cpu", {"varout": str(tmp_path / "out.var")})
    assert abs(out["T"] - 300.0) < 1e-9
    assert abs(out["beta"] - 1.0 / (300.0 * 1.98773e-3)) < 1e-12
    assert out["len"] == 5
    assert [r.lower() for r in out["residues"]] == ["tyr", "gly", "gly", "phe", "met"]
    assert abs(out["phi1"] - (-86.24)) < 1e-9  # enkefa.var: 1 : phi : -86.24
    assert abs(out["psi2"] - 156.18) < 1e-9    # enkefa.var: 2 : psi : 156.18 (first gly)
    assert out["itysol"] == 1 and out["nvr"] == 24
    assert out["origin"] == [[1.0, 2.0, 3.0], [0.1, 0.2, 0.3]]
    rgy, ee, urgy, uee = out["rgyr"]
    assert abs(rgy - out["rgyr_np"][0]) < 1e-9 and abs(ee - out["rgyr_np"][1]) < 1e-9
    assert urgy == rgy and uee == ee
    assert out["lambda_descriptor"] == "None"      # a plain callable runs the host-driven trial loop
    assert abs(out["beta_from_can_weight"] - out["beta"]) < 1e-12
    # saveVariables writes every value of the .var file it was read from
    ref = {(r, n): v for r, n, v in read_var(os.path.join(GOLDEN, "enkefa.var"))}
    got = {(r, n): v for r, n, v in read_var(str(tmp_path / "out.var"))}
    for k, v in ref.items():
        assert k in got and abs(((got[k] - v + 180.0) % 360.0) - 180.0) < 1e-6, k


def test_muca_update_matches_oracle_recursion():
    """smmp_muca_update (product) = oracle_muca_update (pinned by the closed
    form in test_oracle.py) on random histograms, several updates in a row."""
    from _oracle import F64P, I32P, lib as olib
    from smmp_amd import _ffi
    rng = np.random.default_rng(11)
    kmin, kmax, ebin, beta = -30, 45, 1.5, 1.2
    n = kmax - kmin + 1
    st_p = [np.zeros(n) for _ in range(5)]
    st_o = [np.zeros(n) for _ in range(4)]
    for _ in range(6):
        h = rng.integers(0, 9, size=n).astype(np.int32)
        h[rng.random(n) < 0.3] = 0
        hp, ho = h.copy(), h.copy()
        xh, g1, b, al, ent = st_p
        assert _ffi.lib.smmp_muca_update(kmin, kmax, ebin, beta, _ffi.iptr(hp), _ffi.dptr(xh), _ffi.dptr(g1),
                                         _ffi.dptr(b), _ffi.dptr(al), _ffi.dptr(ent)) == 0
        xo, go, bo, ao = st_o
        olib.oracle_muca_update(kmin, kmax, ebin, beta, ho.ctypes.data_as(I32P), xo.ctypes.data_as(F64P),
                                go.ctypes.data_as(F64P), bo.ctypes.data_as(F64P), ao.ctypes.data_as(F64P))
        assert np.array_equal(xh, xo) and np.array_equal(g1, go) and np.array_equal(b, bo) and np.array_equal(al, ao)
        assert not hp.any()  # the update clears the sweep histogram
        ks = np.arange(kmin, kmax + 1)
        assert np.allclose(ent, (beta + b) * ks * ebin + al)
    b2, a2 = b.copy(), al.copy()
    assert _ffi.lib.smmp_muca_final(kmin, kmax, ebin, beta, _ffi.dptr(b2), _ffi.dptr(a2)) == 0
    assert np.allclose(b2, b + beta) and a2[-1] == 0.0
    for i in range(n - 2, -1, -1):
        assert a2[i] == a2[i + 1] + (b2[i + 1] - b2[i]) * (kmin + i + 0.5) * ebin


def test_strand_direction_zimmer_and_sheet(tmp_path):
    """Protein.strand/directionVector/zimmer and Universe.sheet
    (protein.py:184-221, 265-270; universe.py:66-88) on two extended chains."""
    out = run("sheet_cpu", {"dir": str(tmp_path)})
    per_chain = out["nphi"] // 2
    assert out["strand"] == [per_chain, per_chain]   # every free phi/psi at (-150, 150)
    for d in out["dir"]:
        assert abs(np.linalg.norm(d) - 1.0) < 1e-12
    assert abs(np.dot(out["dir"][0], out["dir"][1]) - 1.0) < 1e-9  # same shape, shifted chain
    # Zimmerman: phi = -150 (<= -110), psi = 150 (> 110) is 'E'; gly 1 has no free phi ('-')
    for z in out["zimmer"]:
        assert len(z) == 7 and set(z[1:]) == {"E"}
    assert out["sheet"] == [[0, 1, 1]]               # parallel strands with mmhb(1, 2) = 3
