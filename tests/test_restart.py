"""Parallel-tempering restart files (partem_s.f:97-111, 237-255): par_R.in
and conf_NNNN.var as the reference writes them, read and written by the
drop-in layer; checked against the files of a reference run
(tests/golden/partem_restart_len10.npz)."""
import os
import sys

import numpy as np

from conftest import golden, make_ctx
from _oracle import OracleSystem

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "smmp-svn_amd", "pysmmp"))
import smmp  # noqa: E402


def test_par_r_reader_and_writer_match_the_reference(tmp_path):
    d = golden("partem_restart_len10.npz")
    ref = bytes(d["par_r"])
    p = tmp_path / "par_R.in"
    p.write_bytes(ref)
    nstart, eol = smmp.read_par_r(str(p), 4)
    assert nstart == int(d["nswp"][0]) == 25
    assert np.array_equal(eol, d["eol_final"])  # replica k's energy is element k of record 2k
    q = tmp_path / "ours.in"
    smmp.write_par_r(str(q), 25, d["eol_final"])
    assert q.read_bytes() == ref


def test_conf_files_round_trip_and_hold_the_final_conformations(tmp_path):
    """Each conf_NNNN.var of the reference run read back (redvar) gives a
    conformation whose energy is the replica's final energy up to the 3
    decimals of degrees the file keeps, and written again (outvar) gives the
    same lines (angles equal modulo 360 degrees)."""
    d = golden("partem_restart_len10.npz")
    c = make_ctx(tag="len10", itysol=1)
    o = OracleSystem(c)
    for k in range(1, 5):
        assert smmp.conf_file_name(k) == f"conf_{k:04d}.var"
        f = tmp_path / smmp.conf_file_name(k)
        f.write_bytes(bytes(d[f"conf_{k:04d}"]))
        v = c.redvar(str(f))
        e = o.energy(v)[0]
        assert abs(e - d["eol_final"][k - 1]) < 0.05, (k, e, d["eol_final"][k - 1])
        keep = c.vlvr()
        c.set_vlvr(v)
        out = tmp_path / f"again_{k}.var"
        c.outvar(0, str(out))
        c.set_vlvr(keep)
        ours, theirs = out.read_text().splitlines(), f.read_text().splitlines()
        assert len(ours) == len(theirs)
        for x, y in zip(ours, theirs):  # same fields; angles equal modulo 360 (180 = -180)
            fx, fy = x.rsplit(":", 1), y.rsplit(":", 1)
            assert fx[0] == fy[0]
            assert abs((float(fx[1]) - float(fy[1]) + 180.0) % 360.0 - 180.0) < 1e-9, (x, y)
