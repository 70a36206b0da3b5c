"""Energy kernels on the GPU against the CPU oracle (FP64 and mixed precision).

Tolerances: FP64 mode agrees with the FP64 oracle to 1e-10 relative on every
component (the reference's own bar is 1e-6); the mixed mode (FP64 coordinate
differences, FP32 pair terms, FP64 totals) to 5e-6 of the pair-term
magnitude |eyel|+|eyvw|+|eyhb|.
"""
import os

import numpy as np
import pytest

from conftest import LIBDIR, TOPO_TAGS, GOLDEN, golden, make_ctx

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-energy]", *a, flush=True)


def comp_tol(ref: np.ndarray, rtol: float) -> np.ndarray:
    scale = np.abs(ref[2]) + np.abs(ref[3]) + np.abs(ref[4]) + np.abs(ref[5]) + np.abs(ref[6]) + 1.0
    return rtol * scale


def test_known_answer_enkefa():
    """README:389-393: E(enkefa.var) = 1035.4 = 19.91 + 114.2 + 900.8 + 0.4511."""
    from _oracle import OracleSystem
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa.var"))
    say("smmp_energy enkefa.var")
    e = c.energy()
    v = c.view
    comps = [float(v("eny_r." + k)[0]) for k in ("eyel", "eyvw", "eyhb", "eyvr")]
    assert abs(e - 1035.4) < 0.05
    assert np.allclose(comps, [19.91, 114.2, 900.8, 0.4511], rtol=0, atol=[0.005, 0.05, 0.05, 0.00005])
    o = OracleSystem(c).energy(c.vlvr())
    assert abs(e - o[0]) <= 1e-10 * abs(o[0])
    assert np.allclose(comps, o[2:6], rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("tag", TOPO_TAGS)
def test_batch_fp64_vs_oracle(tag):
    from smmp_amd import _ffi
    from _oracle import OracleSystem
    f = golden(f"topo_{tag}.npz")
    c = make_ctx(tag=tag, itysol=1)
    osys = OracleSystem(c)
    base = c.vlvr()
    confs = []
    for k in range(f["conf_vlvr"].shape[0]):
        vl = base.copy()
        vl[osys.idvr - 1] = f["conf_vlvr"][k]
        confs.append(vl)
    say("energy_batch fp64", tag, "nat", osys.nat, "nconf", len(confs))
    out = _ffi.energy_batch([c], np.zeros(len(confs), np.int32), np.concatenate(confs), 0)
    for k, vl in enumerate(confs):
        ref = osys.energy(vl)
        tol = comp_tol(ref, 1e-10)
        assert np.all(np.abs(out[k, :7] - ref[:7]) <= tol), (tag, k, out[k, :7], ref[:7])
        # solvent-accessible area: sums of multiples of 4 pi r^2 / 122
        assert abs(out[k, 7] - ref[7]) <= 1e-9 * max(1.0, ref[7])
        # and the Fortran reference itself
        fr = f["conf_energy"][k]
        assert np.all(np.abs(out[k, :7] - fr[:7]) <= comp_tol(fr, 1e-9))


def test_batch_many_systems_uneven():
    """C3 shape in miniature: sequences of uneven length in one launch."""
    from smmp_amd import _ffi, workloads
    from _oracle import OracleSystem
    seqs = [workloads.random_sequence(L, 100 + L) for L in (10, 23, 41, 60)]
    ctxs = [make_ctx(seq=s, itysol=1) for s in seqs]
    oracles = [OracleSystem(c) for c in ctxs]
    rng = np.random.default_rng(5)
    conf_sys, flat, refs = [], [], []
    for k in range(12):
        s = int(rng.integers(0, len(ctxs)))
        vl = ctxs[s].vlvr()
        idv = oracles[s].idvr - 1
        vl[idv] = rng.uniform(-np.pi, np.pi, size=len(idv))
        conf_sys.append(s)
        flat.append(vl)
        refs.append(oracles[s].energy(vl))
    say("energy_batch multi-system", [len(s) for s in seqs])
    out = _ffi.energy_batch(ctxs, conf_sys, np.concatenate(flat), 0)
    for k, ref in enumerate(refs):
        assert np.all(np.abs(out[k, :7] - ref[:7]) <= comp_tol(ref, 1e-10)), (k, out[k], ref)


def test_mixed_precision_error_bound():
    from smmp_amd import _ffi
    from _oracle import OracleSystem
    c = make_ctx(tag="c5_36mer", itysol=1)
    osys = OracleSystem(c)
    rng = np.random.default_rng(9)
    confs = []
    for k in range(16):
        vl = c.vlvr()
        vl[osys.idvr - 1] = rng.uniform(-np.pi, np.pi, size=osys.nvr)
        confs.append(vl)
    say("energy_batch mixed c5_36mer")
    mixed = _ffi.energy_batch([c], np.zeros(16, np.int32), np.concatenate(confs), 1)
    for k, vl in enumerate(confs):
        ref = osys.energy(vl)
        scale = abs(ref[2]) + abs(ref[3]) + abs(ref[4])
        assert abs(mixed[k, 0] - ref[0]) <= 5e-6 * scale + 1e-6, (k, mixed[k, 0], ref[0])
        assert abs(mixed[k, 6] - ref[6]) <= 1e-9 * max(1.0, abs(ref[6]))  # solvent stays FP64
        assert mixed[k, 5] == pytest.approx(ref[5], rel=1e-12, abs=1e-12)  # torsions FP64


def test_single_energy_sets_coordinates():
    from _oracle import OracleSystem
    c = make_ctx(tag="len10", itysol=1)
    osys = OracleSystem(c)
    vl = c.vlvr()
    vl[osys.idvr - 1] = np.linspace(-3.0, 3.0, osys.nvr)
    c.view("var_r.vlvr")[:len(vl)] = vl
    say("smmp_energy coordinates len10")
    e = c.energy()
    ref = osys.energy(vl)
    assert abs(e - ref[0]) <= 1e-10 * max(1.0, abs(ref[0]))
    nat = osys.nat
    xyz = np.stack([c.view("atm_r.xat")[2:2 + nat], c.view("atm_r.yat")[2:2 + nat], c.view("atm_r.zat")[2:2 + nat]], 1)
    assert np.abs(xyz - osys.coords()).max() < 1e-11


def _uneven_set(lengths, per, seed):
    from smmp_amd import workloads
    from _oracle import OracleSystem
    ctxs = [make_ctx(seq=workloads.random_sequence(L, seed + L)) for L in lengths]
    rng = np.random.default_rng(seed)
    cs, flat, refs = [], [], []
    for s, c in enumerate(ctxs):
        o = OracleSystem(c)
        for _ in range(per):
            vl = c.vlvr()
            vl[c.idvr - 1] = rng.uniform(-np.pi, np.pi, size=c.nvr)
            cs.append(s)
            flat.append(vl)
            refs.append(o.energy(vl))
    return ctxs, np.asarray(cs, dtype=np.int32), np.concatenate(flat), refs


def test_resident_batch_both_lds_classes_vs_oracle():
    """Resident batch over 10- to 60-residue systems (both LDS classes, two
    streams): FP64 against the oracle, repeated runs identical, the mixed
    mode within its bound, and energy_batch (kept batch) equal to it."""
    from smmp_amd import _ffi
    ctxs, cs, flat, refs = _uneven_set((10, 17, 33, 48, 60), 3, 300)
    say("resident batch", len(cs), "conformations")
    B = _ffi.EnergyBatch(ctxs, cs)
    assert B.nvals == flat.size
    B.load(flat)
    B.run(0)
    o1 = B.read()
    B.run(0)
    o2 = B.read()
    assert np.array_equal(o1, o2)
    for k, ref in enumerate(refs):
        assert np.all(np.abs(o1[k, :7] - ref[:7]) <= comp_tol(ref, 1e-10)), (k, o1[k], ref)
        assert abs(o1[k, 7] - ref[7]) <= 1e-9 * max(1.0, ref[7])
    B.run(1)
    om = B.read()
    for k, ref in enumerate(refs):
        mag = abs(ref[2]) + abs(ref[3]) + abs(ref[4])
        assert abs(om[k, 0] - ref[0]) <= 5e-6 * mag + 1e-6, (k, om[k, 0], ref[0])
    e1 = _ffi.energy_batch(ctxs, cs, flat, 0)
    e2 = _ffi.energy_batch(ctxs, cs, flat, 0)  # the engine's kept batch
    assert np.array_equal(e1, o1) and np.array_equal(e2, o1)
    B.close()


def test_batch_tuner_shapes_keep_the_energies():
    """smmp_batch_tune: every class gets a measured shape among 256, 512,
    1024 threads with the pair bits in LDS or through the caches; the
    energies of the tuned launches agree with the fixed rule's to rounding
    (the shapes only change the order of the sums) and with the oracle."""
    from smmp_amd import _ffi
    ctxs, cs, flat, refs = _uneven_set((10, 17, 33, 48, 60), 3, 301)
    say("batch tuner", len(cs), "conformations")
    B = _ffi.EnergyBatch(ctxs, cs)
    B.load(flat)
    B.run(0)
    fixed = B.read()
    for prec in (0, 1):
        shapes = B.tune(prec)
        assert 1 <= len(shapes) <= 2
        assert all(t in (256, 512, 1024) and b in (0, 1) for t, b in shapes)
    B.run(0)
    tuned = B.read()
    assert np.all(np.abs(tuned[:, :7] - fixed[:, :7]) <= 1e-12 * np.maximum(1.0, np.abs(fixed[:, :7])))
    for k, ref in enumerate(refs):
        assert np.all(np.abs(tuned[k, :7] - ref[:7]) <= comp_tol(ref, 1e-10)), k
    B.run(1)
    om = B.read()
    for k, ref in enumerate(refs):
        mag = abs(ref[2]) + abs(ref[3]) + abs(ref[4])
        assert abs(om[k, 0] - ref[0]) <= 5e-6 * mag + 1e-6
    B.close()


def test_resident_batch_follows_a_topology_change(tmp_path):
    """A context re-initialised after the batch was made: the batch refuses
    to run on the old layout, takes the new one, and matches smmp_energy."""
    from smmp_amd import _ffi, workloads
    ctxs, cs, flat, refs = _uneven_set((12, 20), 2, 700)
    B = _ffi.EnergyBatch(ctxs, cs)
    B.load(flat)
    B.run(0)
    B.read()
    c = ctxs[0]
    seqf = workloads.write_seq_file(str(tmp_path / "n.seq"), workloads.random_sequence(15, 4), "n")
    c.init_energy(LIBDIR)
    c.init_molecule(1, "nh2", "cooh", seqf, " ")
    nml = int(c.view("mol_i.ntlml")[0])
    g = c.view("mol_r.gbpr")
    for m in range(1, nml):  # chains added to the context start at the origin: move them apart
        g[0, m] += 40.0 * m
    say("topology change: chains", nml, "nvt", c.nvt)
    with pytest.raises(_ffi.SmmpError):
        B.run(0)
    with pytest.raises(_ffi.SmmpError):
        B.load(flat)  # sized for the old systems
    rows = []
    rng = np.random.default_rng(9)
    for s in cs:
        v = ctxs[s].vlvr()
        v[ctxs[s].idvr - 1] = rng.uniform(-np.pi, np.pi, size=ctxs[s].nvr)
        rows.append(v)
    B.load(np.concatenate(rows))
    B.run(0)
    out = B.read()
    for k, s in enumerate(cs):
        ctxs[s].set_vlvr(rows[k])
        e = ctxs[s].energy()
        assert abs(out[k, 0] - e) <= 1e-10 * max(1.0, abs(e)), (k, out[k, 0], e)
    B.close()


def test_energy_batch_after_a_topology_change(tmp_path):
    """smmp_energy_batch called again with the same contexts after one of
    them changed its topology: the first call with values sized for the new
    systems succeeds (the kept batch is rebuilt before the size check) and
    matches smmp_energy."""
    from smmp_amd import _ffi, workloads
    ctxs, cs, flat, refs = _uneven_set((12, 20), 2, 701)
    out0 = _ffi.energy_batch(ctxs, cs, flat, 0)
    ref0 = np.array([r[0] for r in refs])
    assert np.all(np.abs(out0[:, 0] - ref0) <= 1e-10 * np.maximum(1.0, np.abs(ref0)))
    c = ctxs[0]
    seqf = workloads.write_seq_file(str(tmp_path / "n.seq"), workloads.random_sequence(15, 4), "n")
    c.init_energy(LIBDIR)
    c.init_molecule(1, "nh2", "cooh", seqf, " ")
    g = c.view("mol_r.gbpr")
    for m in range(1, int(c.view("mol_i.ntlml")[0])):
        g[0, m] += 40.0 * m
    say("energy_batch after a topology change: nvt", c.nvt)
    rows = []
    rng = np.random.default_rng(10)
    for s in cs:
        v = ctxs[s].vlvr()
        v[ctxs[s].idvr - 1] = rng.uniform(-np.pi, np.pi, size=ctxs[s].nvr)
        rows.append(v)
    out = _ffi.energy_batch(ctxs, cs, np.concatenate(rows), 0)
    for k, s in enumerate(cs):
        ctxs[s].set_vlvr(rows[k])
        e = ctxs[s].energy()
        assert abs(out[k, 0] - e) <= 1e-10 * max(1.0, abs(e)), (k, out[k, 0], e)
