"""Pins the CPU oracle: against the reference's golden vectors (energies,
coordinates, trajectories), the README known answer, and what the geometry
and the algorithms fix in closed form."""
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, TOPO_TAGS, golden, make_ctx

GASC = 1.98773e-3


@pytest.mark.parametrize("tag", TOPO_TAGS)
def test_energies_and_coordinates_match_reference(tag):
    from _oracle import OracleSystem
    f = golden(f"topo_{tag}.npz")
    c = make_ctx(tag=tag, itysol=1)
    o = OracleSystem(c)
    for k in range(f["conf_vlvr"].shape[0]):
        vl = c.vlvr()
        vl[o.idvr - 1] = f["conf_vlvr"][k]
        e = o.energy(vl)
        ref = f["conf_energy"][k]
        scale = np.abs(ref[2:7]).sum() + 1.0
        assert np.all(np.abs(e[:7] - ref) <= 1e-12 * scale), (tag, k, e[:7], ref)
        assert e[7] == pytest.approx(f["conf_asa"][k][0], rel=1e-13)
        if k < 2:
            assert np.abs(o.coords().ravel() - f["conf_xyz"][k]).max() < 1e-11


def test_readme_known_answer():
    """README:389-393: Met-enkephalin at enkefa.var, ECEPP/3, vacuum."""
    from _oracle import OracleSystem
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa.var"))
    e = OracleSystem(c).energy(c.vlvr())
    assert round(e[0], 1) == 1035.4
    assert round(e[2], 2) == 19.91 and round(e[3], 1) == 114.2 and round(e[4], 1) == 900.8
    assert round(e[5], 4) == 0.4511


def test_canonical_trajectory_matches_reference():
    from _oracle import OracleMT, OracleSystem
    r = golden("enkefa_runs.npz")
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa_c1.var"))
    o = OracleSystem(c)
    mt = OracleMT(81236)
    vl, tr = o.metropolis(mt, o.vlvr0, 1.0 / (300.0 * GASC), 40)
    e = np.array([t[0] for t in tr])
    a = np.array([t[1] for t in tr])
    assert np.array_equal(a, r["canon_sweeps"][:, 1])
    assert np.abs(e - r["canon_sweeps"][:, 0]).max() < 1e-10
    assert np.array_equal(vl[o.idvr - 1], r["canon_vlvr"])
    assert mt.s.mti == int(r["canon_mti"][0])


def test_canonical_trajectory_with_solvent_matches_reference():
    from _oracle import OracleMT, OracleSystem
    r = golden("enkefa_runs.npz")
    c = make_ctx(tag="enkefa", itysol=1, varfile=os.path.join(GOLDEN, "enkefa.var"))
    o = OracleSystem(c)
    vl, tr = o.metropolis(OracleMT(4357), o.vlvr0, 1.0 / (500.0 * GASC), 10)
    assert np.array_equal([t[1] for t in tr], r["canon_sol_sweeps"][:, 1])
    assert np.abs(np.array([t[0] for t in tr]) - r["canon_sol_sweeps"][:, 0]).max() < 1e-10


def test_partem_matches_reference():
    from _oracle import OracleMT, OracleSystem
    p = golden("partem_len10.npz")
    c = make_ctx(tag="len10", itysol=1)
    o = OracleSystem(c)
    rc, eol, ipoi, coor, traj = o.partem(OracleMT(int(p["seed"][0])), p["temps"], int(p["nequi"][0]),
                                         int(p["nswp"][0]))
    assert rc == 0
    assert np.abs(traj - p["time_d"][:, 3].reshape(12, 4)).max() <= 5e-4  # time.d prints 3 decimals
    assert np.abs(eol - p["eol_final"]).max() < 1e-10


@pytest.mark.parametrize("tag", ["len10", "c2_20mer"])
def test_partem_from_extended_chain_matches_reference(tag):
    """partem_s with switch < 0 (every free torsion at 180 degrees)."""
    from _oracle import OracleMT, OracleSystem
    p = golden(f"partem_ext_{tag}.npz")
    c = make_ctx(tag=tag, itysol=1)
    o = OracleSystem(c)
    nswp, R = int(p["nswp"][0]), len(p["temps"])
    rc, eol, ipoi, coor, traj = o.partem(OracleMT(int(p["seed"][0])), p["temps"], int(p["nequi"][0]), nswp, -1)
    assert rc == 0
    assert np.abs(traj - p["time_d"][:, 3].reshape(nswp, R)).max() <= 5e-4
    assert np.abs(eol - p["eol_final"]).max() < 1e-9


def test_chain_reproduces_internal_coordinates():
    """Rebuilt bond lengths, bond angles and torsions equal the inputs."""
    from _oracle import OracleSystem
    c = make_ctx(tag="allres", itysol=0)
    o = OracleSystem(c)
    rng = np.random.default_rng(3)
    vl = c.vlvr()
    vl[o.idvr - 1] = rng.uniform(-np.pi, np.pi, o.nvr)
    o.energy(vl)
    x = o.coords()
    par = c.view("atm_i.iowat")[:o.nat] - 1
    bl = c.view("atm_r.blat")[:o.nat]
    ba = c.view("atm_r.baat")[:o.nat]
    for a in range(3, o.nat):
        p = par[a]
        assert abs(np.linalg.norm(x[a] - x[p]) - bl[a]) < 1e-12
        pp = par[p]
        if pp >= 0 and pp != a and p != 0:
            u, v = x[pp] - x[p], x[a] - x[p]
            ang = math.acos(np.dot(u, v) / np.linalg.norm(u) / np.linalg.norm(v))
            assert abs(ang - ba[a]) < 1e-9


def test_rigid_motion_leaves_energy_unchanged():
    """The global position/orientation (gbpr) moves the whole chain: the
    ECEPP terms are invariant; the solvent term is invariant under
    translation only (its 122 points are fixed in the laboratory frame)."""
    from _oracle import OracleSystem
    c = make_ctx(tag="len10", itysol=1)
    o = OracleSystem(c)
    vl = c.vlvr()
    vl[o.idvr - 1] = np.linspace(-3, 3, o.nvr)
    e0 = o.energy(vl)
    for k, v in enumerate([3.0, -2.0, 7.5]):
        o.t.gbpr[k] = v
    o.reset()
    e1 = o.energy(vl)
    assert abs(e1[0] - e0[0]) <= 1e-10 * max(1.0, abs(e0[0]))
    assert e1[7] == pytest.approx(e0[7], rel=1e-12)
    for k, v in enumerate([0.4, -0.3, 1.1]):
        o.t.gbpr[3 + k] = v
    o.reset()
    e2 = o.energy(vl)
    assert abs(e2[1] - e0[1]) <= 1e-10 * (np.abs(e0[2:6]).sum() + 1.0)


def test_solvent_areas_are_point_multiples():
    """Each atom contributes (4 pi r^2 / 122) x (free points): the total
    accessible area of a two-atom configuration equals the closed form."""
    from _oracle import OracleSystem
    c = make_ctx(tag="len10", itysol=1)
    o = OracleSystem(c)
    vl = c.vlvr()
    vl[o.idvr - 1] = 0.7
    e = o.energy(vl)
    rv = c.view("solvent.rvdw")[:o.nat]
    heavy = rv[rv > 0]
    full = (4 * np.pi * heavy ** 2).sum()
    assert 0 < e[7] < full


def test_difang_addang_closed_form():
    from _oracle import lib
    for a1, a2 in [(0.1, 3.0), (-3.0, 3.1), (3.1, -3.1), (2 * np.pi, 0.5), (-7.0, 2.0)]:
        d = lib.oracle_difang(a1, a2)
        assert -np.pi < d <= np.pi and abs(math.remainder(a2 - a1 - d, 2 * np.pi)) < 1e-12
        s = lib.oracle_addang(a1, a2)
        assert -np.pi < s <= np.pi and abs(math.remainder(a1 + a2 - s, 2 * np.pi)) < 1e-12


def test_weights_closed_form():
    from _oracle import F64P, Weight, lib
    b = np.array([0.5, 0.6, 0.7, 0.8])
    al = np.array([1.0, 2.0, 3.0, 4.0])
    w = Weight(kind=2, beta=0.0, kmin=-1, kmax=2, ebin=10.0, b=b.ctypes.data_as(F64P), alpha=al.ctypes.data_as(F64P))
    # k = int(clamp(x/ebin + sign(1/2, x))) (mulcan_par_mod.f90:408-409)
    for x, k in [(-24.0, -1), (-4.0, 0), (4.0, 0), (6.0, 1), (14.9, 1), (25.0, 2), (1e6, 2), (-1e6, -1)]:
        assert lib.oracle_weight_eval(w, x) == pytest.approx(b[k + 1] * x + al[k + 1])
    w1 = Weight(kind=1, beta=0.3, kmin=-1, kmax=2, ebin=10.0, b=b.ctypes.data_as(F64P), alpha=al.ctypes.data_as(F64P))
    assert lib.oracle_weight_eval(w1, 6.0) == pytest.approx((0.3 + b[2]) * 6.0 + al[2])
    w0 = Weight(kind=0, beta=0.25)
    assert lib.oracle_weight_eval(w0, 8.0) == 2.0


def test_muca_recursion_closed_form():
    """Berg's recursion on a hand-computed histogram (mulcan_par_mod.f90:122-155)."""
    from _oracle import F64P, I32P, lib
    kmin, kmax, ebin, beta = 0, 4, 2.0, 1.0
    h = np.array([4, 8, 2, 6, 3], dtype=np.int32)
    xh, g1, b, al = np.zeros(5), np.zeros(5), np.zeros(5), np.zeros(5)
    lib.oracle_muca_update(kmin, kmax, ebin, beta, h.ctypes.data_as(I32P), xh.ctypes.data_as(F64P),
                           g1.ctypes.data_as(F64P), b.ctypes.data_as(F64P), al.ctypes.data_as(F64P))
    exp_b = np.zeros(5)
    for i in (2, 1, 0):
        g0 = h[i + 1] * h[i] / (h[i] + h[i + 1])
        exp_b[i] = (g0 / g0) * (np.log(np.float32(h[i + 1])) - np.log(np.float32(h[i]))) / ebin
    exp_b[4] = exp_b[3] = exp_b[2]
    exp_al = np.zeros(5)
    for i in (3, 2, 1, 0):
        exp_al[i] = exp_al[i + 1] + (exp_b[i + 1] - exp_b[i]) * (i + 0.5) * ebin
    assert np.allclose(b, exp_b, rtol=1e-6) and np.allclose(al, exp_al, rtol=1e-6, atol=1e-9)
    assert np.array_equal(xh, h.astype(float))


def test_muca_recursion_two_updates_hand_computed():
    """Two successive updates of Berg's recursion (mulcan_par_mod.f90:
    122-155) from b = g1 = 0, with every branch written out bin by bin:
    counts in both neighbours (g1 accumulates across the updates, so the
    second update's step is weighted by g0 / (g0_1 + g0_2)), a neighbour
    with too few counts (b kept), and a sparse bin (xhist < 2) that takes
    max(0, b of the next bin) — once a negative one (-> 0), once a positive
    one computed earlier in the same pass.  Logarithms are REAL (float32)
    as in the reference."""
    from _oracle import F64P, I32P, lib
    kmin, kmax, ebin, beta = -1, 5, 1.5, 0.7
    L = lambda n: float(np.log(np.float32(n)))  # noqa: E731  log(float(n))
    h1 = np.array([0, 9, 4, 6, 2, 7, 3], dtype=np.int32)   # bins -1..5
    h2 = np.array([1, 3, 30, 5, 8, 2, 1], dtype=np.int32)
    xh, g1, b, al = np.zeros(7), np.zeros(7), np.zeros(7), np.zeros(7)
    run = lambda h: lib.oracle_muca_update(  # noqa: E731
        kmin, kmax, ebin, beta, h.copy().ctypes.data_as(I32P), xh.ctypes.data_as(F64P), g1.ctypes.data_as(F64P),
        b.ctypes.data_as(F64P), al.ctypes.data_as(F64P))
    ix = lambda k: k - kmin  # noqa: E731
    # ---- first update (bins 3, 2, 1, 0 from their counts; bin -1 sparse)
    run(h1)
    g1_3, g1_2, g1_1, g1_0 = 2 * 7 / 9, 6 * 2 / 8, 4 * 6 / 10, 9 * 4 / 13
    b3 = (L(7) - L(2)) / ebin
    b2 = (L(2) - L(6)) / ebin
    b1 = (L(6) - L(4)) / ebin
    b0 = (L(4) - L(9)) / ebin          # negative
    bm1 = 0.0                          # max(0, b0 < 0)
    exp1 = np.array([bm1, b0, b1, b2, b3, b3, b3])  # b(kmax) = b(kmax-1) = b(kmax-2)
    assert np.allclose(b, exp1, rtol=1e-6, atol=1e-12)
    assert np.allclose(g1[[ix(0), ix(1), ix(2), ix(3)]], [g1_0, g1_1, g1_2, g1_3], rtol=1e-12)
    a = [0.0] * 7                      # alpha(kmax) = 0, then downwards
    for k in (4, 3, 2, 1, 0, -1):
        a[ix(k)] = a[ix(k + 1)] + (exp1[ix(k + 1)] - exp1[ix(k)]) * (k + 0.5) * ebin
    assert np.allclose(al, a, rtol=1e-6, atol=1e-9)
    assert np.array_equal(xh, h1.astype(float))
    # ---- second update: steps weighted by the accumulated g1
    run(h2)
    s3 = 8 * 2 / 10
    b3 += s3 / (g1_3 + s3) * (L(2) - L(8)) / ebin
    s2 = 5 * 8 / 13
    b2 += s2 / (g1_2 + s2) * (L(8) - L(5)) / ebin
    s1 = 30 * 5 / 35
    b1 += s1 / (g1_1 + s1) * (L(5) - L(30)) / ebin
    s0 = 3 * 30 / 33
    b0 += s0 / (g1_0 + s0) * (L(30) - L(3)) / ebin
    assert b0 > 0
    bm1 = b0                           # sparse bin: max(0, b0 > 0), b0 of this same pass
    exp2 = np.array([bm1, b0, b1, b2, b3, b3, b3])
    assert np.allclose(b, exp2, rtol=1e-6, atol=1e-12)
    assert np.allclose(g1[[ix(0), ix(1), ix(2), ix(3)]], [g1_0 + s0, g1_1 + s1, g1_2 + s2, g1_3 + s3], rtol=1e-12)
    a = [0.0] * 7
    for k in (4, 3, 2, 1, 0, -1):
        a[ix(k)] = a[ix(k + 1)] + (exp2[ix(k + 1)] - exp2[ix(k)]) * (k + 0.5) * ebin
    assert np.allclose(al, a, rtol=1e-6, atol=1e-9)
    assert np.array_equal(xh, (h1 + h2).astype(float))


def test_muca_recursion_product_matches_oracle_two_updates():
    """The product's recursion (smmp_muca_update, host code behind the
    C-ABI) equals the oracle's on the same two updates, bin for bin."""
    from _oracle import F64P, I32P, lib
    from smmp_amd import _ffi
    kmin, kmax, ebin, beta = -1, 5, 1.5, 0.7
    st_o = [np.zeros(7) for _ in range(4)]
    st_p = [np.zeros(7) for _ in range(4)]
    for h in (np.array([0, 9, 4, 6, 2, 7, 3], dtype=np.int32), np.array([1, 3, 30, 5, 8, 2, 1], dtype=np.int32)):
        xh, g1, b, al = st_o
        lib.oracle_muca_update(kmin, kmax, ebin, beta, h.copy().ctypes.data_as(I32P), xh.ctypes.data_as(F64P),
                               g1.ctypes.data_as(F64P), b.ctypes.data_as(F64P), al.ctypes.data_as(F64P))
        xh, g1, b, al = st_p
        hp = h.copy()
        rc = _ffi.lib.smmp_muca_update(kmin, kmax, ebin, beta, _ffi.iptr(hp), _ffi.dptr(xh), _ffi.dptr(g1),
                                       _ffi.dptr(b), _ffi.dptr(al), None)
        assert rc == 0 and not hp.any()  # the sweep histogram is cleared
    for o, p in zip(st_o, st_p):
        assert np.array_equal(o, p)
