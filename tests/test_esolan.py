"""The analytic solvent term (itysol < 0, esolan.f) of the CPU oracle, pinned
against the Fortran reference's own outputs (tests/golden/esolan_*.npz:
start and random conformations of three sequences, and a conformation whose
north-point test draws a random rotation), against closed forms of one and
two spheres, against a brute-force surface quadrature of random clusters,
and against the invariance of the areas under the random rotation."""
import numpy as np
import pytest

from conftest import golden, make_ctx
from _oracle import OracleMT, OracleSystem, Spheres, force_rotation, lib


def _ctx(tag, d):
    if tag == "enkefa":
        return make_ctx(tag="enkefa", itysol=-1)
    return make_ctx(seq=list(d["seq_in"]), itysol=-1)


@pytest.mark.parametrize("tag", ["enkefa", "c2_20mer", "len10"])
def test_esolan_matches_fortran(tag):
    """Start conformation and random ones (vlvr = addang(pi, axvr (u - 1/2))
    from the reference's twister): every energy component bit for bit and
    the twister's position after each energy (no random rotation drawn)."""
    d = golden(f"esolan_{tag}.npz")
    c = _ctx(tag, d)
    o = OracleSystem(c)
    v0 = o.vlvr0.copy()
    e = o.energy(v0)
    assert np.array_equal(o.coords().ravel(), d["xyz0"])
    assert np.array_equal(e[:7], d["energy"][0])
    mt = OracleMT(int(d["seed"][0]))
    for k, want in enumerate(d["conf_vlvr"]):
        v = v0.copy()
        for j in o.idvr:
            v[j - 1] = lib.oracle_addang(np.pi, o.axvr[j - 1] * (mt.grnd() - 0.5))
        assert np.array_equal(v[o.idvr - 1], want)
        e = o.energy(v, mt)
        assert np.array_equal(e[:7], d["energy"][k + 1]), (k, e[6], d["energy"][k + 1][6])
        assert mt.s.mti == d["mti"][k + 1]


def test_esolan_random_rotation_matches_fortran():
    """A conformation where a neighbour's surface passes through an atom's
    north point: the reference draws (u_gamma, u_phi) and rotates; the
    oracle draws the same two numbers and gets the same energy, bit for bit,
    and the same random conformations after it."""
    d = golden("esolan_rotation_enkefa.npz")
    c = make_ctx(tag="enkefa", itysol=-1)
    # the reference applies these variables in its first energy() after setup
    o = OracleSystem(c, initial_torsions=True)
    v = o.vlvr0.copy()
    v[d["var_idx"] - 1] = d["var_rad"]
    mt = OracleMT(int(d["seed"][0]))
    e = o.energy(v, mt)
    assert mt.s.mti == d["mti"][0] == 2
    assert np.array_equal(o.coords().ravel(), d["xyz0"])
    assert np.array_equal(e[:7], d["energy"][0])
    for k, want in enumerate(d["conf_vlvr"]):
        w = v.copy()
        for j in o.idvr:
            w[j - 1] = lib.oracle_addang(np.pi, o.axvr[j - 1] * (mt.grnd() - 0.5))
        assert np.array_equal(w[o.idvr - 1], want)
        e = o.energy(w, mt)
        assert np.array_equal(e[:7], d["energy"][k + 1])
        assert mt.s.mti == d["mti"][k + 1]
    # without a generator the draw cannot be made and the oracle says so
    o2 = OracleSystem(c, initial_torsions=True)
    assert np.isnan(o2.energy(v, None)[6])


def test_closed_forms():
    r1, r2 = 1.7, 1.4
    e, a = Spheres([[0, 0, 0]], [r1]).esolan()
    assert abs(a[0] - 4 * np.pi * r1 * r1) < 1e-12
    for dist in (0.2, 0.5, 1.1, 2.1, 2.9, 3.05):
        u = np.array([0.3, -0.5, 0.8]) / np.sqrt(0.98)
        xyz = np.array([[0.0, 0.0, 0.0], dist * u])
        _, a = Spheres(xyz, [r1, r2]).esolan()
        if dist + r2 <= r1:  # sphere 2 inside sphere 1
            want = [4 * np.pi * r1 * r1, 0.0]
        else:
            h1 = (dist * dist + r1 * r1 - r2 * r2) / (2 * dist)
            h2 = (dist * dist + r2 * r2 - r1 * r1) / (2 * dist)
            want = [2 * np.pi * r1 * (r1 + h1), 2 * np.pi * r2 * (r2 + h2)]
        assert np.allclose(a, want, rtol=1e-12, atol=1e-12), (dist, a, want)
    # apart: the full spheres
    _, a = Spheres([[0, 0, 0], [0, 0, 3.2]], [r1, r2]).esolan()
    assert np.allclose(a, [4 * np.pi * r1 * r1, 4 * np.pi * r2 * r2], rtol=1e-14)
    # the energy is the sigma-weighted sum
    e, a = Spheres([[0, 0, 0], [0.2, 0.1, 2.0]], [r1, r2], sigma=[0.5, -2.0]).esolan()
    assert abs(e - (0.5 * a[0] - 2.0 * a[1])) < 1e-12


def _fib_sphere(n):
    k = np.arange(n) + 0.5
    z = 1.0 - 2.0 * k / n
    r = np.sqrt(1.0 - z * z)
    phi = k * np.pi * (3.0 - np.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_clusters_against_quadrature(seed):
    """Random clusters of overlapping spheres: every atom's analytic area
    against the fraction of a fine point set on its sphere that lies
    outside all other spheres."""
    rng = np.random.default_rng(seed)
    n = 14
    xyz = rng.uniform(-3.0, 3.0, (n, 3))
    r = rng.uniform(1.4, 2.2, n)
    _, a = Spheres(xyz, r).esolan(OracleMT(seed + 1))
    pts = _fib_sphere(60000)
    for i in range(n):
        p = xyz[i] + r[i] * pts
        d2 = ((p[:, None, :] - xyz[None, :, :]) ** 2).sum(-1)
        inside = (d2 < r[None, :] ** 2)
        inside[:, i] = False
        frac = 1.0 - inside.any(1).mean()
        want = 4 * np.pi * r[i] ** 2 * frac
        assert abs(a[i] - want) <= 2.5e-3 * 4 * np.pi * r[i] ** 2, (i, a[i], want)


@pytest.mark.parametrize("u", [(0.13, 0.61), (0.5, 0.02), (0.97, 0.88)])
def test_rotation_leaves_the_areas(u):
    """The rotation of esolan.f:219-232 forced for every atom: the areas of
    a molecule and of a random cluster stay the same to rounding."""
    c = make_ctx(tag="len10", itysol=-1)
    o = OracleSystem(c)
    rng = np.random.default_rng(3)
    v = o.vlvr0.copy()
    v[o.idvr - 1] = rng.uniform(-np.pi, np.pi, o.nvr)
    o.energy(v)
    e0, a0 = o.esolan_areas()
    xyz = rng.uniform(-3.0, 3.0, (12, 3))
    s = Spheres(xyz, rng.uniform(1.4, 2.2, 12))
    f0, b0 = s.esolan()
    force_rotation(*u)
    try:
        e1, a1 = o.esolan_areas()
        f1, b1 = s.esolan()
    finally:
        force_rotation(-1.0, -1.0)
    assert np.allclose(a1, a0, rtol=1e-9, atol=1e-9)
    assert np.allclose(b1, b0, rtol=1e-9, atol=1e-9)
    assert abs(e1 - e0) < 1e-8 * max(1.0, abs(e0))


def test_batched_sampler_limits_with_the_analytic_solvent():
    """The batched sampler's analytic-solvent path has no delayed acceptance
    and no verify mode; it says so before touching the device."""
    from smmp_amd import _ffi
    c = make_ctx(tag="enkefa", itysol=-1)
    temps = np.array([300.0, 400.0])
    with pytest.raises(_ffi.SmmpError, match="analytic solvent"):
        c.sample(temps, np.array([1, 2, 3], dtype=np.uint32), 2, np.tile(c.vlvr(), (2, 1)), verify=True)
    with pytest.raises(_ffi.SmmpError, match="analytic solvent"):
        c.sample(temps, np.array([1, 2, 3], dtype=np.uint32), 2, np.tile(c.vlvr(), (2, 1)), da=True)


GASC = 1.98773e-3
CDR = 1.745329251994330e-2


def canon_start(o, d):
    """The reference's canon start: the variable file applied with a rebuild
    from the set-up torsions, the variables read back from the torsions
    (vlvr(a) = toat(iatvr(a)))."""
    v = o.vlvr0.copy()
    v[d["var_idx"] - 1] = d["var_deg"] * CDR
    o.setvar(v)
    toat = np.ctypeslib.as_array(o.c.toat, (o.nat,))
    return toat[o.keep["iatvr"] - 1].copy()


def test_metropolis_with_esolan_matches_fortran():
    """canon with itysol -1: every trial's energy runs esolan on the trials'
    twister; sweep energies, acceptances, final variables and the twister's
    position bit for bit."""
    d = golden("canon_esolan_enkefa.npz")
    c = make_ctx(tag="enkefa", itysol=-1)
    o = OracleSystem(c, initial_torsions=True)
    v = canon_start(o, d)
    mt = OracleMT(int(d["seed"][0]))
    e0 = o.energy(v, mt)[0]
    assert e0 == d["e0"][0]
    vl, tr = o.metropolis(mt, v, 1.0 / (float(d["temp"][0]) * GASC), len(d["sweeps"]))
    for k, (eol, acz, bad) in enumerate(tr):
        assert eol == d["sweeps"][k][0] and acz == d["sweeps"][k][1] and not bad, (k, eol, d["sweeps"][k])
    assert np.array_equal(vl[o.idvr - 1], d["vlvr"])
    assert mt.s.mti == d["mti"][0]


def test_partem_with_esolan_matches_fortran():
    d = golden("partem_esolan_len10.npz")
    c = make_ctx(seq=list(d["seq_in"]), itysol=-1)
    o = OracleSystem(c)
    nswp, R = int(d["nswp"][0]), len(d["temps"])
    rc, eol, ipoi, coor, traj = o.partem(OracleMT(int(d["seed"][0])), d["temps"], int(d["nequi"][0]), nswp,
                                         int(d["switch"][0]))
    assert rc == 0
    assert np.abs(traj - d["time_d"][:, 3].reshape(nswp, R)).max() <= 5e-4  # time.d prints 3 decimals
    assert np.abs(eol - d["eol_final"]).max() < 1e-9
