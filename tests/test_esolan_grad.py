"""Derivatives with the analytic solvent (itysol -1): gradient.f with
opeshe.f and opesol.f over esolan.f's Cartesian gradient (gradan), on the
CPU.

* The oracle's energies and derivatives against the Fortran's own
  (tests/golden/grad_esolan.npz, written by tests/golden/make_golden.py
  'grad_esolan' from the reference compiled here), with the twister replayed:
  the random conformations and esolan's rotations share one stream, so the
  vlvr sequence pins the draws too.
* gradan against central differences of the oracle's analytic surface on
  bare sphere sets (with and without the random rotation of esolan.f:196-232
  forced on every atom) and the variable derivatives against central
  differences of the oracle's energy: the mathematics, independent of the
  reference.
* The product's minimisers driven by the oracle's energy and derivatives
  reproduce the Fortran's minimisations with the analytic solvent.
"""
import numpy as np
import pytest

from conftest import golden, make_ctx


def wrap(x):
    return np.where(np.abs(x) > np.pi, x - np.copysign(2 * np.pi, x), x)


def replay(tag):
    """(context, oracle, twister) of a fixture run, and its conformations with
    the draws that made them checked against the stream."""
    from _oracle import OracleMT, OracleSystem
    g = golden("grad_esolan.npz")
    c = make_ctx(tag=tag, itysol=-1)
    return g, c, OracleSystem(c), OracleMT(int(g[f"{tag}_seed"][0]))


@pytest.mark.parametrize("tag", ["enkefa", "len10"])
def test_oracle_gradient_with_esolan_matches_reference(tag):
    g, c, o, mt = replay(tag)
    idv = c.idvr - 1
    ax = c.view("var_r.axvr")
    for k in range(len(g[f"{tag}_vlvr"])):
        vl = g[f"{tag}_vlvr"][k]
        if k > 0:  # the driver's draws (addang(pi, axvr (u - 1/2))) from the same stream
            drawn = np.array([np.mod(np.pi + ax[j] * (mt.grnd() - 0.5) + np.pi, 2 * np.pi) - np.pi for j in idv])
            assert np.abs(drawn - vl[idv]).max() < 1e-14, k
        e, gd = o.gradient(vl, mt)
        ge, gg = g[f"{tag}_energy"][k], g[f"{tag}_grad"][k]
        # gradient.f leaves eysm = ECEPP + eysl
        got = np.array([e[0] + e[5], e[1], e[2], e[3], e[4]])
        assert np.abs(got - ge).max() <= 1e-12 * max(1.0, np.abs(ge).max()), (k, got, ge)
        assert np.abs(gd - gg).max() <= 1e-11 * max(1.0, np.abs(gg).max()), (k, np.abs(gd - gg).max())
        assert e[5] != 0.0


def _sphere_fd(S, h=1e-5):
    n = S.n
    fd = np.zeros((n, 3))
    for j in range(n):
        for k in range(3):
            S.move(j, k, h)
            ep, _ = S.esolan()
            S.move(j, k, -2 * h)
            em, _ = S.esolan()
            S.move(j, k, h)
            fd[j, k] = (ep - em) / (2 * h)
    return fd


@pytest.mark.parametrize("rotate", [False, True])
def test_oracle_gradan_is_the_derivative(rotate):
    """gradan of random overlapping sphere sets against central differences
    of the oracle's own surface term (sigma of both signs)."""
    from _oracle import Spheres, force_rotation
    rng = np.random.default_rng(3 if not rotate else 4)
    if rotate:
        force_rotation(0.3, 0.7)
    try:
        worst, done = 0.0, 0
        for _ in range(12):
            n = 12
            S = Spheres(rng.uniform(0.0, 6.0, (n, 3)), rng.uniform(1.5, 3.0, n), rng.uniform(-0.02, 0.02, n))
            e0, gan = S.esolan_grad()
            if not np.isfinite(e0):
                continue
            fd = _sphere_fd(S)
            worst = max(worst, np.abs(fd - gan).max() / max(1e-12, np.abs(fd).max()))
            done += 1
            # a plausible slip fails: the atom's own row left out
            assert np.abs(fd - gan).max() < 0.5 * np.abs(fd).max()
        assert done >= 8
        assert worst < 1e-4, worst
    finally:
        force_rotation(-1.0, -1.0)


def test_oracle_solvent_derivatives_by_the_variables():
    """dE/dv with the analytic solvent against central differences of the
    oracle's energy at Met-enkephalin's enkefa.var conformation."""
    from _oracle import OracleMT
    g, c, o, _ = replay("enkefa")
    vl = g["enkefa_vlvr"][0]
    e, gd = o.gradient(vl, OracleMT(5))
    h = 1e-6
    fd = np.zeros_like(gd)
    for v in range(len(vl)):
        a, b = vl.copy(), vl.copy()
        a[v] += h
        b[v] -= h
        ea, eb = o.gradient(a, OracleMT(5))[0], o.gradient(b, OracleMT(5))[0]
        fd[v] = ((ea[0] + ea[5]) - (eb[0] + eb[5])) / (2 * h)
    assert np.abs(fd - gd).max() <= 1e-7 * np.abs(gd).max(), np.abs(fd - gd).max()
    # the solvent part matters here: the vacuum derivatives alone fail it
    from _oracle import OracleSystem
    _, gv = OracleSystem(make_ctx(tag="enkefa", itysol=0)).gradient(vl)
    assert np.abs(fd - gv).max() > 1e3 * np.abs(fd - gd).max()


@pytest.mark.parametrize("tag", ["enkefa_qn", "enkefa_cg", "len10_qn"])
def test_minimisations_with_esolan_match_reference(tag):
    from smmp_amd import _ffi
    from _oracle import OracleMT, OracleSystem
    gz = golden("grad_esolan.npz")
    c = make_ctx(tag="len10" if tag.startswith("len10") else "enkefa", itysol=-1)
    o = OracleSystem(c)
    imin, maxit, eps, seed = gz[f"min_{tag}_args"]
    mt = OracleMT(int(seed))
    nvt = c.nvt

    def fg(x):
        e, gd = o.gradient(x, mt)
        return e[0] + e[5], gd
    for k in range(len(gz[f"min_{tag}_start"])):
        v0 = gz[f"min_{tag}_start"][k]
        if k > 0:  # the driver draws this start from the stream
            for _ in range(c.nvr):
                mt.grnd()
        assert len(v0) == nvt
        if int(imin) == 1:
            x, f, gr, n = _ffi.minqsn(fg, v0, acur=float(eps), maxfun=int(maxit))
        else:
            x, f, gr, n, _ = _ffi.mincjg(fg, v0, acur=float(eps), maxfun=int(maxit))
        ref = gz[f"min_{tag}_energy"][k]  # energy() at the last evaluated point (total)
        assert n == int(gz[f"min_{tag}_nfun"][k])
        assert np.abs(wrap(x) - gz[f"min_{tag}_vlvr"][k]).max() < 1e-6, k  # libm rounding, carried by the updates
        # the best point's energy against energy() at the last evaluated point
        # of a run cut at maxit calls: near each other, not the same point
        assert abs(f - ref[0]) <= 1e-7 * max(1.0, abs(f)) or int(imin) == 2, (k, f, ref)
