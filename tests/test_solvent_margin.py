"""The solvent kernels decide most contact tests and point tests from a
reciprocal-of-radius, dot-product form with rounding margins (FP32 in the
point loop, FP64 for what that leaves undecided), and hand the
rest to the reference's own FP64 expressions (enysol.f: d = (x_i - x_j)/r_i,
contact if |d|^2 <= (1 + r_j/r_i)^2, point p occluded if
(dx+px)^2 + (dy+py)^2 + (dz+pz)^2 <= (r_j/r_i)^2, left to right).  This test
replays the kernels' FP64 operation sequence on the CPU (FMA by exact
rational arithmetic, 1/M perturbed like a hardware reciprocal estimate) on
geometries built to sit on, or within a few ulps of, both boundaries, and
checks that a decision taken by the fast form never contradicts the
reference."""
from fractions import Fraction

import numpy as np

U32 = 3.552713678800501e-15    # 32 * 2^-53
U128 = 1.4210854715202004e-14  # 128 * 2^-53


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def reference(ax, ay, az, rj, trad, px, py, pz):
    dx, dy, dz = ax / trad, ay / trad, az / trad
    dd = dx * dx + dy * dy + dz * dz
    akrad = rj / trad
    dr = 1.0 + akrad
    dr = dr * dr
    contact = dd <= dr
    r2b = akrad * akrad
    sx, sy, sz = dx + px, dy + py, dz + pz
    sdd = sx * sx
    sdd = sdd + sy * sy
    sdd = sdd + sz * sz
    return contact, contact and sdd <= r2b


F24 = 1.430511474609375e-06  # 24 * 2^-24


def fmaf(a, b, c):
    return np.float32(float(Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))))


def fast32(ax, ay, az, rj, trad, px, py, pz, rcp_err):
    """The FP32 point-loop level: occluded / free / None (undecided)."""
    rinv = 1.0 / trad
    dxa, dya, dza = ax * rinv, ay * rinv, az * rinv
    dda = dxa * dxa + dya * dya + dza * dza
    aka = rj * rinv
    r2a = aka * aka
    t = r2a - dda
    im = np.float32((1.0 / (F24 * (r2a + 21.0 * dda + 20.0))) * (1.0 + rcp_err))
    c = np.float32(-t * float(im))
    f = np.float32
    pxf, pyf, pzf = f(px), f(py), f(pz)
    pe = f(f(f(pxf * pxf) + f(pyf * pyf)) + f(pzf * pzf))
    v = fmaf(f(2.0 * dxa), pxf, fmaf(f(2.0 * dya), pyf, fmaf(f(2.0 * dza), pzf, pe)))
    s = fmaf(v, im, c)
    return True if s <= -1.0 else (False if s > 1.0 else None)


def fast(ax, ay, az, rj, trad, px, py, pz, rcp_err):
    """(contact or None, occluded or None); None = left to the reference."""
    rinv = 1.0 / trad
    dxa, dya, dza = ax * rinv, ay * rinv, az * rinv
    dda = dxa * dxa + dya * dya + dza * dza
    aka = rj * rinv
    dra = 1.0 + aka
    dra = dra * dra
    r2a = aka * aka
    mc = U32 * (dda + dra)
    if dda <= dra - mc:
        contact = True
    elif dda > dra + mc:
        contact = False
    else:
        contact = None
    t = r2a - dda
    im = (1.0 / (U128 * (r2a + 3.0 * dda + 9.0))) * (1.0 + rcp_err)
    pe = px * px + py * py + pz * pz
    v = fma(2.0 * dxa, px, fma(2.0 * dya, py, fma(2.0 * dza, pz, pe)))
    s = fma(v, im, -t * im)
    occl = True if s <= -1.0 else (False if s > 1.0 else None)
    return contact, occl


def _unit(v):
    return v / np.linalg.norm(v)


def test_fast_forms_never_contradict_reference():
    rng = np.random.default_rng(2024)
    decided = undecided = 0
    dec32 = [0, 0]  # FP32 level: decided, undecided
    for _ in range(1500):
        trad = rng.uniform(2.4, 3.5)   # r_vdw + 1.4 A water probe
        rj = rng.uniform(2.4, 3.5)
        akrad = rj / trad
        # a neighbour direction and distance, inside contact range
        n = _unit(rng.normal(size=3))
        dist = rng.uniform(0.2, 1.0 + akrad) * trad
        a = n * dist
        d = a / trad
        # a point p on the unit sphere with |d + p| = akrad (the boundary),
        # when that circle exists, else a random point
        dn = np.linalg.norm(d)
        c = (akrad ** 2 - dn ** 2 - 1.0) / (2.0 * dn)  # cos of the angle between p and d
        perp = _unit(np.cross(d, rng.normal(size=3)))
        if abs(c) < 1.0:
            p0 = c * d / dn + np.sqrt(1.0 - c * c) * perp
        else:
            p0 = _unit(rng.normal(size=3))
        for k in range(6):
            # nudge by a few ulps to tens of ulps, and the norm by up to 1e-6 (tes.dat precision)
            p = p0 + rng.normal(size=3) * (k * 4e-16) * (10 ** rng.integers(0, 3))
            if k == 5:
                p = p * (1.0 + rng.uniform(-1e-6, 1e-6))
            ax, ay, az = (float(x) for x in a)
            px, py, pz = (float(x) for x in p)
            ref_c, ref_o = reference(ax, ay, az, rj, trad, px, py, pz)
            fc, fo = fast(ax, ay, az, rj, trad, px, py, pz, rng.uniform(-1e-6, 1e-6))
            if fc is not None:
                assert fc == ref_c
            f32 = fast32(ax, ay, az, rj, trad, px, py, pz, rng.uniform(-1e-6, 1e-6))
            if ref_c and f32 is not None:
                dec32[0] += 1
                assert f32 == ref_o, ("fp32", a, p, rj, trad)
            elif ref_c:
                dec32[1] += 1
            if fo is None:
                undecided += 1
                continue
            decided += 1
            if ref_c:  # occlusion only counts for neighbours in contact
                assert fo == ref_o, (a, p, rj, trad)
        # points at +-f M32 from the boundary: where an undersized FP32 margin would show
        if abs(c) < 1.0:
            dda = float(np.dot(d, d))
            m32 = F24 * (akrad ** 2 + 21.0 * dda + 20.0)
            nrm = _unit(d + p0)
            for f in (-4.0, -2.0, -1.1, -0.9, 0.9, 1.1, 2.0, 4.0):
                p = p0 + nrm * (f * m32 / (2.0 * akrad))
                ax, ay, az = (float(x) for x in a)
                px, py, pz = (float(x) for x in p)
                ref_c, ref_o = reference(ax, ay, az, rj, trad, px, py, pz)
                f32 = fast32(ax, ay, az, rj, trad, px, py, pz, rng.uniform(-1e-6, 1e-6))
                if ref_c and f32 is not None:
                    dec32[0] += 1
                    assert f32 == ref_o, ("fp32 edge", f, a, p, rj, trad)
    # and the contact boundary itself
    for _ in range(1500):
        trad = rng.uniform(2.4, 3.5)
        rj = rng.uniform(2.4, 3.5)
        n = _unit(rng.normal(size=3))
        a = n * (trad + rj) * (1.0 + rng.normal() * 1e-15)
        ax, ay, az = (float(x) for x in a)
        ref_c, _ = reference(ax, ay, az, rj, trad, 1.0, 0.0, 0.0)
        fc, _ = fast(ax, ay, az, rj, trad, 1.0, 0.0, 0.0, 0.0)
        if fc is not None:
            assert fc == ref_c
    assert undecided > 0 and decided > 0  # both outcomes occur near the boundary
    assert dec32[0] > 0 and dec32[1] > 0
