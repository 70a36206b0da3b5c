"""Derivatives (k_gradient, smmp_gradient / smmp_gradient_batch) and local
minimisation (smmp_minim, smmp_minim_batch) on the GPU, against the oracle's
derivatives (pinned to the Fortran's and to finite differences on the CPU)
and the Fortran's minimisations (tests/golden/minim_vacuum.npz), in vacuum
and with the analytic solvent (k_esolan's gradient form, k_sol_project;
tests/golden/grad_esolan.npz)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden, has_gpu, make_ctx

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not has_gpu(), reason="needs a GPU")]


def wrap(x):
    return np.where(np.abs(x) > np.pi, x - np.copysign(2 * np.pi, x), x)


def random_confs(c, n, seed):
    rng = np.random.default_rng(seed)
    v = np.tile(c.vlvr(), (n, 1))
    idv = c.idvr - 1
    v[:, idv] = rng.uniform(-np.pi, np.pi, (n, len(idv)))
    return v


@pytest.mark.parametrize("tag,epsd", [("enkefa", 0), ("enkefa", 1), ("len10", 0), ("c2_20mer", 0),
                                      ("c5_36mer", 0), ("len60", 0)])
def test_gradient_batch_vs_oracle(tag, epsd):
    """FP64 derivatives of every variable within 1e-10 of the largest (the
    oracle sums forces in the reference's order, the kernel per atom row;
    rsqrt refined twice against the oracle's sqrt and division), energies
    within 1e-12 relative at the start and 1e-10 at random starts (a clash
    at r ~ 0.1 A turns the chain build's 1e-13 A rounding into 1e-12 of
    energies of 1e16 through r^-12; the energy kernel's tests allow the
    same)."""
    from _oracle import OracleSystem
    print(f"[gpu] gradient batch {tag} epsd={epsd}", flush=True)
    c = make_ctx(tag=tag, itysol=0)
    c.view("epar_l.epsd")[0] = epsd
    o = OracleSystem(c)
    V = np.vstack([c.vlvr()[None, :], random_confs(c, 15, 7)])
    if tag == "enkefa":
        V = np.vstack([V, golden("grad_vacuum.npz")["enkefa_vlvr"]])
    e, g = c.gradient_batch(V)
    for k in range(len(V)):
        oe, og = o.gradient(V[k])
        tol = 1e-12 if k == 0 else 1e-10
        assert np.abs(e[k] - oe).max() <= tol * max(1.0, np.abs(oe).max()), (k, e[k], oe)
        assert np.abs(g[k] - og).max() <= 1e-10 * max(1.0, np.abs(og).max()), (k, np.abs(g[k] - og).max())


def test_gradient_against_reference_fixture():
    """The Fortran's own derivatives (gradient.f) of the fixture's
    conformations, directly."""
    print("[gpu] gradient vs the Fortran fixture", flush=True)
    gv = golden("grad_vacuum.npz")
    for tag, ctag, epsd in (("enkefa", "enkefa", 0), ("enkefa_epsd", "enkefa", 1), ("len10", "len10", 0)):
        c = make_ctx(tag=ctag, itysol=0)
        c.view("epar_l.epsd")[0] = epsd
        e, g = c.gradient_batch(gv[f"{tag}_vlvr"])
        for k in range(len(e)):
            ref = gv[f"{tag}_grad"][k]
            assert np.abs(e[k] - gv[f"{tag}_energy"][k]).max() <= 1e-12 * max(1.0, np.abs(gv[f"{tag}_energy"][k]).max())
            assert np.abs(g[k] - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max()), (tag, k)


def test_gradient_into_state_and_batch_agree():
    print("[gpu] smmp_gradient state", flush=True)
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa.var"))
    eysm, gd = c.gradient()
    e, g = c.gradient_batch(c.vlvr()[None, :])
    assert eysm == e[0, 0] and np.array_equal(gd, g[0])
    assert abs(eysm - 1035.4) < 0.05  # README: 'Energy BEFORE minimization' 0.10354E+04
    assert float(c.view("eny_r.eyhb")[0]) == e[0, 3]


def test_minim_reproduces_readme_and_reference():
    """smmp_minim (quasi-Newton) on Met-enkephalin from enkefa.var: the
    README's -12.429 kcal/mol and the Fortran's minimum (its energies and
    variables), from the GPU's derivatives."""
    print("[gpu] minim enkefa quasi-Newton", flush=True)
    m = golden("minim_vacuum.npz")
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa.var"))
    assert np.abs(c.vlvr() - m["enkefa_qn_start"][0]).max() < 1e-12
    n = c.minim(1, 10000, 1e-9)
    ref = m["enkefa_qn_energy"][0]  # eysm (total of energy()), eyel, eyvw, eyhb, eyvr, eysl
    assert abs(float(c.view("eny_r.eysm")[0]) - ref[0]) < 1e-9
    for k, name in enumerate(("eyel", "eyvw", "eyhb", "eyvr"), 1):
        assert abs(float(c.view(f"eny_r.{name}")[0]) - ref[k]) < 1e-7, name
    assert float(f"{float(c.view('eny_r.eysm')[0]):.4E}") == -12.429
    assert np.abs(c.vlvr() - m["enkefa_qn_vlvr"][0]).max() < 1e-7
    assert 80 <= n <= 95  # the Fortran: 87 (the last calls sit at rounding level)


def test_minim_conjugate_gradients():
    print("[gpu] minim enkefa conjugate gradients", flush=True)
    m = golden("minim_vacuum.npz")
    c = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa.var"))
    n = c.minim(2, 60, 1e-9)
    assert n == 60
    assert np.abs(c.vlvr() - m["enkefa_cg60_vlvr"][0]).max() < 1e-8
    c2 = make_ctx(tag="enkefa", itysol=0, varfile=os.path.join(GOLDEN, "enkefa.var"))
    c2.minim(2, 10000, 1e-9)
    assert abs(float(c2.view("eny_r.eysm")[0]) - m["enkefa_cg_energy"][0][0]) < 1e-7


def test_minim_batch_equals_single_minimisations():
    """Sixteen starts minimised side by side take the same steps as each
    alone: bit-identical variables, energies and call counts (each
    conformation's block computes the same sums whatever else the launch
    holds)."""
    print("[gpu] minim batch vs single", flush=True)
    c = make_ctx(tag="len10", itysol=0)
    V = random_confs(c, 16, 11)
    for imin in (1, 2):
        Vb, fb, nb = c.minim_batch(imin, V, 150, 1e-9)
        for k in (0, 5, 15):
            c1 = make_ctx(tag="len10", itysol=0)
            c1.set_vlvr(V[k])
            n1 = c1.minim(imin, 150, 1e-9)
            assert n1 == nb[k], (imin, k)
            assert np.array_equal(c1.vlvr(), Vb[k]), (imin, k)
        assert np.all(nb <= 150) and np.all(np.isfinite(fb))
        # every result is a point the minimiser evaluated: its energy there
        e, _ = c.gradient_batch(Vb)
        assert np.abs(e[:, 0] - fb).max() <= 1e-9 * max(1.0, np.abs(fb).max())


def test_minim_batch_full_size_36mer():
    """The 36-residue system (C5's sequence): 64 starts, 300 calls each;
    every energy falls and each result's derivative norm is what the
    oracle gives there."""
    from _oracle import OracleSystem
    print("[gpu] minim batch 36-mer", flush=True)
    c = make_ctx(tag="c5_36mer", itysol=0)
    V = random_confs(c, 64, 5)
    e0, _ = c.gradient_batch(V)
    Vb, fb, nb = c.minim_batch(1, V, 300, 1e-9)
    assert np.all(fb <= e0[:, 0])
    assert np.median(fb) < np.median(e0[:, 0]) * 1e-3 or np.median(fb) < 0.0
    o = OracleSystem(c)
    for k in (0, 63):
        oe, og = o.gradient(Vb[k])
        assert abs(oe[0] - fb[k]) <= 1e-10 * max(1.0, abs(fb[k]))


def test_gradient_refuses_solvents():
    """itysol > 0: gradient.f stops ('Set itysol < 0'); itysol < 0: the
    device derivatives of the analytic solvent are switched off."""
    from smmp_amd import _ffi
    for ity in (1, -1):
        c = make_ctx(tag="enkefa", itysol=ity)
        with pytest.raises(_ffi.SmmpError, match="itysol"):
            c.gradient_batch(c.vlvr()[None, :])


# ---- the analytic solvent (itysol -1): opesol.f over esolan.f's gradient
#
# Skipped: the first run of the device form (k_esolan's gradient form)
# faulted the GPU (memory aperture violation in pass 1, enkefa, 14
# conformations); a CPU emulation of the kernel under ASan with the same
# input ran clean, so no cause has been shown and the path is switched off
# (smmp_gradient refuses itysol < 0).  The oracle's derivatives are pinned
# on the CPU (tests/test_esolan_grad.py).
SOLVENT_GRAD_OFF = pytest.mark.skip(reason="device derivatives of the analytic solvent switched off: their first run "
                                           "faulted the GPU and no cause has been shown")


@pytest.mark.parametrize("tag", ["enkefa", "len10", "c2_20mer"])
@SOLVENT_GRAD_OFF
def test_gradient_with_esolan_vs_oracle(tag):
    """Energies (eysm carries eysl, as gradient.f leaves it) and derivatives
    of the start and random conformations against the oracle's (pinned to
    the Fortran's and to finite differences): within 1e-9 of the largest
    derivative (the arc terms use the device's atan/sin/cos, the areas agree
    with the reference to 1e-9 relative)."""
    from _oracle import OracleMT, OracleSystem
    print(f"[gpu] gradient batch with esolan {tag}", flush=True)
    c = make_ctx(tag=tag, itysol=-1)
    o = OracleSystem(c)
    V = np.vstack([c.vlvr()[None, :], random_confs(c, 7, 9)])
    if tag in ("enkefa", "len10"):
        V = np.vstack([V, golden("grad_esolan.npz")[f"{tag}_vlvr"]])
    c.sgrnd(77)  # rotations of degenerate north points: conformations in order from the context's twister
    mt = OracleMT(77)
    e, g = c.gradient_batch(V)
    for k in range(len(V)):
        oe, og = o.gradient(V[k], mt)
        ref = np.array([oe[0] + oe[5], oe[1], oe[2], oe[3], oe[4]])
        assert np.abs(e[k] - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max()), (k, e[k], ref)
        assert np.abs(g[k] - og).max() <= 1e-9 * max(1.0, np.abs(og).max()), (k, np.abs(g[k] - og).max())


@SOLVENT_GRAD_OFF
def test_gradient_with_esolan_against_reference_fixture():
    """The Fortran's own derivatives (gradient.f with opesol.f) of the
    fixture's conformations, directly."""
    print("[gpu] gradient with esolan vs the Fortran fixture", flush=True)
    gz = golden("grad_esolan.npz")
    for tag in ("enkefa", "len10"):
        c = make_ctx(tag=tag, itysol=-1)
        c.sgrnd(int(gz[f"{tag}_seed"][0]))  # the driver's stream: its draws of each conformation, then esolan's
        for k in range(len(gz[f"{tag}_vlvr"])):
            if k > 0:
                for _ in range(c.nvr):
                    c.grnd()
            e, g = c.gradient_batch(gz[f"{tag}_vlvr"][k][None, :])
            e, g = e[0], g[0]
            ref, rg = gz[f"{tag}_energy"][k], gz[f"{tag}_grad"][k]
            assert np.abs(e - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max()), (tag, k)
            assert np.abs(g - rg).max() <= 1e-9 * max(1.0, np.abs(rg).max()), (tag, k)


@SOLVENT_GRAD_OFF
def test_gradient_with_esolan_into_state():
    """smmp_gradient: eysm and eysl in the state, eysl equal to energy()'s."""
    print("[gpu] smmp_gradient with esolan", flush=True)
    c = make_ctx(tag="enkefa", itysol=-1, varfile=os.path.join(GOLDEN, "enkefa.var"))
    eysm, gd = c.gradient()
    eysl = float(c.view("eny_r.eysl")[0])
    e, g = c.gradient_batch(c.vlvr()[None, :])
    assert eysm == e[0, 0] and np.array_equal(gd, g[0])
    assert abs(eysm - golden("grad_esolan.npz")["enkefa_energy"][0][0]) < 1e-9
    etot = c.energy()
    assert abs(float(c.view("eny_r.eysl")[0]) - eysl) <= 1e-9 * max(1.0, abs(eysl))
    assert abs(etot - eysm) <= 1e-9 * max(1.0, abs(etot))


@SOLVENT_GRAD_OFF
def test_minim_with_esolan_vs_reference():
    """smmp_minim with the analytic solvent: the Fortran's quasi-Newton run
    from enkefa.var (40 calls) and its conjugate-gradient run, the same call
    counts and variables to rounding carried by the updates."""
    print("[gpu] minim with esolan enkefa", flush=True)
    gz = golden("grad_esolan.npz")
    for tag, imin in (("enkefa_qn", 1), ("enkefa_cg", 2)):
        c = make_ctx(tag="enkefa", itysol=-1, varfile=os.path.join(GOLDEN, "enkefa.var"))
        assert np.abs(c.vlvr() - gz[f"min_{tag}_start"][0]).max() < 1e-12
        n = c.minim(imin, 40, 1e-9)
        assert n == int(gz[f"min_{tag}_nfun"][0]), tag
        assert np.abs(c.vlvr() - gz[f"min_{tag}_vlvr"][0]).max() < 1e-6, tag
        ref = gz[f"min_{tag}_energy"][0]
        assert abs(float(c.view("eny_r.eysm")[0]) - ref[0]) <= 1e-6 * max(1.0, abs(ref[0])), tag


@SOLVENT_GRAD_OFF
def test_minim_batch_with_esolan_equals_single():
    """Eight starts minimised side by side with the analytic solvent take the
    same steps as each alone."""
    print("[gpu] minim batch with esolan vs single", flush=True)
    c = make_ctx(tag="len10", itysol=-1)
    V = random_confs(c, 8, 12)
    Vb, fb, nb = c.minim_batch(1, V, 60, 1e-9)
    for k in (0, 7):
        c1 = make_ctx(tag="len10", itysol=-1)
        c1.set_vlvr(V[k])
        n1 = c1.minim(1, 60, 1e-9)
        assert n1 == nb[k] and np.array_equal(c1.vlvr(), Vb[k]), k
    e, _ = c.gradient_batch(Vb)
    assert np.abs(e[:, 0] - fb).max() <= 1e-9 * max(1.0, np.abs(fb).max())
