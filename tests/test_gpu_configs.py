"""The sampler configurations at their full sizes on the GPU.

C4: multicanonical sampling, 128 independent replicas of the random
36-residue sequence, ECEPP/3 in vacuo, a flat-histogram-shaped weight
b(k)x + alpha(k) over 200 bins (muca_weight2, mulcan_par_mod.f90:398-412),
mixed precision (FP32 pair terms of full recomputations, FP64 chain
rebuild and trial changes), sweeps replayed from a captured HIP graph."""
import numpy as np
import pytest

from conftest import make_ctx

pytestmark = pytest.mark.gpu
GASC = 1.98773e-3


def say(*a):
    print("[gpu-configs]", *a, flush=True)


def muca_table(kmin, kmax, ebin, t_low, t_high):
    """Slopes b(k) = 1/(k_B T_k), T_k geometric from t_low to t_high over the
    bins, alpha continuous (mulcan_par_mod.f90:147-150, alpha(kmax) = 0)."""
    k = np.arange(kmin, kmax + 1)
    b = 1.0 / (t_low * (t_high / t_low) ** ((k - kmin) / (kmax - kmin)) * GASC)
    al = np.zeros_like(b)
    for i in range(len(k) - 2, -1, -1):
        al[i] = al[i + 1] + (b[i + 1] - b[i]) * (k[i] + 0.5) * ebin
    return b, al


def mixed_bound(e):
    """|E_mixed - E_fp64| allowed for one full evaluation with FP32 pair
    terms: 2e-5 of the summed magnitudes of the pair components + 1e-3."""
    return 2e-5 * (abs(e[2]) + abs(e[3]) + abs(e[4])) + 1e-3


def test_c4_multicanonical_128x36_mixed_graph():
    from smmp_amd import _ffi
    from _oracle import OracleSystem
    c = make_ctx(tag="c5_36mer", itysol=0)
    o = OracleSystem(c)
    R, nsw, K = 128, 40, 20
    kmin, kmax, ebin = -100, 99, 2.0                       # 200 bins
    b, al = muca_table(kmin, kmax, ebin, 250.0, 1000.0)
    w = _ffi.Weight(kind=2, beta=0.0, kmin=kmin, kmax=kmax, ebin=ebin, b=_ffi.dptr(b), alpha=_ffi.dptr(al))
    v0 = np.tile(c.vlvr(), (R, 1))
    rng = np.random.default_rng(36)
    v0[:, c.idvr - 1] = np.pi + rng.normal(0.0, 0.25, size=(R, c.nvr))   # near-extended starts
    seeds = np.arange(1000, 1000 + R, dtype=np.uint32)
    temps = np.full(R, 1000.0)
    say(f"C4: {R} replicas x 36-mer ({c.nvr} free torsions), {nsw} sweeps, mixed, graph of {K} sweeps")
    g = c.sample(temps, seeds, nsw, v0, mode=2, weight=w, precision=1, graph_sweeps=K)
    info = c.sampler_info()
    assert info["graph_replays"] == nsw // K
    say("C4: the same run with direct launches")
    d = c.sample(temps, seeds, nsw, v0, mode=2, weight=w, precision=1, graph_sweeps=0)
    for k in ("acc", "eol", "vlvr", "traj", "hist"):
        assert np.array_equal(g[k], d[k]), k                 # graph replay = direct launches, bit for bit
    assert g["hist"].sum() == R * nsw                         # one sweep-end bin per replica and sweep
    assert g["acc"].sum() > 0
    say("C4: final energies against the FP64 oracle")
    worst = 0.0
    for r in range(R):
        ref = o.energy(g["vlvr"][r])
        err = abs(g["eol"][r] - ref[0])
        assert err <= mixed_bound(ref), (r, g["eol"][r], ref[0])
        worst = max(worst, err / mixed_bound(ref))
    say(f"C4: worst |E_mixed - E_fp64| / bound = {worst:.3g}")
