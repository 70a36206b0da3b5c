"""The delayed-acceptance surrogate on the GPU: the BF16 MFMA forward and
backward kernels against the CPU oracle (which rounds to BF16 where the
kernels do), Adam, the sampler's in-kernel screen against the oracle's
restatement of it, delayed acceptance with a zero network against the plain
sampler (identical draws and decisions, also at the full C5 size), the
sampled distribution with an untrained network (KS), and what a trained
network gains over none."""
import os
import sys

import numpy as np
import pytest

from conftest import make_ctx

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_oracle as so  # noqa: E402

pytestmark = pytest.mark.gpu
GASC = 1.98773e-3


def say(*a):
    print("[gpu-surrogate]", *a, flush=True)


def _confs(c, n, seed):
    rng = np.random.default_rng(seed)
    v = np.tile(c.vlvr(), (n, 1))
    v[:, c.idvr - 1] = rng.uniform(-np.pi, np.pi, size=(n, c.nvr))
    return v


def _blocks(L):
    return {"W1": (L.W1, L.B1), "b1": (L.B1, L.W2), "W2": (L.W2, L.B2), "b2": (L.B2, L.W3),
            "W3": (L.W3, L.B3), "b3": (L.B3, L.W4), "w4": (L.W4, L.B4), "b4": (L.B4, L.B4 + 1),
            "emb": (L.EMBO, L.NPARAM)}


def test_layout_matches_oracle():
    from smmp_amd.surrogate import layout
    c = make_ctx(tag="c5_36mer", itysol=1)
    L, R = layout(c), so.Residues.from_context(c)
    assert L["nres"] == 36 and L["D"] == 1296 and L["hid"] == 256
    assert L["n"] == R.lay.NPARAM == 464400
    assert (L["w1"], L["b1"], L["w2"], L["w3"], L["w4"], L["b4"], L["emb"]) == \
        (R.lay.W1, R.lay.B1, R.lay.W2, R.lay.W3, R.lay.W4, R.lay.B4, R.lay.EMBO)


def test_forward_and_gradient_kernels_match_oracle():
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = so.Residues.from_context(c)
    V = _confs(c, 96, 4)
    th = so.init_theta(R.lay, 3)
    tgt = np.random.default_rng(5).normal(0, 1, len(V)).astype(np.float32)
    s = Surrogate(c, theta=th)
    say("smmp_mlp_grad c2_20mer 96 conformations")
    loss, pg, g = s.grad(V, tgt)
    g = g.astype(np.float64)
    lo, go, po = so.loss_grad(R, th.astype(np.float64), V, tgt.astype(np.float64), bf=True)
    assert np.abs(pg - po).max() <= 2e-3 * (np.abs(po).max() + 1.0), (pg[:4], po[:4])
    assert abs(loss - lo) <= 5e-3 * lo
    for name, (a, b) in _blocks(R.lay).items():
        ga, gb = g[a:b], go[a:b]
        rel = np.linalg.norm(ga - gb) / (np.linalg.norm(gb) + 1e-12)
        say(f"gradient {name}: relative L2 difference {rel:.2e}")
        assert rel <= 2e-2, (name, rel)


def test_pair_gradient_kernels_match_oracle():
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = so.Residues.from_context(c)
    V = _confs(c, 64, 9)
    th = so.init_theta(R.lay, 5)
    tgt = np.random.default_rng(6).normal(0, 1, 32).astype(np.float32)
    s = Surrogate(c, theta=th)
    say("smmp_mlp_grad c2_20mer 32 pairs")
    loss, pg, g = s.grad(V, tgt, pairs=True)
    lo, go, po = so.loss_grad(R, th.astype(np.float64), V, tgt.astype(np.float64), bf=True, pairs=True)
    assert abs(loss - lo) <= 5e-3 * lo
    for name, (a, b) in _blocks(R.lay).items():
        if name == "b4":
            assert abs(g[a]) < 1e-4
            continue
        rel = np.linalg.norm(g[a:b] - go[a:b]) / (np.linalg.norm(go[a:b]) + 1e-12)
        assert rel <= 2e-2, (name, rel)


def test_fit_step_is_adam():
    """One epoch with the whole set as one batch is one Adam step with the
    full-batch gradient: theta - lr g / (|g| + eps) at the first step."""
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="enkefa", itysol=1)
    V = _confs(c, 64, 8)
    tgt = np.random.default_rng(2).normal(0, 1, 64)
    s = Surrogate(c, seed=1)
    th0 = s.theta().astype(np.float64)
    _, _, g = s.grad(V, tgt.astype(np.float32))
    lr = 1e-2
    s.fit(V, tgt, epochs=1, batch=64, lr=lr, rescale=False)
    th1 = s.theta().astype(np.float64)
    g = g.astype(np.float64)
    on = np.abs(g) > 1e-6
    want = th0[on] - lr * g[on] / (np.abs(g[on]) + 1e-8)
    assert np.abs(th1[on] - want).max() <= 2e-5 * lr + 1e-7
    assert np.all(np.abs(th1 - th0) <= lr * 1.0001)


def test_screen_in_the_sweep_matches_oracle(tmp_path, monkeypatch):
    """Every screened trial of one replica (traced by the sweep kernel):
    the network's change equals the oracle's restatement of the screen,
    scale * (y(x') - y(x)), on the trajectory the trace itself records."""
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = so.Residues.from_context(c)
    th = so.init_theta(R.lay, 11, 0.5)
    s = Surrogate(c, theta=th)
    s.scale = 200.0  # large enough that the network rejects some trials by itself
    v0 = _confs(c, 1, 12)
    v0[0, c.idvr - 1] = np.pi + np.random.default_rng(13).normal(0, 0.3, c.nvr)  # near-extended
    path = tmp_path / "trace.txt"
    monkeypatch.setenv("SMMP_TRACE_REP", "0")
    monkeypatch.setenv("SMMP_TRACE_FILE", str(path))
    say("smmp_sample c2_20mer one replica, 2 sweeps, traced screen")
    s.sample(np.array([600.0]), np.array([21], dtype=np.uint32), 2, v0, mode=0, groups=1)
    tr = np.loadtxt(path, ndmin=2)
    assert len(tr) == 2 * c.nvr
    x = v0[0].copy()
    states = []
    for jv, vnew, ds, st in tr:
        x1 = x.copy()
        x1[int(jv)] = vnew
        states += [x, x1]
        if st == 2:
            x = x1
    y = so.screen_output(R, th, np.array(states))
    want = s.scale * (y[1::2] - y[0::2])
    err = np.abs(tr[:, 2] - want)
    say(f"screen vs oracle: max |d| {err.max():.3g}, outcomes {np.bincount(tr[:, 3].astype(int), minlength=3)}")
    assert err.max() <= 2e-3 * (np.abs(want).max() + 1.0)
    assert (tr[:, 3] == 0).any() and (tr[:, 3] == 2).any()


def test_zero_network_is_plain_metropolis():
    """Stage 1 passes without a draw when the network's change is zero, and
    stage 2 is then the plain rule: the runs agree bit for bit."""
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 8
    temps = np.geomspace(250.0, 1000.0, R)
    seeds = np.arange(3, 3 + R + 1, dtype=np.uint32)
    v0 = _confs(c, R, 6)
    say("smmp_sample c2_20mer R=8 6 sweeps: plain vs delayed acceptance with a zero network")
    a = c.sample(temps, seeds, 6, v0, mode=1, groups=2)
    z = Surrogate.zero(c)
    b = z.sample(temps, seeds, 6, v0, mode=1, groups=2)
    for k in ("acc", "ipoi", "eol", "vlvr", "traj"):
        assert np.array_equal(a[k], b[k]), k
    st = b["da_stats"]
    assert st[:, 1].sum() == 0 and np.array_equal(st[:, 0], st[:, 2])
    assert np.array_equal(st[:, 3], a["acc"])


def test_c5_full_size_zero_network_and_oracle():
    """C5 at its full size: 128 replicas of the 36-residue sequence,
    ECEPP/3 + OONS, parallel tempering 250-1000 K.  Delayed acceptance with
    a zero network equals the plain sampler bit for bit, and the final
    energies of a random-network run equal the FP64 oracle's."""
    from _oracle import OracleSystem
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c5_36mer", itysol=1)
    R, nsw = 128, 3
    temps = np.geomspace(250.0, 1000.0, R)
    seeds = np.arange(40, 40 + R + 1, dtype=np.uint32)
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, c.idvr - 1] = np.pi + np.random.default_rng(14).normal(0, 0.3, (R, c.nvr))
    say(f"C5: {R} replicas x 36-mer, {nsw} sweeps: plain, zero network, random network")
    a = c.sample(temps, seeds, nsw, v0, mode=1)
    b = Surrogate.zero(c).sample(temps, seeds, nsw, v0, mode=1)
    for k in ("acc", "ipoi", "eol", "vlvr", "traj"):
        assert np.array_equal(a[k], b[k]), k
    s = Surrogate(c, seed=15, init_scale=0.5)
    s.scale = 3.0
    d = s.sample(temps, seeds, nsw, v0, mode=1)
    assert d["da_stats"][:, 1].sum() > 0
    o = OracleSystem(c)
    for r in range(0, R, 8):
        ref = o.energy(d["vlvr"][r])
        assert abs(d["eol"][r] - ref[0]) <= 1e-9 * max(1.0, abs(ref[0])), r


def test_untrained_network_keeps_the_distribution_ks():
    """Any network leaves the stationary distribution exact: energies of a
    delayed-acceptance run with a random network against a plain run."""
    from scipy.stats import ks_2samp
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="enkefa", itysol=1)
    T, nsw, burn, every = 600.0, 240, 80, 20
    R = 64
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, c.idvr - 1] = np.pi
    s = Surrogate(c, seed=7, init_scale=2.0)
    s.scale = 3.0
    say("smmp_sample enkefa R=64 240 sweeps: plain and with a random network")
    a = c.sample(np.full(R, T), np.arange(100, 100 + R, dtype=np.uint32), nsw, v0, mode=0)
    b = s.sample(np.full(R, T), np.arange(500, 500 + R, dtype=np.uint32), nsw, v0, mode=0)
    ea = a["traj"][burn - 1::every, :].ravel()
    eb = b["traj"][burn - 1::every, :].ravel()
    p = ks_2samp(ea, eb).pvalue
    st = b["da_stats"]
    say(f"KS p = {p:.3g}; network rejected {st[:, 1].sum():.0f} of {st[:, 0].sum():.0f} trials")
    assert st[:, 1].sum() > 0
    assert p > 1e-3


def test_trained_network_gains_over_none():
    """Online training on the sampler's own moves, then the same
    delayed-acceptance run with the trained and with a zero network: the
    trained one needs fewer exact energies per accepted move."""
    from smmp_amd.surrogate import Surrogate, move_pairs
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 32
    temps = np.geomspace(250.0, 1000.0, R)
    v0 = _confs(c, R, 11)
    say("smmp_sample c2_20mer R=32: 40 warm-up sweeps, 100 sweeps with snapshots")
    w = c.sample(temps, np.arange(1, R + 2, dtype=np.uint32), 40, v0, mode=1, with_traj=False)
    d = c.sample(temps, np.arange(50, 50 + R + 1, dtype=np.uint32), 100, w["vlvr"], mode=1, with_traj=False,
                 snap_every=1)
    V = d["snap_vlvr"].reshape(-1, c.nvt)
    pv, de, _ = move_pairs(c, np.repeat(V, 8, axis=0), seed=1)
    s = Surrogate(c, seed=3)
    say(f"fit on {len(de)} moves")
    hist = s.fit_moves(pv, de, epochs=12, batch=512, lr=2e-3, clip=10.0)
    say(f"loss: first epoch {hist[0]:.3f}, last {hist[-1]:.3f}")
    assert hist[-1] < 0.9 * hist[0]
    seeds = np.arange(200, 200 + R + 1, dtype=np.uint32)
    t = s.sample(temps, seeds, 20, d["vlvr"], mode=1, with_traj=False)
    z = Surrogate.zero(c).sample(temps, seeds, 20, d["vlvr"], mode=1, with_traj=False)
    per_t = t["da_stats"][:, 2].sum() / max(t["da_stats"][:, 3].sum(), 1.0)
    per_z = z["da_stats"][:, 2].sum() / max(z["da_stats"][:, 3].sum(), 1.0)
    say(f"exact energies per accepted move: trained {per_t:.2f}, none {per_z:.2f}; "
        f"network rejected {t['da_stats'][:, 1].sum() / t['da_stats'][:, 0].sum():.2f} of the trials")
    assert per_t < 0.8 * per_z
