"""The delayed-acceptance surrogate on the GPU: the BF16 MFMA forward and
backward kernels against the CPU oracle (which rounds to BF16 where the
kernels do), Adam, delayed acceptance with a zero network against the plain
sampler (identical draws and decisions), the sampled distribution with an
untrained network (KS), and what a trained network gains over none."""
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


def _blocks():
    return {"W1": (so.W1, so.B1), "b1": (so.B1, so.W2), "W2": (so.W2, so.B2), "b2": (so.B2, so.W3),
            "W3": (so.W3, so.B3), "b3": (so.B3, so.W4), "w4": (so.W4, so.B4), "b4": (so.B4, so.B4 + 1),
            "emb": (so.EMBO, so.NPARAM)}


def test_forward_and_gradient_kernels_match_oracle():
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = so.Residues.from_context(c)
    V = _confs(c, 96, 4)
    th = so.init_theta(3)
    tgt = np.random.default_rng(5).normal(0, 1, len(V)).astype(np.float32)
    s = Surrogate(c, theta=th)
    say("smmp_mlp_grad c2_20mer 96 conformations")
    loss, pg, g = s.grad(V, tgt)
    g = g.astype(np.float64)
    lo, go, po = so.loss_grad(R, th.astype(np.float64), V, tgt.astype(np.float64), bf=True)
    assert np.abs(pg - po).max() <= 2e-3 * (np.abs(po).max() + 1.0), (pg[:4], po[:4])
    assert abs(loss - lo) <= 5e-3 * lo
    for name, (a, b) in _blocks().items():
        ga, gb = g[a:b], go[a:b]
        rel = np.linalg.norm(ga - gb) / (np.linalg.norm(gb) + 1e-12)
        say(f"gradient {name}: relative L2 difference {rel:.2e}")
        assert rel <= 2e-2, (name, rel)


def test_pair_gradient_kernels_match_oracle():
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = so.Residues.from_context(c)
    V = _confs(c, 64, 9)
    th = so.init_theta(5)
    tgt = np.random.default_rng(6).normal(0, 1, 32).astype(np.float32)
    s = Surrogate(c, theta=th)
    say("smmp_mlp_grad c2_20mer 32 pairs")
    loss, pg, g = s.grad(V, tgt, pairs=True)
    lo, go, po = so.loss_grad(R, th.astype(np.float64), V, tgt.astype(np.float64), bf=True, pairs=True)
    assert abs(loss - lo) <= 5e-3 * lo
    for name, (a, b) in _blocks().items():
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
    """Online training on the sampler's own energies, then the same
    delayed-acceptance run with the trained and with a zero network: the
    trained one needs fewer exact energies per accepted move."""
    from smmp_amd.surrogate import Surrogate
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 32
    temps = np.geomspace(250.0, 1000.0, R)
    v0 = _confs(c, R, 11)
    say("smmp_sample c2_20mer R=32: 40 warm-up sweeps, 100 sweeps with snapshots")
    w = c.sample(temps, np.arange(1, R + 2, dtype=np.uint32), 40, v0, mode=1, with_traj=False)
    d = c.sample(temps, np.arange(50, 50 + R + 1, dtype=np.uint32), 100, w["vlvr"], mode=1, with_traj=False,
                 snap_every=1)
    V = d["snap_vlvr"].reshape(-1, c.nvt)
    E = d["snap_e"].ravel()
    s = Surrogate(c, seed=3)
    say(f"fit on {len(E)} energies")
    hist = s.fit(V, E, epochs=30, batch=128, lr=2e-3, clip=float(np.quantile(E, 0.98)))
    say(f"loss: first epoch {hist[0]:.3f}, last {hist[-1]:.3f}")
    assert hist[-1] < 0.6 * hist[0]
    seeds = np.arange(200, 200 + R + 1, dtype=np.uint32)
    t = s.sample(temps, seeds, 20, d["vlvr"], mode=1, with_traj=False)
    z = Surrogate.zero(c).sample(temps, seeds, 20, d["vlvr"], mode=1, with_traj=False)
    per_t = t["da_stats"][:, 2].sum() / max(t["da_stats"][:, 3].sum(), 1.0)
    per_z = z["da_stats"][:, 2].sum() / max(z["da_stats"][:, 3].sum(), 1.0)
    say(f"exact energies per accepted move: trained {per_t:.2f}, none {per_z:.2f}; "
        f"network rejected {t['da_stats'][:, 1].sum() / t['da_stats'][:, 0].sum():.2f} of the trials")
    assert per_t < 0.8 * per_z
