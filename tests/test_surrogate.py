"""The delayed-acceptance surrogate on the CPU: the oracle's network
gradient against finite differences, its BF16 rounding against closed
forms, and the two-stage acceptance rule of the sampler (stage 1 with the
network's energy change, stage 2 with the exact change minus the network's)
against the stationarity it must keep."""
import os
import sys

import numpy as np
import pytest

from conftest import make_ctx

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import surrogate_oracle as so  # noqa: E402


def test_bf16_rounding_closed_forms():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing at 1
    assert so.bf16(np.array([one + ulp / 2]))[0] == one            # tie: to even
    assert so.bf16(np.array([one + 3 * ulp / 2]))[0] == one + 2 * ulp  # tie: to even (up)
    assert so.bf16(np.array([one + ulp / 2 + 2.0 ** -20]))[0] == one + ulp
    assert so.bf16(np.array([-3.0]))[0] == -3.0
    x = np.random.default_rng(0).normal(size=1000)
    r = so.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * 1.0001)


def test_network_gradient_matches_finite_differences():
    c = make_ctx(tag="enkefa", itysol=1)
    R = so.Residues.from_context(c)
    assert list(R.type) == [19, 8, 8, 14, 13]  # tyr gly gly phe met
    rng = np.random.default_rng(0)
    V = np.tile(c.vlvr(), (4, 1)) + rng.normal(0, 1, (4, len(c.vlvr())))
    tgt = np.array([0.3, -1.2, 2.0, 0.1])
    L = R.lay
    assert L.D == 5 * 36 and L.NPARAM == 180 * 256 + 2 * 256 * 256 + 4 * 256 + 16 + 32 * 16
    th = so.init_theta(L, 2).astype(np.float64)
    loss, g, _ = so.loss_grad(R, th, V, tgt, bf=False)
    idx = [L.W1 + 5, L.W1 + 61 * 256 + 3, L.W1 + 179 * 256 + 255, L.B1 + 3, L.W2 + 77, L.B2 + 9, L.W3 + 1000,
           L.B3 + 1, L.W4 + 7, L.B4, L.EMBO + 16 * 8 + 3, L.EMBO + 16 * 13 + 5, L.EMBO + 16 * 0 + 2]
    for i in idx:
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (so.loss_grad(R, tp, V, tgt, bf=False)[0] - so.loss_grad(R, tm, V, tgt, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, g[i], fd)


def test_pair_loss_gradient_matches_finite_differences():
    """The loss on differences of pairs (x, x'), the form the network is
    trained in for the first stage."""
    c = make_ctx(tag="enkefa", itysol=1)
    R = so.Residues.from_context(c)
    rng = np.random.default_rng(1)
    V = np.tile(c.vlvr(), (6, 1)) + rng.normal(0, 1, (6, len(c.vlvr())))
    tgt = np.array([0.5, -0.7, 1.1])
    L = R.lay
    th = so.init_theta(L, 4).astype(np.float64)
    _, g, _ = so.loss_grad(R, th, V, tgt, bf=False, pairs=True)
    for i in (L.W1 + 11, L.B1, L.W2 + 500, L.W3 + 3, L.W4 + 2, L.EMBO + 16 * 19 + 1):
        h = 1e-6
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (so.loss_grad(R, tp, V, tgt, bf=False, pairs=True)[0]
              - so.loss_grad(R, tm, V, tgt, bf=False, pairs=True)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, g[i], fd)
    # the bias of the output cancels in a difference
    assert abs(g[L.B4]) < 1e-12


def test_whole_chain_features_closed_form():
    """Column 36 r + 2 s holds sin, the next cos, of slot s of residue r;
    slots a residue lacks are zero; the embedding columns repeat the type's
    row; one variable moves exactly two columns."""
    c = make_ctx(tag="enkefa", itysol=1)
    R = so.Residues.from_context(c)
    L = R.lay
    th = so.init_theta(L, 7)
    v = c.vlvr()
    X = R.features(v, th, bf=False)[0]
    emb = th[L.EMBO:].reshape(32, 16)
    for r in range(R.nres):
        for s in range(10):
            col = 36 * r + 2 * s
            if s < R.nvar[r]:
                a = v[R.var0[r] + s]
                assert X[col] == np.float32(np.sin(a)) and X[col + 1] == np.float32(np.cos(a))
            else:
                assert X[col] == 0 and X[col + 1] == 0
        assert np.array_equal(X[36 * r + 20:36 * r + 36], emb[R.type[r]].astype(np.float64))
    v2 = v.copy()
    v2[R.var0[2] + 1] += 0.7
    d = np.nonzero(R.features(v2, th, bf=False)[0] != X)[0]
    assert list(d) == [36 * 2 + 2, 36 * 2 + 3]


def test_screen_output_matches_forward():
    """The sampler's screen (integer layer-1 sums, then FP32 layers) gives
    the network's output to within the rounding of the two forms."""
    c = make_ctx(tag="enkefa", itysol=1)
    R = so.Residues.from_context(c)
    th = so.init_theta(R.lay, 3)
    rng = np.random.default_rng(4)
    V = np.tile(c.vlvr(), (5, 1)) + rng.normal(0, 1, (5, len(c.vlvr())))
    y, _ = so.forward(R, th.astype(np.float64), V, bf=True)
    ys = so.screen_output(R, th, V)
    assert np.abs(ys - y).max() <= 2e-3 * (np.abs(y).max() + 1.0), (ys, y)


def _da_kernel(E, S, beta):
    """Transition matrix of one trial: uniform proposal to another state,
    stage 1 with the surrogate S, stage 2 with E - S (the sampler's rule)."""
    n = len(E)
    T = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            d1 = beta * (S[j] - S[i])
            a1 = 1.0 if d1 <= 0 else np.exp(-min(d1, 100.0))
            d2 = beta * (E[j] - E[i]) - beta * (S[j] - S[i])
            a2 = 1.0 if d2 <= 0 else np.exp(-min(d2, 100.0))
            T[i, j] = a1 * a2 / (n - 1)
        T[i, i] = 1.0 - T[i].sum()
    return T


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_two_stage_rule_keeps_the_boltzmann_distribution(seed):
    rng = np.random.default_rng(seed)
    E = rng.normal(0, 2.0, 7)
    beta = 1.3
    pi = np.exp(-beta * (E - E.min()))
    pi /= pi.sum()
    for S in (rng.normal(0, 3.0, 7), E + rng.normal(0, 0.3, 7), np.zeros(7)):
        T = _da_kernel(E, S, beta)
        assert np.abs(pi @ T - pi).max() < 1e-14
    # with a zero network the rule is Metropolis: acceptance min(1, exp(-beta dE))
    T0 = _da_kernel(E, np.zeros(7), beta)
    for i in range(7):
        for j in range(7):
            if i != j:
                assert abs(T0[i, j] * 6 - min(1.0, np.exp(-beta * (E[j] - E[i])))) < 1e-15
    # a wrong sign in stage 2 breaks stationarity (the test can tell)
    S = rng.normal(0, 3.0, 7)
    n = 7
    Tw = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                d1 = beta * (S[j] - S[i])
                d2 = beta * (E[j] - E[i]) + beta * (S[j] - S[i])
                Tw[i, j] = min(1.0, np.exp(-d1)) * min(1.0, np.exp(-d2)) / (n - 1)
        Tw[i, i] = 1.0 - Tw[i].sum()
    assert np.abs(pi @ Tw - pi).max() > 1e-4
