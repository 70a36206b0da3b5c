"""Mixed precision in the sampler: each trial's pair-energy change from FP32
pair terms (new and old geometry) whose differences and sums are FP64; the
chain, the rotations, the solvent decisions and the acceptance test stay
FP64.

A mixed run and an FP64 run from the same state with the same streams take
the same decisions until the first trial whose FP32 rounding crosses the
acceptance threshold; until then every trial's energy change must agree
within the FP32 bound, and the verify mode (every trial recomputed in full)
must hold in mixed mode as in FP64."""
import numpy as np
import pytest

from conftest import make_ctx

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-mixed]", *a, flush=True)


def read_trace(path):
    t = np.loadtxt(path, ndmin=2)
    return t.reshape(-1, 4)


def trial_bound(e_old_components, d):
    """FP32 pair terms: every term is within a few 2^-24 of its magnitude;
    the old pairs' magnitudes are bounded by the whole conformation's, the
    new ones by those and the change d itself (a trial into a clash); the
    solvent and torsion changes are the FP64 ones."""
    el, vw, hb = e_old_components
    return 2e-5 * (abs(el) + abs(vw) + abs(hb)) + 4e-6 * abs(d) + 1e-6


@pytest.mark.parametrize("tag,itysol", [("c5_36mer", 0), ("c2_20mer", 1)])
def test_mixed_trial_changes_match_fp64_until_first_flip(tag, itysol, tmp_path, monkeypatch):
    from _oracle import OracleSystem
    c = make_ctx(tag=tag, itysol=itysol)
    o = OracleSystem(c)
    R, nsw = 8, 3
    v0 = np.tile(c.vlvr(), (R, 1))
    rng = np.random.default_rng(71)
    v0[:, c.idvr - 1] = np.pi + rng.normal(0.0, 0.3, size=(R, c.nvr))   # near-extended starts
    temps = np.geomspace(250.0, 1000.0, R)
    seeds = np.arange(501, 501 + R, dtype=np.uint32)
    # out of the clashes first (FP64): from a clash (|E| >= 1e5) every trial is
    # a full recomputation, whose FP32 terms in mixed mode are as large as E
    say(f"smmp_sample {tag} R={R} 60 sweeps FP64 at 1000 K (start states)")
    v0 = c.sample(np.full(R, 1000.0), seeds + 100, 60, v0, mode=0)["vlvr"]
    monkeypatch.setenv("SMMP_TRACE_REP", "0")
    traces = {}
    for prec in (0, 1):
        path = tmp_path / f"trace{prec}.txt"
        monkeypatch.setenv("SMMP_TRACE_FILE", str(path))
        say(f"smmp_sample {tag} itysol {itysol} R={R} {nsw} sweeps precision {prec}, trace replica 0")
        c.sample(temps, seeds, nsw, v0, mode=0, precision=prec, groups=1)
        traces[prec] = read_trace(path)
    t64, t32 = traces[0], traces[1]
    assert len(t64) == nsw * c.nvr and len(t32) == len(t64)
    comp = o.energy(v0[0])
    assert abs(comp[0]) < 1e5, comp[0]
    same = 0
    for k in range(len(t64)):
        assert t64[k, 0] == t32[k, 0], k                    # same variable drawn
        d64 = t64[k, 1] - t64[k, 2]
        d32 = t32[k, 1] - t32[k, 2]
        if abs(t64[k, 2]) < 1e5:                            # incremental trials (clashes take the full path)
            bound = trial_bound((comp[2], comp[3], comp[4]), d64)
            assert abs(d64 - d32) <= bound, (k, d64, d32, bound)
        if t64[k, 3] != t32[k, 3]:
            say(f"first decision flip at trial {k}: dE fp64 {d64:.9g}, mixed {d32:.9g}")
            break
        same += 1
    say(f"{tag}: {same} of {len(t64)} trials with identical decisions before the first flip")
    assert same >= min(len(t64), 50)


def test_mixed_verify_mode_incremental_equals_full_recomputation():
    """Mixed mode: every trial's incremental change (FP32 pair terms, FP64
    sums) against a full mixed recomputation within the FP32 bound; the
    solvent and torsion parts as tightly as in FP64."""
    c = make_ctx(tag="c2_20mer", itysol=1)
    R = 4
    v0 = np.tile(c.vlvr(), (R, 1))
    rng = np.random.default_rng(4)
    idv = c.idvr - 1
    v0[:2, idv] = rng.uniform(-np.pi, np.pi, size=(2, len(idv)))
    v0[2:, idv] = np.pi + rng.normal(0.0, 0.3, size=(2, len(idv)))
    say("smmp_sample verify c2_20mer R=4 2 sweeps, mixed")
    out = c.sample(np.array([250.0, 400.0, 630.0, 1000.0]), np.arange(31, 35, dtype=np.uint32), 2, v0,
                   mode=0, verify=True, precision=1)
    assert out["rc"] == 0


def test_mixed_parallel_tempering_energies_within_bound():
    """Mixed-precision parallel tempering on the C2 system: the tracked
    energies at the end agree with the FP64 oracle within the mixed bound."""
    from _oracle import OracleSystem
    c = make_ctx(tag="c2_20mer", itysol=1)
    o = OracleSystem(c)
    R = 16
    v0 = np.tile(c.vlvr(), (R, 1))
    v0[:, o.idvr - 1] = np.random.default_rng(8).uniform(-np.pi, np.pi, size=(R, o.nvr))
    say("smmp_sample PT c2_20mer R=16 10 sweeps, mixed")
    out = c.sample(np.geomspace(250.0, 1000.0, R), np.arange(1, R + 2, dtype=np.uint32), 10, v0, mode=1,
                   precision=1)
    for r in range(R):
        ref = o.energy(out["vlvr"][r])
        assert abs(out["eol"][r] - ref[0]) <= 2e-5 * (abs(ref[2]) + abs(ref[3]) + abs(ref[4])) + 1e-3, r
