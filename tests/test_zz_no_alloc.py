"""smmp_energy_batch_noalloc (the entry point that allocates no device
memory) against smmp_energy_batch: identical energies on two calls in a
row, the second with a smaller batch; sizes beyond the scratch and the
analytic solvent are refused."""
import numpy as np
import pytest

from conftest import make_ctx

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)


def _set(lengths, per, seed, itysol=1):
    from smmp_amd import workloads
    ctxs = [make_ctx(seq=workloads.random_sequence(L, seed + L), itysol=itysol) for L in lengths]
    rng = np.random.default_rng(seed)
    cs, flat = [], []
    for s, c in enumerate(ctxs):
        for _ in range(per):
            vl = c.vlvr()
            vl[c.idvr - 1] = rng.uniform(-np.pi, np.pi, size=c.nvr)
            cs.append(s)
            flat.append(vl)
    return ctxs, np.asarray(cs, dtype=np.int32), np.concatenate(flat)


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_energy_batch_two_calls(precision):
    from smmp_amd import _ffi
    big = _set((12, 33, 60), 4, 900)      # both LDS classes (512- and 1024-thread workgroups)
    small = _set((20,), 3, 901)
    for name, (ctxs, cs, flat) in (("first call: 12 conformations of 3 systems", big),
                                   ("second call: 3 conformations of 1 system", small)):
        say(name, "precision", precision)
        na = _ffi.energy_batch_noalloc(ctxs, cs, flat, precision)
        ref = _ffi.energy_batch(ctxs, cs, flat, precision)
        assert np.array_equal(na, ref), np.abs(na - ref).max()
    say("the first batch again after the smaller one")
    ctxs, cs, flat = big
    assert np.array_equal(_ffi.energy_batch_noalloc(ctxs, cs, flat, precision),
                          _ffi.energy_batch(ctxs, cs, flat, precision))


def test_noalloc_refuses_what_it_cannot_hold():
    from smmp_amd import _ffi
    ctxs, cs, flat = _set((60,), 1, 902)
    n = 14000                               # 14000 x 60-residue conformations > 32 MiB of values
    say("oversized batch", n, "conformations")
    with pytest.raises(_ffi.SmmpError) as e:
        _ffi.energy_batch_noalloc(ctxs, np.zeros(n, dtype=np.int32), np.tile(flat, n), 0)
    assert e.value.code == -1 and "scratch" in str(e.value)
    a, acs, aflat = _set((10,), 2, 903, itysol=-1)
    say("analytic solvent refused")
    with pytest.raises(_ffi.SmmpError) as e:
        _ffi.energy_batch_noalloc(a, acs, aflat, 0)
    assert e.value.code == -1
