"""MT19937 of twister.f: oracle, product host stream and a library routine
against the reference's own output."""
import numpy as np
import pytest

from conftest import golden

SEEDS = {"seed_4357": 4357, "seed_81236": 81236, "seed_1": 1, "seed_2147483647": 2147483647, "seed_m5": -5}


@pytest.mark.parametrize("name", list(SEEDS))
def test_oracle_stream_matches_reference(name):
    from _oracle import OracleMT
    ref = golden("mt19937.npz")[name]
    got = OracleMT(SEEDS[name]).u32(len(ref))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("name", list(SEEDS))
def test_numpy_mt19937_with_the_69069_seeding(name):
    """numpy's MT19937 with state key k[0]=seed, k[i]=69069 k[i-1] mod 2^32
    (twister.f:85-88) reproduces the reference: pins the twist and temper."""
    ref = golden("mt19937.npz")[name]
    key = np.zeros(624, dtype=np.uint32)
    key[0] = np.uint32(SEEDS[name] & 0xFFFFFFFF)
    for i in range(1, 624):
        key[i] = np.uint32((69069 * int(key[i - 1])) & 0xFFFFFFFF)
    bg = np.random.MT19937()
    bg.state = {"bit_generator": "MT19937", "state": {"key": key, "pos": 624}}
    got = bg.random_raw(len(ref)).astype(np.uint64)
    assert np.array_equal(got, ref)


def test_grnd_scaling_closed_form():
    """u = y / (2^32 - 1), so 0 and 1 are both reachable (twister.f:172-176)."""
    from _oracle import OracleMT
    ref = golden("mt19937.npz")["seed_4357"]
    mt = OracleMT(4357)
    u = np.array([mt.grnd() for _ in range(50)])
    assert np.array_equal(u, ref[:50].astype(np.float64) / 4294967295.0)


def test_default_seed_without_sgrnd():
    from _oracle import OracleMT
    ref = golden("mt19937.npz")["seed_4357"]
    assert np.array_equal(OracleMT(None).u32(700), ref[:700])


def test_product_host_stream_matches_reference():
    """smmp_sgrnd/smmp_grnd (host side of the context's stream)."""
    from smmp_amd import _ffi
    c = _ffi.Context()
    for name in ("seed_81236", "seed_m5"):
        ref = golden("mt19937.npz")[name][:1500]
        c.sgrnd(SEEDS[name])
        u = np.array([c.grnd() for _ in range(len(ref))])
        assert np.array_equal(np.rint(u * 4294967295.0).astype(np.uint64), ref)
    assert int(c.view("myblock.mti")[0]) == 1500 % 624
