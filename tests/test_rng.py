"""Random-number generators: RANECU and its jump-ahead against the closed
form and the reference's own routines; Philox4x32-10 against the published
known-answer vectors; the per-photon streams' distribution."""
import os

import numpy as np
import pytest
from scipy import stats

M1, A1, M2, A2 = 2147483563, 40014, 2147483399, 40692


def ranecu_closed_form(s1, s2, n):
    """L'Ecuyer (1988) combined MLCG from its definition s <- a s mod m."""
    out = []
    for _ in range(n):
        s1 = (A1 * s1) % M1
        s2 = (A2 * s2) % M2
        z = s1 - s2
        if z < 1:
            z += M1 - 1
        out.append(np.float32(z) * np.float32(1.0 / M1))
    return np.array(out, np.float32)


def test_ranecu_matches_closed_form(oracle):
    for seeds in [(12345, 67890), (1, 1), (M1 - 1, M2 - 1), (2**30, 7)]:
        got = oracle.ranecu_stream(seeds, 2000)
        assert np.array_equal(got, ranecu_closed_form(*seeds, 2000))


def test_init_prng_is_modular_jump(oracle):
    # seed_k = seed_input * a_k^((batch+1)*50000*1000) mod m_k
    for batch in (0, 1, 5, 2303999, 99999999):
        for seed in (271828182, 1, 2000000000):
            leap = (batch + 1) * 50000 * 1000
            want = (seed * pow(A1, leap, M1) % M1, seed * pow(A2, leap, M2) % M2)
            assert oracle.init_prng(batch, 50000, seed) == want


def test_rng_matches_reference_fixture(oracle, golden):
    here, man = golden
    d = np.load(os.path.join(here, man["_rng"]["file"]))
    for seed, batch, s0, s1 in d["init"]:
        assert oracle.init_prng(int(batch), 50000, int(seed)) == (s0, s1)
    for (seed, batch, s0, s1), ref_stream in zip(d["init"], d["streams"]):
        assert np.array_equal(oracle.ranecu_stream((int(s0), int(s1)), 1000), ref_stream)
    lib = oracle.lib()
    for m, a, s, want in d["abmodm"]:
        assert lib.fdo_abmodm(int(m), int(a), int(s)) == want
        assert want == (int(a) * int(s)) % int(m)


def test_ranecu_range_is_open_unit_interval():
    # z in [1, 2147483562]: float(z) * float(1/m1) spans [4.66e-10, 0.99999994],
    # so -ln(u) is finite and u < 1 always
    scale = np.float32(1.0 / M1)
    assert np.float32(1) * scale > 0
    assert np.float32(2147483562) * scale == np.nextafter(np.float32(1), np.float32(0))


# Random123 known-answer vectors for Philox4x32-10
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(oracle, ctr, key, want):
    assert oracle.philox(ctr, key) == want


def test_counter_streams_uniform_and_independent(oracle):
    u = np.concatenate([oracle.counter_stream(7, e, k, 500)
                        for e in range(20) for k in range(100)])
    assert u.min() > 0.0 and u.max() <= 1.0
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    # lag-1 serial correlation and cross-photon correlation near zero
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.01
    a = oracle.counter_stream(7, 3, 10, 20000)
    b = oracle.counter_stream(7, 3, 11, 20000)
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.03
    assert not np.array_equal(a[:16], b[:16])
    # 24-bit resolution: values are multiples of 2^-24
    assert np.all((u * 2**24) == np.round(u * 2**24))


def _xoshiro_u(state, n):
    """xoshiro128+ (Blackman & Vigna 2018) from a 4-word state, 24 high
    bits -> (0, 1], in Python integers."""
    s = [int(v) for v in state]
    if not any(s):
        s[0] = 0x9E3779B9
    m = 0xFFFFFFFF
    out = []
    for _ in range(n):
        r = (s[0] + s[3]) & m
        t = (s[1] << 9) & m
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 11) | (s[3] >> 21)) & m
        out.append(np.float32(((r >> 8) + 1) / 16777216.0))
    return np.array(out, np.float32)



def test_xoshiro128_plus_worked_example():
    """Three draws from the state (1, 2, 3, 4), worked by hand from the
    generator's definition: output s0 + s3, then t = s1 << 9, s2 ^= s0,
    s3 ^= s1, s1 ^= s2, s0 ^= s3, s2 ^= t, s3 = rotl(s3, 11).
    Outputs 5, 7 + 12288 = 12295 and 12295 + 25165824 = 25178119, i.e. the
    uniforms (0 + 1), (48 + 1) and (98352 + 1) times 2^-24."""
    got = _xoshiro_u([1, 2, 3, 4], 3)
    want = np.array([1, 49, 98353], np.float64) / 2**24
    assert np.array_equal(got.astype(np.float64), want)

@pytest.mark.parametrize("event", [0, 5, 2**32 - 1, 2**32, 2**32 + 5, 3 * 2**32 + 77, 2**63 + 1])
def test_counter_stream_key_is_64_bit(oracle, event):
    """A photon's stream state is Philox4x32-10(counter (photon, event >>
    32, 0, 0), key (seed, event mod 2^32)): below 2^32 the high word is 0
    (the 32-bit keying), and events 2^32 apart get different streams."""
    seed, photon = 271828182, 9
    want = _xoshiro_u(oracle.philox([photon, event >> 32, 0, 0], [seed, event & 0xFFFFFFFF]), 16)
    got = oracle.counter_stream(seed, event, photon, 16)
    assert np.array_equal(got, want)
    if event >= 2**32:
        low = oracle.counter_stream(seed, event & 0xFFFFFFFF, photon, 16)
        assert not np.array_equal(got, low)
