"""Philox4x32-10 stream used by the oracle and the device library.

Pins: the published Random123 known-answer vectors for philox4x32-10, and the
round structure checked through its 4x64 sibling against numpy.random.Philox
(an independent implementation of the same SC'11 construction)."""
import numpy as np
import pytest

M32 = (0xD2511F53, 0xCD9E8D57)
W32 = (0x9E3779B9, 0xBB67AE85)
M64 = (0xD2E7470EE14C6C93, 0xCA5A826395121157)
W64 = (0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B)


def philox_py(ctr, key, bits):
    """Straight-from-the-paper Philox4xW-10 in Python integers."""
    mask = (1 << bits) - 1
    M, W = (M32, W32) if bits == 32 else (M64, W64)
    c = list(ctr)
    k = list(key)
    for r in range(10):
        if r > 0:
            k = [(k[0] + W[0]) & mask, (k[1] + W[1]) & mask]
        p0 = M[0] * c[0]
        p1 = M[1] * c[2]
        hi0, lo0 = p0 >> bits, p0 & mask
        hi1, lo1 = p1 >> bits, p1 & mask
        c = [hi1 ^ c[1] ^ k[0], lo1, hi0 ^ c[3] ^ k[1], lo0]
    return c


# Random123 kat_vectors, philox4x32 R=10
KATS_4x32 = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KATS_4x32)
def test_kat_python(ctr, key, expect):
    assert tuple(philox_py(ctr, key, 32)) == expect


@pytest.mark.parametrize("ctr,key,expect", KATS_4x32)
def test_kat_oracle(oracle, ctr, key, expect):
    assert tuple(int(x) for x in oracle.philox4x32_10(ctr, key)) == expect


def test_round_structure_vs_numpy_philox4x64():
    # numpy's Philox is 4x64-10; after advance-free construction with an
    # explicit counter/key the first output block is philox(ctr+1, key).
    key = (0x0123456789ABCDEF, 0x0FEDCBA987654321)
    for c0 in (0, 7, 2**40 + 3):
        ctr = [c0, 5, 9, 11]
        bg = np.random.Philox(counter=ctr, key=list(key))
        got = [int(x) for x in bg.random_raw(4)]
        c1 = [(c0 + 1) & (2**64 - 1), 5, 9, 11]
        assert got == philox_py(c1, key, 64)


def test_random_oracle_vs_python(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        ctr = [int(x) for x in rng.integers(0, 2**32, 4)]
        key = [int(x) for x in rng.integers(0, 2**32, 2)]
        assert [int(x) for x in oracle.philox4x32_10(ctr, key)] == philox_py(ctr, key, 32)


def test_stream_layout_and_uniform(oracle):
    seed = 0x1234_5678_9ABC_DEF0
    tag, idx, step = 2, (7 << 32) + 5, 3
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for slot in range(6):
        ctr = [idx & 0xFFFFFFFF, ((idx >> 32) & 0xFFFFFF) | (tag << 24), step, slot // 2]
        w = philox_py(ctr, key, 32)
        a, b = (w[0], w[1]) if slot % 2 == 0 else (w[2], w[3])
        u = ((a >> 5) * 67108864.0 + (b >> 6)) / 9007199254740992.0
        assert oracle.stream_u01(seed, tag, idx, step, slot) == u
        assert 0.0 <= u < 1.0


def test_uniform_moments(oracle):
    u = np.array([oracle.stream_u01(9, 1, k, 0, 0) for k in range(20000)])
    assert abs(u.mean() - 0.5) < 4 * np.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 0.004
