"""Pins the oracle's Morton codes (restating Core/CudaUtils.cu:278-325) against a
bit-by-bit interleave written independently here, and the 63/30-bit identity."""
import numpy as np
import pytest


def interleave(qx, qy, qz, bits):
    code = 0
    for k in range(bits):
        code |= ((qx >> k) & 1) << (3 * k + 2)
        code |= ((qy >> k) & 1) << (3 * k + 1)
        code |= ((qz >> k) & 1) << (3 * k)
    return code


def quant(c, bits):
    s = np.float32(c) * np.float32(1 << bits)
    s = min(max(s, np.float32(0.0)), np.float32((1 << bits) - 1))
    return int(s)


def test_morton63_matches_interleave(oracle_mod):
    rng = np.random.default_rng(11)
    c = rng.uniform(-0.2, 1.2, size=(2000, 3)).astype(np.float32)
    c[:5] = [[0, 0, 0], [1, 1, 1], [0.5, 0.25, 0.125], [-1, 2, 0.999999], [1e-7, 0.5, 1 - 1e-7]]
    got = oracle_mod.morton63(c)
    for i in range(len(c)):
        want = interleave(quant(c[i, 0], 21), quant(c[i, 1], 21), quant(c[i, 2], 21), 21)
        assert int(got[i]) == want, i


def test_morton30_is_morton63_shifted(oracle_mod):
    # same float coordinate, box [0,1]: the 30-bit code is the top 30 bits of the 63-bit one
    rng = np.random.default_rng(12)
    c = rng.uniform(-0.1, 1.1, size=(50000, 3)).astype(np.float32)
    k63 = oracle_mod.morton63(c)
    k30 = oracle_mod.morton30(c, [0, 0, 0], [1, 1, 1])
    assert np.array_equal(k30, k63 >> np.uint64(33))


def test_morton30_normalised_box(oracle_mod):
    c = np.array([[2.0, -1.0, 5.0], [4.0, 1.0, 5.0], [3.0, 0.0, 5.0]], dtype=np.float32)
    k = oracle_mod.morton30(c, c.min(0), c.max(0))
    # first point maps to (0, 0, 0); second to (1023, 1023, 0); third to (512, 512, 0)
    assert int(k[0]) == 0
    assert int(k[1]) == interleave(1023, 1023, 0, 10)
    assert int(k[2]) == interleave(512, 512, 0, 10)
