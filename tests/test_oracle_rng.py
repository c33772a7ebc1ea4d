"""Pins of the oracle's random-number layer (CPU only).

Philox4x32-10: known-answer vectors of the Random123 distribution (Salmon et al., SC'11,
file kat_vectors) and vectors computed by torch's ATen/core/PhiloxRNGEngine.h
(tests/golden/philox_kat.json, written by tests/golden/make_philox_kat.py).
LCG: the BSD rand() sequence the reference's cuda_rand restates (main.cu:213-216),
published on Rosetta Code "Linear congruential generator" (seed 0).
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# Random123 kat_vectors: philox4x32 10 rounds
R123 = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", R123)
def test_philox_random123_kat(ctr, key, out):
    assert tuple(int(v) for v in O.philox(ctr, key)) == out


def test_philox_torch_vectors():
    kat = json.load(open(os.path.join(GOLDEN, "philox_kat.json")))
    assert len(kat["vectors"]) >= 5
    for v in kat["vectors"]:
        assert list(O.philox(v["ctr"], v["key"])) == v["out"]


def test_stream_consumes_blocks_in_order():
    seed, chain, step, purpose, slot = 0x1234_5678_9abc_def0, 77, 123456, O.P_STRUCT, 5
    words = O.stream(seed, chain, step, purpose, slot, 12)
    key = (seed & 0xffffffff, seed >> 32)
    expect = np.concatenate([O.philox((chain, step, (purpose << 24) | slot, b), key) for b in range(3)])
    assert np.array_equal(words, expect)


def test_lcg_bsd_sequence():
    rx = C.c_uint32(0)
    got = [O.lib().gpo_lcg_next(C.byref(rx)) for _ in range(10)]
    assert got == [12345, 1406932606, 654583775, 1449466924, 229283573, 1109335178, 1051550459, 1293799192,
                   794471793, 551188310]


def test_u01_grid():
    L = O.lib()
    assert L.gpo_u01(0) == 0.5 * 2 ** -23
    assert L.gpo_u01(0xffffffff) == 1.0 - 0.5 * 2 ** -23
    r = np.random.default_rng(0).integers(0, 2 ** 32, 1000, dtype=np.uint64)
    u = np.array([L.gpo_u01(int(v)) for v in r])
    assert np.all((u > 0) & (u < 1))
    assert np.allclose(u, ((r >> 9) + 0.5) * 2.0 ** -23, rtol=0, atol=0)


def test_reference_random_int_uniform_and_in_range():
    # random_int(n): rejection with divisor (2^31-1)/(n+1), main.cu:218-228
    for rng in (0, 1):
        v = O.ref_draws(0, 60000, rng=rng, seed=9, lcg_seed=12345, arg=31)
        assert v.min() >= 0 and v.max() <= 30
        counts = np.bincount(v.astype(int), minlength=31)
        chi2 = ((counts - len(v) / 31) ** 2 / (len(v) / 31)).sum()
        assert chi2 < 70  # 30 dof, p ~ 1e-4


def test_reference_random_float_grid():
    # random_float = random_int(10^6)/10^6 in [0, 0.999999], main.cu:230-232
    v = O.ref_draws(1, 20000, seed=3)
    assert v.min() >= 0 and v.max() <= 0.999999
    k = np.round(v * 1e6)
    assert np.allclose(k / 1e6, v, atol=1e-7)
    assert abs(v.mean() - 0.5) < 0.01


def test_reference_geometric_pmf():
    # random_geometric(0.1) = ceil(log u / log 0.1): P(k) = 0.9 * 0.1^(k-1) = exp(lgeometricpdf)
    v = O.ref_draws(3, 200000, seed=4)
    v = v[np.isfinite(v)]
    for k in (1, 2, 3):
        p = np.mean(v == k)
        expect = np.exp(O.lib().gpo_ref_lgeometricpdf(float(k), 0.1))
        assert abs(p - 0.9 * 0.1 ** (k - 1)) < 4 * np.sqrt(expect / len(v)) + 1e-4
        assert abs(expect - 0.9 * 0.1 ** (k - 1)) < 1e-6


def test_reference_normal_moments():
    for kind in (2, 6):
        v = O.ref_draws(kind, 200000, seed=5)
        v = v[np.isfinite(v)]
        assert abs(v.mean()) < 0.01 and abs(v.std() - 1) < 0.01


def test_reference_host_random_int_range():
    v = O.ref_draws(4, 50000, seed=6, arg=26)   # init draws ops from U{0..NOPS-2}, main.cu:1133
    assert v.min() == 0 and v.max() == 25
