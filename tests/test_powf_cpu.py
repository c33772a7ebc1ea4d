"""The evaluators' pow (gpr_powf, gpuropolis_amd/csrc/gpr_pow.h) on the CPU: the host
build of the same header, against numpy's double pow rounded to float (the correctly
rounded result, up to that double rounding) and against C99 Annex F on the special values.
The device build gives the same bits (test_gpu_units.py)."""
import numpy as np

from powf_host import SPECIAL, host_powf, inputs


def ref(a, b):
    with np.errstate(all="ignore"):
        return np.power(a.astype(np.float64), b.astype(np.float64)).astype(np.float32)


def same(g, r):
    return (g.view(np.uint32) == r.view(np.uint32)) | (np.isnan(g) & np.isnan(r))


def test_special_values_follow_c99():
    A, B = np.meshgrid(SPECIAL, SPECIAL)
    a, b = A.ravel(), B.ravel()
    g, r = host_powf(a, b), ref(a, b)
    bad = np.flatnonzero(~same(g, r))
    assert len(bad) == 0, [(a[i], b[i], g[i], r[i]) for i in bad[:10]]
    # a few by name: signs of zero and infinity, NaN rules
    assert host_powf([-0.0], [-3.0])[0] == -np.inf and host_powf([-0.0], [3.0]).view(np.uint32)[0] == 0x80000000
    assert host_powf([np.nan], [0.0])[0] == 1.0 and host_powf([1.0], [np.nan])[0] == 1.0
    assert np.isnan(host_powf([-8.0], [1.0 / 3.0])[0]) and host_powf([-2.0], [3.0])[0] == -8.0
    assert host_powf([-np.inf], [0.5])[0] == np.inf and host_powf([-1.0], [np.inf])[0] == 1.0


def test_within_one_ulp_and_almost_always_correctly_rounded():
    a, b = inputs(400000)
    g, r = host_powf(a, b), ref(a, b)
    assert np.array_equal(np.isfinite(g), np.isfinite(r))
    assert np.array_equal(np.isnan(g), np.isnan(r))
    fin = np.isfinite(r)
    ulp = np.abs(g[fin].view(np.int32).astype(np.int64) - r[fin].view(np.int32).astype(np.int64))
    assert ulp.max() <= 1, ulp.max()
    assert same(g, r).mean() >= 0.9999, same(g, r).mean()
