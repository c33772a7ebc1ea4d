"""A host build of the evaluators' pow (gpuropolis_amd/csrc/gpr_pow.h, plain C++ with g++)
for the tests, and the inputs they share."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

SHIM = r"""
#include "gpr_pow.h"
extern "C" void host_powf(const float *a, const float *b, float *o, int n) {
    for (int i = 0; i < n; i++) o[i] = gpr_powf(a[i], b[i]);
}
"""


def lib():
    global _LIB
    if _LIB is None:
        d = tempfile.mkdtemp(prefix="gpr_pow_")
        src, out = os.path.join(d, "shim.cpp"), os.path.join(d, "libpow.so")
        with open(src, "w") as f:
            f.write(SHIM)
        subprocess.check_call(["g++", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
                               "-I", os.path.join(ROOT, "gpuropolis_amd", "csrc"), src, "-o", out])
        _LIB = C.CDLL(out)
    return _LIB


def host_powf(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.zeros_like(a)
    vp = C.c_void_p
    lib().host_powf(vp(a.ctypes.data), vp(b.ctypes.data), vp(out.ctypes.data), len(a))
    return out


SPECIAL = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 0.5, -0.5, 2.0, -2.0, 3.0, -3.0, 2.5, -2.5,
                    1e-40, -1e-40, 1e-45, 3e38, -3e38, 1e30, 16777216.0, 16777217.0, -16777215.0, 0.1, -7.0],
                   dtype=np.float32)


def inputs(n, seed=0):
    """Pairs (a, b): the special values crossed with each other, then random powers over
    the whole range of a (both signs), integer and non-integer b, bases near 1 with large
    exponents, results near the underflow threshold, and the full form's x range."""
    rng = np.random.default_rng(seed)
    A, B = np.meshgrid(SPECIAL, SPECIAL)
    pa, pb = [A.ravel()], [B.ravel()]
    pa.append(np.exp(rng.uniform(-80, 80, n)) * rng.choice([1.0, -1.0], n))
    pb.append(np.where(rng.random(n) < 0.3, rng.integers(-20, 20, n), rng.normal(0, 3, n)))
    pa.append(1 + rng.normal(0, 1e-3, n))
    pb.append(rng.normal(0, 3e4, n))
    a = np.exp(rng.uniform(-5, 5, n))
    pa.append(a)
    pb.append(rng.uniform(-150, -100, n) / np.log2(a.astype(np.float32).astype(np.float64)))
    pa.append(rng.uniform(-10, 10, n))
    pb.append(rng.uniform(-10, 10, n))
    return np.concatenate(pa).astype(np.float32), np.concatenate(pb).astype(np.float32)
