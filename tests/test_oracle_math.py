"""Pins of the draw transforms shared by the oracle and the HIP library (CPU only):
against libm and scipy within a few ulp, special cases, exact odd symmetry."""
import math

import numpy as np
from scipy.stats import norm

import oracle_lib as O


def ulps(a, b):
    b32 = np.float32(b)
    return abs(a - b) / float(np.spacing(np.abs(b32)) if b32 != 0 else np.float32(1e-45))


def test_logf_accuracy_and_specials():
    L = O.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-87, 88, 20000)), rng.uniform(0.5, 2, 5000),
                         [1e-40, 1.4e-45, 3.0e38, 1.0, 2.0, 0.5]]).astype(np.float32)
    worst = max(ulps(L.gpo_logf(float(x)), math.log(float(x))) for x in xs)
    assert worst <= 2.0, worst
    assert L.gpo_logf(1.0) == 0.0
    assert L.gpo_logf(0.0) == -math.inf
    assert L.gpo_logf(math.inf) == math.inf
    assert math.isnan(L.gpo_logf(-1.0)) and math.isnan(L.gpo_logf(math.nan))


def test_expf_accuracy_and_specials():
    L = O.lib()
    xs = np.random.default_rng(2).uniform(-87, 88.7, 20000).astype(np.float32)
    worst = max(ulps(L.gpo_expf(float(x)), math.exp(float(x))) for x in xs)
    assert worst <= 2.0, worst
    assert L.gpo_expf(0.0) == 1.0
    assert L.gpo_expf(100.0) == math.inf and L.gpo_expf(-200.0) == 0.0
    # gradual underflow is rounded once
    assert abs(L.gpo_expf(-100.0) - math.exp(-100.0)) <= 2 ** -149


def test_sin2pi():
    L = O.lib()
    vs = np.random.default_rng(3).uniform(0, 1, 20000).astype(np.float32)
    err = max(abs(L.gpo_sin2pi(float(v)) - math.sin(2 * math.pi * float(v))) for v in vs)
    assert err < 2e-7
    assert L.gpo_sin2pi(0.25) == 1.0 and L.gpo_sin2pi(0.0) == 0.0 and L.gpo_sin2pi(0.75) == -1.0


def test_normal_quantile_accuracy_and_symmetry():
    L = O.lib()
    ks = np.concatenate([np.arange(2000), np.random.default_rng(4).integers(0, 2 ** 23, 30000),
                         2 ** 23 - 1 - np.arange(2000)])
    worst, asym = 0.0, 0
    for k in ks:
        u = (int(k) + 0.5) * 2.0 ** -23
        z = L.gpo_normal_icdf(u)
        worst = max(worst, abs(z - norm.ppf(u)) / max(1.0, abs(norm.ppf(u))))
        asym += z != -L.gpo_normal_icdf((2 ** 23 - 1 - int(k) + 0.5) * 2.0 ** -23)
    assert worst < 1e-6, worst
    assert asym == 0           # exact: random-walk proposals are exactly symmetric


def test_normal_draws_moments():
    L = O.lib()
    r = np.random.default_rng(5).integers(0, 2 ** 32, 200000, dtype=np.uint64)
    z = np.array([L.gpo_normal_icdf(L.gpo_u01(int(v))) for v in r[:100000]])
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert abs(np.mean(z ** 4) - 3) < 0.1
