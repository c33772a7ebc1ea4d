"""Proposal metric (preconditioned constant moves) on the device against the oracle.

Random walk with a metric draws exactly the oracle's proposals, so chains follow the
oracle's trajectories bit for bit.  MALA with a metric uses the device adjoint (FP32
sums) where the oracle uses FP64, so device chains shadow the oracle's.  The oracle's
metric moves are pinned by quadrature on a correlated two-constant posterior
(tests/test_oracle_sampler.py)."""
import numpy as np
import pytest

from gpu_util import say

pytestmark = pytest.mark.gpu


def line_problem(nch):
    rng = np.random.default_rng(21)
    x = rng.uniform(5.0, 7.0, 30).astype(np.float32)
    y = (0.8 * x - 1.5 + rng.normal(0.0, 0.2, 30)).astype(np.float32)
    sd = np.full(30, 0.2, dtype=np.float32)
    xd, yd, s = (v.astype(np.float64) for v in (x, y, sd))
    A = np.stack([xd, np.ones_like(xd)], 1) / s[:, None]
    cov = np.linalg.inv(A.T @ A)
    mu = cov @ (A.T @ (yd / s))
    L = np.linalg.cholesky(cov)
    ops = np.zeros((8, nch), dtype=np.uint8)
    ops[:5] = np.array([0, 1, 5, 1, 3], dtype=np.uint8)[:, None]      # x c * c +
    cs = np.zeros((8, nch), dtype=np.float32)
    cs[1] = (mu[0] + 0.1 * rng.normal(size=nch)).astype(np.float32)
    cs[3] = (mu[1] + 0.5 * rng.normal(size=nch)).astype(np.float32)
    return (x, y, sd), L, ops, cs, np.full(nch, 5, dtype=np.int32)


@pytest.mark.parametrize("proposal", ["rw", "mala"])
def test_metric_moves_match_oracle(oracle, gpr, proposal):
    nch, steps = 256, 300
    (x, y, sd), L, ops, cs, ln = line_problem(nch)
    prop = 1 if proposal == "mala" else 0
    step = 0.9
    cfg = oracle.make_cfg(cap=8, seed=13, proposal=prop, sigma=step, eps=step, fixed_structure=1,
                          metric=([1, 3], L))
    o_cs = cs.copy()
    _, _, cnt = oracle.cfg_run_block(cfg, ops.copy(), o_cs, ln.copy(), x, y, sd, 0, steps)
    with gpr.Context(cap=8, seed=13, proposal=prop, sigma=step, eps=step, fixed_structure=1,
                     resample=gpr.RS_NONE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.set_metric(0, [1, 3], L)
        ctx.write_state(ops, cs, ln)
        say(f"metric {proposal}: {nch} chains x {steps} steps")
        ctx.run_block(steps)
        g_cs = ctx.read_state()[1]
        st = ctx.stats()
    acc = st.const_accepted / st.const_proposed
    say(f"acceptance gpu {acc:.3f} oracle {cnt[3] / cnt[2]:.3f}")
    if proposal == "rw":
        same = np.all(g_cs.view(np.uint32) == o_cs.view(np.uint32), axis=0)
        assert same.sum() >= nch - 1
        assert abs(st.const_accepted - cnt[3]) <= 2
    else:
        rel = np.max(np.abs(g_cs[[1, 3]].astype(np.float64) - o_cs[[1, 3]]) /
                     np.maximum(np.abs(o_cs[[1, 3]].astype(np.float64)), 1e-3), axis=0)
        say("chains within 1e-3 of the oracle:", int((rel < 1e-3).sum()))
        assert (rel < 1e-3).sum() >= 0.9 * nch
        assert abs(st.const_accepted - cnt[3]) <= 0.05 * cnt[3] + 5
    assert 0.2 < acc < 0.95


def test_metric_argument_checks(gpr):
    (x, y, sd), L, ops, cs, ln = line_problem(8)
    with gpr.Context(cap=8, fixed_structure=0) as ctx:
        ctx.set_data([(x, y, sd)], [8])
        with pytest.raises(gpr.GprError, match="fixed_structure"):
            ctx.set_metric(0, [1, 3], L)
    with gpr.Context(cap=8, fixed_structure=1) as ctx:
        ctx.set_data([(x, y, sd)], [8])
        with pytest.raises(gpr.GprError, match="distinct"):
            ctx.set_metric(0, [1, 1], L)
        with pytest.raises(gpr.GprError, match="cap"):
            ctx.set_metric(0, [1, 9], L)
        with pytest.raises(gpr.GprError, match="diagonal"):
            ctx.set_metric(0, [1, 3], -L)
        ctx.set_metric(0, [1, 3], L)
        ctx.set_metric(0, [], None)
