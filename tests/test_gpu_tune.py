"""Launch shapes and the autotuner.  Every shape (points per lane per chunk, chains per
workgroup, workgroups per CU) must give the same chains: a batch with one dataset in
each size class follows the oracle's trajectories under every forced shape.  The
autotuner leaves the chains, counters and step clock as it found them and returns a
valid shape for each class present."""
import numpy as np
import pytest

from gpu_util import ll_close, say

pytestmark = pytest.mark.gpu


def batch():
    rng = np.random.default_rng(12)
    out = []
    for n in (20, 100, 250):
        x = rng.uniform(-3, 3, n).astype(np.float32)
        y = (np.exp(0.5 * x) - x + rng.normal(0, 0.1, n)).astype(np.float32)
        out.append((x, y, np.full(n, 0.1, dtype=np.float32)))
    return out


@pytest.mark.parametrize("kc,wpg", [(1, 1), (2, 2), (4, 1), (1, 4), (4, 4), (2, 1)])
def test_forced_shapes_follow_oracle(oracle, gpr, kc, wpg):
    dsets = batch()
    caps = [32, 64, 128]
    nch, steps, seed = 40, 100, 21
    with gpr.Context(cap=128, seed=seed) as ctx:
        ctx.set_data(dsets, [nch] * 3, caps)
        ctx.set_tuning([dict(n_class=c, hi=h, kc=kc, waves_per_wg=wpg, wg_per_cu=4, ms=0.0)
                        for c in (1, 2, 4) for h in (0, 1)])
        ctx.init_chains()
        say(f"forced shape kc={kc} wpg={wpg}")
        ctx.run_block(steps)
        g_ops, g_cs, g_ln, g_pr, _ = ctx.read_state()
    for d, ((x, y, sd), cap) in enumerate(zip(dsets, caps)):
        sl = slice(d * nch, (d + 1) * nch)
        cfg = oracle.make_cfg(cap=cap, seed=seed)
        o_ops, o_cs, o_ln = oracle.cfg_init(cfg, nch, chain0=d * nch)
        p1, _, _ = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps, chain0=d * nch)
        same = (g_ln[sl] == o_ln) & np.all(g_ops[:cap, sl] == o_ops, axis=0) & \
            np.all(g_cs[:cap, sl].view(np.uint32) == o_cs.view(np.uint32), axis=0)
        assert same.sum() >= nch - 1, (d, int(same.sum()))
        assert np.array_equal(g_pr[sl][same], p1[same])


def test_autotune_restores_state_and_picks_valid_shapes(gpr):
    dsets = batch()
    with gpr.Context(cap=128, seed=5) as ctx:
        ctx.set_data(dsets, [512] * 3, [32, 64, 128])
        ctx.init_chains()
        ctx.run_block(30)
        before = ctx.read_state()
        st0 = ctx.stats()
        say("autotune, 3 size classes")
        table, sched = ctx.autotune(20)
        after = ctx.read_state()
        st1 = ctx.stats()
        for u, v in zip(before, after):
            assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
        assert st0.steps == st1.steps and st0.const_proposed == st1.const_proposed
        say("tuned:", table, "schedule", sched)
        assert sorted((t["n_class"], t["hi"]) for t in table) == [(1, 0), (2, 0), (4, 1)]
        for t in table:
            assert t["kc"] in (1, 2, 4) and t["waves_per_wg"] in (1, 2, 4) and t["ms"] > 0
        # the tuned context continues exactly like an untuned one
        ctx.run_block(30)
        tuned = ctx.read_state()
    with gpr.Context(cap=128, seed=5) as ctx:
        ctx.set_data(dsets, [512] * 3, [32, 64, 128])
        ctx.init_chains()
        ctx.run_block(30)
        ctx.run_block(30)
        plain = ctx.read_state()
    # the schedule may have changed: programs, constants and priors agree bit for bit,
    # log-likelihoods to the evaluator tolerance (lane and wave sum in different orders)
    for u, v in zip(tuned[:4], plain[:4]):
        assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    ok, _, _ = ll_close(tuned[4], plain[4], 250)
    assert ok.all()
