"""Batching: many datasets of uneven size and programs of uneven cap in one persistent
launch.  The batched run equals independent per-dataset runs of the oracle whose chain
numbering starts at the dataset's first chain (the Philox streams are keyed by the
global chain index): identical initial programs, the same trajectories, identical
per-dataset MAP resampling."""
import numpy as np
import pytest

from gpu_util import check_paths, kappas, ll_agree, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("sched", ["wave", "lane"])
def test_batched_run_equals_per_dataset_runs(oracle, gpr, sched):
    from gpuropolis_amd import workloads as W
    dsets, caps, _ = W.datasets_B(0, nd=9)
    # n spans the chunking cases (one, two and four points per lane)
    nch, steps, seed, nats = 48, 120, 17, 20.0
    chains = [nch] * len(dsets)
    schedule = gpr.SCHED_WAVE if sched == "wave" else gpr.SCHED_LANE
    with gpr.Context(cap=128, seed=seed, schedule=schedule, resample_nats=nats) as ctx:
        ctx.set_data(dsets, chains, caps)
        ctx.init_chains()
        i_ops, i_cs, i_ln, _, _ = ctx.read_state()
        say(f"batched: {len(dsets)} datasets n={[len(d[0]) for d in dsets]} caps={caps} schedule={sched}")
        ctx.run_block(steps)
        g_ops, g_cs, g_ln, g_pr, g_ll = ctx.read_state()
        ctx.resample()
        r_ops, r_cs, r_ln, r_pr, r_ll = ctx.read_state()
    on_path = total = 0
    for d, ((x, y, sd), cap) in enumerate(zip(dsets, caps)):
        c0 = d * nch
        sl = slice(c0, c0 + nch)
        cfg = oracle.make_cfg(cap=cap, seed=seed)
        o_ops, o_cs, o_ln = oracle.cfg_init(cfg, nch, chain0=c0)
        assert np.array_equal(i_ln[sl], o_ln) and np.array_equal(i_ops[:cap, sl], o_ops)
        assert np.array_equal(i_cs[:cap, sl].view(np.uint32), o_cs.view(np.uint32))
        assert not np.any(i_ops[cap:, sl])
        marg = np.zeros(nch)
        p1, l1, _ = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps, chain0=c0, margin=marg)
        same = (g_ln[sl] == o_ln) & np.all(g_ops[:cap, sl] == o_ops, axis=0) & \
            np.all(g_cs[:cap, sl].view(np.uint32) == o_cs.view(np.uint32), axis=0)
        on_path += int(same.sum())
        total += nch
        check_paths(same, marg, f"dataset {d} (n={len(x)}):")
        c = np.flatnonzero(same)
        assert np.array_equal(g_pr[sl][c], p1[c])
        kap = kappas(oracle, o_ops, o_cs, o_ln, c, x, y, sd)
        ok, _ = ll_agree(g_ll[sl][c], l1[c], len(x), kap, sd)
        assert ok.all(), (d, g_ll[sl][c][~ok], l1[c][~ok])
        # per-dataset MAP resampling of the device state, on the host
        h_ops, h_cs, h_ln = g_ops[:cap, sl].copy(), g_cs[:cap, sl].copy(), g_ln[sl].copy()
        h_pr, h_ll = g_pr[sl].copy(), g_ll[sl].copy()
        src = np.zeros(nch, dtype=np.int32)
        oracle.lib().gpo_cfg_resample(nch, c0, 1, seed, cap, nats, oracle.u8(h_ops), oracle.f32(h_cs),
                                      oracle.i32(h_ln), oracle.f64(h_pr), oracle.f64(h_ll), oracle.i32(src))
        assert np.array_equal(r_ln[sl], h_ln) and np.array_equal(r_ops[:cap, sl], h_ops)
        assert np.array_equal(r_cs[:cap, sl].view(np.uint32), h_cs.view(np.uint32))
    say(f"batched {sched}: {on_path}/{total} chains on the per-dataset oracle path")
