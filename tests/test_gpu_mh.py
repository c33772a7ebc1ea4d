"""Config-semantics MH kernel against the oracle on the same Philox streams.

The draws (structure choices, generated subtrees, constant proposals, accept
uniforms) are bit-identical by construction, so chains must follow the same
trajectories: identical programs and lengths, constants equal, prior/LL within
tolerance.  A chain may leave the shared path only where the oracle's trajectory met an
accept decision within rounding of its threshold (gpu_util.check_paths).
"""
import numpy as np
import pytest

from gpu_util import check_paths, kappas, ll_agree, say

pytestmark = pytest.mark.gpu


def s_data():
    from gpuropolis_amd import workloads as W
    return W.dataset_S(0)


def test_init_matches_oracle(oracle, gpr):
    x, y, sd = s_data()
    for cap in (32, 128):
        cfg = oracle.make_cfg(cap=cap, seed=7)
        o_ops, o_cs, o_len = oracle.cfg_init(cfg, 2048)
        with gpr.Context(cap=cap, seed=7) as ctx:
            ctx.set_data([(x, y, sd)], [2048])
            say(f"init parity cap={cap}")
            ctx.init_chains()
            ops, cs, ln, _, _ = ctx.read_state()
        assert np.array_equal(ln, o_len)
        assert np.array_equal(ops, o_ops)
        assert np.array_equal(cs.view(np.uint32), o_cs.view(np.uint32))


@pytest.mark.parametrize("cap,n,steps,sched", [(32, 32, 300, "wave"), (64, 250, 150, "wave"), (128, 70, 150, "wave"),
                                               (32, 32, 300, "lane"), (128, 70, 100, "lane"),
                                               (64, "L", 150, "wave"), (64, "L", 100, "lane")])
def test_mh_trajectories_match_oracle(oracle, gpr, cap, n, steps, sched):
    """n = "L": the plain large workload's own dataset (n = 250, x on [-5, 5], a random
    target of at most 12 nodes plus 1% noise) at its cap of 64, on both schedules."""
    if n == "L":
        from gpuropolis_amd import workloads as W
        x, y, sd = W.dataset_L(0)[0]
        n = len(x)
    elif n == 32:
        x, y, sd = s_data()
    else:
        rng = np.random.default_rng(cap + n)
        x = rng.uniform(-5, 5, n).astype(np.float32)
        y = (1.5 * np.exp(-0.4 * x) + rng.normal(0, 0.05, n)).astype(np.float32)
        sd = np.full(n, 0.05, dtype=np.float32)
    nch = 64
    cfg = oracle.make_cfg(cap=cap, seed=123)
    ops, cs, ln = oracle.cfg_init(cfg, nch)
    o_ops, o_cs, o_ln = ops.copy(), cs.copy(), ln.copy()
    m0, m1 = np.zeros(nch), np.zeros(nch)
    p0, l0, cnt = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps, margin=m0)
    p1, l1, cnt1 = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, steps, steps, margin=m1)
    schedule = gpr.SCHED_WAVE if sched == "wave" else gpr.SCHED_LANE
    with gpr.Context(cap=cap, seed=123, schedule=schedule) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(ops, cs, ln)
        say(f"mh trajectories cap={cap} n={n} steps={steps} schedule={sched}")
        ctx.run_block(steps)
        ctx.run_block(steps)
        g_ops, g_cs, g_ln, g_pr, g_ll = ctx.read_state()
        st = ctx.stats()
    # on the oracle path: identical program, length and constant bits
    same = (g_ln == o_ln) & np.all(g_ops == o_ops, axis=0) & np.all(g_cs.view(np.uint32) == o_cs.view(np.uint32), axis=0)
    check_paths(same, np.minimum(m0, m1), f"cap={cap} n={n} {sched}:")
    c = np.flatnonzero(same)
    assert np.array_equal(g_pr[c], p1[c])          # prior is bit-reproducible
    kap = kappas(oracle, o_ops, o_cs, o_ln, c, x, y, sd)
    ok, strict = ll_agree(g_ll[c], l1[c], n, kap, sd)
    say("LL checks:", int(ok.sum()), "of", len(c), "agree,", int(strict.sum()), "within 1e-5 of |LL| (well conditioned)")
    assert ok.all(), (g_ll[c][~ok], l1[c][~ok], kap[~ok])
    tot = cnt + cnt1
    ndiv = nch - same.sum()
    assert abs(st.struct_proposed - tot[0]) <= 2 * steps * ndiv // 5 and abs(st.const_proposed - tot[2]) <= 2 * steps * ndiv
    assert abs(st.struct_accepted - tot[1]) <= 3 + nch - same.sum()
    assert abs(st.const_accepted - tot[3]) <= 6 + 2 * (nch - same.sum())


def test_map_resample_matches_oracle(oracle, gpr):
    x, y, sd = s_data()
    nch = 3000
    cfg = oracle.make_cfg(cap=32, seed=5)
    ops, cs, ln = oracle.cfg_init(cfg, nch)
    with gpr.Context(cap=32, seed=5, resample_nats=10.0) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(ops, cs, ln)
        ctx.run_block(50)
        g_ops, g_cs, g_ln, g_pr, g_ll = ctx.read_state()
        say("map resample")
        ctx.resample()
        r_ops, r_cs, r_ln, r_pr, r_ll = ctx.read_state()
    src = np.zeros(nch, dtype=np.int32)
    o_ops, o_cs, o_ln, o_pr, o_ll = g_ops.copy(), g_cs.copy(), g_ln.copy(), g_pr.copy(), g_ll.copy()
    import ctypes as C
    nbad = oracle.lib().gpo_cfg_resample(nch, 0, 1, 5, 32, 10.0, oracle.u8(o_ops), oracle.f32(o_cs),
                                         oracle.i32(o_ln), oracle.f64(o_pr), oracle.f64(o_ll), oracle.i32(src))
    assert nbad > 0
    assert np.array_equal(r_ln, o_ln) and np.array_equal(r_ops, o_ops)
    assert np.array_equal(r_cs.view(np.uint32), o_cs.view(np.uint32))
    assert np.array_equal(r_pr + r_ll, o_pr + o_ll)


@pytest.mark.parametrize("kc", [1, 2])
def test_early_rejection_is_exact(oracle, gpr, kc):
    """n = 250 at kc = 1 or 2 (4 or 2 chunks per evaluation): a proposal whose LL bound
    after a chunk already rejects it stops there.  Chains and accept counts are bit for
    bit those of kc = 4 (one chunk, every point evaluated) and follow the oracle, while the
    evaluated-point counter shows that evaluations stopped early."""
    rng = np.random.default_rng(250)
    x = rng.uniform(-5, 5, 250).astype(np.float32)
    y = (1.5 * np.exp(-0.4 * x) + rng.normal(0, 0.05, 250)).astype(np.float32)
    sd = np.full(250, 0.05, dtype=np.float32)
    nch, steps = 256, 300
    cfg = oracle.make_cfg(cap=64, seed=55)
    ops, cs, ln = oracle.cfg_init(cfg, nch)
    o_ops, o_cs, o_ln = ops.copy(), cs.copy(), ln.copy()
    marg = np.zeros(nch)
    _, _, cnt = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps, margin=marg)
    out = {}
    for k in (4, kc):
        with gpr.Context(cap=64, seed=55, schedule=gpr.SCHED_WAVE) as ctx:
            ctx.set_data([(x, y, sd)], [nch])
            ctx.set_tuning([dict(n_class=c, hi=h, kc=k, waves_per_wg=1, wg_per_cu=8, ms=0.0)
                            for c in (1, 2, 4) for h in (0, 1)])
            ctx.write_state(ops, cs, ln)
            say(f"early rejection n=250 kc={k}")
            ctx.run_block(steps)
            out[k] = (ctx.read_state(), ctx.stats())
    (s4, st4), (sk, stk) = out[4], out[kc]
    for u, v in zip(s4, sk):
        assert np.array_equal(np.asarray(u).view(np.uint8), np.asarray(v).view(np.uint8))
    for f in ("struct_proposed", "struct_accepted", "const_proposed", "const_accepted", "steps"):
        assert getattr(st4, f) == getattr(stk, f), f
    say(f"evaluated node x points: kc=4 {st4.node_evals}, kc={kc} {stk.node_evals}")
    assert stk.node_evals < 0.9 * st4.node_evals
    g_ops, g_cs, g_ln = sk[0], sk[1], sk[2]
    same = (g_ln == o_ln) & np.all(g_ops == o_ops, axis=0) & np.all(g_cs.view(np.uint32) == o_cs.view(np.uint32), axis=0)
    check_paths(same, marg, f"early rejection kc={kc}:")
    assert abs(stk.const_accepted - cnt[3]) <= 6 + 2 * (nch - same.sum())


def test_work_counters_exact_on_fixed_structures(oracle, gpr):
    """With structure moves off and one chunk of points (n <= 64, so no early rejection)
    every block evaluates each chain's program once at its start and once per step when
    it has a constant to move: node_evals = sum len * n * (1 + steps * has_const) and the
    FP32 work count = sum ((len - leaves + 4) * n + constants) over the same evaluations."""
    x, y, sd = s_data()
    n, nch, steps, blocks = len(x), 256, 60, 2
    assert n <= 64
    cfg = oracle.make_cfg(cap=32, seed=5)
    ops, cs, ln = oracle.cfg_init(cfg, nch)
    X, C, ONE = gpr.OP["X"], gpr.OP["CONST"], gpr.OP["ONE"]
    CADD = gpr.OP["CADD"]
    with gpr.Context(cap=32, seed=5, fixed_structure=1, resample=gpr.RS_NONE, schedule=gpr.SCHED_WAVE) as ctx:
        ctx.set_data([(x, y, sd)], [nch])
        ctx.write_state(ops, cs, ln)
        say("work counters on fixed structures")
        for _ in range(blocks):
            ctx.run_block(steps)
        st = ctx.stats()
    evals = flops = 0
    for c in range(nch):
        o = ops[:ln[c], c]
        nconst = int(np.sum((o == C) | (o >= CADD)))
        leaves = int(np.sum(o <= ONE))
        k = blocks * (1 + steps * (nconst > 0))
        evals += k * int(ln[c]) * n
        flops += k * ((int(ln[c]) - leaves + 4) * n + nconst)
    assert (X, C, ONE) == (0, 1, 2)
    assert st.node_evals == evals and st.alg_flops == flops, (st.node_evals, evals, st.alg_flops, flops)
