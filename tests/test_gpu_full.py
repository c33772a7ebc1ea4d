"""The full form on the device against per-dataset oracle runs.

Full form: datasets of uneven size (n log-uniform in [16, 1024], x on [-10, 10], random
targets of at most 24 nodes plus 1% noise), programs of at most 128 tokens, MALA
constant moves with subtree structure moves, all datasets in one persistent launch,
MAP-100 resampling between blocks (the reference's run loop, main.cu:1156-1239, around
MH_simple_kernel, main.cu:679-813; its N cap, main.cu:1059, is what the full form's
1,048,576 chains exceed).

A subset of the full form's 256 datasets runs here: twelve datasets spanning n = 16 to
1017, four of them above 512 points, so the two-register MALA kernel's token slab in
global memory sweeps up to 16 one-point chunks per evaluation.  Each dataset's chains
are numbered from its first chain, as a separate oracle run numbers them.

The MALA gradient is summed in FP32 per 64-point chunk on the device and in FP64 in the
oracle, so a Langevin proposal's mean can differ in its last bits and the two samplers
then shadow each other instead of agreeing bit for bit.  The checks:
* structure: the chain's program (tokens and length) equals the oracle's for at least
  98% of chains (a last-bit difference of a constant can flip an accept decision that
  lies within rounding of its threshold; in practice every chain agrees);
* constants: within 1e-3 relative of the oracle's chain (most are bit-identical);
* priors: bit-exact against the oracle's prior of the device's own final programs;
* likelihoods: against the oracle's evaluation of the device's own final programs, to
  the tolerance of the other trajectory tests and within 1e-5 of |LL| for
  well-conditioned programs;
* resampling: the device's MAP-100 resample equals the oracle's rule applied to the
  device's pre-resample state, bit for bit;
* a second block continues from the resampled state on both sides (the Philox clock).
"""
import numpy as np
import pytest

from gpu_util import kappas, ll_agree, say

pytestmark = pytest.mark.gpu

SEED, EPS, NATS, CAP = 11, 0.01, 100.0, 128


def subset():
    from gpuropolis_amd import workloads as W
    dsets, _ = W.datasets_F(0)
    ns = np.array([len(d[0]) for d in dsets])
    order = np.argsort(ns, kind="stable")
    pick = list(order[[0, 30, 70, 110, 150, 190, 220, 235]]) + list(order[-4:])
    return [dsets[i] for i in pick], ns[pick]


def oracle_block(oracle, cfg, state, dsets, nch, step0, steps):
    ops, cs, ln = state
    out = []
    for d, (x, y, sd) in enumerate(dsets):
        sl = slice(d * nch, (d + 1) * nch)
        o, c, l_ = ops[:, sl].copy(), cs[:, sl].copy(), ln[sl].copy()
        p, ll, cnt = oracle.cfg_run_block(cfg, o, c, l_, x, y, sd, step0, steps, chain0=d * nch)
        out.append((o, c, l_, p, ll, cnt))
    return out


def check_block(oracle, dsets, nch, g, ref, tag):
    g_ops, g_cs, g_ln, g_pr, g_ll = g
    struct_same = const_bits = const_close = total = 0
    strict_n = 0
    for d, (x, y, sd) in enumerate(dsets):
        sl = slice(d * nch, (d + 1) * nch)
        o_ops, o_cs, o_ln, _, _, _ = ref[d]
        ops, cs, ln = g_ops[:, sl], g_cs[:, sl], g_ln[sl]
        same = (ln == o_ln) & np.all(ops == o_ops, axis=0)
        struct_same += int(same.sum())
        bits = same & np.all(cs.view(np.uint32) == o_cs.view(np.uint32), axis=0)
        const_bits += int(bits.sum())
        rel = np.max(np.abs(cs.astype(np.float64) - o_cs) / np.maximum(np.abs(o_cs.astype(np.float64)), 1e-3), axis=0)
        close = same & (rel < 1e-3)
        const_close += int(close.sum())
        total += nch
        # the device's own final state: prior bit-exact, LL to tolerance
        pr = np.array([oracle.tape_prior(ops[:ln[i], i], cs[:ln[i], i], 10) for i in range(nch)])
        assert np.array_equal(g_pr[sl], pr), (tag, d, np.flatnonzero(g_pr[sl] != pr)[:5])
        lr = np.array([oracle.tape_loglik(ops[:ln[i], i], cs[:ln[i], i], x, y, sd, 2) for i in range(nch)])
        kap = kappas(oracle, ops, cs, ln, range(nch), x, y, sd)
        ok, strict = ll_agree(g_ll[sl], lr, len(x), kap, sd)
        strict_n += int(strict.sum())
        assert ok.all(), (tag, d, len(x), g_ll[sl][~ok], lr[~ok], kap[~ok])
    say(f"{tag}: structure equal {struct_same}/{total}, constants bit-equal {const_bits}, within 1e-3 "
        f"{const_close}; LL strict checks {strict_n}")
    assert struct_same >= 0.98 * total
    assert const_close >= 0.95 * total
    return const_close


def test_full_form_subset_matches_oracle(oracle, gpr):
    dsets, ns = subset()
    assert ns.min() == 16 and ns.max() == 1017 and (ns > 512).sum() >= 4
    nch, steps = 48, 150
    chains = [nch] * len(dsets)
    with gpr.Context(cap=CAP, seed=SEED, schedule=gpr.SCHED_WAVE, proposal=gpr.PROP_MALA, eps=EPS,
                     resample=gpr.RS_MAP, resample_nats=NATS) as ctx:
        ctx.set_data(dsets, chains, [CAP] * len(dsets))
        ctx.init_chains()
        init = ctx.read_state()[:3]
        say(f"full form subset: {len(dsets)} datasets n={list(map(int, ns))}, {nch} chains each, cap {CAP}, "
            f"MALA eps {EPS} + subtree moves, {steps} steps, MAP-{NATS:g} resample, then a second block")
        ctx.run_block(steps)
        g1 = ctx.read_state()
        st1 = ctx.stats()
        ctx.resample()
        r1 = ctx.read_state()
        ctx.run_block(steps)
        g2 = ctx.read_state()
    cfg = oracle.make_cfg(cap=CAP, seed=SEED, proposal=1, eps=EPS)
    # initial programs are the oracle's, dataset by dataset
    for d in range(len(dsets)):
        o_ops, o_cs, o_ln = oracle.cfg_init(cfg, nch, chain0=d * nch)
        sl = slice(d * nch, (d + 1) * nch)
        assert np.array_equal(init[2][sl], o_ln) and np.array_equal(init[0][:, sl], o_ops)
        assert np.array_equal(init[1][:, sl].view(np.uint32), o_cs.view(np.uint32))
    ref1 = oracle_block(oracle, cfg, init, dsets, nch, 0, steps)
    check_block(oracle, dsets, nch, g1, ref1, "block 1")
    sp = sum(int(r[5][0]) for r in ref1)
    say("structure moves proposed: device", st1.struct_proposed, "oracle", sp)
    assert st1.struct_proposed == sp
    assert st1.const_accepted > 0 and st1.struct_accepted > 0
    # MAP-100 resampling of the device state, per dataset on the host
    nres = 0
    for d in range(len(dsets)):
        sl = slice(d * nch, (d + 1) * nch)
        h_ops, h_cs, h_ln = g1[0][:, sl].copy(), g1[1][:, sl].copy(), g1[2][sl].copy()
        h_pr, h_ll = g1[3][sl].copy(), g1[4][sl].copy()
        src = np.zeros(nch, dtype=np.int32)
        nres += oracle.lib().gpo_cfg_resample(nch, d * nch, 1, SEED, CAP, NATS, oracle.u8(h_ops), oracle.f32(h_cs),
                                              oracle.i32(h_ln), oracle.f64(h_pr), oracle.f64(h_ll), oracle.i32(src))
        assert np.array_equal(r1[2][sl], h_ln) and np.array_equal(r1[0][:, sl], h_ops)
        assert np.array_equal(r1[1][:, sl].view(np.uint32), h_cs.view(np.uint32))
        assert np.array_equal(r1[3][sl], h_pr) and np.array_equal(r1[4][sl], h_ll)
    say("chains resampled:", nres)
    # the second block from the resampled state, steps [steps, 2 steps)
    ref2 = oracle_block(oracle, cfg, r1[:3], dsets, nch, steps, steps)
    check_block(oracle, dsets, nch, g2, ref2, "block 2")
