"""The NIST StRD real-data workload on the device.

* The reference's 18 NIST datasets (data-sources/NIST/*/data.txt, n = 6 to 236; the
  reference's run-all.sh:26,43 runs them one process each) in one context: all of them
  in the persistent launches of one block, in reference semantics (heap-31 programs,
  the reference's moves, median re-draw rule) and in config semantics (postfix
  programs, MAP-100 resampling), against per-dataset oracle runs numbered from each
  dataset's first chain.
* The 8 datasets with certified fits (data-sources/NIST/*/*.dat): the certified model
  through the device evaluator (gpr_eval_loglik, both schedules) reproduces the
  certified residual sum of squares.
"""
import os

import numpy as np
import pytest

from gpu_util import check_paths, dev, kappa, kappas, lconst, ll_agree, say
from nist_models import MODELS, certified, tape

pytestmark = pytest.mark.gpu

NIST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nist")


def nist_all():
    names = sorted(os.listdir(NIST))
    out = []
    for nm in names:
        a = np.loadtxt(os.path.join(NIST, nm, "data.txt"), comments="#", ndmin=2)
        out.append(tuple(np.ascontiguousarray(a[:, k], dtype=np.float32) for k in range(3)))
    return names, out


def test_nist_batch_reference_semantics(oracle, gpr):
    names, dsets = nist_all()
    assert len(dsets) == 18
    nch, steps, seed = 32, 150, 41
    with gpr.Context(semantics=gpr.SEM_REFERENCE, seed=seed, schedule=gpr.SCHED_WAVE,
                     resample=gpr.RS_QUANTILE) as ctx:
        ctx.set_data(dsets, [nch] * len(dsets))
        ctx.init_chains()
        i_hop, i_hc, _, _, _ = ctx.read_state()
        say(f"NIST batch, reference semantics: 18 datasets n={[len(d[0]) for d in dsets]}, {nch} chains each, "
            f"{steps} steps, median re-draw")
        ctx.run_block(steps)
        g_hop, g_hc, _, g_pr, g_ll = ctx.read_state()
        ctx.resample()
        r_hop, r_hc, _, _, _ = ctx.read_state()
    cfg = oracle.GpoRefCfg(0, seed, 0)
    on_path = strict_n = 0
    for d, (x, y, sd) in enumerate(dsets):
        sl = slice(d * nch, (d + 1) * nch)
        hop, hc = oracle.ref_init(seed, nch, chain0=d * nch)
        assert np.array_equal(i_hop[:, sl], hop) and np.array_equal(i_hc[:, sl].view(np.uint32), hc.view(np.uint32))
        marg = np.zeros(nch)
        p1, l1 = oracle.ref_run_block(cfg, hop, hc, x, y, sd, 0, steps, chain0=d * nch, margin=marg)
        same = np.all(g_hop[:, sl] == hop, axis=0) & np.all(g_hc[:, sl].view(np.uint32) == hc.view(np.uint32), axis=0)
        on_path += int(same.sum())
        check_paths(same, marg, f"{names[d]} (n={len(x)}):")
        c = np.flatnonzero(same)
        assert np.array_equal(g_pr[sl][c], p1[c]), names[d]
        kap = np.array([kappa(oracle, *oracle.ref_decode(hop[:, i].copy(), hc[:, i].copy())[:2], x, y, sd)
                        for i in c])
        ok, strict = ll_agree(g_ll[sl][c], l1[c], len(x), kap, sd)
        strict_n += int(strict.sum())
        assert ok.all(), (names[d], g_ll[sl][c][~ok], l1[c][~ok])
        # the median rule per dataset on the device's pre-resample state
        o_hop, o_hc = g_hop[:, sl].copy(), g_hc[:, sl].copy()
        pr, ll = g_pr[sl].copy(), g_ll[sl].copy()
        cutoff = np.zeros(1)
        oracle.lib().gpo_ref_resample(seed, 1, nch, d * nch, oracle.u8(o_hop), oracle.f32(o_hc), oracle.f64(pr),
                                      oracle.f64(ll), oracle.f64(cutoff), 0)
        assert np.array_equal(r_hop[:, sl], o_hop) and np.array_equal(r_hc[:, sl].view(np.uint32), o_hc.view(np.uint32))
    say(f"reference semantics: {on_path}/{nch * len(dsets)} chains on the per-dataset oracle path, "
        f"{strict_n} LLs within 1e-5 of |LL|")


def test_nist_batch_config_semantics(oracle, gpr):
    names, dsets = nist_all()
    nch, steps, seed, cap, nats = 32, 150, 43, 64, 100.0
    with gpr.Context(cap=cap, seed=seed, schedule=gpr.SCHED_WAVE, resample=gpr.RS_MAP, resample_nats=nats) as ctx:
        ctx.set_data(dsets, [nch] * len(dsets))
        ctx.init_chains()
        i_ops, i_cs, i_ln, _, _ = ctx.read_state()
        say(f"NIST batch, config semantics: 18 datasets, {nch} chains each, cap {cap}, {steps} steps, MAP-100")
        ctx.run_block(steps)
        g_ops, g_cs, g_ln, g_pr, g_ll = ctx.read_state()
        ctx.resample()
        r_ops, r_cs, r_ln, r_pr, r_ll = ctx.read_state()
    cfg = oracle.make_cfg(cap=cap, seed=seed)
    on_path = 0
    for d, (x, y, sd) in enumerate(dsets):
        sl = slice(d * nch, (d + 1) * nch)
        o_ops, o_cs, o_ln = oracle.cfg_init(cfg, nch, chain0=d * nch)
        assert np.array_equal(i_ln[sl], o_ln) and np.array_equal(i_ops[:, sl], o_ops)
        marg = np.zeros(nch)
        p1, l1, _ = oracle.cfg_run_block(cfg, o_ops, o_cs, o_ln, x, y, sd, 0, steps, chain0=d * nch, margin=marg)
        same = (g_ln[sl] == o_ln) & np.all(g_ops[:, sl] == o_ops, axis=0) & \
            np.all(g_cs[:, sl].view(np.uint32) == o_cs.view(np.uint32), axis=0)
        on_path += int(same.sum())
        check_paths(same, marg, f"{names[d]} (n={len(x)}):")
        c = np.flatnonzero(same)
        assert np.array_equal(g_pr[sl][c], p1[c]), names[d]
        kap = kappas(oracle, o_ops, o_cs, o_ln, c, x, y, sd)
        ok, _ = ll_agree(g_ll[sl][c], l1[c], len(x), kap, sd)
        assert ok.all(), (names[d], g_ll[sl][c][~ok], l1[c][~ok])
        h_ops, h_cs, h_ln = g_ops[:, sl].copy(), g_cs[:, sl].copy(), g_ln[sl].copy()
        h_pr, h_ll = g_pr[sl].copy(), g_ll[sl].copy()
        src = np.zeros(nch, dtype=np.int32)
        oracle.lib().gpo_cfg_resample(nch, d * nch, 1, seed, cap, nats, oracle.u8(h_ops), oracle.f32(h_cs),
                                      oracle.i32(h_ln), oracle.f64(h_pr), oracle.f64(h_ll), oracle.i32(src))
        assert np.array_equal(r_ln[sl], h_ln) and np.array_equal(r_ops[:, sl], h_ops)
        assert np.array_equal(r_cs[:, sl].view(np.uint32), h_cs.view(np.uint32))
        assert np.array_equal(r_pr[sl] + r_ll[sl], h_pr + h_ll)
    say(f"config semantics: {on_path}/{nch * len(dsets)} chains on the per-dataset oracle path")


@pytest.mark.parametrize("schedule", ["wave", "lane"])
def test_certified_models_through_device_evaluator(oracle, gpr, schedule):
    """gpr_eval_loglik of each certified model (float32 data and constants) against the
    oracle's FP32-node likelihood, and the residual sum of squares it implies,
    2 sd^2 (lconst - LL), against the certified RSS: within 1e-5 relative plus four
    times the model's sensitivity to one-ulp input changes (kappa), the amount by which
    FP32 evaluation itself moves the RSS."""
    import torch
    names = sorted(MODELS)
    dsets, tapes, certs = [], [], []
    for nm in names:
        cert, b = certified(nm)
        x = np.array(cert["x"], dtype=np.float32)
        y = np.array(cert["y"], dtype=np.float32)
        sd = np.full(len(x), cert["residual_sd"], dtype=np.float32)
        dsets.append((x, y, sd))
        tapes.append(tape(nm, b))
        certs.append(cert)
    cap = 64
    reps = 5      # each model in several waves and lanes
    with gpr.Context(cap=cap) as ctx:
        ctx.set_data(dsets, [1] * len(dsets))
        sched = gpr.SCHED_WAVE if schedule == "wave" else gpr.SCHED_LANE
        for d, (nm, (ops, cs), (x, y, sd), cert) in enumerate(zip(names, tapes, dsets, certs)):
            O = np.zeros((cap, reps), dtype=np.uint8)
            V = np.zeros((cap, reps), dtype=np.float32)
            O[:len(ops)] = ops[:, None]
            V[:len(ops)] = cs[:, None]
            L = np.full(reps, len(ops), dtype=np.int32)
            dO, dV, dL = dev(O), dev(V), dev(L)
            out = torch.zeros(reps, dtype=torch.float64, device="cuda")
            say(f"certified {nm} n={len(x)} schedule={schedule}")
            ctx.eval_loglik_dev(d, dO.data_ptr(), dV.data_ptr(), dL.data_ptr(), reps, cap, sched, out.data_ptr())
            got = out.cpu().numpy()
            assert np.all(got == got[0])
            ref = oracle.tape_loglik(ops, cs, x, y, sd, 2)
            k = kappa(oracle, ops, cs, x, y, sd)
            ok, strict = ll_agree(got[:1], [ref], len(x), [k], sd)
            assert ok.all(), (nm, got[0], ref, k)
            s2 = 2.0 * float(sd[0]) ** 2
            rss = (lconst(sd) - got[0]) * s2
            tol = 1e-5 * cert["rss"] + 4 * k * s2
            say(f"  RSS device {rss:.10g} certified {cert['rss']:.10g} rel {abs(rss - cert['rss']) / cert['rss']:.2e} "
                f"(tolerance {tol / cert['rss']:.2e})")
            assert abs(rss - cert["rss"]) <= tol, (nm, rss, cert["rss"], tol)
