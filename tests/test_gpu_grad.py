"""Adjoint kernel (gpr_grad_logpost): d log posterior / d constants against FP64 central
differences of the log posterior (FP64 nodes and constants, the oracle), against the
oracle's FP64 reverse sweep, and through the torch.autograd wrapper under
torch.autograd.gradcheck.  The log posterior it returns equals prior + the evaluator's
log-likelihood bit for bit."""
import numpy as np
import pytest

from gpu_util import dev, say

pytestmark = pytest.mark.gpu


def pack(tapes, cap):
    N = len(tapes)
    O = np.zeros((cap, N), dtype=np.uint8)
    V = np.zeros((cap, N), dtype=np.float32)
    L = np.zeros(N, dtype=np.int32)
    for i, (o, v) in enumerate(tapes):
        O[:len(o), i] = o
        V[:len(o), i] = v
        L[i] = len(o)
    return O, V, L


def device_grad(gpr, data, O, V, L, cap):
    import torch
    N = L.shape[0]
    with gpr.Context(cap=cap) as ctx:
        ctx.set_data([data], [64])
        d_o, d_c, d_l = dev(O), dev(V), dev(L)
        g = torch.zeros((cap, N), dtype=torch.float32, device="cuda")
        lp = torch.zeros(N, dtype=torch.float64, device="cuda")
        ll = torch.zeros(N, dtype=torch.float64, device="cuda")
        say(f"grad_logpost: {N} tapes, cap {cap}, n {len(data[0])}")
        ctx.grad_logpost_dev(0, d_o.data_ptr(), d_c.data_ptr(), d_l.data_ptr(), N, cap, g.data_ptr(), lp.data_ptr())
        ctx.eval_loglik_dev(0, d_o.data_ptr(), d_c.data_ptr(), d_l.data_ptr(), N, cap, gpr.SCHED_WAVE, ll.data_ptr())
        torch.cuda.synchronize()
        return g.cpu().numpy(), lp.cpu().numpy(), ll.cpu().numpy()


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def test_rational_gradient_matches_fp64_central_differences(oracle, gpr):
    """The differentiation workload: (b1+b2x+b3x^2+b4x^3)/(1+b5x+b6x^2+b7x^3), n = 37."""
    from gpuropolis_amd.workloads import dataset_D, rational_tape
    (x, y, sd), b = dataset_D(0)
    rng = np.random.default_rng(1)
    tapes = [rational_tape(b * (1 + 0.1 * rng.normal(size=7))) for _ in range(48)]
    O, V, L = pack(tapes, 32)
    g, lp, ll = device_grad(gpr, (x, y, sd), O, V, L, 32)
    worst = 0.0
    for i, (o, v) in enumerate(tapes):
        fd, fd2 = oracle.central_diff_grad(o, v, x, y, sd)
        assert rel_err(fd, fd2) < 1e-7                      # the differences have converged
        e = rel_err(g[:len(o), i].astype(np.float64), fd)
        worst = max(worst, e)
        assert e < 1e-4, (i, e)
        pr = oracle.tape_prior(o, v, 10)
        assert lp[i] == pr + ll[i]
    say("rational: worst relative error vs FP64 central differences", worst)


def test_random_programs_match_fp64_adjoint(oracle, gpr):
    """Config-grammar programs from the init distribution plus long sums of products
    (over 64 tokens: the two-register tape path), on dataset S and a mixed-sd dataset."""
    from gpuropolis_amd.workloads import dataset_S
    rng = np.random.default_rng(5)
    x2 = rng.uniform(-2, 2, 150).astype(np.float32)
    y2 = (np.exp(0.3 * x2) + 0.1 * rng.normal(size=150)).astype(np.float32)
    sd2 = rng.uniform(0.05, 0.3, 150).astype(np.float32)
    for data, cap in ((dataset_S(0), 64), ((x2, y2, sd2), 64), ((x2, y2, sd2), 128)):
        x, y, sd = data
        cfg = oracle.make_cfg(cap=min(cap, 64), seed=cap + len(x))
        ops, cs, ln = oracle.cfg_init(cfg, 256)
        tapes = [(ops[:ln[c], c].copy(), cs[:ln[c], c].copy()) for c in range(256)]
        if cap > 64:
            for k in (17, 20, 24, 28):
                toks = []
                for j in range(k):
                    toks += ["X", ("CONST", float(rng.normal(0, 0.5))), "MUL"] + (["ADD"] if j else [])
                tapes.append(oracle.tape_arrays(toks))
        O, V, L = pack(tapes, cap)
        g, lp, ll = device_grad(gpr, data, O, V, L, cap)
        n_ok = n_fin = 0
        for i, (o, v) in enumerate(tapes):
            pr = oracle.tape_prior(o, v, 10)
            assert lp[i] == pr + ll[i] or (np.isneginf(lp[i]) and np.isneginf(ll[i]))
            rc, g64, _ = oracle.tape_grad(o, v, x, y, sd)
            if rc != 0 or not np.isfinite(ll[i]) or not np.any(g64):
                continue
            n_fin += 1
            e = rel_err(g[:len(o), i].astype(np.float64), g64)
            n_ok += e < 1e-4
            if len(o) > 64:
                assert e < 1e-4, (i, len(o), e)
        say(f"n={len(x)} cap={cap}: {n_ok}/{n_fin} programs within 1e-4 of the FP64 adjoint")
        assert n_fin > 50 and n_ok >= 0.97 * n_fin


def test_autograd_wrapper_gradcheck(oracle, gpr):
    import torch
    from gpuropolis_amd.autograd import log_posterior
    from gpuropolis_amd.workloads import dataset_D, rational_tape
    (x, y, sd), b = dataset_D(0)
    rng = np.random.default_rng(2)
    tapes = [rational_tape(b * (1 + 0.05 * rng.normal(size=7))) for _ in range(4)]
    O, V, L = pack(tapes, 32)
    with gpr.Context(cap=32) as ctx:
        ctx.set_data([(x, y, sd)], [64])
        ops, ln = dev(O), dev(L)
        cs = dev(V.astype(np.float64)).requires_grad_(True)
        lp = log_posterior(ctx, 0, ops, cs, ln)
        lp.sum().backward()
        g_dev, _, _ = device_grad(gpr, (x, y, sd), O, V, L, 32)
        assert np.array_equal(cs.grad.cpu().numpy(), g_dev.astype(np.float64))
        # the constants being varied: only constant tokens (others have zero gradient)
        mask = torch.from_numpy(np.isin(O, [oracle.OP["CONST"]])).to("cuda")

        def f(c):
            return log_posterior(ctx, 0, ops, torch.where(mask, c, cs.detach()), ln)

        say("torch.autograd.gradcheck on the rational workload")
        assert torch.autograd.gradcheck(f, (cs.detach().clone().requires_grad_(True),), eps=1e-3, atol=5.0,
                                        rtol=2e-3, nondet_tol=0.0)
