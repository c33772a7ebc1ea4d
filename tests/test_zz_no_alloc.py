"""The no-allocation evaluator (gpr_eval_loglik_noalloc_async) against gpr_eval_loglik:
the same tapes on the same data give the same log-likelihoods, on both schedules, over
two calls in a row, the second with fewer points and fewer tapes (the static scratch
keeps the first call's contents, which the second must not see)."""
import numpy as np
import pytest

from gpu_util import dev, say

pytestmark = pytest.mark.gpu


def tapes(oracle, nprog, cap, seed):
    cfg = oracle.make_cfg(cap=cap, seed=seed)
    ops, cs, ln = oracle.cfg_init(cfg, nprog)
    return ops, cs, ln


@pytest.mark.parametrize("schedule", ["wave", "lane"])
def test_noalloc_eval_matches_context_eval(oracle, gpr, schedule):
    import torch
    rng = np.random.default_rng(5)
    sched = gpr.SCHED_WAVE if schedule == "wave" else gpr.SCHED_LANE
    stream = torch.cuda.current_stream()
    # call 1: 700 points with per-point sd (weights in scratch); call 2: 90 points with one sd
    cases = [(700, 300, 64, rng.uniform(0.1, 1.0, 700)), (90, 40, 32, np.full(90, 0.3))]
    for k, (n, nprog, cap, sdv) in enumerate(cases):
        x = rng.uniform(-3, 3, n).astype(np.float32)
        y = (np.sin(x) + 0.3 * x).astype(np.float32)
        sd = sdv.astype(np.float32)
        O, V, L = tapes(oracle, nprog, cap, 10 + k)
        dO, dV, dL = dev(O), dev(V), dev(L)
        with gpr.Context(cap=cap) as ctx:
            ctx.set_data([(x, y, sd)], [1])
            ref = torch.zeros(nprog, dtype=torch.float64, device="cuda")
            ctx.eval_loglik_dev(0, dO.data_ptr(), dV.data_ptr(), dL.data_ptr(), nprog, cap, sched, ref.data_ptr())
            ref = ref.cpu().numpy()
        dx, dy, dsd = dev(x), dev(y), dev(sd)
        out = torch.full((nprog,), 123.0, dtype=torch.float64, device="cuda")
        say(f"no-alloc eval call {k + 1}: n={n} tapes={nprog} cap={cap} schedule={schedule}")
        gpr.eval_loglik_noalloc_async(0, dx.data_ptr(), dy.data_ptr(), dsd.data_ptr(), n, dO.data_ptr(),
                                      dV.data_ptr(), dL.data_ptr(), nprog, cap, sched, out.data_ptr(),
                                      stream.cuda_stream)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        fin = np.isfinite(ref)
        assert np.array_equal(np.isfinite(got), fin)
        assert fin.sum() > nprog // 4
        rel = np.abs(got[fin] - ref[fin]) / np.maximum(np.abs(ref[fin]), n)
        say(f"  max relative difference {rel.max():.2e} over {int(fin.sum())} finite likelihoods")
        # the normaliser is the same FP64 sum in another order, of logs that are correctly
        # rounded on the device and may differ by an ulp in the host libm's logf
        assert rel.max() <= 1e-9
        # the oracle as well, to the tests' 1e-5 bar
        lo = np.array([oracle.tape_loglik(O[:L[i], i], V[:L[i], i], x, y, sd, 2) for i in range(nprog)])
        ok = ~np.isfinite(lo) | (np.abs(got - lo) <= 1e-5 * np.maximum(np.abs(lo), n))
        assert ok[fin].mean() > 0.97


def test_noalloc_eval_rejects_too_many_points(gpr):
    import torch
    n = gpr.NOALLOC_MAX_POINTS + 1
    z = torch.zeros(8, dtype=torch.float32, device="cuda")
    with pytest.raises(gpr.GprError, match="static scratch"):
        gpr.eval_loglik_noalloc_async(0, z.data_ptr(), z.data_ptr(), z.data_ptr(), n, z.data_ptr(), z.data_ptr(),
                                      z.data_ptr(), 1, 8, gpr.SCHED_WAVE, z.data_ptr(), None)
