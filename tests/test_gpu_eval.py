"""Evaluator + Gaussian log-likelihood on the GPU (both schedules) against the FP64
oracle: per-program LL within 1e-5 * max(|LL|, n), invalid programs -inf on both."""
import numpy as np
import pytest

from gpu_util import dev, ll_close, say

pytestmark = pytest.mark.gpu

NS = [1, 7, 32, 63, 64, 65, 129, 250, 257, 1024]


def programs(oracle, nprog=384, cap=64, seed=11):
    """Tapes in SoA [cap][N]: config-grammar programs from the init distribution,
    decoded reference heaps over all 27 ops, and hand-written edge cases."""
    tapes = []
    cfg = oracle.make_cfg(cap=cap, seed=seed)
    ops, cs, ln = oracle.cfg_init(cfg, nprog // 2)
    for c in range(ops.shape[1]):
        tapes.append((ops[:ln[c], c].copy(), cs[:ln[c], c].copy()))
    rng = np.random.default_rng(seed)
    for _ in range(nprog // 2):
        hop = rng.integers(0, 27, 31).astype(np.uint8)
        hc = (10.0 ** rng.integers(-2, 2, 31) * rng.normal(0, 1, 31)).astype(np.float32)
        o, v, _ = oracle.ref_decode(hop, hc)
        tapes.append((o, v))
    hand = [
        ["X"], [("CONST", 2.5)], ["ONE"], ["X", "X", "MUL", ("CONST", 1.5), "ADD"],
        ["X", "EXP", "LOG"], ["X", "NEG", "EXP"], ["X", ("CONST", 0.0), "DIV"], ["X", "X", "SUB", "LOG"],
        ["X", ("CONST", 2.0), "POW"], ["X", "ABS", ("CONST", 0.5), "POW"], ["X", ("CADD", 3.0), "SQRT"],
        ["X", ("CMUL", -2.0), "SIN"], ["X", "COS", "ATAN"], ["X", ("CPOW", 3.0)], ["X", ("CRPOW", 1.7)],
        ["X", ("CONST", 4.0), "RSUB", "INV"], ["X", "X", "X", "X", "X", "ADD", "ADD", "ADD", "ADD"],
        ["X", ("CONST", 100.0), "MUL", "EXP"], ["X", "TAN", "ASIN", "ACOS"], ["X", "X", "RDIV", "X", "RPOW"],
    ]
    for h in hand:
        tapes.append(oracle.tape_arrays(h))
    # deepest allowed stack: 16 leaves then 15 binary ops
    deep = ["X"] * 16 + ["ADD", "MUL", "SUB", "ADD", "DIV", "ADD", "MUL", "ADD", "SUB", "ADD", "ADD", "MUL", "ADD",
                         "SUB", "ADD"]
    tapes.append(oracle.tape_arrays(deep))
    N = len(tapes)
    cap_all = max(cap, max(len(t[0]) for t in tapes))
    O = np.zeros((cap_all, N), dtype=np.uint8)
    V = np.zeros((cap_all, N), dtype=np.float32)
    L = np.zeros(N, dtype=np.int32)
    for i, (o, v) in enumerate(tapes):
        O[:len(o), i] = o
        V[:len(o), i] = v
        L[i] = len(o)
    return O, V, L


def dataset(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-3, 3, n).astype(np.float32)
    y = (np.sin(x) * 2 + 0.5 * x + rng.normal(0, 0.3, n)).astype(np.float32)
    sd = rng.uniform(0.1, 1.0, n).astype(np.float32)
    return x, y, sd


@pytest.mark.parametrize("schedule", ["wave", "lane"])
def test_eval_loglik_parity(oracle, gpr, schedule):
    import torch
    O, V, L = programs(oracle)
    cap, N = O.shape
    dsets = [dataset(n, 100 + n) for n in NS]
    with gpr.Context(cap=min(cap, 128)) as ctx:
        ctx.set_data(dsets, [1] * len(dsets))
        dO, dV, dL = dev(O), dev(V), dev(L)
        out = torch.zeros(N, dtype=torch.float64, device="cuda")
        sched = gpr.SCHED_WAVE if schedule == "wave" else gpr.SCHED_LANE
        for di, (x, y, sd) in enumerate(dsets):
            say(f"eval parity schedule={schedule} n={len(x)} programs={N}")
            ctx.eval_loglik_dev(di, dO.data_ptr(), dV.data_ptr(), dL.data_ptr(), N, cap, sched, out.data_ptr())
            got = out.cpu().numpy()
            ref = np.array([oracle.tape_loglik(O[:L[i], i], V[:L[i], i], x, y, sd) for i in range(N)])
            r64 = np.array([oracle.tape_loglik(O[:L[i], i], V[:L[i], i], x, y, sd, mode=1) for i in range(N)])
            # kappa: how far FP32 node rounding alone moves this program's LL (FP32-node vs
            # FP64-node oracle).  Well-conditioned programs (kappa << tol) must match to tol;
            # others to tol + 4 kappa, since a 1-ulp libm difference is amplified the same way.
            # and how far a one-ulp move of every x moves it (FP64 nodes): catches programs
            # such as tan(1/log|x|) whose FP32-vs-FP64 gap happens to be small
            xp = np.nextafter(x, np.float32(np.inf))
            r64p = np.array([oracle.tape_loglik(O[:L[i], i], V[:L[i], i], xp, y, sd, mode=1) for i in range(N)])
            both = np.isfinite(ref) & np.isfinite(r64) & np.isfinite(r64p)
            kappa = np.where(both, np.maximum(np.abs(ref - np.where(both, r64, 0)),
                                              np.abs(np.where(both, r64p - r64, 0))), 0.0)
            ok, diff, tol = ll_close(got, ref, len(x))
            within = ok | (np.isfinite(got) & np.isfinite(ref) & (diff <= tol + 4 * kappa))
            well = both & (kappa <= 0.1 * tol)
            bad = np.flatnonzero(~within)
            assert within.all(), (schedule, len(x), bad[:10], got[bad[:5]], ref[bad[:5]], kappa[bad[:5]])
            assert ok[well].all(), (schedule, len(x), np.flatnonzero(well & ~ok)[:10])
            assert well.sum() > N // 4
