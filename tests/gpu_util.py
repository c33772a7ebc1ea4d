"""Helpers shared by the GPU tests: device buffers via torch, oracle comparisons."""
import numpy as np


def dev(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    torch.cuda.synchronize()
    return t


def ll_close(gpu, ref, n, rel=1e-5):
    """|gpu - ref| <= rel * max(|ref|, n); -inf must match -inf.  diff is +inf where one
    side is finite and the other is not, 0 where both are not."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    inv_g = ~np.isfinite(gpu)
    inv_r = ~np.isfinite(ref)
    ok_valid = inv_g == inv_r
    tol = rel * np.maximum(np.abs(np.where(inv_r, 0, ref)), n)
    diff = np.where(inv_r & inv_g, 0.0, np.where(inv_r | inv_g, np.inf, np.abs(gpu - np.where(inv_r | inv_g, 0, ref))))
    return ok_valid & (diff <= tol), diff, tol


def lconst(sd):
    """The likelihood's normaliser -sum_i 0.5 (logf(2 pi) + 2 logf(sd_i)) as the library
    forms it (float logs, FP64 sum)."""
    sd = np.asarray(sd, dtype=np.float32).astype(np.float64)
    logf = np.log(sd).astype(np.float32).astype(np.float64)     # correctly rounded float log
    return float(-np.sum(0.5 * (np.float64(np.float32(np.log(2 * np.pi))) + 2.0 * logf)))


def ll_agree(gpu, ref, n, kap, sd=None, rel=1e-5):
    """Per-chain verdict of GPU log-likelihoods against the oracle's.

    * Finite on both sides or on neither (an invalid program must be -inf on both).
    * Both finite: |gpu - ref| <= rel * max(|ref|, n), or that plus 4 kappa for a program
      whose LL moves by kappa under one-ulp changes of its inputs (finite kappa only).
    * Strict, for well-conditioned programs (kappa <= 1e-7 |ref|) whose LL is not the
      small difference of its two parts (|lconst| + |lconst - ref| <= 50 |ref|, which
      needs sd): |gpu - ref| <= rel * |ref|.
    Returns (ok, strict_mask)."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    kap = np.asarray(kap, dtype=np.float64)
    fin_g, fin_r = np.isfinite(gpu), np.isfinite(ref)
    both = fin_g & fin_r
    ok, diff, tol = ll_close(gpu, ref, n, rel)
    loose = both & np.isfinite(kap) & (diff <= tol + 4 * np.where(np.isfinite(kap), kap, 0.0))
    ok = (fin_g == fin_r) & (ok | loose)
    strict = np.zeros_like(both)
    if sd is not None:
        lc = lconst(sd)
        r = np.where(both, ref, 0.0)
        strict = both & np.isfinite(kap) & (kap <= 1e-7 * np.abs(r)) & \
            (abs(lc) + np.abs(lc - r) <= 50.0 * np.abs(r))
        ok = ok & (~strict | (diff <= rel * np.abs(r)))
    return ok, strict


def say(*a):
    print(*a, flush=True)


def kappa(oracle, ops, cs, x, y, sd):
    """Sensitivity of a program's LL to FP32 rounding: the larger of the FP32-node vs
    FP64-node gap and the LL change under a one-ulp move of every x (FP64 nodes)."""
    r0 = oracle.tape_loglik(ops, cs, x, y, sd, 0)
    r1 = oracle.tape_loglik(ops, cs, x, y, sd, 1)
    xp = np.nextafter(np.asarray(x, np.float32), np.float32(np.inf))
    r1p = oracle.tape_loglik(ops, cs, xp, y, sd, 1)
    vals = [abs(r0 - r1), abs(r1p - r1)]
    if not all(np.isfinite(v) for v in vals):
        return np.inf
    return max(vals)


def kappas(oracle, ops, cs, ln, chains, x, y, sd):
    """kappa of each chain's program in the [row][chain] arrays."""
    return np.array([kappa(oracle, ops[:ln[i], i], cs[:ln[i], i], x, y, sd) for i in chains])


def check_paths(same, margin, tag="", rel=1e-5):
    """Every chain off the oracle's path must be explained: somewhere on the oracle's
    trajectory it met an accept decision within rounding of its threshold, |a - log u|
    below `rel` of the likelihood scale (the oracle's per-chain accept margin), where a
    last-bit difference of a device likelihood (within the tests' relative tolerance) can
    flip the decision.  Such chains stay rare: at most 2 or 2% of the chains."""
    same = np.asarray(same, dtype=bool)
    margin = np.asarray(margin, dtype=np.float64)
    off = np.flatnonzero(~same)
    unexplained = off[~(margin[off] < rel)]
    say(f"{tag} chains on the oracle path: {int(same.sum())} of {len(same)}; off-path {list(off)} with accept "
        f"margins {[float(f'{m:.3g}') for m in margin[off]]}")
    assert len(unexplained) == 0, (tag, list(unexplained), margin[unexplained])
    assert len(off) <= max(2, len(same) // 50), (tag, len(off))
