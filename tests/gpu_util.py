"""Helpers shared by the GPU tests: device buffers via torch, oracle comparisons."""
import numpy as np


def dev(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    torch.cuda.synchronize()
    return t


def ll_close(gpu, ref, n, rel=1e-5):
    """|gpu - ref| <= rel * max(|ref|, n); -inf must match -inf."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    inv_g = ~np.isfinite(gpu)
    inv_r = ~np.isfinite(ref)
    ok_valid = inv_g == inv_r
    tol = rel * np.maximum(np.abs(np.where(inv_r, 0, ref)), n)
    diff = np.abs(np.where(inv_r | inv_g, 0, gpu - ref))
    return ok_valid & (diff <= tol), diff, tol


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
