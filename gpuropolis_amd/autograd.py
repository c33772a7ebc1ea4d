"""The log posterior of postfix tapes as a differentiable torch function of their
constants: forward runs the device evaluator, backward returns the reverse-mode adjoint
(gpr_grad_logpost).  A thin wrapper for gradient checks; the sampler itself calls the
adjoint inside the MH kernel.

    lp = log_posterior(ctx, dataset, ops, cs, length)   # ops uint8 [cap, N], cs [cap, N]
    lp.sum().backward()                                 # cs.grad = d lp / d cs

The constants cross the C ABI as float32 (the kernels' arithmetic type); a float64 cs
is rounded on the way in and its gradient returned as float64.

The arguments are checked before anything reaches the device: ops, cs and length are
tensors on the context's GPU, ops is uint8 of the shape of cs, [cap, N] with
1 <= cap <= 128, length holds N lengths in [1, cap], and every tape is well-formed
postfix whose evaluation stack stays within 16 entries.
"""
import torch

from . import lib as G


def _check(gctx, ops, cs, length):
    dev = torch.device("cuda", gctx.cfg.device)
    for name, t in (("ops", ops), ("cs", cs), ("length", length)):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.device != dev:
            raise G.GprError(f"{name} must be a tensor on the context's device {dev}")
    if ops.dtype != torch.uint8:
        raise G.GprError("ops must be uint8 op codes")
    if ops.dim() != 2 or cs.shape != ops.shape:
        raise G.GprError(f"ops and cs must both be [cap, N]; got {tuple(ops.shape)} and {tuple(cs.shape)}")
    cap, n = ops.shape
    if not 1 <= cap <= G.MAX_CAP or n < 1:
        raise G.GprError(f"cap must be in [1, {G.MAX_CAP}] and N >= 1; got cap {cap}, N {n}")
    if length.dim() != 1 or length.shape[0] != n:
        raise G.GprError(f"length must hold N = {n} entries")
    check_tapes(ops.cpu(), length.cpu())


def check_tapes(ops, length):
    """Host check of tapes ops [cap, N] (uint8) with lengths [N]: lengths in [1, cap],
    known op codes, well-formed postfix, evaluation stack within MAX_DEPTH entries."""
    cap, n = ops.shape
    ln = length.to(torch.int64)
    if int(ln.min()) < 1 or int(ln.max()) > cap:
        raise G.GprError("every length must lie in [1, cap]")
    o = ops.to(torch.int64)
    if int(o.max()) >= len(G.OPS):
        raise G.GprError("bad op code")
    arity = torch.tensor([_arity(k) for k in range(len(G.OPS))], dtype=torch.int64)
    step = 1 - arity[o]                                           # stack change per token, [cap, N]
    live = torch.arange(cap)[:, None] < ln[None, :]
    depth = torch.cumsum(torch.where(live, step, torch.zeros_like(step)), dim=0)
    need = torch.where(live, torch.cat([torch.zeros(1, n, dtype=torch.int64), depth[:-1]]) - arity[o],
                       torch.zeros_like(step))
    if int(need.min()) < 0 or not bool(torch.all(depth[-1] == 1)):
        raise G.GprError("every tape must be well-formed postfix")
    if int(depth.max()) > G.MAX_DEPTH:
        raise G.GprError(f"tape evaluation stacks must stay within {G.MAX_DEPTH} entries")


def _arity(op):
    names = G.OPS
    if names[op] in ("X", "CONST", "ONE"):
        return 0
    if names[op] in ("ADD", "SUB", "MUL", "DIV", "POW", "RSUB", "RDIV", "RPOW"):
        return 2
    return 1


class LogPosterior(torch.autograd.Function):
    @staticmethod
    def forward(fctx, cs, gctx, dataset, ops, length):
        _check(gctx, ops, cs, length)
        cap, n = ops.shape
        ops = ops.contiguous()
        length = length.to(torch.int32).contiguous()
        cs32 = cs.detach().to(torch.float32).contiguous()
        grad = torch.empty_like(cs32)
        lp = torch.empty(n, dtype=torch.float64, device=cs.device)
        # the library runs on its own stream: inputs made on torch's stream must be ready
        torch.cuda.current_stream(cs.device).synchronize()
        gctx.grad_logpost_dev(dataset, ops.data_ptr(), cs32.data_ptr(), length.data_ptr(), n, cap,
                              grad.data_ptr(), lp.data_ptr())
        fctx.save_for_backward(grad)
        fctx.in_dtype = cs.dtype
        return lp

    @staticmethod
    def backward(fctx, g_lp):
        (grad,) = fctx.saved_tensors
        g = g_lp.to(torch.float64)[None, :] * grad.to(torch.float64)
        return g.to(fctx.in_dtype), None, None, None, None


def log_posterior(gctx, dataset, ops, cs, length):
    return LogPosterior.apply(cs, gctx, dataset, ops, length)
