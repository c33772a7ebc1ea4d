"""The log posterior of postfix tapes as a differentiable torch function of their
constants: forward runs the device evaluator, backward returns the reverse-mode adjoint
(gpr_grad_logpost).  A thin wrapper for gradient checks; the sampler itself calls the
adjoint inside the MH kernel.

    lp = log_posterior(ctx, dataset, ops, cs, length)   # ops uint8 [cap, N], cs [cap, N]
    lp.sum().backward()                                 # cs.grad = d lp / d cs

The constants cross the C ABI as float32 (the kernels' arithmetic type); a float64 cs
is rounded on the way in and its gradient returned as float64.
"""
import torch

from . import lib as G


class LogPosterior(torch.autograd.Function):
    @staticmethod
    def forward(fctx, cs, gctx, dataset, ops, length):
        if not cs.is_cuda:
            raise G.GprError("cs must be a device tensor")
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
