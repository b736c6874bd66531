"""torch.autograd front end of the batched 3x2pt projection.

``Projection3x2pt.apply(ctx, lnP, src_dz, lens_dz, b1, ia_A, shear_m)`` maps a
batch of B draws (uploaded beforehand with ``Cosmolike.batch_upload``; the
tensors below override their differentiable inputs) to the B real-space 3x2pt
data vectors, FP64, on the GPU.  The backward pass runs the library's
hand-written adjoint kernels (``clh_batch_vjp``), which recompute the Limber
integrands; ``checkpoint=True`` has the forward store them for the backward
instead (the same gradients, no faster at batch 512).

Shapes (float64, cuda, contiguous): lnP [B, nk*nz] (k-major), src_dz [B, ns],
lens_dz [B, nl], b1 [B, nl], ia_A [B], shear_m [B, ns].  An input given as
None takes the value uploaded with the batch (not that of an earlier call).
"""
from __future__ import annotations

import torch

from .interface import Cosmolike


def _ptr(t):
    return 0 if t is None else t.data_ptr()


class Projection3x2pt(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lib_ctx: Cosmolike, lnP, src_dz, lens_dz, b1, ia_A, shear_m, ord=(0, 1, 2),
                checkpoint=False):
        given = [t.contiguous() if t is not None else None for t in (lnP, src_dz, lens_dz, b1, ia_A, shear_m)]
        for t in given:
            if t is not None and (t.dtype != torch.float64 or not t.is_cuda):
                raise TypeError("inputs must be float64 CUDA tensors")
        dev = next((t.device for t in given if t is not None), torch.device("cuda", torch.cuda.current_device()))
        # every input is passed explicitly: a None takes the uploaded value,
        # never what an earlier call left on the device
        base = lib_ctx.uploaded_params(dev)
        ins = [t if t is not None else b for t, b in zip(given, base)]
        B = lib_ctx.batch_size
        n = lib_ctx.datavector_size()
        out = torch.empty(B, n, dtype=torch.float64, device=dev)
        lib_ctx.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        lib_ctx.batch_set_params(*[_ptr(t) for t in ins])
        # the backward recomputes the Limber integrands (chi-slices are not
        # stored); checkpoint=True keeps them from this forward instead
        lib_ctx.set_vjp_checkpoint(bool(checkpoint) and any(ctx.needs_input_grad[1:7]))
        lib_ctx.batch_forward(out.data_ptr(), ord)
        ctx.lib_ctx = lib_ctx
        ctx.ord = ord
        ctx.generation = lib_ctx.param_generation
        ctx.save_for_backward(*[t if t is not None else torch.empty(0) for t in ins])
        ctx.present = [t is not None for t in ins]
        return out

    @staticmethod
    def backward(ctx, g):
        c = ctx.lib_ctx
        saved = ctx.saved_tensors
        ins = [s if p else None for s, p in zip(saved, ctx.present)]
        g = g.contiguous()
        dev = g.device
        c.set_stream(torch.cuda.current_stream(dev).cuda_stream)
        # restore the inputs if another call changed the batch since this
        # forward; otherwise the library reuses the forward's device state
        if c.param_generation != ctx.generation:
            c.batch_set_params(*[_ptr(t) for t in ins])
        grads = [torch.empty_like(t) if (t is not None and need) else None
                 for t, need in zip(ins, ctx.needs_input_grad[1:7])]
        c.batch_vjp(g.data_ptr(), *[_ptr(t) for t in grads], ord=ctx.ord)
        return (None, *grads, None, None)


def projection_3x2pt(lib_ctx, lnP, src_dz, lens_dz, b1, ia_A, shear_m, ord=(0, 1, 2), checkpoint=False):
    return Projection3x2pt.apply(lib_ctx, lnP, src_dz, lens_dz, b1, ia_A, shear_m, ord, checkpoint)
