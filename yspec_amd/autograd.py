"""torch.autograd wrapper of the GPU forward and its adjoint with respect to
the moment tensor (the source vectors of src/module_solver.f90:145-215,
566-627, 1747-1847 are linear in it).  Thin on purpose: the forward and the
backward are the library's own entry points; torch only carries the tensors,
so finite-difference checks and optimisers can use the gradient."""
import dataclasses

import numpy as np
import torch


class MomentTensorSeismograms(torch.autograd.Function):
    """seis = f(mm): mm a (6,) float64 tensor (Mrr, Mrt, Mrp, Mtt, Mtp, Mpp in
    N m), seis the (nt, 3, nr) seismograms of run `p` (its own moment tensor
    replaced by mm) on the context `g` (an api.YSpec).  The backward uses the
    per-pair records of this forward, so it checks that no other run was set
    on `g` in between."""

    @staticmethod
    def forward(ctx, mm, g, p):
        m = mm.detach().cpu().numpy().astype(np.float64)
        g.set_run(dataclasses.replace(p, mm=m.copy()))
        seis = g.forward()
        ctx.g, ctx.run = g, g.p
        return torch.from_numpy(np.ascontiguousarray(seis))

    @staticmethod
    def backward(ctx, seis_bar):
        if ctx.g.p is not ctx.run:
            raise RuntimeError("another run was set on this context after the forward")
        grad = ctx.g.misfit_gradient_mt(seis_bar.detach().cpu().numpy())
        return torch.from_numpy(grad), None, None


def seismograms(mm, g, p):
    return MomentTensorSeismograms.apply(mm, g, p)
