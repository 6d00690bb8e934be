"""torch.autograd wrappers of the GPU forward and its adjoints:

* MomentTensorSeismograms: with respect to the moment tensor (the source
  vectors of src/module_solver.f90:145-215, 566-627, 1747-1847 are linear
  in it);
* KnotSeismograms: with respect to each deck knot's rho, vp, vs and Qmu,
  through the discrete adjoint of every per-pair solve (ysx_model_vjp) and
  the deck -> table transform (ysx_model_vjp_deck), all on the device.

Thin on purpose: the forward and the backward are the library's own entry
points; torch only carries the (host) tensors, so finite-difference checks
and optimisers can use the gradients."""
import dataclasses

import numpy as np
import torch

from .api import deck_with_knots, knot_gradient


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


class KnotSeismograms(torch.autograd.Function):
    """seis = f(knots): knots an (nknot, 4) float64 tensor of each knot's
    rho, vp, vs and Qmu (SI units) replacing those columns of `deck` (vp and
    vs in both polarisations), seis the (nt, 3, nr) seismograms of run `p`
    on the context `g` (an api.YSpec, whose model and run are replaced).
    The backward takes seis_bar through the seismogram, spectra, per-pair
    solve and deck -> table adjoints; it needs the records of this forward,
    so no other run may be set on `g` in between."""

    @staticmethod
    def forward(ctx, knots, g, deck, p):
        d = deck_with_knots(deck, knots.detach().cpu().numpy().astype(np.float64))
        g.set_model(d)
        g.set_run(p)
        seis = g.forward()
        ctx.g, ctx.run, ctx.deck = g, g.p, d
        return torch.from_numpy(np.ascontiguousarray(seis))

    @staticmethod
    def backward(ctx, seis_bar):
        g = ctx.g
        if g.p is not ctx.run:
            raise RuntimeError("another run was set on this context after the forward")
        g.seismograms_vjp(seis_bar.detach().cpu().numpy())
        g.spectra_vjp(np.zeros(0, np.int32), np.zeros(0, np.int32))
        grad = knot_gradient(g.model_vjp_deck(g.model_vjp()))
        return torch.from_numpy(grad), None, None, None


def knot_seismograms(knots, g, deck, p):
    return KnotSeismograms.apply(knots, g, deck, p)
