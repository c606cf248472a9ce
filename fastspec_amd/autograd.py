"""PyTorch autograd over libfsgpu: the normalised flux of one synthesis as a differentiable
function of the lines' log gf, the temperature of each depth and the EOS number densities.

    F/Fc = NormalisedFlux.apply(loggf, temp, lnfpf, synth, dlines, nlines_lim)

Forward runs the same stages as the reference's synthesis for one list (LineopCalc,
RadTranCalc, the flux sums of IntegConvol, src/fastspec.f90:241-449) on the GPU; backward
runs the library's adjoint kernels (fsg_flux_vjp, fsg_lineop_vjp).  ``temp`` and ``lnfpf``
(the log of FractionPF relative to the context's values, all zeros) are the points at
which dL/dT and dL/d ln FractionPF are taken: the host's EOS and continuum, which map T and
abundances onto the library's inputs, stay outside, so a new atmosphere needs a new
context.  The element-abundance gradient of a trace species is the sum of its slots' rows:
dL/d log10 A = ln 10 sum_depth dL/d ln FractionPF.

    F/Fc (all atmospheres) = NormalisedFluxBatch.apply(loggf, temp, lnfpf, batch, blines, nlim)

is the same over a batch context (many atmospheres in one launch sequence, one shared
list): temp [nb][ntau], lnfpf [nb][ntau][pf_ld], and the gradient in loggf summed over the
atmospheres that use each line.
"""
from __future__ import annotations

import numpy as np
import torch

from .synth import BatchLines, DeviceLines, FastSpecBatch, FastSpecGPU

# Each forward leaves its state in the context (LineOpac, the formal solution, the list's gf
# and the pass-1 buffers) and backward reads that state.  A serial number per context names
# the forward whose state the context holds; backward of an earlier forward (two forwards
# before one backward, as in a loss over two parameter vectors or gradcheck's perturbed
# evaluations) first runs that forward again from its saved gf.
_SERIAL = "_autograd_serial"


def _next_serial(synth) -> int:
    n = getattr(synth, _SERIAL, 0) + 1
    setattr(synth, _SERIAL, n)
    return n


def _gf32(loggf: torch.Tensor) -> np.ndarray:
    """gf = 10**log gf in REAL(4), as the host's reader converts it."""
    return torch.pow(torch.tensor(10.0, dtype=torch.float32), loggf.detach().to("cpu").float()).numpy()


def _forward_scalar(synth: FastSpecGPU, dlines: DeviceLines, gf: np.ndarray, nlines_lim):
    synth.set_gf(dlines, gf)
    synth.reset_lineop()
    synth.lineop_dev(dlines, nlines_lim)
    synth.RadTranCalc(intensities=False)
    return synth.flux()


def _forward_batch(batch: FastSpecBatch, blines: BatchLines, gf: np.ndarray, nlines_lim):
    batch.set_gf(blines, gf)
    batch.reset_lineop()
    batch.lineop_dev(blines, nlines_lim)
    batch.RadTranCalc(intensities=False)
    return batch.flux()


class NormalisedFlux(torch.autograd.Function):
    @staticmethod
    def forward(ctx, loggf, temp, lnfpf, synth: FastSpecGPU, dlines: DeviceLines, nlines_lim=None):
        s = synth.setup
        t_now = torch.as_tensor(s.temp, dtype=torch.float64)
        if not torch.equal(temp.detach().to("cpu", torch.float64), t_now):
            raise ValueError("NormalisedFlux: temp differs from the context's atmosphere; build a new context")
        if bool(torch.any(lnfpf.detach() != 0)):
            raise ValueError("NormalisedFlux: lnfpf must be zero (FractionPF is fixed by the context)")
        gf = _gf32(loggf)
        sfl, sfc = _forward_scalar(synth, dlines, gf, nlines_lim)
        ctx.synth, ctx.dlines, ctx.nlim = synth, dlines, nlines_lim
        ctx.sfl, ctx.sfc, ctx.gf = sfl, sfc, gf
        ctx.serial = _next_serial(synth)
        ctx.meta = [(x.dtype, x.device) for x in (loggf, temp, lnfpf)]
        return torch.as_tensor(sfl / sfc, dtype=torch.float64, device=loggf.device)

    @staticmethod
    def backward(ctx, g):
        g = g.detach().to("cpu", torch.float64).numpy()
        sfl, sfc = ctx.sfl, ctx.sfc
        if getattr(ctx.synth, _SERIAL, 0) != ctx.serial:  # another forward ran since
            _forward_scalar(ctx.synth, ctx.dlines, ctx.gf, ctx.nlim)
            setattr(ctx.synth, _SERIAL, ctx.serial)
        # d(F/Fc): dL/dSfL = g / Fc, dL/dSfC = -g F / Fc^2
        _, _, gt_src = ctx.synth.flux_vjp(g / sfc, -g * sfl / (sfc * sfc))
        g_loggf, g_fpf, gt_line = ctx.synth.lineop_vjp(ctx.dlines, ctx.nlim)
        out = [torch.as_tensor(a).to(dtype=dt, device=dev)
               for a, (dt, dev) in zip((g_loggf, gt_src + gt_line, g_fpf), ctx.meta)]
        return (*out, None, None, None)


def normalised_flux(synth: FastSpecGPU, dlines: DeviceLines, loggf: torch.Tensor, temp=None, lnfpf=None,
                    nlines_lim=None) -> torch.Tensor:
    """F/Fc [nwl] of one synthesis; differentiable in loggf (and temp / lnfpf when given
    with requires_grad)."""
    s = synth.setup
    if temp is None:
        temp = torch.as_tensor(np.array(s.temp), dtype=torch.float64)
    if lnfpf is None:
        lnfpf = torch.zeros(s.fraction_pf.shape, dtype=torch.float64)
    return NormalisedFlux.apply(loggf, temp, lnfpf, synth, dlines, nlines_lim)


class NormalisedFluxBatch(torch.autograd.Function):
    """F/Fc of every atmosphere of a batch context (fsg_init_batch), concatenated over the
    atmospheres' own grids, as a function of the shared list's log gf, the temperatures
    [nb][ntau] and ln FractionPF [nb][ntau][pf_ld] (zeros: the contexts' values)."""

    @staticmethod
    def forward(ctx, loggf, temp, lnfpf, batch: FastSpecBatch, blines: BatchLines, nlines_lim=None):
        t_now = torch.as_tensor(np.stack([su.temp for su in batch.setups]), dtype=torch.float64)
        if not torch.equal(temp.detach().to("cpu", torch.float64), t_now):
            raise ValueError("NormalisedFluxBatch: temp differs from the batch's atmospheres; build a new batch")
        if bool(torch.any(lnfpf.detach() != 0)):
            raise ValueError("NormalisedFluxBatch: lnfpf must be zero (FractionPF is fixed by the batch)")
        gf = _gf32(loggf)
        sfl, sfc = _forward_batch(batch, blines, gf, nlines_lim)
        ctx.batch, ctx.blines, ctx.nlim = batch, blines, nlines_lim
        ctx.sfl, ctx.sfc, ctx.gf = sfl, sfc, gf
        ctx.serial = _next_serial(batch)
        ctx.meta = [(x.dtype, x.device) for x in (loggf, temp, lnfpf)]
        f = np.concatenate([l / c for l, c in zip(sfl, sfc)])
        return torch.as_tensor(f, dtype=torch.float64, device=loggf.device)

    @staticmethod
    def backward(ctx, g):
        g = g.detach().to("cpu", torch.float64).numpy()
        sfl, sfc = ctx.sfl, ctx.sfc
        if getattr(ctx.batch, _SERIAL, 0) != ctx.serial:  # another forward ran since
            _forward_batch(ctx.batch, ctx.blines, ctx.gf, ctx.nlim)
            setattr(ctx.batch, _SERIAL, ctx.serial)
        cut = np.cumsum([len(x) for x in sfl])[:-1]
        gb = np.split(g, cut)
        _, _, gt_src = ctx.batch.flux_vjp([a / c for a, c in zip(gb, sfc)],
                                          [-a * l / (c * c) for a, l, c in zip(gb, sfl, sfc)])
        g_loggf, g_fpf, gt_line = ctx.batch.lineop_vjp(ctx.blines, ctx.nlim)
        g_shared = ctx.blines.to_shared(g_loggf)
        out = [torch.as_tensor(a).to(dtype=dt, device=dev)
               for a, (dt, dev) in zip((g_shared, gt_src + gt_line, g_fpf), ctx.meta)]
        return (*out, None, None, None)


def normalised_flux_batch(batch: FastSpecBatch, blines: BatchLines, loggf: torch.Tensor, temp=None, lnfpf=None,
                          nlines_lim=None) -> torch.Tensor:
    """F/Fc of all atmospheres, concatenated (atmosphere b: batch.setups[b].nwl points);
    differentiable in the shared list's loggf (and temp [nb][ntau] / lnfpf when given with
    requires_grad)."""
    if temp is None:
        temp = torch.as_tensor(np.stack([su.temp for su in batch.setups]), dtype=torch.float64)
    if lnfpf is None:
        lnfpf = torch.zeros((batch.nb,) + batch.setup.fraction_pf.shape, dtype=torch.float64)
    return NormalisedFluxBatch.apply(loggf, temp, lnfpf, batch, blines, nlines_lim)


__all__ = ["NormalisedFlux", "normalised_flux", "NormalisedFluxBatch", "normalised_flux_batch"]
