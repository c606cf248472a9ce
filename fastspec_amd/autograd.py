"""PyTorch autograd over libfsgpu: the normalised flux of one synthesis as a differentiable
function of the lines' log gf, the temperature of each depth and the EOS number densities.
Parameters on a CUDA device stay there: gf goes to the library as a device pointer, F/Fc
comes back in device tensors (fsg_flux_dev) and the gradients through the device-pointer
reverse-mode entry points (the library shares torch's HIP runtime; see _lib).

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


def _gf32(loggf: torch.Tensor):
    """gf = 10**log gf in REAL(4), as the host's reader converts it: a numpy array for host
    tensors, a finished CUDA tensor on the parameters' device for device tensors."""
    ten = torch.tensor(10.0, dtype=torch.float32, device=loggf.device)
    gf = torch.pow(ten, loggf.detach().float()).contiguous()
    if gf.is_cuda:
        torch.cuda.current_stream(gf.device).synchronize()
        return gf
    return gf.numpy()


def _forward_scalar(synth: FastSpecGPU, dlines: DeviceLines, gf, nlines_lim):
    synth.set_gf(dlines, gf)
    synth.reset_lineop()
    synth.lineop_dev(dlines, nlines_lim)
    synth.RadTranCalc(intensities=False)


def _forward_batch(batch: FastSpecBatch, blines: BatchLines, gf, nlines_lim):
    batch.set_gf(blines, gf)
    batch.reset_lineop()
    batch.lineop_dev(blines, nlines_lim)
    batch.RadTranCalc(intensities=False)


def _flux(synth, device):
    """SfL, SfC over the context's grid: CUDA tensors filled on the device (fsg_flux_dev), or
    host arrays."""
    if device.type == "cuda":
        sfl = torch.empty(synth.nwl_total, dtype=torch.float64, device=device)
        sfc = torch.empty_like(sfl)
        synth.flux_device(sfl, sfc)
        return sfl, sfc
    sfl, sfc = np.zeros(synth.nwl_total), np.zeros(synth.nwl_total)
    from . import _lib
    from .abi import ptr
    import ctypes as C
    _lib.check(synth._lib.fsg_flux(synth.ctx, ptr(sfl, C.c_double), ptr(sfc, C.c_double)))
    return torch.from_numpy(sfl), torch.from_numpy(sfc)


def _vjp(synth, dlines, nlim, g, sfl, sfc):
    """dL/d log gf (on g's device), dL/dT and dL/d ln FractionPF (host) from dL/d(F/Fc) = g
    over the context's grid: d(F/Fc) gives dL/dSfL = g / Fc, dL/dSfC = -g F / Fc^2."""
    gsfl = (g / sfc).contiguous()
    gsfc = (-g * sfl / (sfc * sfc)).contiguous()
    if g.is_cuda:
        torch.cuda.current_stream(g.device).synchronize()
        gt_src = synth.flux_vjp_device(gsfl, gsfc)
        gl = torch.empty(dlines.n, dtype=torch.float64, device=g.device)
        gfpf, gt_line = synth.lineop_vjp_device(dlines, gl, nlim)
        return gl, gt_src + gt_line, gfpf
    if isinstance(synth, FastSpecBatch):
        _, _, gt_src = synth.flux_vjp(synth.split(gsfl.numpy()), synth.split(gsfc.numpy()))
    else:
        _, _, gt_src = synth.flux_vjp(gsfl.numpy(), gsfc.numpy())
    gl, gfpf, gt_line = synth.lineop_vjp(dlines, nlim)
    return torch.from_numpy(np.ascontiguousarray(gl)), gt_src + gt_line, gfpf


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
        _forward_scalar(synth, dlines, gf, nlines_lim)
        sfl, sfc = _flux(synth, loggf.device)
        ctx.synth, ctx.dlines, ctx.nlim = synth, dlines, nlines_lim
        ctx.sfl, ctx.sfc, ctx.gf = sfl, sfc, gf
        ctx.serial = _next_serial(synth)
        ctx.meta = [(x.dtype, x.device) for x in (loggf, temp, lnfpf)]
        return (sfl / sfc).to(loggf.device)

    @staticmethod
    def backward(ctx, g):
        if getattr(ctx.synth, _SERIAL, 0) != ctx.serial:  # another forward ran since
            _forward_scalar(ctx.synth, ctx.dlines, ctx.gf, ctx.nlim)
            setattr(ctx.synth, _SERIAL, ctx.serial)
        g = g.detach().to(ctx.sfl.device, torch.float64)
        g_loggf, gt, g_fpf = _vjp(ctx.synth, ctx.dlines, ctx.nlim, g, ctx.sfl, ctx.sfc)
        out = [torch.as_tensor(a).to(dtype=dt, device=dev)
               for a, (dt, dev) in zip((g_loggf, gt, g_fpf), ctx.meta)]
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
        _forward_batch(batch, blines, gf, nlines_lim)
        sfl, sfc = _flux(batch, loggf.device)
        ctx.batch, ctx.blines, ctx.nlim = batch, blines, nlines_lim
        ctx.sfl, ctx.sfc, ctx.gf = sfl, sfc, gf
        ctx.serial = _next_serial(batch)
        ctx.meta = [(x.dtype, x.device) for x in (loggf, temp, lnfpf)]
        # the atmospheres' points of the concatenated grid, in order
        sel = np.concatenate([np.arange(batch.off[b], batch.off[b] + su.nwl) for b, su in enumerate(batch.setups)])
        ctx.sel = torch.as_tensor(sel, device=sfl.device)
        return (sfl / sfc)[ctx.sel].to(loggf.device)

    @staticmethod
    def backward(ctx, g):
        if getattr(ctx.batch, _SERIAL, 0) != ctx.serial:  # another forward ran since
            _forward_batch(ctx.batch, ctx.blines, ctx.gf, ctx.nlim)
            setattr(ctx.batch, _SERIAL, ctx.serial)
        gfull = torch.zeros(ctx.batch.nwl_total, dtype=torch.float64, device=ctx.sfl.device)
        gfull[ctx.sel] = g.detach().to(ctx.sfl.device, torch.float64)
        g_loggf, gt, g_fpf = _vjp(ctx.batch, ctx.blines, ctx.nlim, gfull, ctx.sfl, ctx.sfc)
        # per-line gradients of the batch list summed onto the shared list
        idx = torch.as_tensor(ctx.blines.index, device=g_loggf.device)
        g_shared = torch.zeros(ctx.blines.nshared, dtype=torch.float64, device=g_loggf.device)
        g_shared.index_add_(0, idx, g_loggf)
        out = [torch.as_tensor(a).to(dtype=dt, device=dev)
               for a, (dt, dev) in zip((g_shared, gt, g_fpf), ctx.meta)]
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
