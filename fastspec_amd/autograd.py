"""PyTorch autograd over libfsgpu: the normalised flux of one synthesis as a differentiable
function of the lines' log gf, the temperature of each depth, the EOS number densities and
the continuum opacity.  Parameters on a CUDA device stay there: gf goes to the library as a
device pointer, F/Fc comes back in device tensors (fsg_flux_dev) and the gradients through
the device-pointer reverse-mode entry points (the library shares torch's HIP runtime; see
_lib).

    F/Fc = NormalisedFlux.apply(loggf, temp, lnfpf, contop, synth, dlines, nlines_lim, state)

Forward runs the same stages as the reference's synthesis for one list (LineopCalc,
RadTranCalc, the flux sums of IntegConvol, src/fastspec.f90:241-449) on the GPU; backward
runs the library's adjoint kernels (fsg_flux_vjp, fsg_lineop_vjp).  The atmosphere inputs
are those the host's EOS and continuum produce (EosSolve, ContopCalc; src/synth.f90:334-406,
642-674):

* ``temp`` [ntau], T of each depth;
* ``lnfpf`` [ntau][pf_ld], ln FractionPF relative to the values the context was built with
  (zeros: those values);
* ``contop`` [nwl][ntau], ContOpac (None: the context's).

A forward whose inputs differ from the context's state moves the context there first
(fsg_set_atmosphere), so a model written in torch that maps physical parameters onto these
inputs -- an EOS giving FractionPF(T, abundances), a continuum giving ContOpac(T,
abundances) -- is differentiated end to end by autograd: the library supplies dL/dT at
fixed FractionPF, ContOpac and electron density, dL/d ln FractionPF and dL/dContOpac, and
torch chains the model's Jacobians onto them.  ``state`` holds further EOS results that
enter the forward but carry no gradient (xne, rho, frac_h1, frac_he1, frac_hmol: the
electron and perturber densities of the damping constants, as numpy arrays).  The
element-abundance gradient of a trace species at fixed continuum is the sum of its slots'
rows: dL/d log10 A = ln 10 sum_depth dL/d ln FractionPF.

    F/Fc (all atmospheres) = NormalisedFluxBatch.apply(loggf, temp, lnfpf, contop, batch,
                                                       blines, nlim, state)

is the same over a batch context (many atmospheres in one launch sequence, one shared
list): temp [nb][ntau], lnfpf [nb][ntau][pf_ld], contop over the atmospheres' own grids
concatenated ([sum nwl_b][ntau]) or None, state a list of per-atmosphere dicts, and the
gradient in loggf summed over the atmospheres that use each line.
"""
from __future__ import annotations

import numpy as np
import torch

from .synth import BatchLines, DeviceLines, FastSpecBatch, FastSpecGPU

# Each forward leaves its state in the context (atmosphere, LineOpac, the formal solution,
# the list's gf and the pass-1 buffers) and backward reads that state.  A serial number per
# context names the forward whose state the context holds; backward of an earlier forward
# (two forwards before one backward, as in a loss over two parameter vectors or gradcheck's
# perturbed evaluations) first runs that forward again from its saved inputs.
_SERIAL = "_autograd_serial"
_STATE_KEYS = ("xne", "rho", "frac_h1", "frac_he1", "frac_hmol")


def _next_serial(synth) -> int:
    n = getattr(synth, _SERIAL, 0) + 1
    setattr(synth, _SERIAL, n)
    return n


def _base_fpf(synth, b: int):
    """FractionPF of atmosphere b as the context was built (the origin of lnfpf)."""
    return synth.fraction_pf0[b]


def _gf32(loggf: torch.Tensor):
    """gf = 10**log gf in REAL(4), as the host's reader converts it: a numpy array for host
    tensors, a finished CUDA tensor on the parameters' device for device tensors."""
    ten = torch.tensor(10.0, dtype=torch.float32, device=loggf.device)
    gf = torch.pow(ten, loggf.detach().float()).contiguous()
    if gf.is_cuda:
        torch.cuda.current_stream(gf.device).synchronize()
        return gf
    return gf.numpy()


def _host(x):
    return None if x is None else np.ascontiguousarray(x.detach().to("cpu", torch.float64).numpy())


def _atm_targets(synth, b, temp, lnfpf, contop, state):
    """The set-up arrays one atmosphere must hold for these inputs."""
    fpf = (_base_fpf(synth, b).astype(np.float64) * np.exp(lnfpf)).astype(np.float32)
    out = dict(temp=temp, fraction_pf=fpf, contop=contop)
    out.update({k: v for k, v in (state or {}).items() if k in _STATE_KEYS})
    return {k: v for k, v in out.items() if v is not None}


def _apply_atmosphere(synth, b, targets) -> None:
    """fsg_set_atmosphere with the arrays that differ from the context's current state."""
    su = synth.setups[b] if isinstance(synth, FastSpecBatch) else synth.setup
    dt = {k: np.float32 if k.startswith("frac") else np.float64 for k in targets}
    change = {k: v for k, v in targets.items()
              if not np.array_equal(np.asarray(v, dt[k]), getattr(su, k))}
    if change:
        synth.set_atmosphere(b, **change)


def _forward_scalar(synth: FastSpecGPU, dlines: DeviceLines, gf, nlines_lim, targets):
    _apply_atmosphere(synth, 0, targets)
    synth.set_gf(dlines, gf)
    synth.reset_lineop()
    synth.lineop_dev(dlines, nlines_lim)
    synth.RadTranCalc(intensities=False)


def _forward_batch(batch: FastSpecBatch, blines: BatchLines, gf, nlines_lim, targets):
    for b, t in enumerate(targets):
        _apply_atmosphere(batch, b, t)
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


def _vjp(synth, dlines, nlim, g, sfl, sfc, want_cont: bool):
    """dL/d log gf (on g's device), dL/dT and dL/d ln FractionPF (host) and, with want_cont,
    dL/dContOpac over the context's grid [nwl_total][ntau] (on g's device), from dL/d(F/Fc) =
    g: d(F/Fc) gives dL/dSfL = g / Fc, dL/dSfC = -g F / Fc^2."""
    gsfl = (g / sfc).contiguous()
    gsfc = (-g * sfl / (sfc * sfc)).contiguous()
    ntau = synth.setup.ntau
    if g.is_cuda:
        torch.cuda.current_stream(g.device).synchronize()
        gc = torch.empty((ntau, synth.nwl_total), dtype=torch.float64, device=g.device) if want_cont else None
        gt_src = synth.flux_vjp_device(gsfl, gsfc, gc)
        gl = torch.empty(dlines.n, dtype=torch.float64, device=g.device)
        gfpf, gt_line = synth.lineop_vjp_device(dlines, gl, nlim)
        return gl, gt_src + gt_line, gfpf, (gc.t() if want_cont else None)
    if isinstance(synth, FastSpecBatch):
        _, gc, gt_src = synth.flux_vjp(synth.split(gsfl.numpy()), synth.split(gsfc.numpy()), want_opac=want_cont)
        if want_cont:
            gc = synth.join(gc)
    else:
        _, gc, gt_src = synth.flux_vjp(gsfl.numpy(), gsfc.numpy(), want_opac=want_cont)
    gl, gfpf, gt_line = synth.lineop_vjp(dlines, nlim)
    gc = torch.from_numpy(gc) if want_cont else None
    return torch.from_numpy(np.ascontiguousarray(gl)), gt_src + gt_line, gfpf, gc


def _grads(ctx, values):
    """Gradients cast to each input's dtype and device (None where no gradient is needed)."""
    out = []
    for k, (a, (dt, dev)) in enumerate(zip(values, ctx.meta)):
        out.append(None if a is None or not ctx.needs_input_grad[k] else torch.as_tensor(a).to(dtype=dt, device=dev))
    return out


class NormalisedFlux(torch.autograd.Function):
    @staticmethod
    def forward(ctx, loggf, temp, lnfpf, contop, synth: FastSpecGPU, dlines: DeviceLines, nlines_lim=None,
                state=None):
        targets = _atm_targets(synth, 0, _host(temp), _host(lnfpf), _host(contop), state)
        gf = _gf32(loggf)
        _forward_scalar(synth, dlines, gf, nlines_lim, targets)
        sfl, sfc = _flux(synth, loggf.device)
        ctx.synth, ctx.dlines, ctx.nlim, ctx.targets = synth, dlines, nlines_lim, targets
        ctx.sfl, ctx.sfc, ctx.gf = sfl, sfc, gf
        ctx.serial = _next_serial(synth)
        ctx.meta = [(x.dtype, x.device) if x is not None else (None, None) for x in (loggf, temp, lnfpf, contop)]
        return (sfl / sfc).to(loggf.device)

    @staticmethod
    def backward(ctx, g):
        if getattr(ctx.synth, _SERIAL, 0) != ctx.serial:  # another forward ran since
            _forward_scalar(ctx.synth, ctx.dlines, ctx.gf, ctx.nlim, ctx.targets)
            setattr(ctx.synth, _SERIAL, ctx.serial)
        g = g.detach().to(ctx.sfl.device, torch.float64)
        g_loggf, gt, g_fpf, g_cont = _vjp(ctx.synth, ctx.dlines, ctx.nlim, g, ctx.sfl, ctx.sfc,
                                          ctx.needs_input_grad[3])
        return (*_grads(ctx, (g_loggf, gt, g_fpf, g_cont)), None, None, None, None)


def normalised_flux(synth: FastSpecGPU, dlines: DeviceLines, loggf: torch.Tensor, temp=None, lnfpf=None,
                    contop=None, nlines_lim=None, **state) -> torch.Tensor:
    """F/Fc [nwl] of one synthesis; differentiable in loggf and in temp [ntau], lnfpf
    [ntau][pf_ld] and contop [nwl][ntau] when given with requires_grad.  Keyword arrays xne,
    rho, frac_h1, frac_he1, frac_hmol set the non-differentiated EOS results."""
    if temp is None:
        temp = torch.as_tensor(np.array(synth.setup.temp), dtype=torch.float64)
    if lnfpf is None:
        lnfpf = torch.zeros(synth.setup.fraction_pf.shape, dtype=torch.float64)
    return NormalisedFlux.apply(loggf, temp, lnfpf, contop, synth, dlines, nlines_lim, state)


class NormalisedFluxBatch(torch.autograd.Function):
    """F/Fc of every atmosphere of a batch context (fsg_init_batch), concatenated over the
    atmospheres' own grids, as a function of the shared list's log gf, the temperatures
    [nb][ntau], ln FractionPF [nb][ntau][pf_ld] (relative to the batch's set-ups) and
    ContOpac over the atmospheres' own grids ([sum nwl_b][ntau], or None)."""

    @staticmethod
    def forward(ctx, loggf, temp, lnfpf, contop, batch: FastSpecBatch, blines: BatchLines, nlines_lim=None,
                state=None):
        t, lf, co = _host(temp), _host(lnfpf), _host(contop)
        nw = np.cumsum([0] + [su.nwl for su in batch.setups])
        targets = [_atm_targets(batch, b, t[b], lf[b], None if co is None else co[nw[b]:nw[b + 1]],
                                state[b] if state else None) for b in range(batch.nb)]
        gf = _gf32(loggf)
        _forward_batch(batch, blines, gf, nlines_lim, targets)
        sfl, sfc = _flux(batch, loggf.device)
        ctx.batch, ctx.blines, ctx.nlim, ctx.targets = batch, blines, nlines_lim, targets
        ctx.sfl, ctx.sfc, ctx.gf = sfl, sfc, gf
        ctx.serial = _next_serial(batch)
        ctx.meta = [(x.dtype, x.device) if x is not None else (None, None) for x in (loggf, temp, lnfpf, contop)]
        # the atmospheres' points of the concatenated grid, in order
        sel = np.concatenate([np.arange(batch.off[b], batch.off[b] + su.nwl) for b, su in enumerate(batch.setups)])
        ctx.sel = torch.as_tensor(sel, device=sfl.device)
        return (sfl / sfc)[ctx.sel].to(loggf.device)

    @staticmethod
    def backward(ctx, g):
        if getattr(ctx.batch, _SERIAL, 0) != ctx.serial:  # another forward ran since
            _forward_batch(ctx.batch, ctx.blines, ctx.gf, ctx.nlim, ctx.targets)
            setattr(ctx.batch, _SERIAL, ctx.serial)
        gfull = torch.zeros(ctx.batch.nwl_total, dtype=torch.float64, device=ctx.sfl.device)
        gfull[ctx.sel] = g.detach().to(ctx.sfl.device, torch.float64)
        g_loggf, gt, g_fpf, g_cont = _vjp(ctx.batch, ctx.blines, ctx.nlim, gfull, ctx.sfl, ctx.sfc,
                                          ctx.needs_input_grad[3])
        if g_cont is not None:
            g_cont = g_cont[ctx.sel.to(g_cont.device)]
        # per-line gradients of the batch list summed onto the shared list
        idx = torch.as_tensor(ctx.blines.index, device=g_loggf.device)
        g_shared = torch.zeros(ctx.blines.nshared, dtype=torch.float64, device=g_loggf.device)
        g_shared.index_add_(0, idx, g_loggf)
        return (*_grads(ctx, (g_shared, gt, g_fpf, g_cont)), None, None, None, None)


def normalised_flux_batch(batch: FastSpecBatch, blines: BatchLines, loggf: torch.Tensor, temp=None, lnfpf=None,
                          contop=None, nlines_lim=None, state=None) -> torch.Tensor:
    """F/Fc of all atmospheres, concatenated (atmosphere b: batch.setups[b].nwl points);
    differentiable in the shared list's loggf and in temp [nb][ntau], lnfpf and contop when
    given with requires_grad.  state: per-atmosphere dicts of non-differentiated EOS results
    (xne, rho, frac_h1, frac_he1, frac_hmol) or None."""
    if temp is None:
        temp = torch.as_tensor(np.stack([su.temp for su in batch.setups]), dtype=torch.float64)
    if lnfpf is None:
        lnfpf = torch.zeros((batch.nb,) + batch.setup.fraction_pf.shape, dtype=torch.float64)
    return NormalisedFluxBatch.apply(loggf, temp, lnfpf, contop, batch, blines, nlines_lim, state)


__all__ = ["NormalisedFlux", "normalised_flux", "NormalisedFluxBatch", "normalised_flux_batch"]
