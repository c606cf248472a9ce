"""A new atmosphere state on an existing context (fsg_set_atmosphere) and the gradient chain
through the host's EOS and continuum that it allows in torch.

* A context built on one atmosphere and moved to another gives the same line opacity and
  flux, bit for bit, as a context built on the other (scalar and batch contexts).
* With a smooth model of the EOS and continuum written in torch -- FractionPF and ContOpac
  as functions of T and of an element abundance -- autograd's dL/dT and dL/d log A through
  NormalisedFlux match centred finite differences of the whole chain; dL/dContOpac matches
  a directional difference.
"""
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64

pytestmark = pytest.mark.gpu

REL = 1e-3


def _atmospheres():
    a = W.build_setup(W.grey_atmosphere(5777.0, 4.44), 5000.0, 5008.0, 0.01, 0)
    b = W.build_setup(W.grey_atmosphere(6100.0, 4.2), 5000.0, 5008.0, 0.01, 0)
    lines = W.atomic_lines(240, 5000.0, 5008.0, 31)
    return a, b, lines


STATE = ("temp", "xne", "rho", "rhox", "contop", "fraction_pf", "frac_h1", "frac_he1", "frac_hmol")


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_moved_context_equals_fresh_context(prec):
    from fastspec_amd.synth import FastSpecGPU
    a, b, lines = _atmospheres()
    out = []
    for setup, move in ((b, None), (a, b)):
        with FastSpecGPU(setup, precision=prec) as g:
            if move is not None:
                g.set_atmosphere(**{k: getattr(move, k) for k in STATE})
                assert np.array_equal(g.setup.temp, b.temp)
            dl = g.upload(lines)
            st = g.lineop_dev(dl)
            g.RadTranCalc(intensities=False)
            out.append((st["nlines_use"], g.get_lineop(), *g.flux()))
            dl.free()
    assert out[0][0] == out[1][0]
    for x, y in zip(out[0][1:], out[1][1:]):
        assert np.array_equal(x, y)


def test_moved_batch_atmosphere_equals_fresh_batch():
    from fastspec_amd.synth import FastSpecBatch
    a, b, lines = _atmospheres()
    c = W.build_setup(W.grey_atmosphere(5400.0, 4.6), 5000.0, 5006.0, 0.01, 0)
    idx = [np.arange(lines.n), np.arange(0, lines.n, 2)]
    res = []
    for setups, move in (([a, b], None), ([a, c], b)):
        with FastSpecBatch(setups, precision=FSG_FP32) as bt:
            if move is not None:
                # the second atmosphere's grid stays (6 A); its state becomes b's on that grid
                bsub = W.build_setup(W.grey_atmosphere(6100.0, 4.2), 5000.0, 5006.0, 0.01, 0)
                bt.set_atmosphere(1, **{k: getattr(bsub, k) for k in STATE})
            sh = bt.upload(lines)
            bl = bt.batch_lines(sh, idx)
            bt.lineop_dev(bl)
            bt.RadTranCalc(intensities=False)
            sfl, sfc = bt.flux()
            res.append([x / y for x, y in zip(sfl, sfc)])
            bl.free()
            sh.free()
    # atmosphere 0 is untouched by moving atmosphere 1
    assert np.array_equal(res[0][0], res[1][0])
    # atmosphere 1 against a fresh single context on the 6 A grid with b's state
    from fastspec_amd.synth import FastSpecGPU
    bsub = W.build_setup(W.grey_atmosphere(6100.0, 4.2), 5000.0, 5006.0, 0.01, 0)
    with FastSpecGPU(bsub, precision=FSG_FP32) as g:
        dl = g.upload(lines.take(idx[1]))
        g.lineop_dev(dl)
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
        dl.free()
    assert np.max(np.abs(res[1][1] - sfl / sfc)) <= 1e-6


# ---- the chain through a torch model of the EOS and continuum ------------------------------

@pytest.fixture(scope="module")
def smooth():
    """EpsOpac = 1e-12: every line retained and every window spanning the grid, so the
    forward is smooth in T, FractionPF and ContOpac."""
    a, _, lines = _atmospheres()
    setup = dataclasses.replace(a, eps_opac=1e-12)
    wvec = np.random.default_rng(11).normal(size=setup.nwl)
    return setup, lines, wvec


def _model(torch, setup, t, dlogA, slots):
    """A smooth stand-in for the host's EOS and continuum: ln FractionPF of the line
    species' slots falls with T as a Saha-like factor (slope -40 in ln T) and rises with the
    abundance; ContOpac scales as (T/T0)^3 at each depth and with the abundance."""
    t0 = torch.as_tensor(setup.temp, dtype=torch.float64)
    x = (t - t0) / t0
    ln10 = float(np.log(10.0))
    cols = [(-40.0 * x + ln10 * dlogA) if j in slots else torch.zeros_like(x)
            for j in range(setup.fraction_pf.shape[1])]
    lnfpf = torch.stack(cols, dim=1)
    c0 = torch.as_tensor(setup.contop, dtype=torch.float64)
    contop = c0 * ((t / t0) ** 3)[None, :] * (1.0 + 0.1 * dlogA)
    return lnfpf, contop


def _chain_loss(torch, g, dl, setup, wvec, loggf, t, dlogA, slots):
    from fastspec_amd.autograd import normalised_flux
    lnfpf, contop = _model(torch, setup, t, dlogA, slots)
    f = normalised_flux(g, dl, loggf, temp=t, lnfpf=lnfpf, contop=contop)
    return (torch.as_tensor(wvec) * f).sum()


def test_temperature_and_abundance_chain_vs_finite_differences(smooth):
    torch = pytest.importorskip("torch")
    from fastspec_amd.synth import FastSpecGPU
    setup, lines, wvec = smooth
    slots = sorted({int(setup.spc_slot[s - 1]) - 1 for s in np.unique(lines.spc)})
    loggf = torch.as_tensor(np.log10(lines.gf.astype(np.float64)))
    with FastSpecGPU(setup, precision=FSG_FP64) as g:
        dl = g.upload(lines)
        t = torch.tensor(setup.temp, dtype=torch.float64, requires_grad=True)
        dA = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
        loss = _chain_loss(torch, g, dl, setup, wvec, loggf, t, dA, slots)
        loss.backward()
        gt = t.grad.numpy().copy()
        gA = float(dA.grad)

        def L(tv, av):
            with torch.no_grad():
                return float(_chain_loss(torch, g, dl, setup, wvec, loggf, torch.as_tensor(tv, dtype=torch.float64),
                                         torch.tensor(av, dtype=torch.float64), slots))

        # FractionPF is REAL(4): steps must move it by much more than its rounding (a 0.2 K
        # step moves ln FractionPF by ~1e-3), so they are taken in the deeper half of the
        # model, whose layers lie tens of K apart (the top layers lie mK apart, inside
        # CentDer's monotonicity switch)
        h = 0.2
        deep = np.arange(setup.ntau // 2, setup.ntau)
        for k in deep[np.argsort(-np.abs(gt[deep]))[:6]]:
            tp, tm = setup.temp.copy(), setup.temp.copy()
            tp[k] += h
            tm[k] -= h
            fd = (L(tp, 0.0) - L(tm, 0.0)) / (2 * h)
            assert abs(gt[k] - fd) <= REL * abs(fd) + 1e-6 * np.abs(gt).max(), (k, gt[k], fd)
        ha = 1e-3
        fa = (L(setup.temp, ha) - L(setup.temp, -ha)) / (2 * ha)
        assert abs(gA - fa) <= REL * abs(fa), (gA, fa)
        # the finite differences left the context elsewhere; the base state gives the first loss
        assert L(setup.temp, 0.0) == loss.item()
        dl.free()


def test_continuum_gradient_directional(smooth):
    torch = pytest.importorskip("torch")
    from fastspec_amd.autograd import normalised_flux
    from fastspec_amd.synth import FastSpecGPU
    setup, lines, wvec = smooth
    direction = np.random.default_rng(3).uniform(-1.0, 1.0, setup.contop.shape) * setup.contop
    with FastSpecGPU(setup, precision=FSG_FP64) as g:
        dl = g.upload(lines)
        loggf = torch.as_tensor(np.log10(lines.gf.astype(np.float64)))
        c = torch.tensor(setup.contop, dtype=torch.float64, requires_grad=True)
        loss = (torch.as_tensor(wvec) * normalised_flux(g, dl, loggf, contop=c)).sum()
        loss.backward()
        an = float(np.sum(c.grad.numpy() * direction))

        def L(eps):
            with torch.no_grad():
                cc = torch.as_tensor(setup.contop + eps * direction)
                return float((torch.as_tensor(wvec) * normalised_flux(g, dl, loggf, contop=cc)).sum())

        eps = 1e-6
        fd = (L(eps) - L(-eps)) / (2 * eps)
        assert abs(an - fd) <= REL * abs(fd), (an, fd)
        dl.free()


def test_chain_on_device_equals_host(smooth):
    """The same chain with CUDA tensors: device-pointer entry points, dL/dContOpac on the
    device; gradients equal the host path's."""
    torch = pytest.importorskip("torch")
    from fastspec_amd.autograd import normalised_flux
    from fastspec_amd.synth import FastSpecGPU
    setup, lines, wvec = smooth
    grads = []
    with FastSpecGPU(setup, precision=FSG_FP64) as g:
        dl = g.upload(lines)
        for dev in ("cpu", "cuda"):
            loggf = torch.tensor(np.log10(lines.gf.astype(np.float64)), device=dev, requires_grad=True)
            t = torch.tensor(setup.temp * 1.01, device=dev, requires_grad=True)
            c = torch.tensor(setup.contop * 1.02, device=dev, requires_grad=True)
            lnf = torch.full(setup.fraction_pf.shape, 0.05, dtype=torch.float64, device=dev, requires_grad=True)
            f = normalised_flux(g, dl, loggf, temp=t, lnfpf=lnf, contop=c)
            (torch.as_tensor(wvec, device=dev) * f).sum().backward()
            grads.append([x.grad.detach().cpu().numpy() for x in (loggf, t, lnf, c)])
        dl.free()
    # gf = 10**log gf is rounded to REAL(4) by torch's pow on each device, which may differ
    # in the last bit: agreement to a few float32 ulps
    for h, d in zip(*grads):
        assert np.max(np.abs(h - d)) <= 1e-6 * np.max(np.abs(h))
