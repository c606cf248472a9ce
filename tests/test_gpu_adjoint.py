"""Reverse mode (fsg_flux_vjp, fsg_lineop_vjp and the torch autograd wrapper) against centred
finite differences of the forward synthesis: of the library's FP64 path and of the FP64
oracle (the reference's algorithm).  Loss: L = sum_lambda a(lambda) F/Fc with seeded a."""
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64
from oracle_bridge import Oracle

pytestmark = pytest.mark.gpu

REL = 1e-3  # gradients vs centred finite differences


def _gpu(setup, prec=FSG_FP64):
    from fastspec_amd.synth import FastSpecGPU
    return FastSpecGPU(setup, precision=prec)


@pytest.fixture(scope="module")
def case():
    atm = W.grey_atmosphere(5777.0, 4.44)
    setup = W.build_setup(atm, 5000.0, 5008.0, 0.01, 0)
    lines = W.atomic_lines(240, 5000.0, 5008.0, 31)
    a = np.random.default_rng(5).normal(size=setup.nwl)
    return setup, lines, a


def _loss_gpu(setup, lines, a, prec=FSG_FP64):
    with _gpu(setup, prec) as g:
        g.LineopCalc(lines)
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
    return float(np.sum(a * sfl / sfc))


def _loss_oracle(setup, lines, a):
    o = Oracle(setup)
    o.lineop(lines)
    o.radtran()
    sfl, sfc = o.flux()
    return float(np.sum(a * sfl / sfc))


def _backward(setup, lines, a, prec=FSG_FP64):
    """All gradients of L from one forward + reverse pass."""
    with _gpu(setup, prec) as g:
        dl = g.upload(lines)
        g.lineop_dev(dl)
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
        glop, gcont, gt = g.flux_vjp(a / sfc, -a * sfl / sfc**2, want_opac=True)
        ggf, gfpf, gtl = g.lineop_vjp(dl)
        lop = g.get_lineop()
        dl.free()
    return dict(glop=glop, gcont=gcont, gt=gt + gtl, ggf=ggf, gfpf=gfpf, lop=lop, L=float(np.sum(a * sfl / sfc)))


def _rel(x, y):
    return float(np.max(np.abs(x - y)) / np.max(np.abs(y)))


def test_opacity_adjoints_directional(case):
    """dL/dLineOpac and dL/dContOpac along seeded directions (the DELO-Bezier reverse sweep)."""
    setup, lines, a = case
    b = _backward(setup, lines, a)
    rng = np.random.default_rng(7)
    # LineOpac direction: relative perturbations of the accumulated line opacity
    d = rng.normal(size=b["lop"].shape) * b["lop"]
    h = 1e-6
    with _gpu(setup) as g:
        fd = []
        for sgn in (1, -1):
            g.set_lineop(b["lop"] + sgn * h * d)
            g.RadTranCalc(intensities=False)
            sfl, sfc = g.flux()
            fd.append(np.sum(a * sfl / sfc))
    fd = (fd[0] - fd[1]) / (2 * h)
    an = float(np.sum(b["glop"] * d))
    assert abs(an - fd) <= REL * abs(fd), (an, fd)
    # the same directional derivative with the oracle's forward
    o = Oracle(setup)
    fo = []
    for sgn in (1, -1):
        o.set_lineop(b["lop"] + sgn * h * d)
        o.radtran()
        sfl, sfc = o.flux()
        fo.append(np.sum(a * sfl / sfc))
    fo = (fo[0] - fo[1]) / (2 * h)
    assert abs(an - fo) <= REL * abs(fo), (an, fo)
    # ContOpac direction (continuum enters both the line and the continuum solution)
    dc = rng.normal(size=setup.contop.shape) * setup.contop
    fc = []
    for sgn in (1, -1):
        s2 = dataclasses.replace(setup, contop=setup.contop + sgn * h * dc)
        with _gpu(s2) as g:
            g.set_lineop(b["lop"])
            g.RadTranCalc(intensities=False)
            sfl, sfc = g.flux()
        fc.append(np.sum(a * sfl / sfc))
    fc = (fc[0] - fc[1]) / (2 * h)
    anc = float(np.sum(b["gcont"] * dc))
    assert abs(anc - fc) <= REL * abs(fc), (anc, fc)


def test_loggf_gradient_vs_finite_differences(case):
    setup, lines, a = case
    b = _backward(setup, lines, a)
    ggf = b["ggf"]
    assert np.count_nonzero(ggf) > 50
    idx = np.argsort(-np.abs(ggf))[:24]

    def fd_of(i, loss):
        # gf is REAL(4): the step is taken between the two rounded values
        out, lg = [], []
        for sgn in (1, -1):
            l2 = lines.take(np.arange(lines.n))
            gf = l2.gf.copy()
            gf[i] = np.float32(gf[i] * 10.0 ** (sgn * 1e-3))
            lg.append(np.log10(np.float64(gf[i])))
            l2.gf = gf
            out.append(loss(setup, l2, a))
        return (out[0] - out[1]) / (lg[0] - lg[1])

    for i in idx:
        fd = fd_of(i, _loss_gpu)
        assert abs(ggf[i] - fd) <= REL * abs(fd) + 1e-9 * np.abs(ggf).max(), (i, ggf[i], fd)
    # three of them against the oracle's forward (the reference's algorithm)
    for i in idx[:3]:
        fo = fd_of(i, _loss_oracle)
        assert abs(ggf[i] - fo) <= REL * abs(fo), (i, ggf[i], fo)


@pytest.fixture(scope="module")
def smooth_case(case):
    """The same synthesis with EpsOpac = 1e-12: every line is retained and every window spans
    the grid, so the forward is smooth in T and FractionPF (no window edge or weak-line
    flag moves under a finite-difference step)."""
    setup, lines, a = case
    return dataclasses.replace(setup, eps_opac=1e-12), lines, a


def test_source_function_temperature_gradient(case):
    """dL/dT through the Planck function alone: LineOpac held fixed with fsg_set_lineop."""
    setup, lines, a = case
    b = _backward(setup, lines, a)

    def loss(t):
        with _gpu(dataclasses.replace(setup, temp=t)) as g:
            g.set_lineop(b["lop"])
            g.RadTranCalc(intensities=False)
            sfl, sfc = g.flux()
        return float(np.sum(a * sfl / sfc))

    with _gpu(setup) as g:
        g.set_lineop(b["lop"])
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
        _, _, gsrc = g.flux_vjp(a / sfc, -a * sfl / sfc**2)
    # near the top of a grey atmosphere neighbouring layers differ by ~5 mK, and CentDer's
    # monotonicity switch sits that close: the step must stay well inside it
    h = 1e-4
    for k in np.argsort(-np.abs(gsrc))[:10]:
        tp, tm = setup.temp.copy(), setup.temp.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss(tp) - loss(tm)) / (2 * h)
        assert abs(gsrc[k] - fd) <= REL * abs(fd), (k, gsrc[k], fd)


def test_temperature_gradient_vs_finite_differences(smooth_case):
    """dL/dT per depth at fixed EOS, continuum and electron densities: the Planck function of
    both solutions and the temperature dependence of the line opacity."""
    setup, lines, a = smooth_case
    b = _backward(setup, lines, a)
    gt = b["gt"]
    h = 1e-4  # K, inside the spacing of the top layers (see the source-function test)
    ks = np.argsort(-np.abs(gt))[:12]
    for k in ks:
        fd = []
        for sgn in (1, -1):
            t = setup.temp.copy()
            t[k] += sgn * h
            fd.append(_loss_gpu(dataclasses.replace(setup, temp=t), lines, a))
        fd = (fd[0] - fd[1]) / (2 * h)
        assert abs(gt[k] - fd) <= REL * abs(fd) + 1e-6 * np.abs(gt).max(), (k, gt[k], fd)
    k = ks[0]
    fo = []
    for sgn in (1, -1):
        t = setup.temp.copy()
        t[k] += sgn * h
        fo.append(_loss_oracle(dataclasses.replace(setup, temp=t), lines, a))
    fo = (fo[0] - fo[1]) / (2 * h)
    assert abs(gt[k] - fo) <= REL * abs(fo), (gt[k], fo)


def test_fractionpf_gradient_vs_finite_differences(smooth_case):
    """dL/d ln FractionPF of each line species along the abundance direction of a trace
    element (all depths scaled together)."""
    setup, lines, a = smooth_case
    b = _backward(setup, lines, a)
    g = b["gfpf"]
    h = 1e-3
    for slot in np.unique(lines.spc):
        col = slot - 1
        fd, dlog = [], []
        for sgn in (1, -1):
            f = setup.fraction_pf.copy()
            f[:, col] = (f[:, col] * np.exp(sgn * h)).astype(np.float32)
            # FractionPF is REAL(4): the directional step is the rounded one, per depth
            dlog.append(np.log(f[:, col].astype(np.float64) / setup.fraction_pf[:, col]))
            fd.append(_loss_gpu(dataclasses.replace(setup, fraction_pf=f), lines, a))
        an_p, an_m = float(np.sum(g[:, col] * dlog[0])), float(np.sum(g[:, col] * dlog[1]))
        assert abs((an_p - an_m) - (fd[0] - fd[1])) <= REL * abs(fd[0] - fd[1]), (slot, an_p - an_m, fd[0] - fd[1])
    # slots without lines have zero gradient
    used = set(int(x) - 1 for x in np.unique(lines.spc))
    for col in range(g.shape[1]):
        if col not in used:
            assert np.all(g[:, col] == 0.0)


def test_fp32_forward_gradients_close_to_fp64(case):
    setup, lines, a = case
    b64 = _backward(setup, lines, a, FSG_FP64)
    b32 = _backward(setup, lines, a, FSG_FP32)
    assert _rel(b32["ggf"], b64["ggf"]) < 1e-3
    assert _rel(b32["gt"], b64["gt"]) < 1e-3


def test_torch_autograd(case):
    torch = pytest.importorskip("torch")
    from fastspec_amd.autograd import normalised_flux
    setup, lines, a = case
    # host tensors: the synthesis itself stays on the device inside the library
    loggf = torch.tensor(np.log10(lines.gf.astype(np.float64)), dtype=torch.float64, requires_grad=True)
    temp = torch.tensor(setup.temp, dtype=torch.float64, requires_grad=True)
    with _gpu(setup) as g:
        dl = g.upload(lines)
        f = normalised_flux(g, dl, loggf, temp=temp)
        loss = (torch.as_tensor(a, device=f.device) * f).sum()
        loss.backward()
        dl.free()
    b = _backward(setup, lines, a)
    # gf = 10**loggf is re-rounded to REAL(4) on the device: agreement to float32 precision
    assert abs(loss.item() - b["L"]) < 1e-5 * np.sum(np.abs(a))
    assert _rel(loggf.grad.cpu().numpy(), b["ggf"]) < 1e-4
    assert _rel(temp.grad.numpy(), b["gt"]) < 1e-4


def _loss_reference(setup, lines, a, windows):
    """L from the reference's own compiled kernels (oracle/_ref: XlineSetup/XlineDepth,
    LineOpacity1, the RadTranCalc loop of LocPrfPP) and the flux sums, with the windows and
    flags given (held fixed, as the reverse mode holds them)."""
    from oracle_bridge import Ref
    flags, il, ir = windows
    ref = Ref(setup)
    ref.set_lines(lines)
    cop, vv, av, _ = ref.xline()
    ref.set_windows(flags & ~8, il, ir)
    ref.set_block(cop, vv, av)
    lop = ref.lineop1(1, setup.nwl, 1, lines.n)
    spl, spc = ref.radtran(1, setup.nwl, lop)
    sfl, sfc = spl @ setup.wmu, spc @ setup.wmu
    return float(np.sum(a * sfl / sfc))


@pytest.mark.ref
def test_gradients_vs_reference_code_finite_differences(smooth_case):
    """dL/d log gf and dL/dT of the library's reverse mode against centred finite
    differences of the reference's own kernels, compiled from its sources."""
    setup, lines, a = smooth_case
    b = _backward(setup, lines, a)
    ro = Oracle(setup).pass1(lines)
    win = (ro["flags"], ro["ileft"], ro["iright"])
    l0 = _loss_reference(setup, lines, a, win)
    assert abs(l0 - b["L"]) <= 1e-9 * np.sum(np.abs(a))  # same forward
    ggf = b["ggf"]
    for i in np.argsort(-np.abs(ggf))[:4]:
        out, lg = [], []
        for sgn in (1, -1):
            l2 = lines.take(np.arange(lines.n))
            gf = l2.gf.copy()
            gf[i] = np.float32(gf[i] * 10.0 ** (sgn * 1e-3))
            lg.append(np.log10(np.float64(gf[i])))
            l2.gf = gf
            out.append(_loss_reference(setup, l2, a, win))
        fd = (out[0] - out[1]) / (lg[0] - lg[1])
        assert abs(ggf[i] - fd) <= REL * abs(fd), (i, ggf[i], fd)
    gt, h = b["gt"], 1e-4
    for k in np.argsort(-np.abs(gt))[:3]:
        out = []
        for sgn in (1, -1):
            t = setup.temp.copy()
            t[k] += sgn * h
            out.append(_loss_reference(dataclasses.replace(setup, temp=t), lines, a, win))
        fd = (out[0] - out[1]) / (2 * h)
        assert abs(gt[k] - fd) <= REL * abs(fd), (k, gt[k], fd)


def test_torch_two_forwards_then_backward(case):
    """Two forwards on one context before one backward: each output's gradient is taken at
    its own log gf (the context is brought back to the first forward's state)."""
    torch = pytest.importorskip("torch")
    from fastspec_amd.autograd import normalised_flux
    setup, lines, a = case
    x0 = np.log10(lines.gf.astype(np.float64))
    x1 = x0.copy()
    strong = np.argsort(-x0)[:30]
    x1[strong] += 0.4
    t0 = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    t1 = torch.tensor(x1, dtype=torch.float64, requires_grad=True)
    w = torch.as_tensor(a)
    with _gpu(setup) as g:
        dl = g.upload(lines)
        f0 = normalised_flux(g, dl, t0)
        f1 = normalised_flux(g, dl, t1)
        ((w * f0).sum() + 2.0 * (w * f1).sum()).backward()
        dl.free()
    for x, t, scale in ((x0, t0, 1.0), (x1, t1, 2.0)):
        l2 = lines.take(np.arange(lines.n))
        l2.gf = torch.pow(torch.tensor(10.0, dtype=torch.float32), torch.tensor(x).float()).numpy()
        b = _backward(setup, l2, scale * a)
        assert _rel(t.grad.numpy(), b["ggf"]) < 1e-10


def test_reverse_mode_multi_chunk_equals_one_chunk(case):
    """The reverse pass rebuilds each chunk's records from the list (per-block
    checkpointing): chunks of 37 lines give the one-chunk gradients (FP64 sums in another
    order only)."""
    setup, lines, a = case
    one = _backward(setup, lines, a)
    with _gpu(setup) as g:
        dl = g.upload(lines)
        st = g.lineop_dev(dl, 37)
        assert st["nblocks"] == -(-st["nlines_use"] // 37) and st["nblocks"] >= 3
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
        _, _, gt = g.flux_vjp(a / sfc, -a * sfl / sfc**2)
        ggf, gfpf, gtl = g.lineop_vjp(dl, 37)
        dl.free()
    assert _rel(ggf, one["ggf"]) < 1e-10
    assert _rel(gt + gtl, one["gt"]) < 1e-10
    # a species' dL/d ln FractionPF sums the signed contributions of all its lines, added per
    # chunk: the order changes the last digits (measured 1.8e-10)
    assert _rel(gfpf, one["gfpf"]) < 1e-9


@pytest.fixture(scope="module")
def c5_atmospheres():
    """Three atmospheres of the differentiable batch (coolest, middle, hottest of the 256)
    with their own windows and the shared list's lines that reach them."""
    lines, models = W.config_c5()
    teff = np.array([s.meta["atmosphere"].teff for s, _ in models])
    pick = [int(np.argmin(teff)), int(np.argsort(teff)[128]), int(np.argmax(teff))]
    out = []
    for m in pick:
        setup, idx = models[m]
        sub = lines.take(idx)
        a = np.random.default_rng(100 + m).normal(size=setup.nwl)
        out.append((setup, sub, a))
    return out


def _backward_chunks(setup, lines, a, nlim):
    with _gpu(setup) as g:
        dl = g.upload(lines)
        g.lineop_dev(dl, nlim)
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
        _, _, gt = g.flux_vjp(a / sfc, -a * sfl / sfc**2)
        ggf, gfpf, gtl = g.lineop_vjp(dl, nlim)
        dl.free()
    return dict(gt=gt + gtl, ggf=ggf, gfpf=gfpf, L=float(np.sum(a * sfl / sfc)))


def _loss_chunks(setup, lines, a, nlim):
    with _gpu(setup) as g:
        g.LineopCalc(lines, NlinesLim=nlim)
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
    return float(np.sum(a * sfl / sfc))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_c5_atmosphere_gradients_vs_finite_differences(c5_atmospheres, k):
    """One atmosphere of the differentiable batch, its lines in chunks of a third of the list
    (forward and reverse): dL/d log gf, dL/dT and dL/d ln FractionPF against centred finite
    differences of the forward."""
    setup, lines, a = c5_atmospheres[k]
    nlim = max(lines.n // 3, 1)
    b = _backward_chunks(setup, lines, a, nlim)
    ggf = b["ggf"]
    for i in np.argsort(-np.abs(ggf))[:6]:
        out, lg = [], []
        for sgn in (1, -1):
            l2 = lines.take(np.arange(lines.n))
            gf = l2.gf.copy()
            gf[i] = np.float32(gf[i] * 10.0 ** (sgn * 1e-3))
            lg.append(np.log10(np.float64(gf[i])))
            l2.gf = gf
            out.append(_loss_chunks(setup, l2, a, nlim))
        fd = (out[0] - out[1]) / (lg[0] - lg[1])
        assert abs(ggf[i] - fd) <= REL * abs(fd) + 1e-9 * np.abs(ggf).max(), (i, ggf[i], fd)
    # T and FractionPF on the smooth version (every line retained, windows spanning the grid)
    smooth = dataclasses.replace(setup, eps_opac=1e-12)
    bs = _backward_chunks(smooth, lines, a, nlim)
    gt, h = bs["gt"], 1e-4
    for kk in np.argsort(-np.abs(gt))[:4]:
        out = []
        for sgn in (1, -1):
            t = smooth.temp.copy()
            t[kk] += sgn * h
            out.append(_loss_chunks(dataclasses.replace(smooth, temp=t), lines, a, nlim))
        fd = (out[0] - out[1]) / (2 * h)
        assert abs(gt[kk] - fd) <= REL * abs(fd) + 1e-6 * np.abs(gt).max(), (kk, gt[kk], fd)
    g = bs["gfpf"]
    for slot in np.unique(lines.spc):
        col = slot - 1
        fd, dlog = [], []
        for sgn in (1, -1):
            f = smooth.fraction_pf.copy()
            f[:, col] = (f[:, col] * np.exp(sgn * 1e-3)).astype(np.float32)
            dlog.append(np.log(f[:, col].astype(np.float64) / smooth.fraction_pf[:, col]))
            fd.append(_loss_chunks(dataclasses.replace(smooth, fraction_pf=f), lines, a, nlim))
        an = float(np.sum(g[:, col] * dlog[0])) - float(np.sum(g[:, col] * dlog[1]))
        assert abs(an - (fd[0] - fd[1])) <= REL * abs(fd[0] - fd[1]), (slot, an, fd[0] - fd[1])
