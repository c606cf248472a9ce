"""GPU parity of IntegConvol / DiskInteg / ConvolGau (fsg_integ_convol) against the FP64
oracle, on the intensities the library computed itself, through the C-ABI."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64
from oracle_bridge import Oracle, oracle_lib

pytestmark = pytest.mark.gpu

REL = 1e-13  # the same sums in the same order, FP64 on both sides


def _gpu(setup, prec=FSG_FP64):
    from fastspec_amd.synth import FastSpecGPU
    return FastSpecGPU(setup, precision=prec)


def _setup(wmode, nmu=7, step=None, quad=1):
    atm = W.grey_atmosphere(5777.0, 4.44, ntau=32)
    if wmode == 0:
        s = W.build_setup(atm, 5000.0, 5006.0, step or 0.01, 0, nmu=nmu)
    else:
        s = W.build_setup(atm, 5000.0, 5006.0, step or 0.6, 1, nmu=nmu)
    if quad == 0 and nmu > 1:  # equal-area annuli: their nodes and weights
        x, w = np.zeros(nmu), np.zeros(nmu)
        dp = C.POINTER(C.c_double)
        oracle_lib().fso_mu_quadrature_r(0, nmu, 1, x.ctypes.data_as(dp), w.ctypes.data_as(dp), None)
        s = dataclasses.replace(s, xmu=x, wmu=w)
    return s


def _lines(setup, seed):
    w1 = setup.wfirst + setup.wave[-1]
    return W.atomic_lines(120, setup.wfirst, w1, seed)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


CASES = [(0.0, 0.0, 0.0), (15.0, 0.0, 0.0), (0.0, 3.5, 0.0), (0.0, 0.0, 4.0), (40.0, 2.5, 6.0), (0.3, 1.0, 0.0)]


@pytest.mark.parametrize("wmode,nmu,quad", [(0, 7, 1), (1, 5, 0), (1, 1, 1), (0, 3, 0)])
@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_diskinteg_matches_oracle(wmode, nmu, quad, prec):
    setup = _setup(wmode, nmu, quad=quad)
    lines = _lines(setup, 11 + nmu)
    o = Oracle(setup)
    with _gpu(setup, prec) as g:
        g.LineopCalc(lines)
        g.RadTranCalc(intensities=False)
        spl, spc = g.get_spectra()
        f0l, f0c = g.flux()
        o.set_spectra(spl, spc)
        for vsini, vmacro, vgau in CASES:
            sfl, sfc = g.IntegConvol(2, vsini, vmacro, vgau)
            rl, rc = o.integ_convol(2, vsini, vmacro, vgau, quad)
            assert _rel(sfl, rl) <= REL and _rel(sfc, rc) <= REL, (vsini, vmacro, vgau, _rel(sfl, rl))
        # no broadening: the fused flux sums of the formal solution
        sfl, sfc = g.IntegConvol(2, 0.0, 0.0, 0.0)
    assert _rel(sfl, f0l) <= 1e-13 and _rel(sfc, f0c) <= 1e-13


def test_long_rotation_kernels_take_the_cached_path():
    """Kernels longer than the LDS-staged limit (3072 taps): 100 km/s at 0.05 km/s steps."""
    setup = _setup(1, 3, step=0.05)
    lines = _lines(setup, 5)
    o = Oracle(setup)
    with _gpu(setup) as g:
        g.LineopCalc(lines)
        g.RadTranCalc(intensities=False)
        spl, spc = g.get_spectra()
        o.set_spectra(spl, spc)
        sfl, sfc = g.IntegConvol(2, 100.0, 2.0, 3.0)
    rl, rc = o.integ_convol(2, 100.0, 2.0, 3.0, 1)
    assert _rel(sfl, rl) <= REL and _rel(sfc, rc) <= REL


def test_intensity_broadening_matches_oracle():
    setup = _setup(0, 4)
    lines = _lines(setup, 21)
    o = Oracle(setup)
    with _gpu(setup) as g:
        g.LineopCalc(lines)
        g.RadTranCalc(intensities=False)
        spl, spc = g.get_spectra()
        o.set_spectra(spl, spc)
        gl, gc = g.IntegConvol(1, 0.0, 0.0, 5.0)
    o.integ_convol(1, 0.0, 0.0, 5.0, 1)
    ol, oc = o.spectra()
    assert _rel(gl, ol) <= REL and _rel(gc, oc) <= REL
    assert _rel(gl, spl) > 1e-4  # the lines did broaden


def test_polarised_spectra():
    setup = _setup(0, 3)
    w1 = setup.wfirst + setup.wave[-1]
    lines = W.magnetic_lines(20, setup.wfirst, w1, 7)
    bfield = np.array([2500.0, 30.0, 40.0])
    o = Oracle(setup)
    with _gpu(setup) as g:
        g.MagLineopCalc(lines, bfield, 1)
        g.MagRadTranCalc(bfield, 1, intensities=False)
        spz, spc = g.get_spectra(magnetic=True)
        o.set_spectra(None, spc, spz)
        sfl, sfc = g.IntegConvol(2, 10.0, 2.0, 3.0)  # disk integration of Stokes I
        rl, rc = o.integ_convol(2, 10.0, 2.0, 3.0, 1)
        assert _rel(sfl, rl) <= REL and _rel(sfc, rc) <= REL
        gz, gc = g.IntegConvol(1, 0.0, 0.0, 4.0)  # all four Stokes spectra and SpC
    o.integ_convol(1, 0.0, 0.0, 4.0, 1)
    oz, oc = o.spectra(magnetic=True)
    assert _rel(gz, oz) <= REL and _rel(gc, oc) <= REL
    assert np.max(np.abs(gz[:, 3, :])) > 0


def test_errors():
    from fastspec_amd._lib import FsgError
    setup = _setup(0, 3)
    with _gpu(setup) as g:
        with pytest.raises(FsgError, match="no formal solution"):
            g.IntegConvol(2, 1.0, 0.0, 0.0)
        g.LineopCalc(_lines(setup, 3))
        g.RadTranCalc(intensities=False)
        with pytest.raises(FsgError, match="invalid value"):
            g.IntegConvol(2, -1.0, 0.0, 0.0)
        with pytest.raises(FsgError, match="CImode"):
            g.IntegConvol(3, 1.0, 0.0, 0.0)
        with pytest.raises(FsgError, match="not polarised"):
            g.get_spectra(magnetic=True)
