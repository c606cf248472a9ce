"""Pins the oracle's convolution and disk integration (IntegConvol, DiskInteg, ConvolGau,
SemiArea, MuQuadrature radii; src/convol.f90, src/utils.f90) against the reference's own
compiled routines (oracle/_ref) and against closed forms: normalised kernels, the projected
area of the whole disk for the rotational kernels, the Gaussian and the radial-tangential
mixture at disk centre."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from oracle_bridge import Oracle, oracle_lib, ref_convolgau, ref_diskinteg, ref_lib

dp = C.POINTER(C.c_double)
VC = 299792.458


def _p(a):
    return a.ctypes.data_as(dp)


def _setup(wmode: int, nmu: int = 7, step: float | None = None):
    atm = W.grey_atmosphere(5777.0, 4.44, ntau=24)
    if wmode == 0:
        return W.build_setup(atm, 5000.0, 5012.0, step or 0.01, 0, nmu=nmu)
    return W.build_setup(atm, 5000.0, 5012.0, step or 0.6, 1, nmu=nmu)


def _spectra(setup, seed=0):
    """Smooth intensities with absorption lines, different per mu: [nwl][nmu]."""
    rng = np.random.default_rng(seed)
    x = np.arange(setup.nwl, dtype=np.float64)
    spc = np.empty((setup.nwl, setup.nmu))
    spl = np.empty_like(spc)
    for m in range(setup.nmu):
        c = 1.0 + 0.3 * setup.xmu[m] + 1e-4 * x
        depth = np.zeros(setup.nwl)
        for _ in range(12):
            x0, w, d = rng.uniform(0, setup.nwl), rng.uniform(0.5, 8.0), rng.uniform(0.1, 0.8)
            depth += d * np.exp(-0.5 * ((x - x0) / w) ** 2)
        spc[:, m] = c
        spl[:, m] = c * np.exp(-depth * (0.5 + setup.xmu[m]))
    return spl, spc


@pytest.mark.ref
def test_semi_area_and_radii_match_reference():
    lib, ref = oracle_lib(), ref_lib()
    for r in (0.0, 0.3, 1.0, 17.5):
        for v in (-30.0, -1.0, -0.2, 0.0, 0.11, 0.9, 1.0, 25.0):
            assert lib.fso_semi_area(r, v) == pytest.approx(ref.ref_semi_area(r, v), abs=1e-15, rel=1e-15)
    for quad in (0, 1, 2):
        for n in (1, 2, 3, 7, 12):
            x = np.sqrt(np.linspace(1.0, 0.05, n)) if n > 1 else np.ones(1)
            wo, ro = np.zeros(n), np.zeros(n + 1)
            wr, rr = np.zeros(n), np.zeros(n + 1)
            xo, xr = x.copy(), x.copy()
            lib.fso_mu_quadrature_r(quad, n, 0, _p(xo), _p(wo), _p(ro))
            ref.ref_mu_quadrature_r(quad, n, _p(xr), _p(wr), _p(rr))
            assert np.allclose(wo, wr, rtol=1e-14, atol=1e-15) and np.allclose(ro, rr, rtol=1e-14, atol=1e-15)
            # radii: increasing from 0 to 1, annulus areas equal to the weights
            assert ro[0] == 0.0 and ro[-1] == 1.0 and np.all(np.diff(ro) > 0)
            assert np.allclose(np.diff(ro**2), wo, rtol=1e-12)


def test_semi_area_closed_forms():
    lib = oracle_lib()
    r = 2.5
    assert lib.fso_semi_area(r, r) - lib.fso_semi_area(r, -r) == pytest.approx(np.pi * r * r / 2, rel=1e-15)
    assert lib.fso_semi_area(r, 0.0) == 0.0
    assert lib.fso_semi_area(r, 1.0) == pytest.approx(-lib.fso_semi_area(r, -1.0), rel=1e-15)  # odd in V
    assert lib.fso_semi_area(0.0, 1.0) == 0.0


@pytest.mark.ref
@pytest.mark.parametrize("wmode", [0, 1])
@pytest.mark.parametrize("fwhm", [0.05, 3.0, 40.0, 1e5])
def test_convolgau_matches_reference(wmode, fwhm):
    setup = _setup(wmode)
    spl, _ = _spectra(setup, 1)
    s = np.ascontiguousarray(spl[:, 0])
    wlast = setup.wfirst + setup.wave[-1]
    mine = s.copy()
    oracle_lib().fso_convolgau(setup.nwl, _p(mine), fwhm, wmode, setup.wfirst, wlast, setup.dstep)
    theirs = ref_convolgau(setup, s, fwhm)
    assert np.max(np.abs(mine - theirs)) <= 1e-14 * np.max(np.abs(theirs))


def test_convolgau_is_a_normalised_gaussian():
    setup = _setup(1, step=0.5)  # 0.5 km/s per point
    n = setup.nwl
    d = np.zeros(n)
    d[n // 2] = 1.0
    fwhm = 6.0
    wlast = setup.wfirst + setup.wave[-1]
    oracle_lib().fso_convolgau(n, _p(d), fwhm, 1, setup.wfirst, wlast, setup.dstep)
    sig = fwhm / 0.5 / (2 * np.sqrt(2 * np.log(2)))
    k = np.arange(-int(10 * sig), int(10 * sig) + 1)
    g = np.exp(np.maximum(-20.0, -0.5 * (k / sig) ** 2))
    g /= g.sum()
    assert np.allclose(d[n // 2 + k], g[::-1], rtol=1e-13, atol=1e-17)
    assert d.sum() == pytest.approx(1.0, rel=1e-14)
    # a constant spectrum stays constant (edges padded with the end values)
    c = np.full(n, 2.5)
    oracle_lib().fso_convolgau(n, _p(c), fwhm, 1, setup.wfirst, wlast, setup.dstep)
    assert np.allclose(c, 2.5, rtol=1e-14)


@pytest.mark.ref
@pytest.mark.parametrize("wmode,quad,nmu", [(0, 1, 7), (1, 0, 5), (1, 2, 6), (0, 0, 1), (1, 1, 3)])
@pytest.mark.parametrize("vsini,vmacro,vgau", [(0.0, 0.0, 0.0), (12.0, 0.0, 0.0), (0.0, 4.0, 0.0),
                                               (35.0, 3.0, 5.0), (0.2, 0.0, 2.0), (8.0, 2.5, 0.0)])
def test_diskinteg_matches_reference(wmode, quad, nmu, vsini, vmacro, vgau):
    setup = _setup(wmode, nmu)
    spl, spc = _spectra(setup, 2)
    o = Oracle(setup)
    o.set_spectra(spl, spc)
    sfl, sfc = o.integ_convol(2, vsini, vmacro, vgau, quad)
    rl, rc = ref_diskinteg(setup, quad, vsini, vmacro, vgau, spl, spc)
    assert np.max(np.abs(sfl - rl)) <= 1e-13 * np.max(np.abs(rl))
    assert np.max(np.abs(sfc - rc)) <= 1e-13 * np.max(np.abs(rc))


@pytest.mark.parametrize("quad", [0, 1, 2])
def test_rotation_kernels_sum_to_the_disk(quad):
    """A dark line as a single point in a unit continuum, the same at every mu, no
    macroturbulence: the disk-integrated line is sum_i W_i Rkern_i, the projected area of the
    whole disk between the velocities of each pixel's edges over half the disk area."""
    setup = _setup(1, 7, step=0.5)
    n = setup.nwl
    x0 = n // 2
    spl = np.ones((n, setup.nmu))
    spl[x0, :] = 0.0
    spc = np.ones((n, setup.nmu))
    o = Oracle(setup)
    o.set_spectra(spl, spc)
    vsini = 20.0
    sfl, sfc = o.integ_convol(2, vsini, 0.0, 0.0, quad)
    assert np.allclose(sfc, 1.0, rtol=1e-14)
    lib = oracle_lib()
    dv = setup.dstep
    k = np.arange(-50, 51)
    area = np.array([lib.fso_semi_area(vsini, abs(v) * dv + 0.5 * dv) - lib.fso_semi_area(vsini, abs(v) * dv - 0.5 * dv)
                     for v in k])
    prof = area / (np.pi * vsini**2 / 2)
    assert np.allclose(1.0 - sfl[x0 + k], prof, rtol=1e-12, atol=1e-14)


def test_macroturbulence_at_disk_centre():
    """One angle at mu = 1: the tangential part is a delta, the radial part a Gaussian of
    sigma = Vmacro / Dv / sqrt 2 points; each carries half of the weight."""
    setup = dataclasses.replace(_setup(1, 1, step=0.5), xmu=np.array([1.0]), wmu=np.array([1.0]))
    n = setup.nwl
    x0 = n // 2
    spl = np.zeros((n, 1))
    spl[x0, 0] = 1.0
    o = Oracle(setup)
    o.set_spectra(spl, np.ones((n, 1)))
    vmacro = 3.0
    sfl, _ = o.integ_convol(2, 0.0, vmacro, 0.0, 1)
    sig = vmacro / setup.dstep / np.sqrt(2.0)
    nmk = max(3, min(int(10 * sig), (n - 3) // 2))
    k = np.arange(-nmk, nmk + 1)
    g = np.exp(np.maximum(-20.0, -0.5 * (k / sig) ** 2))
    g /= g.sum()
    expect = 0.5 * g
    expect[nmk] += 0.5
    assert np.allclose(sfl[x0 + k], expect[::-1], rtol=1e-13, atol=1e-17)


def test_intensity_mode_broadens_every_spectrum():
    setup = _setup(0, 4)
    spl, spc = _spectra(setup, 3)
    o = Oracle(setup)
    o.set_spectra(spl, spc)
    o.integ_convol(1, 0.0, 0.0, 4.0, 1)
    gl, gc = o.spectra()
    wlast = setup.wfirst + setup.wave[-1]
    for m in range(setup.nmu):
        e = np.ascontiguousarray(spl[:, m])
        oracle_lib().fso_convolgau(setup.nwl, _p(e), 4.0, 0, setup.wfirst, wlast, setup.dstep)
        assert np.array_equal(gl[:, m], e)
    # polarised: the four Stokes spectra and the continuum
    spz = np.stack([spl, 0.1 * spl, -0.05 * spl, 0.02 * spc], axis=1)
    o.set_spectra(None, spc, spz)
    o.integ_convol(1, 0.0, 0.0, 4.0, 1)
    gz, gc2 = o.spectra(magnetic=True)
    e = np.ascontiguousarray(spz[:, 2, 1])
    oracle_lib().fso_convolgau(setup.nwl, _p(e), 4.0, 0, setup.wfirst, wlast, setup.dstep)
    assert np.array_equal(gz[:, 2, 1], e)
    assert np.array_equal(gc2, gc)
