"""Closed-form pins of two oracle parts that the reference's compiled kernels cannot pin here
(their driver routines are in synth.f90, which needs the HDF5 Fortran module):

* LinStrCalc (src/synth.f90:1774-1822): central line depth of one line-centre column;
* the plane-parallel field geometry of MagRadTranCalc (src/synth.f90:1930-1994) and what it
  does to the Stokes spectra.

Each check is chosen so that a plausible slip fails it: the flux weights dropped, the
single-angle branch taken for several angles, line and continuum swapped, sin and cos or
the sign of a projection exchanged, the azimuth not doubled."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from oracle_bridge import Oracle, oracle_lib

dp = C.POINTER(C.c_double)

# hc/k [A K] from the exact SI constants, and the reference's 2hc^2 factor of the Planck
# function in its units (the source function is an input here, not what is tested)
HC_K = 6.62607015e-27 * 299792.458e13 / 1.380649e-16


def _planck_c2(wl):
    return np.exp(50.7649141 - 5.0 * np.log(wl))


def _setup_with(temp, rhox, xmu, wmu):
    atm = W.grey_atmosphere(ntau=len(temp))
    s = W.build_setup(atm, 5000.0, 5001.0, 0.01, 0, nmu=len(xmu))
    return dataclasses.replace(s, temp=np.asarray(temp, np.float64), rhox=np.asarray(rhox, np.float64),
                               xmu=np.asarray(xmu, np.float64), wmu=np.asarray(wmu, np.float64))


def _planck(temp, wl):
    return _planck_c2(wl) / np.expm1(HC_K / wl / temp)


def _linear_source(tau, t_top, t_bottom, wl=None):
    """Coefficients a, b of a source S = a + b tau_c that runs from the Planck function at
    t_top (tau = 0) to that at t_bottom (the slab's bottom): physical intensities, so that
    the temperatures are stellar ones."""
    wl = WL if wl is None else wl
    a = _planck(t_top, wl)
    b = (_planck(t_bottom, wl) - a) / tau[-1]
    return a, b


def _temp_for_source(src, wl):
    """T at which the Planck function at wl equals src (inverse Planck)."""
    return HC_K / wl / np.log1p(_planck_c2(wl) / src)


def _lin_str(setup, wl, lop, cop):
    o = Oracle(setup)
    a = np.ascontiguousarray(lop, np.float64)
    b = np.ascontiguousarray(cop, np.float64)
    v = oracle_lib().fso_lin_str(o.ctx, wl, a.ctypes.data_as(dp), b.ctypes.data_as(dp))
    o.close()
    return float(v)


def _slab(ntau=72, kc=0.8):
    """Column masses with optical-depth steps between 0.01 and 1 (no Taylor branch issues,
    no saturation) down to tau_c = 60, and a constant continuum opacity."""
    tau = np.concatenate([[0.0], np.cumsum(np.geomspace(0.01, 1.9, ntau - 1))])
    tau = tau / tau[-1] * 60.0
    rhox = 1e-3 + tau / kc
    return rhox, np.full(ntau, kc), tau


def _intensity_linear(a, b, mu, T):
    """Exact emergent intensity of S = a + b t in a slab of optical thickness T with
    I = S(T) entering at the bottom: a + b mu - b mu exp(-T/mu)."""
    return a + b * mu - b * mu * np.exp(-T / mu)


WL = 5123.4


@pytest.mark.parametrize("nmu", [7, 3])
@pytest.mark.parametrize("eta", [0.3, 4.0, 50.0])
def test_linstr_constant_ratio_linear_source(nmu, eta):
    """Line/continuum ratio eta constant with depth and S linear in tau_c: the emergent
    intensities are exact for the DELO-Bezier scheme, so the line depth is
    1 - sum W (a + b' mu - b' mu e^{-T_l/mu}) / sum W (a + b mu - b mu e^{-T_c/mu}), with
    b' = b / (1 + eta), T_l = (1 + eta) T_c."""
    xmu, wmu = W.gauss_mu(nmu)
    rhox, kc, tau = _slab()
    a, b = _linear_source(tau, 4500.0, 9000.0)
    src = a + b * tau
    setup = _setup_with(_temp_for_source(src, WL), rhox, xmu, wmu)
    got = _lin_str(setup, WL, eta * kc, kc)
    Tc = tau[-1]
    ic = _intensity_linear(a, b, xmu, Tc)
    il = _intensity_linear(a, b / (1 + eta), xmu, (1 + eta) * Tc)
    exp = 1.0 - np.sum(wmu * il) / np.sum(wmu * ic)
    assert abs(got - exp) <= 2e-7 * max(1.0, abs(exp)), (got, exp)
    # the flux weights matter: an unweighted mean over mu, or the first angle alone,
    # differ far beyond the tolerance
    assert abs(got - (1.0 - np.mean(il) / np.mean(ic))) > 1e-4
    assert abs(got - (1.0 - il[0] / ic[0])) > 1e-4
    assert 0.0 < got < 1.0


def test_linstr_single_angle_branch():
    """Nmu = 1: the depth at that angle, 1 - I_l(mu)/I_c(mu), whatever its weight."""
    rhox, kc, tau = _slab()
    a, b = _linear_source(tau, 4000.0, 12000.0)
    eta = 2.5
    src = a + b * tau
    mu = np.array([0.61])
    setup = _setup_with(_temp_for_source(src, WL), rhox, mu, np.array([0.37]))
    got = _lin_str(setup, WL, eta * kc, kc)
    exp = 1.0 - (_intensity_linear(a, b / (1 + eta), mu, (1 + eta) * tau[-1]) /
                 _intensity_linear(a, b, mu, tau[-1]))[0]
    assert abs(got - exp) <= 2e-7, (got, exp)


def test_linstr_zero_line_and_constant_source():
    xmu, wmu = W.gauss_mu(7)
    rhox, kc, tau = _slab()
    a, b = _linear_source(tau, 5000.0, 8000.0)
    src = a + b * tau
    setup = _setup_with(_temp_for_source(src, WL), rhox, xmu, wmu)
    assert _lin_str(setup, WL, np.zeros_like(kc), kc) == 0.0
    # a constant source (isothermal slab): every intensity equals S, the depth is zero
    # whatever the line opacity and its depth dependence
    setup = _setup_with(np.full(len(tau), 5400.0), rhox, xmu, wmu)
    lop = 3.0 * kc * np.linspace(0.1, 2.0, len(tau))
    assert abs(_lin_str(setup, WL, lop, kc)) <= 1e-7


def test_linstr_emission_line_is_negative():
    """S decreasing inward (a temperature inversion): the line forms higher, where S is
    larger, and the central depth turns negative; exact value as above with b < 0."""
    xmu, wmu = W.gauss_mu(7)
    rhox, kc, tau = _slab()
    a, b = _linear_source(tau, 7000.0, 5000.0)
    eta = 1.0
    src = a + b * tau
    assert src.min() > 0
    setup = _setup_with(_temp_for_source(src, WL), rhox, xmu, wmu)
    got = _lin_str(setup, WL, eta * kc, kc)
    il = _intensity_linear(a, b / (1 + eta), xmu, (1 + eta) * tau[-1])
    ic = _intensity_linear(a, b, xmu, tau[-1])
    exp = 1.0 - np.sum(wmu * il) / np.sum(wmu * ic)
    assert exp < 0 and abs(got - exp) <= 2e-7 * max(1.0, abs(exp))


# ---- plane-parallel field geometry ---------------------------------------------------------

def _geometry(xmu, bfield, bframe):
    m = len(xmu)
    out = [np.zeros(m) for _ in range(5)]
    x = np.ascontiguousarray(xmu, np.float64)
    b = np.ascontiguousarray(bfield, np.float64)
    rc = oracle_lib().fso_field_geometry(m, x.ctypes.data_as(dp), b.ctypes.data_as(dp), len(b), bframe,
                                         *(o.ctypes.data_as(dp) for o in out))
    assert rc == 0
    return dict(zip(("SPhi2", "CPhi2", "CPhi", "S2Psi", "C2Psi"), out))


XMU = W.gauss_mu(7)[0]


@pytest.mark.parametrize("incl,azim", [(45.0, 30.0), (0.0, 0.0), (90.0, 45.0), (120.0, -70.0)])
def test_bframe1_observer_frame(incl, azim):
    """BFRAME 1: the same projection at every mu; cos gamma = cos(inclination), the azimuth
    enters doubled."""
    g = _geometry(XMU, [2000.0, incl, azim], 1)
    r = np.deg2rad
    for k, v in dict(CPhi=np.cos(r(incl)), CPhi2=np.cos(r(incl)) ** 2, SPhi2=np.sin(r(incl)) ** 2,
                     S2Psi=np.sin(2 * r(azim)), C2Psi=np.cos(2 * r(azim))).items():
        assert np.allclose(g[k], v, rtol=0, atol=1e-14), k
    assert np.allclose(g["SPhi2"] + g["CPhi2"], 1.0, atol=1e-15)


def test_bframe0_radial_field():
    """A radial surface field seen at heliocentric angle theta (mu = cos theta): gamma =
    theta, so cos gamma = mu; the sky projection lies along the radius vector, psi = 0."""
    g = _geometry(XMU, [1500.0, 0.0, 0.0], 0)
    assert np.allclose(g["CPhi"], XMU, atol=1e-15)
    assert np.allclose(g["SPhi2"], 1 - XMU ** 2, atol=1e-15)
    assert np.allclose(g["S2Psi"], 0.0, atol=1e-15) and np.allclose(g["C2Psi"], 1.0, atol=1e-15)
    # a negative radial field: gamma = pi - theta
    g = _geometry(XMU, [-1500.0, 0.0, 0.0], 0)
    assert np.allclose(g["CPhi"], -XMU, atol=1e-15)


def test_bframe0_meridional_and_tangential():
    """The second component is perpendicular to both the radius and the line of sight:
    gamma = 90 deg at every mu and psi = 90 deg.  The third lies in the plane of the radius
    and the line of sight: transverse at disk centre (gamma = 90, psi = 0) and tilted by
    theta away from the sky plane elsewhere, |cos gamma| = sin theta."""
    g = _geometry(XMU, [0.0, 800.0, 0.0], 0)
    assert np.allclose(g["CPhi"], 0.0, atol=1e-15) and np.allclose(g["SPhi2"], 1.0, atol=1e-15)
    assert np.allclose(g["S2Psi"], 0.0, atol=1e-15) and np.allclose(g["C2Psi"], -1.0, atol=1e-15)
    g = _geometry(np.array([1.0, 0.5]), [0.0, 0.0, 800.0], 0)
    assert abs(g["CPhi"][0]) <= 1e-15 and abs(g["C2Psi"][0] - 1.0) <= 1e-15
    assert abs(abs(g["CPhi"][1]) - np.sqrt(0.75)) <= 1e-15
    assert abs(g["C2Psi"][1] - 1.0) <= 1e-15


def test_bframe0_is_a_rotation():
    """For any surface vector, the observer-frame field at each mu is a rotation of it:
    cos gamma equals the projection on the line of sight of the vector rotated by theta
    about the meridional axis, and the angles are consistent (sin^2 + cos^2 = 1, the
    doubled azimuth on the unit circle)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        b = rng.normal(size=3) * 1000.0
        g = _geometry(XMU, b, 0)
        mod = np.linalg.norm(b)
        st = np.sqrt(1 - XMU ** 2)
        # rotate (radial, meridional, third) so that the radial axis makes angle theta with
        # the line of sight: line-of-sight component r mu - t sin(theta)
        los = (b[0] * XMU - b[2] * st) / mod
        assert np.allclose(g["CPhi"], los, atol=1e-14)
        assert np.allclose(g["SPhi2"] + g["CPhi2"], 1.0, atol=1e-14)
        assert np.allclose(g["S2Psi"] ** 2 + g["C2Psi"] ** 2, 1.0, atol=1e-13)
        # the sky-plane components (r sin(theta) + t mu, m) give the azimuth
        bx, by = b[0] * st + b[2] * XMU, b[1]
        psi = np.arctan2(by, bx)
        assert np.allclose(g["C2Psi"], np.cos(2 * psi), atol=1e-13)
        assert np.allclose(g["S2Psi"], np.sin(2 * psi), atol=1e-13)


def test_bframe2_per_mu_angles():
    m = len(XMU)
    ang = np.empty(1 + 2 * m)
    ang[0] = 3000.0
    ang[1::2] = np.linspace(10.0, 170.0, m)
    ang[2::2] = np.linspace(-80.0, 80.0, m)
    g = _geometry(XMU, ang, 2)
    assert np.allclose(g["CPhi"], np.cos(np.deg2rad(ang[1::2])), atol=1e-15)
    assert np.allclose(g["C2Psi"], np.cos(2 * np.deg2rad(ang[2::2])), atol=1e-14)
    assert np.allclose(g["S2Psi"], np.sin(2 * np.deg2rad(ang[2::2])), atol=1e-14)


# ---- the geometry's effect on the Stokes spectra (the oracle's MagRadTranCalc) ------------------

@pytest.fixture(scope="module")
def mag_case():
    atm = W.grey_atmosphere(5777.0, 4.44)
    setup = W.build_setup(atm, 6000.0, 6004.0, 0.005, 0)
    lines = W.magnetic_lines(12, 6000.5, 6003.5, 4242)
    return setup, lines


def _stokes(setup, lines, bfield, bframe):
    o = Oracle(setup)
    o.maglineop(lines, bfield, bframe)
    spz, spc = o.magradtran(bfield, bframe)
    o.close()
    return spz, spc   # spz [nwl][4][nmu]: I, Q, U, V


def test_stokes_longitudinal_and_transverse(mag_case):
    setup, lines = mag_case
    # B along the line of sight: no linear polarisation, V present
    spz, spc = _stokes(setup, lines, [2000.0, 0.0, 30.0], 1)
    assert np.max(np.abs(spz[:, 1:3, :])) <= 1e-12 * spc.max()
    assert np.max(np.abs(spz[:, 3, :])) > 1e-3 * spc.max()
    # B in the plane of the sky: no circular polarisation, Q present
    spz, spc = _stokes(setup, lines, [2000.0, 90.0, 0.0], 1)
    assert np.max(np.abs(spz[:, 3, :])) <= 1e-12 * spc.max()
    assert np.max(np.abs(spz[:, 1, :])) > 1e-4 * spc.max()
    assert np.max(np.abs(spz[:, 2, :])) <= 1e-12 * spc.max()


def test_stokes_azimuth_rotation(mag_case):
    """Turning the field azimuth by 45 deg turns Q into U (Q ~ cos 2 psi, U ~ sin 2 psi) and
    leaves I and V unchanged; by 90 deg it reverses Q."""
    setup, lines = mag_case
    z0, c0 = _stokes(setup, lines, [2000.0, 60.0, 0.0], 1)
    z45, _ = _stokes(setup, lines, [2000.0, 60.0, 45.0], 1)
    z90, _ = _stokes(setup, lines, [2000.0, 60.0, 90.0], 1)
    tol = 1e-12 * c0.max()
    assert np.max(np.abs(z45[:, 0] - z0[:, 0])) <= tol and np.max(np.abs(z45[:, 3] - z0[:, 3])) <= tol
    assert np.max(np.abs(z45[:, 2] - z0[:, 1])) <= tol
    assert np.max(np.abs(z45[:, 1] + z0[:, 2])) <= tol
    assert np.max(np.abs(z90[:, 1] + z0[:, 1])) <= tol
    assert np.max(np.abs(z0[:, 1])) > 1e-4 * c0.max()


def test_stokes_bframe0_equals_bframe2_of_its_projection(mag_case):
    """A surface vector (BFRAME 0) gives the spectra of BFRAME 2 with the per-mu angles of
    its rotation into the observer frame, computed independently here."""
    setup, lines = mag_case
    b = np.array([1200.0, -700.0, 900.0])
    st = np.sqrt(1 - setup.xmu ** 2)
    bz = b[0] * setup.xmu - b[2] * st
    bx, by = b[0] * st + b[2] * setup.xmu, b[1]
    mod = np.linalg.norm(b)
    gam = np.degrees(np.arccos(bz / mod))
    psi = np.degrees(np.arctan2(by, bx))
    ang = np.empty(1 + 2 * setup.nmu)
    ang[0] = mod
    ang[1::2] = gam
    ang[2::2] = psi
    z0, c0 = _stokes(setup, lines, b, 0)
    z2, _ = _stokes(setup, lines, ang, 2)
    assert np.max(np.abs(z0 - z2)) <= 1e-9 * c0.max()
