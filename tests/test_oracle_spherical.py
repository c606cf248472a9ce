"""Pins the oracle's spherical transfer (SphereInit, LocPrfSph/DeloScalar, StokesPrfSph and
the ray field geometry) against the reference's compiled solver and closed forms."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from oracle_bridge import Oracle, Ref, oracle_lib

DP = C.POINTER(C.c_double)
PLK = 6.62607015e-27 * 299792.458 * 1e13 / 1.380649e-16


def giant(wl_hi=6004.0, radius=7e12, nmu=7):
    """An extended grey giant (Teff 4000 K, log g 1.5) on 100 R_sun: the two outermost
    Gauss angles are grazing rays, the last one tangent in the second layer."""
    atm = W.grey_atmosphere(4000.0, 1.5)
    return W.build_setup(atm, 6000.0, wl_hi, 0.005, 0, radius=radius, nmu=nmu)


def planck(setup, p):
    lam = setup.wave[p] + setup.wfirst
    return np.exp(50.7649141 - 5 * np.log(lam)) / (np.exp(PLK / lam / setup.temp) - 1)


@pytest.fixture(scope="module")
def sph():
    return giant()


def test_ray_geometry_invariants(sph):
    g = Oracle(sph).sphere_geometry()
    N = sph.ntau
    R = sph.radius + sph.height
    P = R[0] * np.sqrt(1 - sph.xmu**2)
    assert np.array_equal(g["graz"], (P > R[-1]).astype(np.int32))
    assert list(g["graz"]) == [0, 0, 0, 0, 0, 1, 1]
    assert g["ntau_mu"][-1] == 4
    for im in range(sph.nmu):
        n = g["ntau_mu"][im]
        rm, mp, ml = g["rhox_mu"][im, :n], g["map"][im, :n], g["muloc"][im, :n]
        assert np.all(np.diff(rm) > 0)                       # column mass grows along the ray
        assert abs(rm[0] - sph.rhox[0] / sph.xmu[im]) < 1e-15 * rm[0]
        if g["graz"][im]:
            h = n // 2
            assert np.array_equal(mp[h:], mp[:h][::-1])     # far half mirrors the near half
            assert np.allclose(ml[h:], -ml[:h][::-1])
            assert R[h - 1] > P[im] >= R[h]                  # tangent shell below layer h
        else:
            assert n == N and np.array_equal(mp, np.arange(1, N + 1))
            # far from the limb the ray's column mass is close to rhox / mu
            if sph.xmu[im] > 0.8:
                assert np.allclose(rm, sph.rhox / sph.xmu[im], rtol=5e-3)


@pytest.mark.ref
def test_ray_geometry_equals_reference(sph):
    g = Oracle(sph).sphere_geometry()
    r = Ref(sph).sphere(sph.radius, sph.height)
    assert np.array_equal(g["ntau_mu"], r["ntau_mu"])
    assert np.array_equal(g["graz"], r["graz"])
    for im in range(sph.nmu):
        n = g["ntau_mu"][im]
        assert np.array_equal(g["map"][im, :n], r["map"][im, :n])
        assert np.allclose(g["rhox_mu"][im, :n], r["rhox_mu"][im, :n], rtol=1e-14, atol=0)
        assert np.allclose(g["muloc"][im, :n], r["muloc"][im, :n], rtol=1e-14, atol=1e-16)


@pytest.mark.ref
def test_locprfsph_equals_reference(sph):
    orc = Oracle(sph)
    ref = Ref(sph)
    ref.sphere(sph.radius, sph.height)
    rng = np.random.default_rng(3)
    for p in (0, 200, 600):
        src = planck(sph, p)
        for scale in (0.0, 1.0, 1e3):
            xki = sph.contop[p] * (1 + scale * rng.random(sph.ntau))
            a, b = orc.locprfsph(xki, src), ref.locprfsph(xki, src)
            assert np.max(np.abs(a - b) / b) < 1e-13, (p, scale)


def test_locprfsph_closed_forms(sph):
    orc = Oracle(sph)
    g = orc.sphere_geometry()
    xki = sph.contop[0]
    # constant S: rays reaching the deepest layer return S exactly
    I = orc.locprfsph(xki, np.full(sph.ntau, 2.5))
    normal = g["graz"] == 0
    assert np.allclose(I[normal], 2.5, rtol=1e-13)
    # grazing rays start from zero: the optically thick one (tangent deep down) still
    # returns S, the one tangent in the second layer sees little matter: 0 < I << S
    assert abs(I[-2] - 2.5) < 1e-9 and 0 < I[-1] < 0.1
    # a large radius approaches the plane-parallel solution, the difference falling in
    # proportion to the atmosphere's extension over the radius
    src = planck(sph, 0)
    ipp = np.zeros(sph.nmu)
    args = [np.ascontiguousarray(x, np.float64) for x in (xki, src, sph.rhox)]
    oracle_lib().fso_locprfpp(ipp.ctypes.data_as(DP), sph.xmu.ctypes.data_as(DP), sph.nmu,
                              *(a.ctypes.data_as(DP) for a in args), sph.ntau)
    err = []
    for radius in (1e15, 1e16):
        orf = Oracle(dataclasses.replace(sph, radius=radius))
        assert orf.sphere_geometry()["graz"].sum() == 0
        err.append(np.max(np.abs(orf.locprfsph(xki, src) - ipp) / ipp))
    assert err[0] < 5e-3 and 7 < err[0] / err[1] < 13, err


def test_mu_errors(sph):
    bad = dataclasses.replace(sph, xmu=np.array([0.9, 1e-6]), wmu=np.array([0.5, 0.5]))
    with pytest.raises(RuntimeError, match="grazing ray tangent in the top model layer"):
        Oracle(bad)
    zero = dataclasses.replace(sph, xmu=np.array([0.9, 0.0]), wmu=np.array([0.5, 0.5]))
    with pytest.raises(RuntimeError, match="mu = 0 rays are unsupported"):
        Oracle(zero)


@pytest.fixture(scope="module")
def sph_mag():
    setup = giant(wl_hi=6003.0)
    lines = W.magnetic_lines(25, 6000.0, 6003.0, 77)
    return setup, lines


@pytest.mark.ref
@pytest.mark.parametrize("bframe,bfield", [(1, [2000.0, 45.0, 30.0]), (0, [1500.0, -800.0, 1200.0])])
def test_stokes_sph_equals_reference(sph_mag, bframe, bfield):
    setup, lines = sph_mag
    bfield = np.array(bfield)
    orc = Oracle(setup)
    orc.maglineop(lines, bfield, bframe)
    spz, spc = orc.magradtran(bfield, bframe)
    mlop = orc.get_maglineop()
    geo = orc.sph_field_geometry(bfield, bframe)
    ref = Ref(setup)
    ref.sphere(setup.radius, setup.height)
    for p in range(0, setup.nwl, 53):
        st = ref.stokesprfsph(p + 1, mlop[p], np.zeros(setup.ntau), planck(setup, p), geo)
        assert np.max(np.abs(st - spz[p]) / spc[p].max()) < 1e-12, p
    # the polarisation signal is there
    assert np.max(np.abs(spz[:, 3, :]) / spc) > 1e-3


def test_stokes_sph_zero_field(sph_mag):
    setup, lines = sph_mag
    b = np.zeros(3)
    orc = Oracle(setup)
    orc.maglineop(lines, b, 0)
    spz, spc = orc.magradtran(b, 0)
    assert np.max(np.abs(spz[:, 1:, :])) == 0.0 or np.max(np.abs(spz[:, 1:, :]) / spz[:, :1, :]) < 1e-12


def test_sph_field_geometry_bframe0_follows_local_mu(sph):
    orc = Oracle(sph)
    g = orc.sphere_geometry()
    geo = orc.sph_field_geometry(np.array([1000.0, 0.0, 0.0]), 0)  # radial field
    for im in range(sph.nmu):
        n = g["ntau_mu"][im]
        # a radial field has cos(gamma) = local ray-to-radial cosine
        assert np.allclose(geo[im, :n, 2], g["muloc"][im, :n], rtol=1e-14)
        assert np.allclose(geo[im, :n, 0] + geo[im, :n, 1], 1.0, rtol=1e-14)
