"""GPU parity of the polarised path (MagLineopCalc + MagRadTranCalc + Stokes I flux) against
the FP64 oracle on identical synthetic inputs, through the C-ABI."""
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64, LineList
from cases import mixed_list, with_he_lines, with_zeeman
from oracle_bridge import Oracle

pytestmark = pytest.mark.gpu

FP64_OPAC_REL = 1e-7    # FP32 table entries may differ by one ulp from the CPU-built tables
FP64_STOKES = 1e-8      # |dI|, |dQ|, |dU|, |dV| over the continuum intensity
FP64_FLUX_ABS = 1e-8
FP32_OPAC_REL = 5e-4    # relative to |absorption| + |dispersion| + continuum opacity
FP32_STOKES = 1e-4
FP32_FLUX_ABS = 1e-4    # the bound on normalised flux

B_INCLINED = np.array([2000.0, 45.0, 30.0])  # BFRAME 1: |B| [G], inclination, azimuth [deg]


def _gpu(setup, prec, batched=None):
    from fastspec_amd.synth import FastSpecGPU
    return FastSpecGPU(setup, precision=prec, batched=batched)


def _narrow_c4(nlines, wl_hi, seed):
    atm = W.grey_atmosphere(5777.0, 4.44)
    setup = W.build_setup(atm, 6000.0, wl_hi, 0.005, 0)
    lines = W.magnetic_lines(nlines, 6000.0, wl_hi, seed)
    return setup, lines


def _oracle(setup, zlist, bfield, bframe, scalar=None, nlim=None, linstr=False):
    orc = Oracle(setup)
    if scalar is not None:
        orc.lineop(scalar)
    ro = orc.maglineop(zlist, bfield, bframe, nlines_lim=nlim, linstr=linstr)
    mlop = orc.get_maglineop()
    spz, spc = orc.magradtran(bfield, bframe)
    sfl, sfc = orc.flux()
    return dict(setup=setup, lines=zlist, ro=ro, mlop=mlop, spz=spz, spc=spc, sfl=sfl, sfc=sfc,
                bfield=bfield, bframe=bframe, scalar=scalar, nlim=nlim)


def _check(ref, g, opac_rel, stokes_tol, flux_tol):
    setup = ref["setup"]
    if ref["scalar"] is not None:
        g.LineopCalc(ref["scalar"])
    rg = g.MagLineopCalc(ref["lines"], ref["bfield"], ref["bframe"], NlinesLim=ref["nlim"])
    ro = ref["ro"]
    assert np.array_equal(rg["flags"], ro["flags"])
    assert np.array_equal(rg["ileft"], ro["ileft"])
    assert np.array_equal(rg["iright"], ro["iright"])
    for k in ("nlines_use", "nwide", "evals"):
        assert rg["stats"][k] == ro["stats"][k], k
    mlop = g.get_maglineop()
    # each absorption/dispersion pair (pi, blue, red) is measured against its absorption
    # profile plus the continuum: dispersion profiles change sign at every component
    scale = np.abs(ref["mlop"][:, :, :3]) + np.abs(ref["mlop"][:, :, 3:]) + setup.contop[:, :, None]
    den = np.concatenate([scale, scale], axis=2)
    err = np.abs(mlop - ref["mlop"]) / den
    opac = float(err.max())
    assert opac <= opac_rel, (opac, err.max(axis=(0, 1)))
    spz, spc = g.MagRadTranCalc(ref["bfield"], ref["bframe"])
    assert np.max(np.abs(spc - ref["spc"]) / ref["spc"]) <= 1e-9
    st = float(np.max(np.abs(spz - ref["spz"]) / ref["spc"][:, None, :]))
    assert st <= stokes_tol, st
    sfl, sfc = g.flux()
    fe = float(np.max(np.abs(sfl / sfc - ref["sfl"] / ref["sfc"])))
    assert fe <= flux_tol, fe
    # the polarisation is real: V reaches a sizeable fraction of the continuum
    assert np.max(np.abs(ref["spz"][:, 3, :]) / ref["spc"]) > 1e-3
    return dict(opac=opac, stokes=st, flux=fe)


@pytest.fixture(scope="module")
def case_a():
    setup, lines = _narrow_c4(150, 6025.0, W.SEED_BASE + 4)
    return _oracle(setup, lines, B_INCLINED, 1)


@pytest.mark.parametrize("batched", [0, 1])
def test_mag_fp64_matches_oracle(case_a, batched):
    with _gpu(case_a["setup"], FSG_FP64, batched) as g:
        print("fp64", batched, _check(case_a, g, FP64_OPAC_REL, FP64_STOKES, FP64_FLUX_ABS))


@pytest.mark.parametrize("batched", [0, 1])
def test_mag_fp32_matches_oracle(case_a, batched):
    with _gpu(case_a["setup"], FSG_FP32, batched) as g:
        print("fp32", batched, _check(case_a, g, FP32_OPAC_REL, FP32_STOKES, FP32_FLUX_ABS))


@pytest.fixture(scope="module")
def small_setup_lines():
    return _narrow_c4(60, 6010.0, 901)


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_mag_bframe0_local_vector(small_setup_lines, prec):
    setup, lines = small_setup_lines
    ref = _oracle(setup, lines, np.array([1500.0, -800.0, 1200.0]), 0)
    tol = (FP64_OPAC_REL, FP64_STOKES, FP64_FLUX_ABS) if prec == FSG_FP64 else (FP32_OPAC_REL, FP32_STOKES,
                                                                                FP32_FLUX_ABS)
    with _gpu(setup, prec) as g:
        print(_check(ref, g, *tol))


def test_mag_bframe2_angles_per_mu(small_setup_lines):
    setup, lines = small_setup_lines
    m = setup.nmu
    ang = np.empty(1 + 2 * m)
    ang[0] = 3000.0
    ang[1::2] = np.linspace(10.0, 80.0, m)
    ang[2::2] = np.linspace(-40.0, 60.0, m)
    ref = _oracle(setup, lines, ang, 2)
    with _gpu(setup, FSG_FP64) as g:
        print(_check(ref, g, FP64_OPAC_REL, FP64_STOKES, FP64_FLUX_ABS))


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_mag_with_scalar_list_in_lineop(small_setup_lines, prec):
    """An atomic list in LineOpac enters PhiI of the Stokes solution."""
    setup, lines = small_setup_lines
    scalar = W.atomic_lines(300, 6000.0, 6010.0, 902)
    ref = _oracle(setup, lines, B_INCLINED, 1, scalar=scalar)
    tol = (FP64_OPAC_REL, FP64_STOKES, FP64_FLUX_ABS) if prec == FSG_FP64 else (FP32_OPAC_REL, FP32_STOKES,
                                                                                FP32_FLUX_ABS)
    with _gpu(setup, prec) as g:
        print(_check(ref, g, *tol))


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_mag_mixed_he_motype1(prec):
    """Molecules, ABO and default damping, He I Stark lines and the tau scale in one list."""
    atm = W.grey_atmosphere()
    setup = W.build_setup(atm, 6000.0, 6008.0, 0.005, 0, motype=1)
    base = W.atomic_lines(60, 6000.0, 6008.0, 903)
    setup, he = with_he_lines(setup, base, nhe_lines=8, seed=904)
    lines = with_zeeman(mixed_list(he, seed=905), seed=906)
    ref = _oracle(setup, lines, B_INCLINED, 1)
    assert np.any((ref["ro"]["flags"] >> 8) > 0), "no He I line retained"
    tol = (FP64_OPAC_REL, FP64_STOKES, FP64_FLUX_ABS) if prec == FSG_FP64 else (FP32_OPAC_REL, FP32_STOKES,
                                                                                FP32_FLUX_ABS)
    with _gpu(setup, prec) as g:
        print(_check(ref, g, *tol))


def test_mag_blocks(small_setup_lines):
    """NlinesLim below the list size: no PACK, MarkWideLines on the full list."""
    setup, lines = small_setup_lines
    ref = _oracle(setup, lines, B_INCLINED, 1, nlim=lines.n // 4)
    with _gpu(setup, FSG_FP64) as g:
        print(_check(ref, g, FP64_OPAC_REL, FP64_STOKES, FP64_FLUX_ABS))
        # a second list in chunks of 7 lines adds the same opacity again
        a = g.get_maglineop()
        r = g.MagLineopCalc(lines, B_INCLINED, 1, NlinesLim=7)
        assert r["stats"]["nblocks"] == -(-r["stats"]["nlines_use"] // 7)
        b = g.get_maglineop()
        assert np.max(np.abs(b - 2 * a) / (np.abs(a) + setup.contop[:, :, None])) <= 1e-11


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_mag_linstr(small_setup_lines, prec):
    setup, lines = small_setup_lines
    orc = Oracle(setup)
    ro = orc.maglineop(lines, B_INCLINED, 1, linstr=True)
    with _gpu(setup, prec) as g:
        rg = g.MagLineopCalc(lines, B_INCLINED, 1, linstr=True)
    keep = (ro["flags"] & 1) == 0
    assert np.all(rg["linstr"][~keep] == 0)
    d = np.abs(rg["linstr"] - ro["linstr"])
    assert d.max() <= 2e-6, d.max()


def test_mag_device_list_and_reset(small_setup_lines):
    """fsg_maglineop_dev on an uploaded list, twice after a reset, equals one host call."""
    setup, lines = small_setup_lines
    with _gpu(setup, FSG_FP32) as g:
        g.MagLineopCalc(lines, B_INCLINED, 1)
        a = g.get_maglineop()
        g.reset_lineop()
        assert np.all(g.get_maglineop() == 0)
        dl = g.upload(lines)
        st = g.maglineop_dev(dl, B_INCLINED, 1)
        b = g.get_maglineop()
        dl.free()
        assert st["nlines_use"] > 0 and st["ms_opacity"] > 0
        assert np.array_equal(a, b)


def test_mag_errors(small_setup_lines):
    from fastspec_amd._lib import FsgError
    setup, lines = small_setup_lines
    with _gpu(setup, FSG_FP64) as g:
        with pytest.raises(FsgError, match="BFRAME 2 needs"):
            g.MagRadTranCalc(np.array([1000.0, 10.0, 20.0]), 2)
        with pytest.raises(FsgError, match="BFRAME"):
            g.MagLineopCalc(lines, B_INCLINED, 3)
        plain = dataclasses.replace(lines, pi_off=None, spl_pi=None, str_pi=None, si_off=None, spl_si=None,
                                    str_si=None)
        with pytest.raises(FsgError, match="Zeeman"):
            g.MagLineopCalc(plain, B_INCLINED, 1)
        bad = dataclasses.replace(lines, si_off=lines.si_off.copy())
        bad.si_off[1:] += 1  # odd sigma count in the first line
        with pytest.raises(FsgError, match="blue/red"):
            g.MagLineopCalc(bad, B_INCLINED, 1)
        empty = LineList(wl=np.zeros(0), elo=np.zeros(0), gf=np.zeros(0), grad=np.zeros(0),
                         spc=np.zeros(0, np.int32), gqst=np.zeros(0), gvdw=np.zeros(0),
                         is_mol=np.zeros(0, np.uint8), list_type=lines.list_type, pi_off=np.zeros(1, np.int32),
                         spl_pi=np.zeros(0), str_pi=np.zeros(0), si_off=np.zeros(1, np.int32),
                         spl_si=np.zeros(0), str_si=np.zeros(0))
        r = g.MagLineopCalc(empty, B_INCLINED, 1)
        assert r["stats"]["nlines_in"] == 0
