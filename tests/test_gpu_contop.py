"""Continuum opacity on the device (fsg_contop) against the reference's own Xcont, compiled
from its sources (oracle/_ref, src/contop.f90:267-314), on the configurations' grids: the
plain configuration (5000-5100 A, solar), a slice of the cool-dwarf log grid (3200 K), the
H- bound-free threshold in the infrared, the near-UV of an 8000 K model and of a 15000 K
model (He II populated), the far UV of the 15000 K model (220-950 A: the He II, He I and
H I edges, and the hydrogenic He I levels below 240 A), and COOLOP's ranges on cool models:
the C I, Si I, Mg I and Al I edges in the far UV of a 3500 K model, Fe I, CH, OH and NH in
its visible, and the collision-induced and H2- free-free absorption in the red and infrared
of a 3000 K model; LUKEOP's and HOTOP's ions on the solar, hot and far-UV grids.  Bound:
1e-12 relative per
(wavelength, depth), every source alone and all together; HOP without its levels 1-6
(Ifop(1) = -1); add mode, Motype 1 and the refusal of sources the host keeps; and a
synthesis that uses the device continuum against the oracle's with the reference
continuum."""
import dataclasses

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64
from oracle_bridge import Oracle, ref_air2vac_array, ref_xcont

pytestmark = [pytest.mark.gpu, pytest.mark.ref]

SOURCES = tuple(range(1, 14))
REL = 1e-12
# H2PLOP's exponent is a quartic in ln(frequency) whose terms, ~1e4, cancel to ~-90: one ulp
# of the logarithm (device and host libm may differ by one) moves H2+ opacity by ~2e-12
# relative; it is a minor source, and the sums stay within REL
REL_H2PLOP = 1e-11


def _grids():
    c1, _ = W.config_c1(10)
    c3, _ = W.config_c3_setup(10000.0, 10020.0)
    ir = W.build_setup(W.grey_atmosphere(4500.0, 4.5), 16300.0, 16600.0, 0.05, 0)
    uv = W.build_setup(W.grey_atmosphere(8000.0, 4.0), 3000.0, 3100.0, 0.05, 0)
    hot = W.build_setup(W.grey_atmosphere(15000.0, 4.0), 3000.0, 3100.0, 0.05, 0)
    fuv = W.build_setup(W.grey_atmosphere(15000.0, 4.0), 220.0, 950.0, 0.25, 0)
    cool_fuv = W.build_setup(W.grey_atmosphere(3500.0, 4.5), 900.0, 2100.0, 0.4, 0)
    cool_vis = W.build_setup(W.grey_atmosphere(3500.0, 4.5), 4000.0, 6500.0, 1.0, 0)
    cool_red = W.build_setup(W.grey_atmosphere(3000.0, 5.0), 6500.0, 12000.0, 2.0, 0)
    cool_ir = W.build_setup(W.grey_atmosphere(3000.0, 5.0), 12000.0, 200000.0, 50.0, 0)
    return dict(c1=c1, c3=c3, ir_edge=ir, uv=uv, hot=hot, fuv=fuv, cool_fuv=cool_fuv, cool_vis=cool_vis,
                cool_red=cool_red, cool_ir=cool_ir)


GRIDS = _grids()


def _ref(setup, pops, sources):
    return ref_xcont(setup.temp, setup.xne, setup.rho, pops, sources, ref_air2vac_array(setup.wavelengths))


def _gpu(setup, pops, sources, **kw):
    from fastspec_amd.synth import FastSpecGPU
    with FastSpecGPU(setup, precision=FSG_FP64) as g:
        return g.ContopCalc(**pops, sources=sources, **kw)


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_all_device_sources_match_reference(name):
    s = GRIDS[name]
    pops = W.continuum_populations(s)
    got = _gpu(s, pops, SOURCES)
    ref = _ref(s, pops, SOURCES)
    err = float(np.max(np.abs(got - ref) / ref))
    print(f"\nPARITY contop config={name} nwl={s.nwl} rel={err:.3e}", flush=True)
    assert err <= REL


# every source on the solar grid, the helium and hydrogen bound-free sources also on the hot
# models (He II is empty in a solar model), COOLOP on the cool grids and on the hot near UV
# (CH and OH above 9000 K and NH above 10000 K add nothing)
EACH = ([(src, "c1") for src in SOURCES if src != 6] + [(src, g) for src in (1, 5, 6) for g in ("hot", "fuv")] +
        [(9, g) for g in ("cool_fuv", "cool_vis", "cool_red", "cool_ir", "uv", "c3")] +
        [(src, g) for src in (10, 11) for g in ("c1", "uv", "hot", "fuv", "cool_fuv", "ir_edge")])


@pytest.mark.parametrize("src,grid", EACH)
def test_each_source_matches_reference(src, grid):
    s = GRIDS[grid]
    pops = W.continuum_populations(s)
    got = _gpu(s, pops, (src,))
    ref = _ref(s, pops, (src,))
    assert np.max(ref) > 0.0
    assert np.all((got == 0.0) == (ref == 0.0))
    nz = ref > 0.0
    err = float(np.max(np.abs(got[nz] - ref[nz]) / ref[nz]))
    print(f"\nPARITY contop source={src} grid={grid} rel={err:.3e}", flush=True)
    assert err <= (REL_H2PLOP if src == 2 else REL)


@pytest.mark.parametrize("grid", ["c1", "fuv"])
def test_hop_without_levels_1_to_6(grid):
    """Ifop(1) = -1 (a hydrogen line list supplies the edges of levels 1-6): less opacity
    wherever those levels' edges lie redward, the reference's to 1e-12."""
    s = GRIDS[grid]
    pops = W.continuum_populations(s)
    from fastspec_amd.synth import FastSpecGPU
    with FastSpecGPU(s, precision=FSG_FP64) as g:
        got = g.ContopCalc(**pops, sources=(1,), hop_nolev=True)
    ref = ref_xcont(s.temp, s.xne, s.rho, pops, (1,), ref_air2vac_array(s.wavelengths), hop_nolev=True)
    full = _ref(s, pops, (1,))
    assert np.all(ref <= full) and np.any(ref < full)
    err = float(np.max(np.abs(got - ref) / ref))
    print(f"\nPARITY contop source=1 nolev grid={grid} rel={err:.3e}", flush=True)
    assert err <= REL


def test_add_mode_and_motype1():
    s = GRIDS["c1"]
    pops = W.continuum_populations(s)
    ref = _ref(s, pops, SOURCES)
    # the host's part (here: the workload's own continuum) plus the device's
    got = _gpu(s, pops, SOURCES, add=True)
    assert np.max(np.abs(got - (s.contop + ref)) / (s.contop + ref)) <= REL
    # Motype 1: ContOpac in units of ContOpacStd
    m1 = W.build_setup(W.grey_atmosphere(5777.0, 4.44), 5000.0, 5100.0, 0.01, 0, motype=1)
    got1 = _gpu(m1, pops, SOURCES)
    ref1 = _ref(m1, pops, SOURCES) / m1.contop_std[None, :]
    assert np.max(np.abs(got1 - ref1) / ref1) <= REL


def test_missing_populations_refused():
    from fastspec_amd._lib import FsgError
    s = GRIDS["c1"]
    pops = W.continuum_populations(s)
    with pytest.raises(FsgError, match="not evaluated on the device"):
        _gpu(s, pops, SOURCES + (14,))
    for key in ("fpf_c1", "fpf_ch", "frac_nh", "frac_he1"):
        with pytest.raises(FsgError, match="COOLOP needs"):
            _gpu(s, dict(pops, **{key: None}), (9,))
    for key in ("fpf_c2", "fpf_si2", "fpf_ca2"):
        with pytest.raises(FsgError, match="LUKEOP needs"):
            _gpu(s, dict(pops, **{key: None}), (10,))
    for key in ("hot_frac", "hot_fpf"):
        with pytest.raises(FsgError, match="HOTOP needs"):
            _gpu(s, dict(pops, **{key: None}), (11,))
    with pytest.raises(FsgError, match="population"):
        _gpu(s, dict(pops, frac_hmin=None), (3,))
    with pytest.raises(FsgError, match="population"):
        _gpu(s, dict(pops, fpf_he3=None), (6,))
    with pytest.raises(FsgError, match="HOP"):
        _gpu(s, pops, (2,), hop_nolev=True)


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_synthesis_with_device_continuum(prec):
    """ContopCalc then LineopCalc and the formal solution: the spectrum of the oracle run on
    the reference's continuum."""
    from fastspec_amd.synth import FastSpecGPU
    setup, lines = W.config_c1(3000)
    pops = W.continuum_populations(setup)
    ref_cont = _ref(setup, pops, SOURCES)
    orc = Oracle(dataclasses.replace(setup, contop=ref_cont))
    orc.lineop(lines)
    orc.radtran()
    sfl, sfc = orc.flux()
    orc.close()
    with FastSpecGPU(setup, precision=prec) as g:
        g.ContopCalc(**pops)
        g.LineopCalc(lines)
        g.RadTranCalc(intensities=False)
        gl, gc = g.flux()
    assert np.max(np.abs(gc - sfc) / sfc) <= 1e-9
    assert np.max(np.abs(gl / gc - sfl / sfc)) <= (1e-8 if prec == FSG_FP64 else 1e-4)
