"""GPU parity: libfsgpu.so (FP64 plain path and FP32 batched path) against the FP64 oracle
on identical synthetic inputs, through the C-ABI."""
import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64
from oracle_bridge import Oracle

pytestmark = pytest.mark.gpu

# tolerances (vs the FP64 oracle)
FP64_OPAC_REL = 1e-7     # FP32 table entries may differ by one ulp from the CPU-built table
FP64_FLUX_ABS = 1e-8
FP32_FLUX_ABS = 1e-4     # the north-star bound on normalised flux
FP32_OPAC_REL = 5e-4     # relative to line + continuum opacity


def _gpu(setup, prec, batched=None):
    from fastspec_amd.synth import FastSpecGPU
    return FastSpecGPU(setup, precision=prec, batched=batched)


@pytest.fixture(scope="module")
def c1_small():
    setup, lines = W.config_c1(3000)
    orc = Oracle(setup)
    ro = orc.lineop(lines)
    lop = orc.get_lineop()
    spl, spc = orc.radtran()
    sfl, sfc = orc.flux()
    return dict(setup=setup, lines=lines, ro=ro, lop=lop, spl=spl, spc=spc, sfl=sfl, sfc=sfc)


def _check(ref, g, opac_rel, flux_abs, windows_exact=True):
    setup = ref["setup"]
    rg = g.LineopCalc(ref["lines"])
    ro = ref["ro"]
    if windows_exact:
        assert np.array_equal(rg["flags"], ro["flags"])
        assert np.array_equal(rg["ileft"], ro["ileft"])
        assert np.array_equal(rg["iright"], ro["iright"])
    assert rg["stats"]["nlines_use"] == ro["stats"]["nlines_use"]
    assert rg["stats"]["nwide"] == ro["stats"]["nwide"]
    assert rg["stats"]["evals"] == ro["stats"]["evals"]
    lop = g.get_lineop()
    den = ref["lop"] + setup.contop
    rel = np.max(np.abs(lop - ref["lop"]) / den)
    assert rel <= opac_rel, rel
    spl, spc = g.RadTranCalc()
    sfl, sfc = g.flux()
    assert np.max(np.abs(spc - ref["spc"]) / ref["spc"]) <= 1e-9
    fe = np.max(np.abs(sfl / sfc - ref["sfl"] / ref["sfc"]))
    assert fe <= flux_abs, fe
    ie = np.max(np.abs(spl - ref["spl"]) / ref["spl"])
    return dict(opac=rel, flux=fe, intensity=ie)


@pytest.mark.parametrize("batched", [0, 1])
def test_fp64_matches_oracle(c1_small, batched):
    with _gpu(c1_small["setup"], FSG_FP64, batched) as g:
        e = _check(c1_small, g, FP64_OPAC_REL, FP64_FLUX_ABS)
        print("fp64", batched, e)


@pytest.mark.parametrize("batched", [0, 1])
def test_fp32_matches_oracle(c1_small, batched):
    with _gpu(c1_small["setup"], FSG_FP32, batched) as g:
        e = _check(c1_small, g, FP32_OPAC_REL, FP32_FLUX_ABS)
        print("fp32", batched, e)


# ---- more cases --------------------------------------------------------------------------
from cases import mixed_list, molecular_list, with_he_lines  # noqa: E402


def _oracle_run(setup, lines, nlim=None, linstr=False):
    orc = Oracle(setup)
    ro = orc.lineop(lines, nlines_lim=nlim, linstr=linstr)
    lop = orc.get_lineop()
    spl, spc = orc.radtran()
    sfl, sfc = orc.flux()
    return dict(setup=setup, lines=lines, ro=ro, lop=lop, spl=spl, spc=spc, sfl=sfl, sfc=sfc)


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_linstr_matches_oracle(prec):
    setup, lines = W.config_c1(800)
    ref = _oracle_run(setup, lines, linstr=True)
    with _gpu(setup, prec) as g:
        rg = g.LineopCalc(lines, linstr=True)
        keep = (ref["ro"]["flags"] & 1) == 0
        assert np.all(rg["linstr"][~keep] == 0)
        d = np.abs(rg["linstr"] - ref["ro"]["linstr"])
        assert d.max() <= 2e-6, d.max()


@pytest.mark.parametrize("prec,tol", [(FSG_FP64, FP64_FLUX_ABS), (FSG_FP32, FP32_FLUX_ABS)])
def test_motype1_tau_scale(prec, tol):
    atm = W.grey_atmosphere()
    setup = W.build_setup(atm, 5000.0, 5020.0, 0.01, 0, motype=1)
    lines = W.atomic_lines(1500, 5000.0, 5020.0, 77)
    ref = _oracle_run(setup, lines)
    with _gpu(setup, prec) as g:
        _check(ref, g, FP64_OPAC_REL if prec == FSG_FP64 else FP32_OPAC_REL, tol)


@pytest.mark.parametrize("prec,tol", [(FSG_FP64, FP64_FLUX_ABS), (FSG_FP32, FP32_FLUX_ABS)])
def test_he_stark_lines(prec, tol):
    setup, lines = W.config_c1(600)
    setup, lines = with_he_lines(setup, lines)
    ref = _oracle_run(setup, lines)
    assert np.any((ref["ro"]["flags"] >> 8) & 0xff), "no He I line matched the table"
    with _gpu(setup, prec) as g:
        _check(ref, g, FP64_OPAC_REL if prec == FSG_FP64 else FP32_OPAC_REL, tol)


@pytest.mark.parametrize("kind", ["mixed", "molecular"])
@pytest.mark.parametrize("prec,tol", [(FSG_FP64, FP64_FLUX_ABS), (FSG_FP32, FP32_FLUX_ABS)])
def test_list_types(kind, prec, tol):
    setup, lines = W.config_c1(1200)
    lines = mixed_list(lines) if kind == "mixed" else molecular_list(lines)
    ref = _oracle_run(setup, lines)
    with _gpu(setup, prec) as g:
        _check(ref, g, FP64_OPAC_REL if prec == FSG_FP64 else FP32_OPAC_REL, tol)


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_multi_block_windows_and_opacity(prec):
    """A block size below the list length: the reference does not PACK and marks wide
    lines against the block size; the accumulated opacity is unchanged."""
    setup, lines = W.config_c1(2000)
    nlim = 300
    ref = _oracle_run(setup, lines, nlim=nlim)
    with _gpu(setup, prec) as g:
        rg = g.LineopCalc(lines, NlinesLim=nlim)
        assert np.array_equal(rg["flags"], ref["ro"]["flags"])
        assert np.array_equal(rg["ileft"], ref["ro"]["ileft"])
        assert rg["stats"]["nwide"] == ref["ro"]["stats"]["nwide"]
        # at most NlinesLim lines have resident coefficients: the retained lines go through
        # the accumulation in ceil(NlinesUse / NlinesLim) chunks
        assert rg["stats"]["nblocks"] == -(-rg["stats"]["nlines_use"] // nlim)
        lop = g.get_lineop()
        tol = FP64_OPAC_REL if prec == FSG_FP64 else FP32_OPAC_REL
        assert np.max(np.abs(lop - ref["lop"]) / (ref["lop"] + setup.contop)) <= tol


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_chunks_across_width_bins_and_he_lines(prec):
    """Line chunks of 37 records cut through width bins and the He I group; every cut
    leaves the accumulated opacity unchanged."""
    setup, lines = W.config_c1(600)
    setup, lines = with_he_lines(setup, lines, nhe_lines=20, seed=41)
    ref = _oracle_run(setup, lines)
    tol = FP64_OPAC_REL if prec == FSG_FP64 else FP32_OPAC_REL
    with _gpu(setup, prec) as g:
        rg = g.LineopCalc(lines, NlinesLim=37)
        assert np.array_equal(rg["flags"], ref["ro"]["flags"])
        assert rg["stats"]["evals"] == ref["ro"]["stats"]["evals"]
        assert rg["stats"]["nblocks"] == -(-rg["stats"]["nlines_use"] // 37)
        lop = g.get_lineop()
        assert np.max(np.abs(lop - ref["lop"]) / (ref["lop"] + setup.contop)) <= tol


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_lists_accumulate_and_opacity_file_round_trip(prec):
    setup, lines = W.config_c1(1000)
    a = lines.take(np.arange(0, lines.n, 2))
    b = lines.take(np.arange(1, lines.n, 2))
    orc = Oracle(setup)
    orc.lineop(a)
    orc.lineop(b)
    ref = orc.get_lineop()
    tol = FP64_OPAC_REL if prec == FSG_FP64 else FP32_OPAC_REL
    with _gpu(setup, prec) as g:
        g.LineopCalc(a)
        g.LineopCalc(b)
        lop = g.get_lineop()
        assert np.max(np.abs(lop - ref) / (ref + setup.contop)) <= tol
        # OPACITY keyword: a saved table is loaded before further lists accumulate
        g.reset_lineop()
        assert np.all(g.get_lineop() == 0)
        g.set_lineop(lop)
        lop2 = g.get_lineop()
        if prec == FSG_FP64:
            assert np.array_equal(lop2, lop)
        else:
            assert np.allclose(lop2, lop, rtol=1e-7, atol=0)


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_empty_and_all_weak_lists(prec):
    setup, lines = W.config_c1(50)
    weak = lines.take(np.arange(30))
    weak.gf[:] = np.float32(1e-30)
    with _gpu(setup, prec) as g:
        r = g.LineopCalc(lines.take(np.arange(0)))
        assert r["stats"]["nlines_use"] == 0
        r = g.LineopCalc(weak)
        assert np.all(r["flags"] & 1) and r["stats"]["nlines_use"] == 0 and r["stats"]["evals"] == 0
        assert np.all(g.get_lineop() == 0)
        spl, spc = g.RadTranCalc()
        assert np.allclose(spl, spc, rtol=0, atol=0)


def test_error_paths():
    from fastspec_amd._lib import FsgError
    setup, lines = W.config_c1(50)
    with _gpu(setup, FSG_FP32) as g:
        with pytest.raises(FsgError, match="no formal solution"):
            g.flux()
        bad = lines.take(np.arange(10)[::-1])  # unsorted wavelengths
        with pytest.raises(FsgError, match="not sorted"):
            g.LineopCalc(bad)
        bad = lines.take(np.arange(10))
        bad.spc[3] = 99
        with pytest.raises(FsgError, match="species"):
            g.LineopCalc(bad)
