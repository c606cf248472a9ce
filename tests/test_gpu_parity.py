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
