"""The cool-dwarf molecular configuration on the GPU: a 20 A slice of the 200,000,000-line
diatomic list (its lines in 10000-10020 A, about 218,000), Teff 3200 K, R = 3x10^5 log grid,
accumulated in chunks of NlinesLim lines (the memory-capped blocks of the full run), FP64 and
FP32 against the oracle; and the same slice streamed from a molecular sidecar (W, E, LGGF,
GRAD columns) equal to the in-memory list."""
import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64, FSG_LIST_MOLECULAR
from oracle_bridge import Oracle

pytestmark = pytest.mark.gpu

SLICE = (10000.0, 10020.0)
NLIM = 50_000


@pytest.fixture(scope="module")
def c3_slice():
    setup, lines = W.config_c3(wl_lo=SLICE[0], wl_hi=SLICE[1])
    assert lines.list_type == FSG_LIST_MOLECULAR and lines.n > 4 * NLIM
    orc = Oracle(setup)
    ro = orc.lineop(lines, nlines_lim=NLIM)
    lop = orc.get_lineop()
    orc.radtran()
    sfl, sfc = orc.flux()
    orc.close()
    return dict(setup=setup, lines=lines, ro=ro, lop=lop, sfl=sfl, sfc=sfc)


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_c3_slice_chunks_match_oracle(c3_slice, prec):
    from fastspec_amd.synth import FastSpecGPU
    ref = c3_slice
    setup = ref["setup"]
    with FastSpecGPU(setup, precision=prec) as g:
        rg = g.LineopCalc(ref["lines"], NlinesLim=NLIM)
        lop = g.get_lineop()
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
    ro = ref["ro"]
    assert np.array_equal(rg["flags"], ro["flags"])
    assert np.array_equal(rg["ileft"], ro["ileft"])
    assert np.array_equal(rg["iright"], ro["iright"])
    for k in ("nlines_use", "nwide", "evals"):
        assert rg["stats"][k] == ro["stats"][k], k
    assert rg["stats"]["nblocks"] == -(-rg["stats"]["nlines_use"] // NLIM) >= 4
    opac = float(np.max(np.abs(lop - ref["lop"]) / (ref["lop"] + setup.contop)))
    flux = float(np.max(np.abs(sfl / sfc - ref["sfl"] / ref["sfc"])))
    p = "fp64" if prec == FSG_FP64 else "fp32"
    print(f"\nPARITY config=c3_slice precision={p} opacity={opac:.3e} flux={flux:.3e} "
          f"evals={rg['stats']['evals']:.3e} nblocks={rg['stats']['nblocks']}", flush=True)
    tol = (1e-7, 1e-8) if prec == FSG_FP64 else (5e-4, 1e-4)
    assert opac <= tol[0] and flux <= tol[1]
    # the molecular forest matters: it removes a sizeable part of the flux
    assert np.min(ref["sfl"] / ref["sfc"]) < 0.8


def test_c3_slice_streamed_from_sidecar(c3_slice, tmp_path):
    from fastspec_amd import linelist as LL
    from fastspec_amd.synth import FastSpecGPU
    setup = c3_slice["setup"]
    cols = W.diatomic_slice(W.C3_NLINES, *SLICE)
    path = LL.write_sidecar(tmp_path / "feh", cols, name="FeH")
    spc = int(c3_slice["lines"].spc[0])
    w0, w1 = SLICE
    # the library's host reader (the same conversions as the streaming reader; 10**x as the
    # reference's compiler evaluates it)
    host = LL.read(path, w0, w1, mol_slot=spc, nthreads=8)
    assert host.n == c3_slice["lines"].n and np.array_equal(host.wl, c3_slice["lines"].wl)
    with FastSpecGPU(setup, precision=FSG_FP32) as g:
        sl = LL.StreamedLines(g, path, w0, w1, mol_slot=spc, nthreads=8)
        assert sl.list_type == FSG_LIST_MOLECULAR and sl.n == host.n
        g.lineop_dev(sl, NLIM)
        a = g.get_lineop()
        sl.free()
        g.reset_lineop()
        dl = g.upload(host)
        g.lineop_dev(dl, NLIM)
        b = g.get_lineop()
        dl.free()
    assert np.array_equal(a, b)
