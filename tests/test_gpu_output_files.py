"""GPU syntheses in the reference's output files: FastSpecGPU.WriteOut / WriteLinStr write
fs_lineop.bin, fs_contop.bin, fs_spec.bin, fs_flux.bin and fs_linst.bin in the records of
WriteOut and WriteLinStr (src/ioutils.f90:1751-1875); the files are read back with the
byte-layout readers (which parse the reference's shipped example outputs, see
test_formats.py) and compared with the oracle's results for the same inputs."""
import numpy as np
import pytest

from fastspec_amd import formats as F
from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64
from oracle_bridge import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_scalar_synthesis_files(tmp_path, prec):
    from fastspec_amd.synth import FastSpecGPU
    setup, lines = W.config_c1(3000)
    half = lines.n // 2
    la, lb = lines.take(np.arange(half)), lines.take(np.arange(half, lines.n))
    orc = Oracle(setup)
    ra = orc.lineop(la, linstr=True)
    rb = orc.lineop(lb, linstr=True)
    lop = orc.get_lineop()
    spl, spc = orc.radtran()
    sfl, sfc = orc.flux()
    orc.close()
    species = [f"S{k}" for k in range(len(setup.spc_ion))]
    pre = str(tmp_path / "fs")
    with FastSpecGPU(setup, precision=prec) as g:
        ga = g.LineopCalc(la, linstr=True)
        gb = g.LineopCalc(lb, linstr=True)
        g.RadTranCalc(intensities=False)
        written = g.WriteOut(pre, out_flags=(1, 1, 1, 1, 1), xna=np.full(setup.ntau, 3.0))
        na = g.WriteLinStr(pre + "_linst.bin", la, ga, species)
        nb = g.WriteLinStr(pre + "_linst.bin", lb, gb, species, append=True)
    assert [p.rsplit("_", 1)[1] for p in written] == ["lineop.bin", "contop.bin", "spec.bin", "flux.bin"]
    lam = setup.wavelengths
    fp64 = prec == FSG_FP64
    lo = F.read_lineop(pre + "_lineop.bin", expect_wave=lam)
    assert np.array_equal(lo["wave"], lam) and np.array_equal(lo["xna"], np.full(setup.ntau, 3.0))
    assert np.array_equal(lo["temp"], setup.temp) and np.array_equal(lo["rhox"], setup.rhox)
    assert np.max(np.abs(lo["lineop"] - lop) / (lop + setup.contop)) <= (1e-7 if fp64 else 5e-4)
    co = F.read_lineop(pre + "_contop.bin", expect_wave=lam)
    assert np.array_equal(co["lineop"], setup.contop)
    s = F.read_spec(pre + "_spec.bin")
    assert not s["magnetic"] and np.array_equal(s["xmu"], setup.xmu)
    assert np.max(np.abs(s["spl"] - spl) / spl) <= (1e-9 if fp64 else 1e-6)
    assert np.max(np.abs(s["spc"] - spc) / spc) <= 1e-9
    f = F.read_flux(pre + "_flux.bin")
    assert np.array_equal(f["wave"], lam)
    assert np.max(np.abs(f["sfl"] / f["sfc"] - sfl / sfc)) <= (1e-8 if fp64 else 1e-4)
    ls = F.read_linst(pre + "_linst.bin")
    assert ls["blocks"] == 2 and na == ra["stats"]["nlines_use"] and nb == rb["stats"]["nlines_use"]
    keep = [(r["flags"] & 1) == 0 for r in (ra, rb)]
    assert np.array_equal(ls["wave"], np.concatenate([la.wl[keep[0]], lb.wl[keep[1]]]))
    spcs = np.concatenate([la.spc[keep[0]], lb.spc[keep[1]]]) - 1
    assert list(ls["species"]) == [species[k] for k in spcs]
    assert np.array_equal(ls["ion"], setup.spc_ion[spcs])
    depth = np.concatenate([ra["linstr"][keep[0]], rb["linstr"][keep[1]]])
    assert np.max(np.abs(ls["depth"] - depth)) <= 2e-6


def test_stokes_synthesis_file(tmp_path):
    """After MagRadTranCalc the spectrum file holds I, Ic, Q, U, V (src/ioutils.f90:1796-1797)."""
    from fastspec_amd.synth import FastSpecGPU
    setup, lines, bfield, bframe = W.config_c4()
    sub = lines.take(np.arange(0, lines.n, 25))
    orc = Oracle(setup)
    orc.maglineop(sub, bfield, bframe)
    spz, spc = orc.magradtran(bfield, bframe)
    orc.close()
    pre = str(tmp_path / "fs")
    with FastSpecGPU(setup, precision=FSG_FP64) as g:
        g.MagLineopCalc(sub, bfield, bframe)
        g.MagRadTranCalc(bfield, bframe, intensities=False)
        g.WriteOut(pre, out_flags=(0, 0, 0, 1, 0))
    s = F.read_spec(pre + "_spec.bin")
    assert s["magnetic"] and np.array_equal(s["wave"], setup.wavelengths)
    assert np.max(np.abs(s["spc"] - spc) / spc) <= 1e-9
    assert np.max(np.abs(s["spz"] - spz) / spc[:, None, :]) <= 1e-8
    assert np.max(np.abs(spz[:, 3, :]) / spc) > 1e-4  # V is there to be written
