"""The Fortran host binding (fastspec_amd/fortran/fsgpu_host.f90), compiled against the
reference's own module definitions, drives libfsgpu.so through ISO_C_BINDING."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fortran_case import read_out, write_case
from oracle_bridge import Oracle, ref_available

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "tests" / "fortran" / "_build" / "host_driver"


def _driver():
    if not DRIVER.exists():
        if not ref_available():
            pytest.skip("reference modules (oracle/_ref) not built here")
        subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "fortran")], check=True)
    return DRIVER


@pytest.mark.ref
def test_driver_reports_library_errors_through_errstop(tmp_path):
    """Without a usable device the binding stops with the library's message (ErrStop)."""
    drv = _driver()
    setup, lines = W.config_c1(50)
    write_case(tmp_path / "case.bin", setup, lines)
    import os
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1")
    r = subprocess.run([str(drv), str(tmp_path / "case.bin")], cwd=tmp_path, capture_output=True, text=True,
                       env=env, timeout=120)
    assert "no HIP device" in r.stdout or "HIP error" in r.stdout, r.stdout + r.stderr
    assert not (tmp_path / "out.bin").exists()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_fortran_host_matches_oracle(tmp_path, prec):
    drv = _driver()
    setup, lines = W.config_c1(600)
    write_case(tmp_path / "case.bin", setup, lines)
    r = subprocess.run([str(drv), str(tmp_path / "case.bin"), str(prec)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and "done" in r.stdout, r.stdout + r.stderr
    out = read_out(tmp_path / "out.bin", setup.nmu)
    orc = Oracle(setup)
    ro = orc.lineop(lines, linstr=True)
    spl, _ = orc.radtran()
    sfl, sfc = orc.flux()
    assert out["nlines_use"] == ro["stats"]["nlines_use"]
    assert np.array_equal(out["flags"], ro["flags"])
    assert np.max(np.abs(out["linstr"] - ro["linstr"])) <= 2e-6
    tol = 1e-8 if prec == 0 else 1e-4
    assert np.max(np.abs(out["sfl"] / out["sfc"] - sfl / sfc)) <= tol
    assert np.max(np.abs(out["spl"] - spl) / spl) <= (1e-9 if prec == 0 else 1e-6)
    _check_output_files(tmp_path, setup, lines, orc, ro, prec)


def _check_output_files(tmp_path, setup, lines, orc, ro, prec):
    """The GPU synthesis in the reference's output files (WriteOut, WriteLinStr records),
    read back with the byte-layout readers that parse the reference's shipped examples."""
    from fastspec_amd import formats as F
    lam = setup.wavelengths
    lop_ref = orc.get_lineop()
    spl, spc = orc.radtran()
    sfl, sfc = orc.flux()
    s = F.read_spec(tmp_path / "fs_spec.bin")
    assert not s["magnetic"] and np.array_equal(s["wave"], lam) and np.array_equal(s["xmu"], setup.xmu)
    assert np.max(np.abs(s["spl"] - spl) / spl) <= (1e-9 if prec == 0 else 1e-6)
    assert np.max(np.abs(s["spc"] - spc) / spc) <= 1e-9
    f = F.read_flux(tmp_path / "fs_flux.bin")
    assert np.array_equal(f["wave"], lam)
    assert np.max(np.abs(f["sfl"] / f["sfc"] - sfl / sfc)) <= (1e-8 if prec == 0 else 1e-4)
    lo = F.read_lineop(tmp_path / "fs_lineop.bin", expect_wave=lam)
    assert np.array_equal(lo["temp"], setup.temp) and np.array_equal(lo["rho"], setup.rho)
    assert np.array_equal(lo["rhox"], setup.rhox) and np.array_equal(lo["xne"], setup.xne)
    rel = np.max(np.abs(lo["lineop"] - lop_ref) / (lop_ref + setup.contop))
    assert rel <= (1e-7 if prec == 0 else 5e-4)
    ls = F.read_linst(tmp_path / "fs_linst.bin")
    keep = (ro["flags"] & 1) == 0
    assert ls["blocks"] == 1 and len(ls["wave"]) == int(keep.sum()) == ro["stats"]["nlines_use"]
    # Wlcent + Wfirst, Wlcent = Wl - Wfirst as the host keeps it
    assert np.array_equal(ls["wave"], (lines.wl[keep] - setup.wfirst) + setup.wfirst)
    assert set(ls["species"]) == {"Fe"}
    assert np.array_equal(ls["ion"], setup.spc_ion[lines.spc[keep] - 1])
    assert np.max(np.abs(ls["depth"] - ro["linstr"][keep])) <= 2e-6


@pytest.mark.gpu
def test_fortran_host_integ_convol(tmp_path):
    """GpuIntegConvol(2) with rotation, macroturbulence and the instrumental profile against
    the oracle's IntegConvol(2) after its own FP64 synthesis."""
    drv = _driver()
    setup, lines = W.config_c1(300)
    write_case(tmp_path / "case.bin", setup, lines)
    r = subprocess.run([str(drv), str(tmp_path / "case.bin"), "0", "12.0", "3.0", "4.0"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "done" in r.stdout, r.stdout + r.stderr
    out = read_out(tmp_path / "out.bin", setup.nmu)
    orc = Oracle(setup)
    orc.lineop(lines)
    spl, _ = orc.radtran()
    sfl, sfc = orc.integ_convol(2, 12.0, 3.0, 4.0, 1)
    f0l, f0c = orc.flux()
    assert np.max(np.abs(out["sfl"] / out["sfc"] - sfl / sfc)) <= 1e-8
    assert np.max(np.abs(out["sfc"] - sfc) / sfc) <= 1e-9
    assert np.max(np.abs(out["spl"] - spl) / spl) <= 1e-9  # CImode 2 leaves the intensities alone
    assert np.max(np.abs(sfl / sfc - f0l / f0c)) > 1e-3  # and the lines did broaden
