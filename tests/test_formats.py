"""Byte layouts of fs_spec.bin, fs_flux.bin, fs_linst.bin and fs_lineop.bin: round trips
and the example outputs shipped with the reference."""
from pathlib import Path

import numpy as np
import pytest

from fastspec_amd import formats as F

EX = Path("/root/reference/examples")
rng = np.random.default_rng(11)


def test_spec_round_trip_scalar(tmp_path):
    nwl, nmu = 37, 5
    wave = np.linspace(5000, 5010, nwl)
    xmu = np.linspace(1, 0.1, nmu)
    spl, spc = rng.random((nwl, nmu)), rng.random((nwl, nmu))
    F.write_spec(tmp_path / "s.bin", wave, xmu, spl, spc)
    raw = (tmp_path / "s.bin").read_bytes()
    assert len(raw) == 8 + 8 * nmu + 8 * nwl + 16 * nwl * nmu
    assert np.frombuffer(raw, "<i4", 2).tolist() == [nwl, nmu]
    # SpL(Nmu, Nwl): mu fastest -> the first Nmu values after the grid are lambda_1's
    first = np.frombuffer(raw, "<f8", nmu, 8 + 8 * nmu + 8 * nwl)
    assert np.array_equal(first, spl[0])
    r = F.read_spec(tmp_path / "s.bin")
    assert not r["magnetic"] and np.array_equal(r["spl"], spl) and np.array_equal(r["spc"], spc)


def test_spec_round_trip_magnetic(tmp_path):
    nwl, nmu = 11, 3
    spz, spc = rng.random((nwl, 4, nmu)), rng.random((nwl, nmu))
    F.write_spec(tmp_path / "m.bin", np.arange(nwl) + 6000.0, np.array([0.9, 0.5, 0.1]), spc=spc, spz=spz)
    r = F.read_spec(tmp_path / "m.bin")
    assert r["magnetic"] and np.array_equal(r["spz"], spz) and np.array_equal(r["spc"], spc)


def test_flux_and_lineop_round_trip(tmp_path):
    nwl, ntau = 23, 7
    wave = np.linspace(3000, 3001, nwl)
    sfl, sfc = rng.random(nwl), rng.random(nwl)
    F.write_flux(tmp_path / "f.bin", wave, sfl, sfc)
    r = F.read_flux(tmp_path / "f.bin")
    assert np.array_equal(r["sfl"], sfl) and np.array_equal(r["wave"], wave)
    lop = rng.random((nwl, ntau))
    atm = [rng.random(ntau) for _ in range(5)]
    F.write_lineop(tmp_path / "l.bin", wave, lop, *atm)
    r = F.read_lineop(tmp_path / "l.bin", expect_wave=wave)
    assert np.array_equal(r["lineop"], lop) and np.array_equal(r["rho"], atm[4])
    with pytest.raises(ValueError, match="ReadOpac"):
        F.read_lineop(tmp_path / "l.bin", expect_wave=wave + 1.0)


def test_linst_blocks_append(tmp_path):
    p = tmp_path / "ls.bin"
    F.write_linst(p, ["Fe", "Ti"], [1, 2], [5000.1, 5000.2], [0.5, 0.25])
    F.write_linst(p, ["OH"], [1], [9000.0], [0.125], append=True)
    r = F.read_linst(p)
    assert r["blocks"] == 2
    assert r["species"].tolist() == ["Fe", "Ti", "OH"]
    assert r["ion"].tolist() == [1, 2, 1]
    assert r["depth"].dtype == np.float32 and r["depth"].tolist() == [0.5, 0.25, 0.125]


@pytest.mark.skipif(not EX.exists(), reason="reference examples not present")
def test_reads_reference_example_outputs():
    """The outputs shipped with the reference parse with these layouts."""
    f = F.read_flux(EX / "atomic-synthesis" / "fs_flux.bin")
    assert f["wave"].shape == (182196,)
    fn = f["sfl"] / f["sfc"]
    assert 0.1 < fn.min() < fn.max() <= 1.0 + 1e-9
    assert np.all(np.diff(f["wave"]) > 0)
    ls = F.read_linst(EX / "atomic-synthesis" / "fs_linst.bin")
    assert ls["blocks"] == 1 and ls["wave"].shape == (44380,)
    assert np.all((ls["depth"] >= 0) & (ls["depth"] <= 1))
    s = F.read_spec(EX / "magnetic-synthesis" / "fs_spec.bin")
    assert s["magnetic"] and s["wave"].shape == (39194,) and s["xmu"].shape == (7,)
    # Gauss-Legendre limb angles (default DISKQUAD 1), descending
    z, _ = np.polynomial.legendre.leggauss(7)
    assert np.allclose(np.sort(s["xmu"]), np.sort(0.5 * (1 + z)), atol=1e-12)
    assert np.max(np.abs(s["spz"][:, 3, :] / s["spc"])) > 0.3   # strong Stokes V
