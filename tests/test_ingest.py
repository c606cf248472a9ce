"""Line-list ingest from sidecars (fastspec_amd/linelist.py, fsg_lines_scan / fsg_lines_read,
host code of libfsgpu.so) against the numpy restatement of ReadHDF in oracle/readhdf.py,
which is pinned here against the reference's compiled Vac2Air and bound searches, brute
force and the REAL(4) rounding of 10**x."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "oracle"))

import readhdf as R  # noqa: E402
from fastspec_amd import linelist as LL  # noqa: E402
from fastspec_amd import workloads as W  # noqa: E402
from fastspec_amd._lib import FsgError  # noqa: E402
from fastspec_amd.abi import FSG_LIST_ATOMIC, FSG_LIST_MIXED, FSG_LIST_MOLECULAR  # noqa: E402
from oracle_bridge import ref_lib  # noqa: E402

SPECIES = {(e, i): (s, int(m)) for e, i, s, m in W.ATOMIC_SPECIES_TABLE}


@pytest.mark.ref
def test_vac2air_and_bounds_match_reference():
    ref = ref_lib()
    for w in (1000.0, 1999.9, 2000.0, 2000.1, 3000.0, 5000.0, 10000.0, 25000.0):
        assert float(R.vac2air(w)) == ref.ref_vac2air(w)
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 100):
        a = np.sort(np.round(rng.uniform(0, 10, n), 1))  # with repeated values
        for x1, x2 in ((-1.0, 11.0), (a[0], a[-1]), (3.3, 3.3), (5.0, 4.0), (11.0, 12.0), (-2.0, -1.0)):
            i1, i2 = C.c_int(), C.c_int()
            ref.ref_bounds(n, a.ctypes.data_as(C.POINTER(C.c_double)), x1, x2, C.byref(i1), C.byref(i2))
            assert (R.lower_bound(a, x1), R.upper_bound(a, x2)) == (i1.value, i2.value)
            # brute force: first a >= x1, last a <= x2 (1-based)
            lb = next((k + 1 for k in range(n) if a[k] >= x1), n + 1)
            ub = max((k + 1 for k in range(n) if a[k] <= x2), default=0)
            assert (lb, ub) == (i1.value, i2.value)


@pytest.mark.ref
def test_pow10_matches_reference_compiler():
    ref = ref_lib()
    x = np.random.default_rng(9).uniform(-12.0, 20.0, 5000).astype(np.float32)
    got = R._p10(x)
    y = C.c_float()
    theirs = np.empty_like(x)
    for i, v in enumerate(x):
        ref.ref_pow10(float(v), C.byref(y))
        theirs[i] = y.value
    assert np.array_equal(got, theirs)


def test_conversion_rounding_and_branches():
    x = np.linspace(-9.0, 2.0, 2001).astype(np.float32)
    p = R._p10(x)
    exact = np.power(10.0, x.astype(np.float64))
    ulp = np.spacing(p.astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(p.astype(np.float64) - exact) <= ulp)
    cols = {"W": np.array([5000.0, 5001.0, 5002.0, 5003.0]), "EL": np.array([b"Fe", b"Fe", b"Ti", b"Ca"], "S3"),
            "ION": np.array([1, 2, 1, 1], np.int16), "E": np.ones(4, np.float32), "LGGF": np.zeros(4, np.float32),
            "GRAD": np.array([0.0, 20.0, 19.5, -1.0], np.float32),
            "GQST": np.array([0.0, -0.5, 3.0, -6.0], np.float32),
            "GVDW": np.array([250.2, -7.5, 0.0, -8.0], np.float32)}
    c = R.convert_atomic(cols, 0, 4, SPECIES)
    # damping stored as log10 is made linear; 0 and the values outside the log ranges pass
    assert c["grad"].tolist() == [0.0, 20.0, float(R._p10(19.5)), -1.0]
    assert c["gqst"].tolist() == [0.0, float(R._p10(-0.5)), 3.0, float(R._p10(-6.0))]
    assert c["gvdw"][0] == np.float32(250.2) and c["gvdw"][2] == 0.0  # ABO packed value kept
    assert c["gf"].tolist() == [1.0] * 4 and c["spc"].tolist() == [1, 2, 3, 4]


def _sidecar(tmp_path, n=3000, seed=7, wl_vac=0, extra_ion=True):
    cols = W.atomic_columns(n, 5000.0, 5030.0, seed)
    if extra_ion:  # a few lines above the EOS ion limit, which the reader drops
        cols["ION"] = cols["ION"].copy()
        cols["ION"][::97] = 6
    path = LL.write_sidecar(tmp_path / "lines.fsl", cols, wl_vac=wl_vac)
    return path, cols


def _species_with_ion6():
    sp = list(W.ATOMIC_SPECIES_TABLE)
    return sp


@pytest.mark.parametrize("win", [(4990.0, 5040.0), (5010.0, 5011.5), (5012.345, 5012.345), (5040.0, 5050.0),
                                 (4000.0, 4999.0), (5020.0, 5000.0)])
@pytest.mark.parametrize("wl_vac", [0, 1])
def test_library_reader_matches_oracle(tmp_path, win, wl_vac):
    path, cols = _sidecar(tmp_path, wl_vac=wl_vac)
    lt, first, count = LL.scan(path, *win)
    f_o, c_o = R.select(cols["W"], cols["W"].min(), cols["W"].max(), wl_vac, *win)
    assert (first, count) == (f_o, c_o) and lt == FSG_LIST_ATOMIC
    got = LL.read(path, *win, species=W.ATOMIC_SPECIES_TABLE, nthreads=4)
    ref = R.convert_atomic(cols, f_o, c_o, SPECIES, wl_vac)
    assert got.n == len(ref["wl"])
    for k in ("wl", "elo", "gf", "grad", "gqst", "gvdw", "spc"):
        assert np.array_equal(getattr(got, k), ref[k]), k
    if count > 200:
        assert got.n < count  # the ion filter dropped lines


def test_thread_counts_give_the_same_list(tmp_path):
    path, _ = _sidecar(tmp_path, n=200_000, seed=3)
    a = LL.read(path, 5000.0, 5030.0, species=W.ATOMIC_SPECIES_TABLE, nthreads=1)
    b = LL.read(path, 5000.0, 5030.0, species=W.ATOMIC_SPECIES_TABLE, nthreads=13)
    for k in ("wl", "elo", "gf", "grad", "gqst", "gvdw", "spc"):
        assert np.array_equal(getattr(a, k), getattr(b, k))


def test_columns_reproduce_the_benchmark_lists(tmp_path):
    """The sidecar of atomic_columns read by the library is atomic_lines of the same draws;
    the benchmark generator converts with powf, the reader like the reference (exp10f), so
    the converted columns agree to one ulp."""
    cols = W.atomic_columns(5000, 3000.0, 10000.0, 11, log_uniform=True)
    path = LL.write_sidecar(tmp_path / "c2.fsl", cols)
    got = LL.read(path, 0.0, 1e6, species=W.ATOMIC_SPECIES_TABLE)
    ref = W.atomic_lines(5000, 3000.0, 10000.0, 11, log_uniform=True)
    for k in ("wl", "elo", "spc"):
        assert np.array_equal(getattr(got, k), getattr(ref, k)), k
    for k in ("gf", "grad", "gqst", "gvdw"):
        a, b = getattr(got, k), getattr(ref, k)
        assert np.all(np.abs(a - b) <= np.spacing(b)), k


def test_molecular_and_mixed(tmp_path):
    rng = np.random.default_rng(2)
    n = 500
    mcols = {"W": np.sort(rng.uniform(6000, 6010, n)), "E": rng.uniform(0, 2, n).astype(np.float32),
             "LGGF": rng.normal(-4, 1, n).astype(np.float32), "GRAD": rng.normal(8, 0.3, n).astype(np.float32)}
    mp = LL.write_sidecar(tmp_path / "mol.fsl", mcols, name="FeH")
    got = LL.read(mp, 6002.0, 6008.0, mol_slot=3)
    f, c = R.select(mcols["W"], mcols["W"].min(), mcols["W"].max(), 0, 6002.0, 6008.0)
    ref = R.convert_molecular(mcols, f, c, 3)
    assert got.list_type == FSG_LIST_MOLECULAR
    for k in ("wl", "elo", "gf", "grad", "spc"):
        assert np.array_equal(getattr(got, k), ref[k]), k
    # an atomic list with a molecule in its species table becomes a mixed list
    cols = W.atomic_columns(300, 5000.0, 5010.0, 4)
    cols["EL"] = cols["EL"].copy()
    cols["EL"][5] = b"CH"
    path = LL.write_sidecar(tmp_path / "mixed.fsl", cols)
    sp = list(W.ATOMIC_SPECIES_TABLE) + [("CH", 1, 6, True)]
    got = LL.read(path, 0.0, 1e4, species=sp)
    assert got.list_type == FSG_LIST_MIXED and got.is_mol.sum() == 1 and got.spc[5] == 6


def test_errors(tmp_path):
    path, cols = _sidecar(tmp_path, n=100)
    with pytest.raises(FsgError, match="not in the species table"):
        LL.read(path, 0.0, 1e4, species=W.ATOMIC_SPECIES_TABLE[:2])
    bad = dict(cols)
    bad["W"] = cols["W"][::-1].copy()
    p2 = LL.write_sidecar(tmp_path / "unsorted.fsl", bad)
    with pytest.raises(FsgError, match="not sorted"):
        LL.scan(p2, 0.0, 1e4)
    with pytest.raises(FsgError, match="no header"):
        LL.scan(tmp_path / "nothing", 0.0, 1e4)
    (tmp_path / "x.fsl").mkdir()
    (tmp_path / "x.fsl" / "header.txt").write_text("format=other\n")
    with pytest.raises(FsgError, match="not a FastSpec line-list sidecar"):
        LL.scan(tmp_path / "x.fsl", 0.0, 1e4)


def test_zeeman_sidecar_needs_the_stream_reader(tmp_path):
    cols = W.atomic_columns(50, 6000.0, 6010.0, 5)
    cols.update(JLO=np.ones(50, np.float32), JUP=np.ones(50, np.float32), GLO=np.ones(50, np.float32),
                GUP=np.full(50, 1.5, np.float32))
    path = LL.write_sidecar(tmp_path / "z.fsl", cols, name="atom Zeeman")
    assert LL.scan(path, 0.0, 1e4)[2] == 50
    with pytest.raises(FsgError, match="Zeeman list"):
        LL.read(path, 0.0, 1e4, species=W.ATOMIC_SPECIES_TABLE)
    del cols["GUP"]
    (path / "GUP.bin").unlink()
    hdr = (path / "header.txt").read_text().replace(",GUP:f4", "")
    (path / "header.txt").write_text(hdr)
    with pytest.raises(FsgError, match="incomplete Zeeman columns"):
        LL.scan(path, 0.0, 1e4)


def test_isotopic_fractions(tmp_path):
    cols = W.atomic_columns(2000, 5000.0, 5020.0, 8)
    iso = np.zeros(2000, np.int16)
    iso[::3] = np.where(cols["EL"][::3] == b"Fe", 56, np.where(cols["EL"][::3] == b"Ti", 48, 0))
    iso[1::7] = np.where(cols["EL"][1::7] == b"Fe", 54, 0)
    cols["ISOT"] = iso
    table = [("Fe", 54, 0.05845), ("Fe", 56, 0.91754), ("Ti", 48, 0.7372)]
    path = LL.write_sidecar(tmp_path / "iso.fsl", cols)
    got = LL.read(path, 5001.0, 5019.0, species=W.ATOMIC_SPECIES_TABLE, isotopes=table)
    f, c = R.select(cols["W"], cols["W"].min(), cols["W"].max(), 0, 5001.0, 5019.0)
    ref = R.convert_atomic(cols, f, c, SPECIES, isotopes={(e, a): v for e, a, v in table})
    for k in ("wl", "elo", "gf", "grad", "gqst", "gvdw", "spc"):
        assert np.array_equal(getattr(got, k), ref[k]), k
    plain = R.convert_atomic({k: v for k, v in cols.items() if k != "ISOT"}, f, c, SPECIES)
    assert np.count_nonzero(got.gf != plain["gf"]) > 100  # the scaled lines differ
    with pytest.raises(FsgError, match="not in the isotopic table"):
        LL.read(path, 5001.0, 5019.0, species=W.ATOMIC_SPECIES_TABLE, isotopes=table[1:])
