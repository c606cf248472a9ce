"""The HDF5 line lists of the reference's converters read by the library (libhdf5 loaded at
run time, chunks decoded by the reader's threads): the same lines, bit for bit, as the
raw-column sidecar of the same columns -- itself pinned against the reference's compiled
ReadHDF steps (tests/test_ingest.py) -- for windows that cut through chunks, air and vacuum
wavelengths, atomic and molecular lists.  Fixtures: tests/golden/*.h5, written by
tests/golden/make_hdf5_fixtures.py from the package's seeded generators."""
from pathlib import Path

import numpy as np
import pytest

from fastspec_amd import linelist as LL
from fastspec_amd import workloads as W
from fastspec_amd._lib import FsgError

GOLDEN = Path(__file__).resolve().parent / "golden"
sys_path = str(GOLDEN)
import sys  # noqa: E402

sys.path.insert(0, sys_path)
from make_hdf5_fixtures import FIXTURES, columns  # noqa: E402

FSG_FEH_SLOT = 7


def _fixture(fn):
    for f, kind, n, rng, seed, chunk, vac in FIXTURES:
        if f == fn:
            return GOLDEN / f, kind, columns(kind, n, rng, seed), rng, vac
    raise KeyError(fn)


def _read(path, w0, w1, kind, nthreads=8):
    if kind == "molecular":
        return LL.read(path, w0, w1, mol_slot=FSG_FEH_SLOT, nthreads=nthreads)
    return LL.read(path, w0, w1, species=W.ATOMIC_SPECIES_TABLE, nthreads=nthreads)


def _equal(a, b):
    assert a.n == b.n and a.list_type == b.list_type
    for k in ("wl", "elo", "gf", "grad", "gqst", "gvdw", "spc", "is_mol"):
        x, y = getattr(a, k), getattr(b, k)
        assert (x is None) == (y is None), k
        if x is not None:
            assert np.array_equal(x, y), k


@pytest.mark.parametrize("fn", ["atomic_ref.h5", "atomic_chunked_vac.h5", "feh_chunked.h5"])
def test_hdf5_equals_sidecar(fn, tmp_path):
    path, kind, cols, (lo, hi), vac = _fixture(fn)
    name = "FeH" if kind == "molecular" else "atom"
    side = LL.write_sidecar(tmp_path / "side", cols, name=name, wl_vac=vac)
    span = hi - lo
    windows = [(lo - 10, hi + 10), (lo + 0.137 * span, lo + 0.61 * span), (lo + 0.5 * span, lo + 0.5 * span + 0.3),
               (hi + 5, hi + 50), (lo - 50, lo - 5)]
    for w0, w1 in windows:
        assert LL.scan(path, w0, w1) == LL.scan(side, w0, w1), (w0, w1)
        for nt in (1, 7):
            _equal(_read(path, w0, w1, kind, nt), _read(side, w0, w1, kind, nt))
    # the whole file read back: the stored columns (W as stored, in air)
    full = _read(path, lo - 10, hi + 10, kind)
    assert full.n == len(cols["W"]) - (0 if kind == "molecular" else int(np.sum(cols["ION"] > 5)))


def test_hdf5_against_readhdf_restatement():
    """The numpy restatement of ReadHDF's selection and conversions (oracle/readhdf.py)."""
    sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
    import readhdf as R
    path, kind, cols, (lo, hi), vac = _fixture("atomic_chunked_vac.h5")
    w0, w1 = lo + 23.4, lo + 151.2
    w = np.asarray(cols["W"], np.float64)
    first, count = R.select(w, w[0], w[-1], vac, w0, w1)
    species = {(el, ion): (slot, int(mol)) for el, ion, slot, mol in W.ATOMIC_SPECIES_TABLE}
    ref = R.convert_atomic(cols, first, count, species, wl_vac=vac)
    got = _read(path, w0, w1, kind)
    assert count > 1000 and got.n == len(ref["wl"])
    for k in ("wl", "elo", "gf", "grad", "gqst", "gvdw", "spc"):
        assert np.array_equal(getattr(got, k), np.asarray(ref[k])), k


def test_hdf5_zeeman_list_needs_the_stream_reader():
    path, *_ = _fixture("zeeman_chunked.h5")
    with pytest.raises(FsgError, match="Zeeman"):
        LL.read(path, 5990.0, 6020.0, species=W.ATOMIC_SPECIES_TABLE)


def test_hdf5_errors(tmp_path):
    with pytest.raises(FsgError):
        LL.scan(tmp_path / "missing.h5", 5000.0, 5100.0)
    bad = tmp_path / "not_a_list.h5"
    bad.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)  # the signature without a file
    with pytest.raises(FsgError, match="ReadHDF"):
        LL.scan(bad, 5000.0, 5100.0)
