"""HDF5 line lists streamed to HBM (fsg_lines_stream on an .h5 file: chunks decoded by the
reader's host threads into pinned buffers) give the device list that the sidecar of the same
columns gives: equal line opacity (scalar list) and equal MagLineOpac (Zeeman list)."""
import sys
from pathlib import Path

import numpy as np
import pytest

from fastspec_amd import linelist as LL
from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
from make_hdf5_fixtures import FIXTURES, columns  # noqa: E402


def _fixture(fn):
    for f, kind, n, rng, seed, chunk, vac in FIXTURES:
        if f == fn:
            return GOLDEN / f, kind, columns(kind, n, rng, seed), rng, vac
    raise KeyError(fn)


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_hdf5_stream_equals_sidecar_stream(prec, tmp_path):
    from fastspec_amd.synth import FastSpecGPU
    path, kind, cols, (lo, hi), vac = _fixture("atomic_chunked_vac.h5")
    side = LL.write_sidecar(tmp_path / "side", cols, name="atom", wl_vac=vac)
    atm = W.grey_atmosphere()
    setup = W.build_setup(atm, lo + 20.0, lo + 80.0, 0.01, 0)
    out = []
    for src in (path, side):
        with FastSpecGPU(setup, precision=prec) as g:
            sl = LL.StreamedLines(g, src, lo + 15.0, lo + 85.0, species=W.ATOMIC_SPECIES_TABLE, nthreads=5)
            g.lineop_dev(sl)
            out.append((sl.n, g.get_lineop()))
            sl.free()
    assert out[0][0] == out[1][0] > 5000
    assert np.array_equal(out[0][1], out[1][1]) and out[0][1].max() > 0


def test_hdf5_zeeman_stream_equals_sidecar_stream(tmp_path):
    from fastspec_amd.synth import FastSpecGPU
    path, kind, cols, (lo, hi), vac = _fixture("zeeman_chunked.h5")
    side = LL.write_sidecar(tmp_path / "side", cols, name="atom Zeeman", wl_vac=vac)
    atm = W.grey_atmosphere()
    setup = W.build_setup(atm, lo, hi, 0.005, 0)
    bfield = np.array([2000.0, 45.0, 30.0])
    out = []
    for src in (path, side):
        with FastSpecGPU(setup, precision=FSG_FP64) as g:
            sl = LL.StreamedLines(g, src, lo - 1.0, hi + 1.0, species=W.ATOMIC_SPECIES_TABLE, nthreads=3)
            g.maglineop_dev(sl, bfield, 1)
            out.append((sl.n, g.get_maglineop()))
            sl.free()
    assert out[0][0] == out[1][0] > 400
    assert np.array_equal(out[0][1], out[1][1]) and np.abs(out[0][1]).max() > 0
