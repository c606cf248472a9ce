"""Line lists streamed from sidecars to the device (fsg_lines_stream) synthesise exactly
what the same lines read on the host and uploaded (fsg_lines_read + fsg_lines_upload) do;
the host reader itself is checked against the ReadHDF restatement in tests/test_ingest.py."""
import os

import numpy as np
import pytest

from fastspec_amd import linelist as LL
from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64

pytestmark = pytest.mark.gpu


def _setup():
    atm = W.grey_atmosphere(5777.0, 4.44, ntau=40)
    return W.build_setup(atm, 5000.0, 5030.0, 0.01, 0, nmu=3)


def _synth(g, dl):
    g.reset_lineop()
    st = g.lineop_dev(dl)
    g.RadTranCalc(intensities=False)
    sfl, sfc = g.flux()
    return st, sfl, sfc


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
@pytest.mark.parametrize("chunk", [None, "997"])
def test_stream_matches_upload(tmp_path, prec, chunk):
    from fastspec_amd.synth import FastSpecGPU
    setup = _setup()
    cols = W.atomic_columns(6000, 4995.0, 5035.0, 21)
    cols["ION"] = cols["ION"].copy()
    cols["ION"][::53] = 6  # dropped by the ion filter
    path = LL.write_sidecar(tmp_path / "lines.fsl", cols, wl_vac=1)
    win = (4998.0, 5032.0)
    host = LL.read(path, *win, species=W.ATOMIC_SPECIES_TABLE)
    old = os.environ.get("FSG_INGEST_CHUNK")
    if chunk:
        os.environ["FSG_INGEST_CHUNK"] = chunk
    try:
        with FastSpecGPU(setup, precision=prec) as g:
            sl = LL.StreamedLines(g, path, *win, species=W.ATOMIC_SPECIES_TABLE, nthreads=5)
            assert sl.n == host.n
            a = _synth(g, sl)
            up = g.upload(host)
            b = _synth(g, up)
            sl.free()
            up.free()
    finally:
        if chunk:
            if old is None:
                del os.environ["FSG_INGEST_CHUNK"]
            else:
                os.environ["FSG_INGEST_CHUNK"] = old
    assert a[0]["evals"] == b[0]["evals"] and a[0]["nlines_use"] == b[0]["nlines_use"]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_stream_molecular_and_empty(tmp_path):
    from fastspec_amd.synth import FastSpecGPU
    setup = _setup()
    rng = np.random.default_rng(4)
    n = 2000
    mcols = {"W": np.sort(rng.uniform(5000, 5030, n)), "E": rng.uniform(0, 2, n).astype(np.float32),
             "LGGF": rng.normal(-2, 1, n).astype(np.float32),
             "GRAD": np.power(10.0, rng.normal(8, 0.3, n)).astype(np.float32)}
    path = LL.write_sidecar(tmp_path / "mol.fsl", mcols, name="FeH")
    host = LL.read(path, 5000.0, 5030.0, mol_slot=1)
    with FastSpecGPU(setup, precision=FSG_FP64) as g:
        sl = LL.StreamedLines(g, path, 5000.0, 5030.0, mol_slot=1)
        a = _synth(g, sl)
        up = g.upload(host)
        b = _synth(g, up)
        assert np.array_equal(a[1], b[1])
        empty = LL.StreamedLines(g, path, 7000.0, 8000.0, mol_slot=1)  # culled on WMIN/WMAX
        assert empty.n == 0
        for x in (sl, up, empty):
            x.free()
