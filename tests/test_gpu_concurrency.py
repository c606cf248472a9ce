"""Contexts driven from several host threads at once (each with its own streams) give the
results of the same contexts run one after another: the library keeps no shared state
between contexts except the autotuner's choice table."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64

pytestmark = pytest.mark.gpu


def _cases():
    out = []
    for k in range(8):
        atm = W.grey_atmosphere(4800.0 + 150.0 * k, 4.0 + 0.05 * k, ntau=48)
        w0 = 5000.0 + 20.0 * k
        setup = W.build_setup(atm, w0, w0 + 8.0, 0.01, 0, nmu=3)
        out.append((setup, W.atomic_lines(1500, w0, w0 + 8.0, 300 + k)))
    return out


def _synth(case, prec):
    from fastspec_amd.synth import FastSpecGPU
    setup, lines = case
    with FastSpecGPU(setup, precision=prec) as g:
        dl = g.upload(lines)
        for _ in range(2):  # twice on the same context: the second call reuses its buffers
            g.reset_lineop()
            st = g.lineop_dev(dl)
            g.RadTranCalc(intensities=False)
            sfl, sfc = g.flux()
        dl.free()
    return sfl / sfc, st["evals"]


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_threads_match_sequential(prec):
    cases = _cases()
    seq = [_synth(c, prec) for c in cases]
    with ThreadPoolExecutor(max_workers=4) as ex:
        par = list(ex.map(lambda c: _synth(c, prec), cases))
    for (fs, es), (fp, ep) in zip(seq, par):
        assert es == ep
        assert np.array_equal(fs, fp)
