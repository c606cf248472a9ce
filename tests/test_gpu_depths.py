"""The FP32 opacity kernels at depth counts that change their shape: the matrix-core kernel
with NQ = 1, 2, 4, 6 and 8 blocks of 16 depths (the workgroup's load split and ring waits
depend on NQ), and the vector-unit fallback above 128 depths; FP64 alongside.  Against the
FP64 oracle, as tests/test_gpu_parity.py."""
import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64
from oracle_bridge import Oracle
from test_gpu_parity import FP32_FLUX_ABS, FP32_OPAC_REL, FP64_FLUX_ABS, FP64_OPAC_REL, _check, _gpu

pytestmark = pytest.mark.gpu


def _case(ntau):
    atm = W.grey_atmosphere(5777.0, 4.44, ntau=ntau)
    setup = W.build_setup(atm, 5000.0, 5012.0, 0.01, 0, nmu=3)
    lines = W.atomic_lines(1200, 5000.0, 5012.0, 40 + ntau)
    orc = Oracle(setup)
    ro = orc.lineop(lines)
    lop = orc.get_lineop()
    spl, spc = orc.radtran()
    sfl, sfc = orc.flux()
    return dict(setup=setup, lines=lines, ro=ro, lop=lop, spl=spl, spc=spc, sfl=sfl, sfc=sfc)


@pytest.mark.parametrize("ntau", [12, 30, 60, 90, 128, 140])
@pytest.mark.parametrize("prec", [FSG_FP32, FSG_FP64])
def test_depth_counts(ntau, prec):
    ref = _case(ntau)
    tol = (FP64_OPAC_REL, FP64_FLUX_ABS) if prec == FSG_FP64 else (FP32_OPAC_REL, FP32_FLUX_ABS)
    with _gpu(ref["setup"], prec, 1) as g:
        _check(ref, g, *tol)
