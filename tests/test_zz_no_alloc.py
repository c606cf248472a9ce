"""fsg_integ_convol_noalloc, the convolution entry point that allocates nothing, against
fsg_integ_convol: identical results through the C-ABI on two calls in a row (the second on
a smaller grid), both modes, and a clear refusal above the static work space."""
import ctypes as C

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP64

pytestmark = pytest.mark.gpu


def _context(wl_hi, step=0.01, nlines=300):
    from fastspec_amd.synth import FastSpecGPU
    atm = W.grey_atmosphere(5777.0, 4.44)
    setup = W.build_setup(atm, 5000.0, wl_hi, step, 0)
    g = FastSpecGPU(setup, precision=FSG_FP64)
    if nlines:
        g.LineopCalc(W.atomic_lines(nlines, 5000.0, wl_hi, 404))
    g.RadTranCalc(intensities=False)
    return g


def _convolve(wl_hi, entry, cimode, vsini, vmacro, vgau):
    from fastspec_amd import _lib
    g = _context(wl_hi)
    try:
        _lib.check(getattr(g._lib, entry)(g.ctx, cimode, vsini, vmacro, vgau))
        sfl, sfc = g.flux()
        spl, spc = g.get_spectra()
    finally:
        g.close()
    return sfl, sfc, spl, spc


@pytest.mark.parametrize("cimode,args", [(2, (12.0, 4.0, 3.0)), (2, (0.0, 0.0, 0.0)), (1, (0.0, 0.0, 2.5))])
def test_noalloc_equals_allocating_entry_point(cimode, args):
    # first call on a 10,001-point grid, second on a 2,001-point one: the second reuses the
    # array the first left behind
    for wl_hi in (5100.0, 5020.0):
        ref = _convolve(wl_hi, "fsg_integ_convol", cimode, *args)
        got = _convolve(wl_hi, "fsg_integ_convol_noalloc", cimode, *args)
        for r, o in zip(ref, got):
            assert np.array_equal(r, o)
        assert np.all(np.isfinite(got[0])) and np.all(got[1] > 0)


def test_noalloc_refuses_beyond_static_work_space():
    from fastspec_amd._lib import FsgError, check
    # 320,001 points at 7 mu need 32 * 7 * 320,001 bytes > 64 MiB
    g = _context(5032.0, step=0.0001, nlines=0)
    try:
        before = g.flux()
        with pytest.raises(FsgError, match="work space"):
            check(g._lib.fsg_integ_convol_noalloc(g.ctx, 2, 10.0, 3.0, 2.0))
        after = g.flux()
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
        # the allocating entry point still serves it
        check(g._lib.fsg_integ_convol(g.ctx, 2, 10.0, 3.0, 2.0))
    finally:
        g.close()
