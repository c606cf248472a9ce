"""Pins the oracle's magnetic path (MagLineopCalc, FVoigt, StokesPrfPP, Solve4x4,
ZeemanPattern) against the reference's compiled kernels, numpy and closed forms."""
import ctypes as C

import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import LineList
from oracle_bridge import Oracle, Ref, oracle_lib, ref_lib

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int32)


@pytest.fixture(scope="module")
def c4():
    setup, lines, bfield, bframe = W.config_c4(400)
    return setup, lines, bfield, bframe


def test_zeeman_normal_triplet_and_sums():
    P, S = W.zeeman_pattern(0.0, 1.0, 1.0, 1.2)
    assert [g for g, _ in P] == [0.0] and abs(sum(r for _, r in S) - 1) < 1e-12
    for jlo, jup in ((1.0, 2.0), (2.5, 2.5), (3.0, 2.0), (6.0, 6.0)):
        P, S = W.zeeman_pattern(jlo, jup, 1.1, 1.7)
        assert abs(sum(r for _, r in P) - 1) < 1e-12 and abs(sum(r for _, r in S) - 1) < 1e-12
        sh = sorted(g for g, _ in P)
        assert np.allclose(sh, [-x for x in sh[::-1]])           # pi shifts symmetric
        assert len(P) + 2 * len(S) <= 3 * (2 * jup + 1)


@pytest.mark.ref
def test_zeeman_pattern_equals_reference():
    lib, ref = oracle_lib(), ref_lib()
    rng = np.random.default_rng(1)
    for _ in range(200):
        jlo = 0.5 * rng.integers(0, 13)
        jup = max(0.0, jlo + rng.integers(-1, 2))
        if jlo == 0 and jup == 0:
            continue
        g = rng.uniform(0.5, 2.5, 2).astype(np.float32)
        outs = []
        for fn in (lib.fso_zeeman_pattern, ref.ref_zeeman):
            npi, nsi = C.c_int32(), C.c_int32()
            a = [np.zeros(40, np.float32) for _ in range(4)]
            fn(C.c_float(jlo), C.c_float(jup), C.c_float(g[0]), C.c_float(g[1]), C.byref(npi),
               a[0].ctypes.data_as(fp), a[1].ctypes.data_as(fp), C.byref(nsi), a[2].ctypes.data_as(fp),
               a[3].ctypes.data_as(fp))
            outs.append((npi.value, nsi.value, [x.copy() for x in a]))
        assert outs[0][0] == outs[1][0] and outs[0][1] == outs[1][1]
        for x, y in zip(outs[0][2], outs[1][2]):
            assert np.array_equal(x, y)
        # the workload generator's pattern agrees too
        P, S = W.zeeman_pattern(float(jlo), float(jup), float(g[0]), float(g[1]))
        assert np.allclose([p for p, _ in P], outs[0][2][0][: outs[0][0]], atol=1e-6)


@pytest.mark.ref
def test_solve4x4_equals_reference_and_numpy():
    lib, ref = oracle_lib(), ref_lib()
    rng = np.random.default_rng(2)
    for _ in range(300):
        A = rng.normal(size=(4, 4)) + np.eye(4) * rng.uniform(0, 3)
        b = rng.normal(size=4)
        x1, x2 = np.zeros(4), np.zeros(4)
        lib.fso_solve4x4(np.ascontiguousarray(A).ctypes.data_as(dp), b.ctypes.data_as(dp), x1.ctypes.data_as(dp))
        At = np.asfortranarray(A)
        ref.ref_solve4x4(At.ctypes.data_as(dp), b.ctypes.data_as(dp), x2.ctypes.data_as(dp))
        assert np.array_equal(x1, x2)
        assert np.allclose(x1, np.linalg.solve(A, b), rtol=1e-9, atol=1e-9)


@pytest.mark.ref
def test_magnetic_windows_equal_reference(c4):
    setup, lines, bfield, bframe = c4
    orc = Oracle(setup)
    ro = orc.maglineop(lines, bfield, bframe)
    ref = Ref(setup)
    ref.set_lines(lines)
    cop, vv, av, _ = ref.xline()
    bsplit = 4.668644778e-13 * abs(bfield[0])
    wave = setup.wave
    wlc = lines.wl - setup.wfirst
    checked = 0
    for i in range(lines.n):
        if ro["flags"][i] & 1:
            continue
        iwl = ref.lib.ref_waveidx(wlc[i])
        dW = (wlc[i] - wave[iwl - 1]) / (wave[iwl] - wave[iwl - 1])
        contl = setup.contop[iwl - 1] * (1 - dW) + setup.contop[iwl] * dW
        il, ir = ref.mag_window(i + 1, bsplit, cop[i], vv[i], av[i], contl)
        assert (il, ir) == (ro["ileft"][i], ro["iright"][i]), i
        checked += 1
    assert checked > 50


@pytest.mark.ref
def test_maglineop_equals_reference_maglineopacity1(c4):
    setup, lines, bfield, bframe = c4
    orc = Oracle(setup)
    ro = orc.maglineop(lines, bfield, bframe)
    mlop = orc.get_maglineop()
    ref = Ref(setup)
    ref.set_lines(lines)
    cop, vv, av, _ = ref.xline()
    # Zeeman shifts scaled as in pass 1, for every line
    bsplit = 4.668644778e-13 * abs(bfield[0])
    for i in range(lines.n):
        ref.mag_window(i + 1, bsplit, cop[i], vv[i], av[i], np.full(setup.ntau, 1e300))
    ref.set_windows(ro["flags"] & ~8, ro["ileft"], ro["iright"])
    ref.set_block(cop, vv, av)
    rl = ref.maglineop1(1, setup.nwl, 1, lines.n)
    den = np.abs(rl).max(axis=2, keepdims=True) + 1e-300
    assert np.max(np.abs(mlop - rl) / den) < 1e-12


@pytest.mark.ref
def test_stokes_prf_equals_reference(c4):
    setup, lines, bfield, bframe = c4
    orc = Oracle(setup)
    orc.maglineop(lines, bfield, bframe)
    spz, spc = orc.magradtran(bfield, bframe)
    mlop = orc.get_maglineop()
    ref = Ref(setup)
    lib = oracle_lib()
    M = setup.nmu
    geo = [np.zeros(M) for _ in range(5)]
    b = np.ascontiguousarray(bfield, np.float64)
    lib.fso_field_geometry(M, setup.xmu.ctypes.data_as(dp), b.ctypes.data_as(dp), 3, bframe,
                           *(g.ctypes.data_as(dp) for g in geo))
    pl = 6.62607015e-27 * 299792.458 * 1e13 / 1.380649e-16
    for p in range(0, setup.nwl, 997):
        lam = setup.wave[p] + setup.wfirst
        src = np.exp(50.7649141 - 5 * np.log(lam)) / (np.exp(pl / lam / setup.temp) - 1)
        st = ref.stokesprfpp(p + 1, mlop[p], np.zeros(setup.ntau), src, geo)
        assert np.max(np.abs(st - spz[p]) / np.abs(st[0]).max()) < 1e-12, p


def test_zero_field_stokes_reduces_to_unpolarised(c4):
    """B = 0: only pi components at zero shift and sigma at zero shift; Q = U = V = 0 and
    Stokes I equals the scalar solution within the solvers' difference."""
    setup, lines, _, _ = c4
    orc = Oracle(setup)
    b0 = np.array([0.0, 45.0, 30.0])
    orc.maglineop(lines, b0, 1)
    spz, spc = orc.magradtran(b0, 1)
    # Q, U, V vanish up to the rounding of the REAL(4) component strengths
    assert np.max(np.abs(spz[:, 1:, :]) / spz[:, :1, :]) < 1e-7
    assert np.all(spz[:, 0, :] <= spc * (1 + 1e-12))
