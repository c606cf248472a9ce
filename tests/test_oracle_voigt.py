"""Pins the oracle's Voigt and Faraday-Voigt functions (src/voigt.f90) against the
Faddeeva function of scipy, closed forms, and the reference's own compiled routines."""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate, special

from oracle_bridge import oracle_lib, ref_lib

rng = np.random.default_rng(7)
oracle_lib().fso_voigt_init()
oracle_lib().fso_fvoigt_init()


def H(a, v):
    return oracle_lib().fso_hvoigt(a, v)


def Hx(a, v):
    return oracle_lib().fso_hvoigt_exact(a, v)


def W(a, v):
    re, im = C.c_double(), C.c_double()
    oracle_lib().fso_wvoigt_exact(a, v, C.byref(re), C.byref(im))
    return re.value, im.value


def FV(a, v):
    h, f = C.c_double(), C.c_double()
    oracle_lib().fso_fvoigt(a, v, C.byref(h), C.byref(f))
    return h.value, f.value


def test_exact_voigt_matches_faddeeva():
    a = 10 ** rng.uniform(-4, 1.2, 400)
    v = rng.uniform(-12, 12, 400)
    ref = special.wofz(np.abs(v) + 1j * a).real
    got = np.array([Hx(x, y) for x, y in zip(a, v)])
    assert np.max(np.abs(got - ref) / ref) < 1e-9


def test_exact_complex_voigt_matches_faddeeva():
    a = 10 ** rng.uniform(-4, 1.2, 300)
    v = rng.uniform(-12, 12, 300)
    w = special.wofz(np.abs(v) + 1j * a)
    w = np.where(v < 0, np.conj(w), w)
    got = np.array([W(x, y) for x, y in zip(a, v)])
    assert np.max(np.abs(got[:, 0] - w.real) / np.abs(w)) < 1e-9
    assert np.max(np.abs(got[:, 1] - w.imag) / np.abs(w)) < 1e-9


def test_a_zero_limits():
    v = np.linspace(0, 7.9, 60)
    assert np.allclose([Hx(0.0, x) for x in v], np.exp(-v * v), rtol=1e-15, atol=0)
    # L(0, v) = 2/sqrt(pi) Dawson(v)
    got = np.array([W(0.0, x)[1] for x in v])
    assert np.allclose(got, 2 / np.sqrt(np.pi) * special.dawsn(v), rtol=1e-9, atol=1e-12)


def test_table_lookup_at_v_zero_closed_form():
    # H(a, 0) = exp(a^2) erfc(a); bilinear interpolation in a between sqrt-spaced nodes
    a = 10 ** rng.uniform(-3, 0.8, 200)
    got = np.array([H(x, 0.0) for x in a])
    assert np.max(np.abs(got - special.erfcx(a)) / special.erfcx(a)) < 2e-4


def test_table_lookup_accuracy_in_box():
    # bilinear interpolation of the FP32 table: relative error bounded by h^2/8 |H''|
    a = 10 ** rng.uniform(-2, 0.85, 2000)
    v = rng.uniform(-7.99, 7.99, 2000)
    ref = special.wofz(np.abs(v) + 1j * a).real
    got = np.array([H(x, y) for x, y in zip(a, v)])
    bound = 8e-6 * (np.abs(4 * v * v - 2) + 2) + 1e-6
    assert np.all(np.abs(got - ref) / ref <= bound)


def test_asymptotic_branches_outside_box():
    a = np.concatenate([10 ** rng.uniform(-3, 0.8, 300), rng.uniform(8.0, 30.0, 100)])
    v = np.concatenate([rng.uniform(8.0, 200.0, 300) * rng.choice([-1, 1], 300), rng.uniform(-20, 20, 100)])
    ref = special.wofz(np.abs(v) + 1j * a).real
    got = np.array([H(x, y) for x, y in zip(a, v)])
    assert np.max(np.abs(got - ref) / ref) < 1e-4


def test_normalisation_integral():
    for a in (0.01, 0.1, 1.0):
        val, _ = integrate.quad(lambda v: H(a, v), -400, 400, limit=2000, points=[0.0])
        tail = 2 * a / (np.pi ** 0.5) / 400  # Lorentz tail beyond |v| = 400
        assert abs(val + tail - np.sqrt(np.pi)) / np.sqrt(np.pi) < 2e-4


def test_faraday_voigt_odd_and_matches_faddeeva():
    a = 10 ** rng.uniform(-3, 0.8, 300)
    v = rng.uniform(0.01, 30, 300)
    for x, y in zip(a[:40], v[:40]):
        h1, f1 = FV(x, y)
        h2, f2 = FV(x, -y)
        assert h1 == h2 and f1 == -f2
    w = special.wofz(v + 1j * a)
    got = np.array([FV(x, y) for x, y in zip(a, v)])
    assert np.max(np.abs(got[:, 0] - w.real) / np.abs(w)) < 5e-4
    assert np.max(np.abs(got[:, 1] - w.imag) / np.abs(w)) < 5e-4


def test_voigt_line_state_equals_full_lookup():
    lib = oracle_lib()
    a = 10 ** rng.uniform(-4, 0.9, 300)
    v = rng.uniform(-10, 10, 300)
    for x, y in zip(a, v):
        assert lib.fso_hvoigt_line(x, y) == lib.fso_hvoigt(x, y)


# ---- against the reference's own compiled routines --------------------------------------

def _ref_arrays(fn, a, v, nout):
    outs = [np.zeros(len(a)) for _ in range(nout)]
    dp = C.POINTER(C.c_double)
    fn(C.c_int(len(a)), a.ctypes.data_as(dp), v.ctypes.data_as(dp), *(o.ctypes.data_as(dp) for o in outs))
    return outs


@pytest.mark.ref
def test_hvoigt_equals_reference():
    a = np.concatenate([10 ** rng.uniform(-4, 1.3, 3000), [0.0, 1e-9, 7.9999, 8.0, 8.5]])
    v = np.concatenate([rng.uniform(-40, 40, 3000), [0.5, 3.0, 7.99, 2.0, 0.0]])
    (ref,) = _ref_arrays(ref_lib().ref_hvoigt, a, v, 1)
    got = np.array([H(x, y) for x, y in zip(a, v)])
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 2e-16


@pytest.mark.ref
def test_hvoigt_exact_equals_reference():
    a = 10 ** rng.uniform(-4, 1.0, 500)
    v = rng.uniform(-10, 10, 500)
    (ref,) = _ref_arrays(ref_lib().ref_hvoigt_exact, a, v, 1)
    got = np.array([Hx(x, y) for x, y in zip(a, v)])
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-14


@pytest.mark.ref
def test_fvoigt_equals_reference():
    a = np.concatenate([10 ** rng.uniform(-4, 1.3, 2000), [0.0]])
    v = np.concatenate([rng.uniform(-40, 40, 2000), [1.0]])
    h, f = _ref_arrays(ref_lib().ref_fvoigt, a, v, 2)
    got = np.array([FV(x, y) for x, y in zip(a, v)])
    assert np.max(np.abs(got[:, 0] - h) / np.abs(h)) <= 2e-16
    assert np.max(np.abs(got[:, 1] - f) / np.maximum(np.abs(f), 1e-300)) <= 2e-16
