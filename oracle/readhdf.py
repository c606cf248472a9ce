"""Test infrastructure only (never imported by the library): a numpy restatement of the
selection and conversions of ReadHDF for atomic and molecular lists, the checker of the
library's line-list sidecar reader (fsg_lines_scan / fsg_lines_read / fsg_lines_stream).

Reference lines restated:
  whole-file cull on WMIN/WMAX after Vac2Air                 src/ioutils.f90:696-727
  full W read, Vac2Air, sort check                           src/ioutils.f90:729-749
  LowerBoundReal8 / UpperBoundReal8 selection                src/ioutils.f90:751-752,
                                                             src/utils.f90:701-737
  molecular LinesM: gf = 10**LGGF, Grad as stored           src/ioutils.f90:806-813
  atomic: ion filter (ION <= Maxion = 5)                     src/ioutils.f90:921-934
  atomic LinesA: gf, GRAD (0 < x < 20), GQST, GVDW (< 0)     src/ioutils.f90:1056-1076
  isotopic fractions of lines with an ISOT mass number       src/ioutils.f90:940-962,1078
  Vac2Air                                                    src/utils.f90:1007-1017
Pinned by tests/test_ingest.py: Vac2Air and the bound searches against the reference's
compiled routines (oracle/_ref); the bounds against brute force; 10.0**x against scalar
REAL(4) code built by the reference's compiler (oracle/_ref ref_pow10) and within one ulp
of the FP64 value; the branch boundaries of the column conventions (python/vald2hdf.py
COLUMN_UNITS: damping stored as log10, GRAD <= 0 and GQST/GVDW >= 0 passed through).
"""
import numpy as np

MAXION = 5


def vac2air(w):
    w = np.asarray(w, np.float64)
    s = 1e4 / w
    s2 = s * s
    n = 1.0 + 8.34254e-5 + 2.406147e-2 / (130.0 - s2) + 1.5998e-4 / (38.9 - s2)
    return np.where(w > 2000.0, w / n, w)


def lower_bound(a, x):
    """first 1-based i with a(i) >= x (size + 1 if none)"""
    lo, hi = 1, len(a) + 1
    while lo < hi:
        m = (lo + hi) // 2
        if a[m - 1] >= x:
            hi = m
        else:
            lo = m + 1
    return lo


def upper_bound(a, x):
    """last 1-based i with a(i) <= x (0 if none)"""
    lo, hi = 1, len(a) + 1
    while lo < hi:
        m = (lo + hi) // 2
        if a[m - 1] <= x:
            lo = m + 1
        else:
            hi = m
    return lo - 1


def select(w_file, wmin, wmax, wl_vac, wstart, wend):
    """(first 0-based, count) of the lines kept by ReadHDF's range selection"""
    if wl_vac == 1:
        wmin, wmax = float(vac2air(wmin)), float(vac2air(wmax))
    if wmax < wstart or wmin > wend or len(w_file) == 0:
        return 0, 0
    w = vac2air(w_file) if wl_vac == 1 else np.asarray(w_file, np.float64)
    if not np.all(w[:-1] <= w[1:]):
        raise ValueError("ReadHDF: wavelength array is not sorted")
    i1, i2 = lower_bound(w, wstart), upper_bound(w, wend)
    if i1 > i2 or i1 > len(w) or i2 < 1:
        return 0, 0
    return i1 - 1, i2 - i1 + 1


_LIBM = None


def _p10(x):
    """REAL(4) 10.0**x: the C library's exp10f, which is what the reference's compiler
    evaluates (flang turns the power of ten into exp10f)"""
    global _LIBM
    if _LIBM is None:
        import ctypes
        _LIBM = ctypes.CDLL("libm.so.6")
        _LIBM.exp10f.restype = ctypes.c_float
        _LIBM.exp10f.argtypes = [ctypes.c_float]
    x = np.atleast_1d(np.asarray(x, np.float32))
    return np.array([_LIBM.exp10f(float(v)) for v in x], np.float32)


def convert_atomic(cols, first, count, species, wl_vac=0, isotopes=None):
    """LinesA of the selected lines; species maps (symbol, ion) -> (slot, is_mol), isotopes
    (symbol, mass number) -> fraction for the lines with an ISOT column entry above 0"""
    sl = slice(first, first + count)
    ion = np.asarray(cols["ION"][sl]).astype(np.int64)
    keep = ion <= MAXION
    el = [bytes(e).rstrip(b"\0 ").lstrip(b" ").decode() for e in np.asarray(cols["EL"][sl])[keep]]
    ion = ion[keep]
    w = np.asarray(cols["W"][sl], np.float64)[keep]
    if wl_vac == 1:
        w = vac2air(w)
    g = np.asarray(cols["GRAD"][sl], np.float32)[keep]
    q = np.asarray(cols["GQST"][sl], np.float32)[keep]
    v = np.asarray(cols["GVDW"][sl], np.float32)[keep]
    slot = np.array([species[(e, int(i))][0] for e, i in zip(el, ion)], np.int32)
    mol = np.array([species[(e, int(i))][1] for e, i in zip(el, ion)], np.uint8)
    gf = _p10(np.asarray(cols["LGGF"][sl])[keep])
    if "ISOT" in cols:  # gf times the REAL(4) isotopic fraction (src/ioutils.f90:940-962, 1078)
        iso = np.asarray(cols["ISOT"][sl]).astype(np.int64)[keep]
        fr = np.array([np.float32(isotopes[(e, int(a))]) if a > 0 else np.float32(1.0) for e, a in zip(el, iso)],
                      np.float32)
        gf = (gf * fr).astype(np.float32)
    return dict(
        wl=w, elo=np.asarray(cols["E"][sl], np.float32)[keep], gf=gf,
        grad=np.where((g < 20.0) & (g > 0.0), _p10(g), g).astype(np.float32),
        gqst=np.where(q < 0.0, _p10(q), q).astype(np.float32),
        gvdw=np.where(v < 0.0, _p10(v), v).astype(np.float32), spc=slot, is_mol=mol)


def convert_molecular(cols, first, count, mol_slot, wl_vac=0):
    """LinesM of the selected lines"""
    sl = slice(first, first + count)
    w = np.asarray(cols["W"][sl], np.float64)
    if wl_vac == 1:
        w = vac2air(w)
    return dict(wl=w, elo=np.asarray(cols["E"][sl], np.float32), gf=_p10(cols["LGGF"][sl]),
                grad=np.asarray(cols["GRAD"][sl], np.float32), spc=np.full(count, mol_slot, np.int32))
