"""Line-list ingest: FastSpec HDF5 line lists and their raw-column sidecars, and the library's
reader of both (fsg_lines_scan, fsg_lines_read, fsg_lines_stream take a sidecar directory or
an .h5 file).

A sidecar directory holds the /Lines group of an HDF5 list as it is laid out by the
reference's converters (python/vald2hdf.py:763-790, python/mol2hdf.py): one raw
little-endian file per column, <COLUMN>.bin, in file order, and a header.txt with the group's
attributes (NAME, WL_VAC, WMIN, WMAX, HFLAG, NL) and the column types.  The library applies
ReadHDF's selection and conversions (src/ioutils.f90:545-1100) while it reads.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Dict, Iterable, Optional, Tuple

import numpy as np

from . import _lib
from .abi import FSG_LIST_ATOMIC, FSG_LIST_MIXED, FSG_LIST_MOLECULAR, CIsotope, CSpecies, LineList, ptr

ATOMIC_COLUMNS = ("W", "EL", "ION", "E", "LGGF", "GRAD", "GQST", "GVDW")
MOLECULAR_COLUMNS = ("W", "E", "LGGF", "GRAD")
ZEEMAN_COLUMNS = ("JLO", "JUP", "GLO", "GUP")
ISOTOPE_COLUMNS = ("ISOT",)
_TYPES = {"W": "<f8", "ION": "<i2", "E": "<f4", "LGGF": "<f4", "GRAD": "<f4", "GQST": "<f4", "GVDW": "<f4",
          "JLO": "<f4", "JUP": "<f4", "GLO": "<f4", "GUP": "<f4", "ISOT": "<i2"}


def _type_code(dt: np.dtype) -> str:
    dt = np.dtype(dt)
    if dt.kind == "S":
        return f"S{dt.itemsize}"
    return {"f": "f", "i": "i"}[dt.kind] + str(dt.itemsize)


def write_sidecar(path, columns: Dict[str, np.ndarray], name: str = "atom", wl_vac: int = 0, hflag: int = 0):
    """Writes a sidecar directory from /Lines columns (W sorted ascending, the stored values:
    log gf and the log damping constants as in the HDF5 file)."""
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    need = MOLECULAR_COLUMNS if not name.startswith("atom") else ATOMIC_COLUMNS
    if name.startswith("atom") and "JLO" in columns:  # a Zeeman list (name 'atom Zeeman')
        need = need + ZEEMAN_COLUMNS
    if name.startswith("atom") and "ISOT" in columns:  # isotope mass numbers (0 = not given)
        need = need + ISOTOPE_COLUMNS
    n = len(columns["W"])
    types = []
    for col in need:
        a = np.asarray(columns[col])
        if col == "EL":
            a = a.astype(a.dtype if a.dtype.kind == "S" else "S3")
        else:
            a = a.astype(_TYPES[col])
        if len(a) != n:
            raise ValueError(f"column {col} has {len(a)} values, W has {n}")
        np.ascontiguousarray(a).tofile(path / f"{col}.bin")
        types.append(f"{col}:{_type_code(a.dtype)}")
    w = np.asarray(columns["W"], np.float64)
    hdr = [
        "format=fastspec-lines-1", f"name={name}", f"wl_vac={int(wl_vac)}", f"hflag={int(hflag)}",
        f"wmin={float(w.min()) if n else 0.0!r}", f"wmax={float(w.max()) if n else 0.0!r}", f"nl={n}",
        "columns=" + ",".join(types),
    ]
    (path / "header.txt").write_text("\n".join(hdr) + "\n")
    return path


def _h5_string_attr(obj, name: str, value: str):
    """A scalar string attribute as the reference's converter writes it: fixed length,
    NUL-terminated ASCII (python/vald2hdf.py:738-758)."""
    import h5py  # noqa: PLC0415 - optional dependency
    raw = value.encode("latin-1")
    tid = h5py.h5t.C_S1.copy()
    tid.set_size(len(raw) + 1)
    tid.set_strpad(h5py.h5t.STR_NULLTERM)
    tid.set_cset(h5py.h5t.CSET_ASCII)
    space = h5py.h5s.create(h5py.h5s.SCALAR)
    aid = h5py.h5a.create(obj.id, name.encode("ascii"), tid, space)
    try:
        aid.write(np.array(raw, dtype=f"S{len(raw) + 1}"), mtype=tid)
    finally:
        aid.close()


def write_hdf5(path, columns: Dict[str, np.ndarray], name: str = "atom", wl_vac: int = 0, hflag: int = 0,
               chunk: int = 1 << 18, mol_id: str = "", mol_ion: int = 1):
    """Writes a line list in the reference's HDF5 layout (python/vald2hdf.py:763-842,
    python/mol2hdf.py): /Lines with W f8 contiguous, the other columns chunked (chunk rows;
    the converters use 262,144) with the shuffle and gzip-4 filters, and the group
    attributes NL, NAME, WL_VAC, HFLAG (atomic) or ID, ION (molecular), WMIN, WMAX.  Needs
    h5py (in this image: the conda Python 3.9)."""
    import h5py  # noqa: PLC0415 - optional dependency
    atomic = name.startswith("atom")
    order = [c for c in (ATOMIC_COLUMNS + ZEEMAN_COLUMNS + ISOTOPE_COLUMNS if atomic else MOLECULAR_COLUMNS)
             if c in columns]
    n = len(columns["W"])
    with h5py.File(path, "w") as f:
        g = f.create_group("Lines")
        for col in order:
            a = np.asarray(columns[col])
            a = a.astype(a.dtype if col == "EL" else _TYPES[col])
            if col == "W" or n == 0:
                d = g.create_dataset(col, data=a)
            else:
                d = g.create_dataset(col, data=a, chunks=(min(n, chunk),), shuffle=True, compression="gzip",
                                     compression_opts=4)
            _h5_string_attr(d, "UNIT", "-")
        g.attrs.create("NL", np.int32(n))
        _h5_string_attr(g, "NAME", name)
        if atomic:
            g.attrs.create("HFLAG", np.int16(hflag))
        else:
            _h5_string_attr(g, "ID", mol_id or name)
            g.attrs.create("ION", np.int16(mol_ion))
        g.attrs.create("WL_VAC", np.int16(wl_vac))
        w = np.asarray(columns["W"], np.float64)
        g.attrs.create("WMIN", np.float64(w[0] if n else 0.0))
        g.attrs.create("WMAX", np.float64(w[-1] if n else 0.0))
    return Path(path)


def hdf5_to_sidecar(h5file, path):
    """Writes the sidecar of an HDF5 line list (needs h5py, which this image does not have)."""
    import h5py  # noqa: PLC0415 - optional dependency

    with h5py.File(h5file, "r") as f:
        g = f["Lines"]
        name = g.attrs["NAME"]
        name = name.decode() if isinstance(name, bytes) else str(name)
        cols = {k: g[k][...] for k in g.keys()}
        return write_sidecar(path, cols, name=name.strip(), wl_vac=int(g.attrs.get("WL_VAC", 0)),
                             hflag=int(g.attrs.get("HFLAG", 0)))


def species_table(entries: Iterable[Tuple[str, int, int, bool]]):
    """ctypes array of fsg_species_entry from (symbol, ion, slot, is_molecule) tuples."""
    ent = list(entries)
    arr = (CSpecies * max(len(ent), 1))()
    for i, (el, ion, slot, mol) in enumerate(ent):
        arr[i].el = el.encode()[:8]
        arr[i].ion = int(ion)
        arr[i].slot = int(slot)
        arr[i].is_mol = int(bool(mol))
    return arr, len(ent)


def isotope_table(entries: Iterable[Tuple[str, int, float]]):
    """ctypes array of fsg_isotope_entry from (element, mass number, fraction) tuples."""
    ent = list(entries)
    arr = (CIsotope * max(len(ent), 1))()
    for i, (el, a, frac) in enumerate(ent):
        arr[i].el = el.encode()[:8]
        arr[i].a = int(a)
        arr[i].frac = float(frac)
    return arr, len(ent)


def scan(path, wstart: float, wend: float):
    """(list type, first, count) of the lines inside [wstart, wend] (fsg_lines_scan)."""
    lib = _lib.load()
    lt, first, count = C.c_int32(), C.c_int64(), C.c_int64()
    _lib.check(lib.fsg_lines_scan(str(path).encode(), wstart, wend, C.byref(lt), C.byref(first), C.byref(count)))
    return lt.value, first.value, count.value


def read(path, wstart: float, wend: float, species=(), mol_slot: int = 0, nthreads: int = 8,
         isotopes=()) -> LineList:
    """The selected, converted lines as a LineList ready for fsg_lines_upload (fsg_lines_read)."""
    lib = _lib.load()
    lt, first, count = scan(path, wstart, wend)
    tab, ns = species_table(species)
    itab, ni = isotope_table(isotopes)
    wl = np.zeros(count)
    elo, gf, grad = (np.zeros(count, np.float32) for _ in range(3))
    mol = lt == FSG_LIST_MOLECULAR
    gq = None if mol else np.zeros(count, np.float32)
    gv = None if mol else np.zeros(count, np.float32)
    spc = np.zeros(count, np.int32)
    im = None if mol else np.zeros(count, np.uint8)
    kept = C.c_int64()
    _lib.check(lib.fsg_lines_read(str(path).encode(), first, count, tab, ns, itab, ni, mol_slot, nthreads,
                                  ptr(wl, C.c_double), ptr(elo, C.c_float), ptr(gf, C.c_float), ptr(grad, C.c_float),
                                  ptr(gq, C.c_float), ptr(gv, C.c_float), ptr(spc, C.c_int32),
                                  ptr(im, C.c_uint8), C.byref(kept)))
    k = kept.value
    if not mol and im[:k].any():
        lt = FSG_LIST_MIXED
    return LineList(wl=wl[:k], elo=elo[:k], gf=gf[:k], grad=grad[:k], gqst=None if mol else gq[:k],
                    gvdw=None if mol else gv[:k], spc=spc[:k], list_type=lt,
                    is_mol=None if lt != FSG_LIST_MIXED else im[:k].copy())


class StreamedLines:
    """A device line list read straight from a sidecar (fsg_lines_stream); usable wherever
    a DeviceLines is (lineop_dev, lineop_vjp, set_gf)."""

    def __init__(self, synth, path, wstart: float, wend: float, species=(), mol_slot: int = 0,
                 nthreads: int = 8, isotopes=()):
        self._lib = _lib.load()
        self.synth = synth
        tab, ns = species_table(species)
        itab, ni = isotope_table(isotopes)
        self.handle = C.c_void_p()
        kept = C.c_int64()
        _lib.check(self._lib.fsg_lines_stream(synth.ctx, str(path).encode(), wstart, wend, tab, ns, itab, ni,
                                              mol_slot, nthreads, C.byref(self.handle), C.byref(kept)))
        from .synth import _LIVE, _adopt
        _LIVE.add(self)  # released at exit before the contexts
        _adopt(synth, self)
        self.n = kept.value
        self.list_type = scan(path, wstart, wend)[0]

    def free(self):
        if self.handle:
            self._lib.fsg_lines_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
