"""FastSpec's output files (unformatted stream, little-endian), written and read in the
reference's byte layout so that existing readers keep working:

  fs_spec.bin    WriteOut   src/ioutils.f90:1776-1798
      i4 Nwl, i4 Nmu, f8 Xmu[Nmu], f8 lambda[Nwl], then f8 arrays (Nmu, Nwl) mu-fastest:
      SpL, SpC                      (scalar)
      I, Ic, Q, U, V                (magnetic: SpZ(:,1,:), SpC, SpZ(:,2..4,:))
  fs_flux.bin    WriteOut   src/ioutils.f90:1808-1816
      i4 Nwl, f8 lambda[Nwl], f8 SfL[Nwl], f8 SfC[Nwl]
  fs_linst.bin   WriteLinStr src/ioutils.f90:1826-1875, one block per line list (appended):
      i4 Nrec, S8 species[Nrec], i4 ion[Nrec], f8 lambda[Nrec], f4 depth[Nrec]
  fs_lineop.bin  WriteOut   src/ioutils.f90:1759-1766 (also read back by the OPACITY keyword)
      i4 Nwl, i4 Ntau, f8 lambda[Nwl], f8 LineOpac(Ntau, Nwl), f8 Rhox, T, Xne, Xna, Rho [Ntau]
  fs_contop.bin  WriteOut   src/ioutils.f90:1773-1786, the same layout with ContOpac

Arrays passed in use this package's host layouts: intensities [nwl][nmu], Stokes
[nwl][4][nmu], LineOpac [nwl][ntau] (the memory order of the Fortran arrays).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

I4, F4, F8 = np.dtype("<i4"), np.dtype("<f4"), np.dtype("<f8")


def _w(f, a, dt):
    f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def write_spec(path, wave, xmu, spl=None, spc=None, spz=None):
    """Intensity spectra; pass spl (scalar) or spz (magnetic, [nwl][4][nmu])."""
    wave = np.asarray(wave, F8)
    xmu = np.asarray(xmu, F8)
    nwl, nmu = wave.shape[0], xmu.shape[0]
    with open(path, "wb") as f:
        _w(f, [nwl, nmu], I4)
        _w(f, xmu, F8)
        _w(f, wave, F8)
        if spz is None:
            _w(f, spl.reshape(nwl, nmu), F8)
            _w(f, spc.reshape(nwl, nmu), F8)
        else:
            spz = np.asarray(spz).reshape(nwl, 4, nmu)
            _w(f, spz[:, 0, :], F8)
            _w(f, spc.reshape(nwl, nmu), F8)
            for q in (1, 2, 3):
                _w(f, spz[:, q, :], F8)


def read_spec(path) -> dict:
    raw = open(path, "rb").read()
    nwl, nmu = np.frombuffer(raw, I4, 2)
    off = 8
    xmu = np.frombuffer(raw, F8, nmu, off)
    off += 8 * nmu
    wave = np.frombuffer(raw, F8, nwl, off)
    off += 8 * nwl
    rest = np.frombuffer(raw, F8, offset=off)
    narr = rest.size // (nwl * nmu)
    if narr not in (2, 5) or rest.size != narr * nwl * nmu:
        raise ValueError(f"{path}: unexpected size for Nwl={nwl}, Nmu={nmu}")
    sp = rest.reshape(narr, nwl, nmu)
    out = {"wave": wave, "xmu": xmu, "spl": sp[0], "spc": sp[1], "magnetic": narr == 5}
    if narr == 5:
        out["spz"] = np.stack([sp[0], sp[2], sp[3], sp[4]], axis=1)
    return out


def write_flux(path, wave, sfl, sfc):
    wave = np.asarray(wave, F8)
    with open(path, "wb") as f:
        _w(f, [wave.shape[0]], I4)
        _w(f, wave, F8)
        _w(f, sfl, F8)
        _w(f, sfc, F8)


def read_flux(path) -> dict:
    raw = open(path, "rb").read()
    (nwl,) = np.frombuffer(raw, I4, 1)
    if len(raw) != 4 + 24 * nwl:
        raise ValueError(f"{path}: unexpected size for Nwl={nwl}")
    a = np.frombuffer(raw, F8, 3 * nwl, 4).reshape(3, nwl)
    return {"wave": a[0], "sfl": a[1], "sfc": a[2]}


def write_linst(path, species: Sequence[str], ion, wave, depth, append: bool = False):
    """One block of the line-strength list (retained lines of one list)."""
    n = len(wave)
    sp = np.array([s.encode()[:8].ljust(8) for s in species], dtype="S8")
    with open(path, "ab" if append else "wb") as f:
        _w(f, [n], I4)
        f.write(sp.tobytes())
        _w(f, ion, I4)
        _w(f, wave, F8)
        _w(f, depth, F4)


def read_linst(path) -> dict:
    raw = open(path, "rb").read()
    off = 0
    out = {"species": [], "ion": [], "wave": [], "depth": [], "blocks": 0}
    while off < len(raw):
        (n,) = np.frombuffer(raw, I4, 1, off)
        off += 4
        out["species"].append(np.frombuffer(raw, "S8", n, off))
        off += 8 * n
        out["ion"].append(np.frombuffer(raw, I4, n, off))
        off += 4 * n
        out["wave"].append(np.frombuffer(raw, F8, n, off))
        off += 8 * n
        out["depth"].append(np.frombuffer(raw, F4, n, off))
        off += 4 * n
        out["blocks"] += 1
    if off != len(raw):
        raise ValueError(f"{path}: trailing bytes")
    for k in ("species", "ion", "wave", "depth"):
        out[k] = np.concatenate(out[k]) if out[k] else np.zeros(0)
    out["species"] = np.char.strip(out["species"].astype("U8")) if len(out["species"]) else out["species"]
    return out


def write_lineop(path, wave, lineop, rhox, temp, xne, xna, rho):
    wave = np.asarray(wave, F8)
    lineop = np.asarray(lineop, F8)
    nwl, ntau = lineop.shape
    with open(path, "wb") as f:
        _w(f, [nwl, ntau], I4)
        _w(f, wave, F8)
        _w(f, lineop, F8)
        for a in (rhox, temp, xne, xna, rho):
            _w(f, a, F8)


def write_contop(path, wave, contop, rhox, temp, xne, xna, rho):
    """The continuum-opacity file (src/ioutils.f90:1773-1786): the line-opacity layout with
    ContOpac in place of LineOpac."""
    write_lineop(path, wave, contop, rhox, temp, xne, xna, rho)


def read_lineop(path, expect_wave: Optional[np.ndarray] = None) -> dict:
    """Reads a line-opacity table; with expect_wave, checks the grid like ReadOpac
    (src/ioutils.f90:1898-1916)."""
    raw = open(path, "rb").read()
    nwl, ntau = np.frombuffer(raw, I4, 2)
    off = 8
    wave = np.frombuffer(raw, F8, nwl, off)
    off += 8 * nwl
    lop = np.frombuffer(raw, F8, nwl * ntau, off).reshape(nwl, ntau)
    off += 8 * nwl * ntau
    atm = np.frombuffer(raw, F8, 5 * ntau, off).reshape(5, ntau)
    if expect_wave is not None:
        if len(expect_wave) != nwl or np.max(np.abs(np.asarray(expect_wave) - wave)) > 1e-6 * np.max(wave):
            raise ValueError("ReadOpac: wavelength grid of the opacity file does not match the synthesis grid")
    return {"wave": wave, "lineop": lop, "rhox": atm[0], "temp": atm[1], "xne": atm[2], "xna": atm[3],
            "rho": atm[4]}
