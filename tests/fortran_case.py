"""Writes the case file read by tests/fortran/host_driver.f90 and reads its output."""
import numpy as np


def write_case(path, setup, lines):
    s = setup
    nsp = int(s.spc_ion.shape[0])
    nslot = int(s.fraction_pf.shape[1])
    frac = s.fraction_pf.copy()
    slots = [nsp + 1, nsp + 2, nsp + 3, nsp + 4, nsp + 5]
    frac[:, slots[0] - 1] = s.frac_h1
    frac[:, slots[1] - 1] = s.frac_he1
    frac[:, slots[2] - 1] = s.frac_hmol
    pot = np.zeros(nslot, np.float32)
    pot[: s.pot_ion.shape[0]] = s.pot_ion
    xam = np.ones(nslot, np.float32)
    xam[: s.xamass.shape[0]] = s.xamass
    cstd = s.contop_std if s.contop_std is not None else np.ones(s.ntau)
    hdr = np.array([s.ntau, s.nwl, s.nmu, s.wmode, s.motype, nslot, nsp, lines.n, lines.list_type] + slots,
                   np.int32)
    dh = np.array([s.wfirst, s.dstep, s.rstep, s.wstart, s.eps_opac, s.vwfudge, s.h2brd, s.vwmol, s.vmicro])
    with open(path, "wb") as f:
        for a, dt in ((hdr, "<i4"), (dh, "<f8"), (s.wave, "<f8"), (s.rhox, "<f8"), (s.temp, "<f8"),
                      (s.xne, "<f8"), (s.rho, "<f8"), (s.contop, "<f8"), (cstd, "<f8"), (s.xmu, "<f8"),
                      (s.wmu, "<f8"), (s.fraction_pf, "<f4"), (frac, "<f4"), (pot, "<f4"), (xam, "<f4"),
                      (s.spc_ion, "<i4"), (s.spc_slot, "<i4"), (lines.wl, "<f8"), (lines.elo, "<f4"),
                      (lines.gf, "<f4"), (lines.grad, "<f4"), (lines.gqst, "<f4"), (lines.gvdw, "<f4"),
                      (lines.spc, "<i4")):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())


def read_out(path, nmu):
    raw = open(path, "rb").read()
    nwl, nuse, dolinstr = np.frombuffer(raw, "<i4", 3)
    off = 12
    sfl = np.frombuffer(raw, "<f8", nwl, off); off += 8 * nwl
    sfc = np.frombuffer(raw, "<f8", nwl, off); off += 8 * nwl
    spl = np.frombuffer(raw, "<f8", nwl * nmu, off).reshape(nwl, nmu); off += 8 * nwl * nmu
    out = dict(nwl=nwl, nlines_use=nuse, sfl=sfl, sfc=sfc, spl=spl)
    if dolinstr:
        n = (len(raw) - off) // 8
        out["flags"] = np.frombuffer(raw, "<i4", n, off); off += 4 * n
        out["linstr"] = np.frombuffer(raw, "<f4", n, off)
    return out
