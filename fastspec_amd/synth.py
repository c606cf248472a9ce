"""Host-side mirror of FastSpec's hot-path drivers over libfsgpu.

The method names follow the reference's module procedures so that a caller of the
reference finds the same stage sequence (src/fastspec.f90:241-461):

    LineopCalc(lines, NlinesLim)        src/synth.f90:747   (non-magnetic line opacity)
    MagLineopCalc(lines, NlinesLim, ..) src/synth.f90:1098  (Zeeman lists, MagLineOpac)
    RadTranCalc()                       src/synth.f90:1825  (scalar DELO-Bezier, SpL/SpC)
    MagRadTranCalc(bfield, bframe)      src/synth.f90:1904  (Stokes DELO-Bezier, SpZ/SpC)
    IntegConvol(2) / DiskInteg flux     src/convol.f90:27   (flux sums, no convolution)

Errors raised by the library carry the reference's ErrStop messages where it has one.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from typing import Optional

import numpy as np

from . import _lib
from .abi import CONTOP_HOP_NOLEV, FSG_FP32, FSG_FP64, CStats, LineList, Setup, ptr

# Contexts and device lists not yet released.  Interpreter teardown runs finalizers in no
# set order and may run them after parts of the runtime are gone, so whatever the caller
# has not closed is released here at exit, device lists before the contexts they live in.
_LIVE: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _release_all():
    live = list(_LIVE)
    for o in live:
        if not isinstance(o, FastSpecGPU):
            o.free()
    for o in live:
        if isinstance(o, FastSpecGPU):
            o.close()


def _adopt(synth, dlist) -> None:
    """Records a device list on its context, which frees it before fsg_finalize (a list
    released after its context would touch freed state)."""
    synth.__dict__.setdefault("_lists", weakref.WeakSet()).add(dlist)


class DeviceLines:
    """A line list resident in HBM (fsg_lines_upload)."""

    def __init__(self, synth: "FastSpecGPU", lines: LineList):
        self._lib = _lib.load()
        self.synth = synth
        self.n = lines.n
        self.list_type = lines.list_type
        cl = lines.cstruct()
        self.handle = C.c_void_p()
        _lib.check(self._lib.fsg_lines_upload(synth.ctx, lines.list_type, C.byref(cl), C.byref(self.handle)))
        _LIVE.add(self)
        _adopt(synth, self)

    def free(self):
        if self.handle:
            self._lib.fsg_lines_free(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FastSpecGPU:
    """One synthesis context on one GPU (fsg_init ... fsg_finalize)."""

    def __init__(self, setup: Setup, precision: int = FSG_FP32, batched: Optional[int] = None,
                 device: int = 0):
        self._lib = _lib.load()
        self.setup = setup
        self.precision = precision
        self._cs = setup.cstruct(precision=precision, batched=batched, device=device)
        self.ctx = C.c_void_p()
        _lib.check(self._lib.fsg_init(C.byref(self._cs), C.byref(self.ctx)))
        _LIVE.add(self)
        self.nwl_total = setup.nwl  # points of the arrays the library returns
        self.fraction_pf0 = [setup.fraction_pf.copy()]  # FractionPF as built (per atmosphere)

    def close(self):
        if self.ctx:
            for dl in list(getattr(self, "_lists", ())):
                dl.free()
            self._lib.fsg_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- atmosphere state -------------------------------------------------------------
    _ATM_F64 = ("temp", "xne", "rho", "rhox", "contop", "contop_std", "he_gam")
    _ATM_F32 = ("fraction_pf", "frac_h1", "frac_he1", "frac_hmol")

    def set_atmosphere(self, atm: int = 0, **arrays) -> None:
        """fsg_set_atmosphere: a new EOS / continuum state of the context (of atmosphere atm
        of a batch) -- any of temp, xne, rho, rhox [ntau], contop [nwl][ntau], contop_std [ntau],
        fraction_pf [ntau][pf_ld], frac_h1, frac_he1, frac_hmol [ntau], he_gam [nhe][ntau];
        the host copy of the set-up follows.  Line opacity must be computed again after it."""
        import dataclasses
        from .abi import CAtmosphere
        unknown = set(arrays) - set(self._ATM_F64 + self._ATM_F32)
        if unknown:
            raise TypeError(f"set_atmosphere: unknown arrays {sorted(unknown)}")
        setups = getattr(self, "setups", None)
        if not 0 <= atm < (len(setups) if setups is not None else 1):
            raise ValueError("set_atmosphere: atmosphere index out of range")
        old = setups[atm] if setups is not None else self.setup
        given = {k: v for k, v in arrays.items() if v is not None}
        new = dataclasses.replace(old, **given)  # converts to the set-up's dtypes
        for k in given:
            ref = getattr(old, k)
            if ref is not None and getattr(new, k).shape != ref.shape:
                raise ValueError(f"set_atmosphere: {k} has shape {getattr(new, k).shape}, expected {ref.shape}")
        ca = CAtmosphere(**{k: ptr(getattr(new, k), C.c_double if k in self._ATM_F64 else C.c_float)
                            for k in given})
        _lib.check(self._lib.fsg_set_atmosphere(self.ctx, int(atm), C.byref(ca)))
        if setups is not None:
            setups[atm] = new
        if atm == 0:
            self.setup = new

    # --- continuum opacity --------------------------------------------------------------
    # HOP H2PLOP HMINOP HRAYOP HE1OP HE2OP HEMIOP HERAOP COOLOP LUKEOP HOTOP ELECOP H2RAOP
    CONTOP_DEVICE_SOURCES = tuple(range(1, 14))

    def ContopCalc(self, sources=CONTOP_DEVICE_SOURCES, hop_nolev: bool = False, add: bool = False,
                   copy_out: bool = True, **pops):
        """ContopCalc's loop (src/synth.f90:663-669) for the continuum sources the device
        evaluates (fsg_contop): sources are Ifop numbers 1-13; pops are the EOS rows [ntau]
        they read, by the names of fsg_contop_eos (FractionPF of H I, H II, He I, He II,
        He III, C I, Mg I, Al I, Si I, Fe I, CH, OH, C II, Mg II, Si II, N I, O I and Ca II;
        Fraction of H-, H2, He II, NH, H I and He I; HOTOP's hot_frac [ntau][8][6] and
        hot_fpf [ntau][21]).  hop_nolev is the reference's Ifop(1) = -1 (HOP without levels 1-6,
        which a hydrogen line list supplies).  add=False replaces the context's ContOpac with
        their sum, add=True adds it to the ContOpac the context holds.  Returns the resulting
        ContOpac [nwl][ntau], which the host copy of the set-up takes too; copy_out=False
        leaves it on the device only (returns None, and the host copy keeps its earlier
        ContOpac)."""
        import dataclasses
        from .abi import CContopEos
        n = self.setup.ntau
        names = {f for f, _ in CContopEos._fields_}
        arrs = {}
        for k, v in pops.items():
            if k not in names:
                raise TypeError(f"ContopCalc: unknown population {k!r}")
            if v is not None:
                a = np.ascontiguousarray(v, np.float32)
                shape = {"hot_frac": (n, 8, 6), "hot_fpf": (n, 21)}.get(k, (n,))
                if a.shape != shape:
                    raise ValueError(f"ContopCalc: {k} must have the shape {shape}")
                arrs[k] = a
        mask = 0
        for src in sources:
            mask |= 1 << (int(src) - 1)
        if hop_nolev:
            mask |= CONTOP_HOP_NOLEV
        eos = CContopEos(**{k: ptr(a, C.c_float) for k, a in arrs.items()})
        out = np.zeros((self.setup.nwl, n)) if copy_out else None
        _lib.check(self._lib.fsg_contop(self.ctx, C.byref(eos), mask, int(bool(add)), ptr(out, C.c_double)))
        if copy_out:
            self.setup = dataclasses.replace(self.setup, contop=out)
        return out

    contop_calc = ContopCalc

    # --- line opacity -------------------------------------------------------------------
    def LineopCalc(self, lines: LineList, NlinesLim: Optional[int] = None, linstr: bool = False) -> dict:
        """Accumulates one list's line opacity into the device-resident LineOpac."""
        n = lines.n
        nlim = n if NlinesLim is None else int(NlinesLim)
        nlim = max(nlim, 1)
        flags = np.zeros(n, np.int32)
        il = np.zeros(n, np.int32)
        ir = np.zeros(n, np.int32)
        ls = np.zeros(n, np.float32) if linstr else None
        st = CStats()
        cl = lines.cstruct()
        _lib.check(self._lib.fsg_lineop(self.ctx, lines.list_type, C.byref(cl), nlim, ptr(ls, C.c_float),
                                        ptr(flags, C.c_int32), ptr(il, C.c_int32), ptr(ir, C.c_int32),
                                        C.byref(st)))
        return dict(flags=flags, ileft=il, iright=ir, linstr=ls, stats=st.as_dict())

    lineop_calc = LineopCalc

    def upload(self, lines: LineList) -> DeviceLines:
        return DeviceLines(self, lines)

    def lineop_dev(self, dlines: DeviceLines, NlinesLim: Optional[int] = None) -> dict:
        nlim = dlines.n if NlinesLim is None else int(NlinesLim)
        st = CStats()
        _lib.check(self._lib.fsg_lineop_dev(self.ctx, dlines.handle, max(nlim, 1), C.byref(st)))
        return st.as_dict()

    def MagLineopCalc(self, lines: LineList, bfield, bframe: int, NlinesLim: Optional[int] = None,
                      linstr: bool = False) -> dict:
        """Accumulates one Zeeman list's six profiles into the device-resident MagLineOpac."""
        n = lines.n
        nlim = max(n if NlinesLim is None else int(NlinesLim), 1)
        b = np.ascontiguousarray(bfield, np.float64)
        flags = np.zeros(n, np.int32)
        il = np.zeros(n, np.int32)
        ir = np.zeros(n, np.int32)
        ls = np.zeros(n, np.float32) if linstr else None
        st = CStats()
        cl = lines.cstruct()
        _lib.check(self._lib.fsg_maglineop(self.ctx, C.byref(cl), nlim, ptr(b, C.c_double), int(bframe),
                                           ptr(ls, C.c_float), ptr(flags, C.c_int32), ptr(il, C.c_int32),
                                           ptr(ir, C.c_int32), C.byref(st)))
        return dict(flags=flags, ileft=il, iright=ir, linstr=ls, stats=st.as_dict())

    mag_lineop_calc = MagLineopCalc

    def maglineop_dev(self, dlines: DeviceLines, bfield, bframe: int, NlinesLim: Optional[int] = None) -> dict:
        nlim = dlines.n if NlinesLim is None else int(NlinesLim)
        b = np.ascontiguousarray(bfield, np.float64)
        st = CStats()
        _lib.check(self._lib.fsg_maglineop_dev(self.ctx, dlines.handle, max(nlim, 1), ptr(b, C.c_double),
                                               int(bframe), C.byref(st)))
        return st.as_dict()

    def get_maglineop(self) -> np.ndarray:
        """MagLineOpac(6, Ntau, Nwl) as an array [nwl][ntau][6]."""
        out = np.zeros((self.setup.nwl, self.setup.ntau, 6))
        _lib.check(self._lib.fsg_get_maglineop(self.ctx, ptr(out, C.c_double)))
        return out

    # --- reverse mode ------------------------------------------------------------------
    def set_gf(self, dlines: DeviceLines, gf) -> None:
        """Replaces the gf column of a device list: a float32 numpy array, or a CUDA tensor
        (float32, contiguous) whose producing work has finished."""
        if hasattr(gf, "data_ptr"):
            assert gf.dtype.itemsize == 4 and gf.is_contiguous() and gf.numel() == dlines.n
            _lib.check(self._lib.fsg_lines_set_gf(dlines.handle, C.c_void_p(gf.data_ptr()), 1))
        else:
            a = np.ascontiguousarray(gf, np.float32)
            assert a.shape == (dlines.n,)
            _lib.check(self._lib.fsg_lines_set_gf(dlines.handle, a.ctypes.data_as(C.c_void_p), 0))

    def flux_vjp(self, gsfl, gsfc, want_opac: bool = False):
        """Reverse mode of RadTranCalc + flux: from dL/dSfL, dL/dSfC returns
        (dL/dLineOpac, dL/dContOpac [nwl][ntau] or None, None, source-function part of dL/dT
        [ntau]); dL/dLineOpac stays on the device for lineop_vjp."""
        s = self.setup
        a = np.ascontiguousarray(gsfl, np.float64)
        b = np.ascontiguousarray(gsfc, np.float64)
        glop = np.zeros((s.nwl, s.ntau)) if want_opac else None
        gcont = np.zeros((s.nwl, s.ntau)) if want_opac else None
        gt = np.zeros(s.ntau)
        _lib.check(self._lib.fsg_flux_vjp(self.ctx, ptr(a, C.c_double), ptr(b, C.c_double), ptr(glop, C.c_double),
                                          ptr(gcont, C.c_double), ptr(gt, C.c_double)))
        return glop, gcont, gt

    # --- device-pointer variants (torch CUDA tensors of this context's device) ---------------
    def flux_device(self, sfl, sfc) -> None:
        """SfL, SfC into float64 CUDA tensors of nwl values (fsg_flux_dev)."""
        for t in (sfl, sfc):
            assert t.is_cuda and t.dtype.itemsize == 8 and t.is_contiguous() and t.numel() == self.nwl_total
        _lib.check(self._lib.fsg_flux_dev(self.ctx, C.c_void_p(sfl.data_ptr()), C.c_void_p(sfc.data_ptr())))

    def flux_vjp_device(self, gsfl, gsfc, gcont=None) -> np.ndarray:
        """Reverse mode of RadTranCalc + flux from dL/dSfL, dL/dSfC in finished float64 CUDA
        tensors (fsg_flux_vjp_dev); dL/dContOpac into gcont (a float64 CUDA tensor of ntau x
        nwl values, wavelength fastest) when given; returns the source-function part of dL/dT
        (host)."""
        for t in (gsfl, gsfc):
            assert t.is_cuda and t.dtype.itemsize == 8 and t.is_contiguous() and t.numel() == self.nwl_total
        if gcont is not None:
            assert gcont.is_cuda and gcont.dtype.itemsize == 8 and gcont.is_contiguous()
            assert gcont.numel() == self.nwl_total * self.setup.ntau
        gt = np.zeros(self._gtemp_shape())
        _lib.check(self._lib.fsg_flux_vjp_dev(self.ctx, C.c_void_p(gsfl.data_ptr()), C.c_void_p(gsfc.data_ptr()),
                                              C.c_void_p(gcont.data_ptr() if gcont is not None else None),
                                              ptr(gt, C.c_double)))
        return gt

    def lineop_vjp_device(self, dlines, gloggf, NlinesLim: Optional[int] = None):
        """dL/d log10 gf into a float64 CUDA tensor of n values (fsg_lineop_vjp_dev); returns
        (dL/d ln FractionPF, line part of dL/dT) on the host."""
        assert gloggf.is_cuda and gloggf.dtype.itemsize == 8 and gloggf.is_contiguous()
        assert gloggf.numel() == dlines.n
        nlim = dlines.n if NlinesLim is None else int(NlinesLim)
        gf = np.zeros(self._gfpf_shape())
        gt = np.zeros(self._gtemp_shape())
        _lib.check(self._lib.fsg_lineop_vjp_dev(self.ctx, dlines.handle, max(nlim, 1), C.c_void_p(gloggf.data_ptr()),
                                                ptr(gf, C.c_double), ptr(gt, C.c_double)))
        return gf, gt

    def _gtemp_shape(self):
        return (self.setup.ntau,)

    def _gfpf_shape(self):
        return (self.setup.ntau, self.setup.fraction_pf.shape[1])

    def lineop_vjp(self, dlines: DeviceLines, NlinesLim: Optional[int] = None):
        """(dL/d log10 gf [n], dL/d ln FractionPF [ntau][pf_ld], line part of dL/dT [ntau])."""
        s = self.setup
        nlim = dlines.n if NlinesLim is None else int(NlinesLim)
        g = np.zeros(max(dlines.n, 1))
        gf = np.zeros((s.ntau, s.fraction_pf.shape[1]))
        gt = np.zeros(s.ntau)
        _lib.check(self._lib.fsg_lineop_vjp(self.ctx, dlines.handle, max(nlim, 1), ptr(g, C.c_double),
                                            ptr(gf, C.c_double), ptr(gt, C.c_double)))
        return g[: dlines.n], gf, gt

    def reset_lineop(self):
        _lib.check(self._lib.fsg_reset_lineop(self.ctx))

    def get_lineop(self) -> np.ndarray:
        out = np.zeros((self.setup.nwl, self.setup.ntau))
        _lib.check(self._lib.fsg_get_lineop(self.ctx, ptr(out, C.c_double)))
        return out

    def set_lineop(self, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        assert a.shape == (self.setup.nwl, self.setup.ntau)
        _lib.check(self._lib.fsg_set_lineop(self.ctx, ptr(a, C.c_double)))

    # --- formal solution and flux -----------------------------------------------------
    def RadTranCalc(self, intensities: bool = True):
        s = self.setup
        spl = np.zeros((s.nwl, s.nmu)) if intensities else None
        spc = np.zeros((s.nwl, s.nmu)) if intensities else None
        _lib.check(self._lib.fsg_radtran(self.ctx, ptr(spl, C.c_double), ptr(spc, C.c_double)))
        self._polarised = False
        return spl, spc

    radtran_calc = RadTranCalc

    def MagRadTranCalc(self, bfield, bframe: int, intensities: bool = True):
        """Stokes I, Q, U, V [nwl][4][nmu] and continuum intensities [nwl][nmu]."""
        s = self.setup
        b = np.ascontiguousarray(bfield, np.float64)
        spz = np.zeros((s.nwl, 4, s.nmu)) if intensities else None
        spc = np.zeros((s.nwl, s.nmu)) if intensities else None
        _lib.check(self._lib.fsg_magradtran(self.ctx, ptr(b, C.c_double), b.shape[0], int(bframe),
                                            ptr(spz, C.c_double), ptr(spc, C.c_double)))
        self._polarised = True
        return spz, spc

    mag_radtran_calc = MagRadTranCalc

    def flux(self):
        s = self.setup
        sfl = np.zeros(s.nwl)
        sfc = np.zeros(s.nwl)
        _lib.check(self._lib.fsg_flux(self.ctx, ptr(sfl, C.c_double), ptr(sfc, C.c_double)))
        return sfl, sfc

    def IntegConvol(self, CImode: int = 2, Vsini: float = 0.0, Vmacro: float = 0.0, Vgau: float = 0.0):
        """IntegConvol (src/convol.f90:27-64) on the device.  CImode 2: disk integration with
        rotation, radial-tangential macroturbulence and Gaussian broadening, returns (SfL,
        SfC); CImode 1: Gaussian broadening of the intensity spectra, returns them as
        get_spectra() does.  Velocities in km/s; Vgau above 1000 is a resolving power."""
        if Vgau > 1000.0:
            Vgau = 299792.458 / Vgau
        _lib.check(self._lib.fsg_integ_convol(self.ctx, CImode, Vsini, Vmacro, Vgau))
        if CImode == 2:
            return self.flux()
        return self.get_spectra()

    def get_spectra(self, magnetic: bool | None = None):
        """Intensities on the device: (SpL, SpC) [nwl][nmu], or (SpZ [nwl][4][nmu], SpC) after
        MagRadTranCalc (or when magnetic is given as True)."""
        if magnetic is None:
            magnetic = getattr(self, "_polarised", False)
        s = self.setup
        spc = np.zeros((s.nwl, s.nmu))
        if magnetic:
            spz = np.zeros((s.nwl, 4, s.nmu))
            _lib.check(self._lib.fsg_get_spectra(self.ctx, None, ptr(spz, C.c_double), ptr(spc, C.c_double)))
            return spz, spc
        spl = np.zeros((s.nwl, s.nmu))
        _lib.check(self._lib.fsg_get_spectra(self.ctx, ptr(spl, C.c_double), None, ptr(spc, C.c_double)))
        return spl, spc

    # --- output files -------------------------------------------------------------------
    def WriteOut(self, prefix: str = "fs", out_flags=(0, 1, 1, 1, 1), xna=None) -> list:
        """WriteOut (src/ioutils.f90:1751-1823): the context's results in the reference's
        output files, record for record -- <prefix>_lineop.bin (flag 2: LineOpac; the
        reference writes no file of MagLineOpac), _contop.bin (flag 3: the ContOpac the context was built with), _spec.bin (flag 4, Nmu
        > 0: the last formal solution, polarised after MagRadTranCalc) and _flux.bin (flag 5,
        Nmu > 1: SfL, SfC as the last flux or IntegConvol left them).  Xna, which the library
        does not hold, is written as given (zeros by default).  Returns the paths written."""
        from . import formats as F
        s = self.setup
        lam = s.wavelengths
        xna = np.zeros(s.ntau) if xna is None else np.asarray(xna, np.float64)
        atm = (s.rhox, s.temp, s.xne, xna, s.rho)
        fl = tuple(out_flags) + (0,) * (5 - len(out_flags))
        out = []
        if fl[1] > 0:
            out.append(f"{prefix}_lineop.bin")
            F.write_lineop(out[-1], lam, self.get_lineop(), *atm)
        if fl[2] > 0:
            out.append(f"{prefix}_contop.bin")
            F.write_contop(out[-1], lam, s.contop, *atm)
        if fl[3] > 0 and s.nmu > 0:
            out.append(f"{prefix}_spec.bin")
            if getattr(self, "_polarised", False):
                spz, spc = self.get_spectra(magnetic=True)
                F.write_spec(out[-1], lam, s.xmu, spc=spc, spz=spz)
            else:
                spl, spc = self.get_spectra(magnetic=False)
                F.write_spec(out[-1], lam, s.xmu, spl, spc)
        if fl[4] > 0 and s.nmu > 1:
            out.append(f"{prefix}_flux.bin")
            F.write_flux(out[-1], lam, *self.flux())
        return out

    write_out = WriteOut

    def WriteLinStr(self, path: str, lines: LineList, result: dict, species, append: bool = False) -> int:
        """WriteLinStr (src/ioutils.f90:1826-1875): one block of the line-strength file for a
        list that LineopCalc/MagLineopCalc ran with linstr=True -- the retained lines (flag
        bit 0 clear) in list order, their species name (species[spc - 1], as Elemc), ion,
        wavelength (Wlcent + Wfirst) and line depth.  The first list creates the file, later
        ones append (append=True).  Returns the record count."""
        from . import formats as F
        if result.get("linstr") is None:
            raise ValueError("WriteLinStr: the list ran without linstr=True")
        keep = np.flatnonzero((result["flags"] & 1) == 0)
        sidx = lines.spc[keep] - 1
        names = [species[k] for k in sidx]
        wfirst = self.setup.wfirst
        F.write_linst(path, names, self.setup.spc_ion[sidx], (lines.wl[keep] - wfirst) + wfirst,
                      result["linstr"][keep], append=append)
        return int(keep.size)

    write_linstr = WriteLinStr

    def device_ptr(self, which: int) -> int:
        p = C.c_void_p()
        _lib.check(self._lib.fsg_device_ptr(self.ctx, which, C.byref(p)))
        return p.value

    def stream(self) -> int:
        p = C.c_void_p()
        _lib.check(self._lib.fsg_stream(self.ctx, C.byref(p)))
        return p.value


class BatchLines:
    """A batch list (fsg_lines_batch): atmosphere b takes the lines indices[b] (ascending) of a
    shared device list."""

    def __init__(self, batch: "FastSpecBatch", shared: DeviceLines, indices):
        self._lib = _lib.load()
        self.batch = batch
        idx = [np.asarray(ix, np.int64) for ix in indices]
        assert len(idx) == batch.nb, "one index array per atmosphere"
        self.index = np.ascontiguousarray(np.concatenate(idx) if idx else np.zeros(0, np.int64))
        self.atm = np.ascontiguousarray(np.concatenate([np.full(len(ix), b, np.int32) for b, ix in enumerate(idx)]))
        self.bounds = np.concatenate([[0], np.cumsum([len(ix) for ix in idx])]).astype(np.int64)
        self.n = int(self.index.shape[0])
        self.nshared = shared.n
        self.list_type = shared.list_type
        self.handle = C.c_void_p()
        _lib.check(self._lib.fsg_lines_batch(batch.ctx, shared.handle, ptr(self.index, C.c_int64),
                                             ptr(self.atm, C.c_int32), self.n, C.byref(self.handle)))
        _LIVE.add(self)
        _adopt(batch, self)

    def per_atmosphere(self, a):
        """Splits an array over the batch list's lines into the atmospheres' parts."""
        return [a[self.bounds[b]:self.bounds[b + 1]] for b in range(len(self.bounds) - 1)]

    def to_shared(self, a):
        """Sums an array over the batch list's lines onto the shared list's lines."""
        return np.bincount(self.index, weights=a, minlength=self.nshared)

    free = DeviceLines.free
    __del__ = DeviceLines.__del__


class FastSpecBatch(FastSpecGPU):
    """Many model atmospheres synthesised in one launch sequence (fsg_init_batch): the
    atmospheres' grids are concatenated, and every stage runs once for all of them.  Per
    atmosphere the results are those of a FastSpecGPU context on its own set-up."""

    def __init__(self, setups, precision: int = FSG_FP32, batched: Optional[int] = None, device: int = 0):
        from .abi import CSetup
        self._lib = _lib.load()
        self.setups = list(setups)
        self.nb = len(self.setups)
        self.setup = self.setups[0]
        self.precision = precision
        self._cs = [su.cstruct(precision=precision, batched=batched, device=device) for su in self.setups]
        arr = (CSetup * self.nb)(*self._cs)
        self.ctx = C.c_void_p()
        _lib.check(self._lib.fsg_init_batch(arr, self.nb, C.byref(self.ctx)))
        _LIVE.add(self)
        self.off = np.zeros(self.nb + 1, np.int64)
        _lib.check(self._lib.fsg_batch_offsets(self.ctx, ptr(self.off, C.c_int64)))
        self.nwl_total = int(self.off[-1])
        self.fraction_pf0 = [su.fraction_pf.copy() for su in self.setups]

    def _gtemp_shape(self):
        return (self.nb, self.setup.ntau)

    def _gfpf_shape(self):
        return (self.nb, self.setup.ntau, self.setup.fraction_pf.shape[1])

    def split(self, a):
        """The atmospheres' parts of an array over the concatenated grid (first axis)."""
        return [a[self.off[b]:self.off[b] + su.nwl] for b, su in enumerate(self.setups)]

    def join(self, parts):
        """An array over the concatenated grid from the atmospheres' parts (zero between)."""
        first = np.asarray(parts[0])
        out = np.zeros((self.nwl_total,) + first.shape[1:], dtype=np.float64)
        for b, su in enumerate(self.setups):
            out[self.off[b]:self.off[b] + su.nwl] = parts[b]
        return out

    def batch_lines(self, shared: DeviceLines, indices) -> BatchLines:
        return BatchLines(self, shared, indices)

    def set_gf(self, dlines, gf) -> None:
        """gf of a batch list from the shared list's gf (numpy float32, or a finished float32
        CUDA tensor)."""
        if isinstance(dlines, BatchLines):
            if hasattr(gf, "data_ptr"):
                import torch
                gf = gf[torch.as_tensor(dlines.index, device=gf.device)].contiguous()
                torch.cuda.current_stream(gf.device).synchronize()
            else:
                gf = np.asarray(gf, np.float32)[dlines.index]
        super().set_gf(dlines, gf)

    def LineopCalc(self, *a, **k):
        raise NotImplementedError("batch contexts take batch lists: batch_lines() and lineop_dev()")

    def RadTranCalc(self, intensities: bool = True):
        """Intensities per atmosphere ([nwl][nmu] each), or None."""
        m = self.setup.nmu
        spl = np.zeros((self.nwl_total, m)) if intensities else None
        spc = np.zeros((self.nwl_total, m)) if intensities else None
        _lib.check(self._lib.fsg_radtran(self.ctx, ptr(spl, C.c_double), ptr(spc, C.c_double)))
        self._polarised = False
        return (self.split(spl), self.split(spc)) if intensities else (None, None)

    radtran_calc = RadTranCalc

    def flux(self, joined: bool = False):
        """(SfL, SfC) per atmosphere: two lists of [nwl] arrays; joined=True: the two arrays
        over the concatenated grid (join's layout, padding points included)."""
        sfl = np.zeros(self.nwl_total)
        sfc = np.zeros(self.nwl_total)
        _lib.check(self._lib.fsg_flux(self.ctx, ptr(sfl, C.c_double), ptr(sfc, C.c_double)))
        if joined:
            return sfl, sfc
        return self.split(sfl), self.split(sfc)

    def _joined(self, g):
        """A per-atmosphere list, or an array already over the concatenated grid."""
        if isinstance(g, np.ndarray) and g.shape == (self.nwl_total,):
            return np.ascontiguousarray(g, np.float64)
        return np.ascontiguousarray(self.join(g))

    def get_lineop(self):
        out = np.zeros((self.nwl_total, self.setup.ntau))
        _lib.check(self._lib.fsg_get_lineop(self.ctx, ptr(out, C.c_double)))
        return self.split(out)

    def flux_vjp(self, gsfl, gsfc, want_opac: bool = False):
        """From per-atmosphere dL/dSfL, dL/dSfC (lists, or arrays over the concatenated grid):
        per-atmosphere dL/dLineOpac and dL/dContOpac (or None) and the source-function part of
        dL/dT [nb][ntau]."""
        n = self.setup.ntau
        a = self._joined(gsfl)
        b = self._joined(gsfc)
        glop = np.zeros((self.nwl_total, n)) if want_opac else None
        gcont = np.zeros((self.nwl_total, n)) if want_opac else None
        gt = np.zeros((self.nb, n))
        _lib.check(self._lib.fsg_flux_vjp(self.ctx, ptr(a, C.c_double), ptr(b, C.c_double), ptr(glop, C.c_double),
                                          ptr(gcont, C.c_double), ptr(gt, C.c_double)))
        if want_opac:
            return self.split(glop), self.split(gcont), gt
        return None, None, gt

    def lineop_vjp(self, dlines: BatchLines, NlinesLim: Optional[int] = None):
        """(dL/d log10 gf [n] over the batch list, dL/d ln FractionPF [nb][ntau][pf_ld], line
        part of dL/dT [nb][ntau])."""
        su = self.setup
        nlim = dlines.n if NlinesLim is None else int(NlinesLim)
        g = np.zeros(max(dlines.n, 1))
        gf = np.zeros((self.nb, su.ntau, su.fraction_pf.shape[1]))
        gt = np.zeros((self.nb, su.ntau))
        _lib.check(self._lib.fsg_lineop_vjp(self.ctx, dlines.handle, max(nlim, 1), ptr(g, C.c_double),
                                            ptr(gf, C.c_double), ptr(gt, C.c_double)))
        return g[: dlines.n], gf, gt


__all__ = ["FastSpecGPU", "FastSpecBatch", "DeviceLines", "BatchLines", "FSG_FP32", "FSG_FP64"]
