"""ctypes bridges to the test-only checkers: the FP64 C oracle (oracle/libfsoracle.so)
and the reference's own Fortran kernels (oracle/_ref/libfsref.so).

Only tests, __graft_entry__.smoke() and bench.py's CPU baseline use this module."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from fastspec_amd.abi import CLines, CSetup, CStats, LineList, Setup, ptr

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "libfsoracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libfsref.so"

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "libfsoracle.so"], check=True)


def _load_oracle() -> C.CDLL:
    if not ORACLE_SO.exists():
        build_oracle()
    lib = C.CDLL(str(ORACLE_SO))
    vp = C.c_void_p
    lib.fso_init.argtypes = [C.POINTER(CSetup), C.POINTER(vp)]
    lib.fso_finalize.argtypes = [vp]
    lib.fso_lineop.argtypes = [vp, C.c_int32, C.POINTER(CLines), C.c_int64, _fp, _ip, _ip, _ip,
                               C.POINTER(CStats)]
    lib.fso_lineop_pass1.argtypes = [vp, C.c_int32, C.POINTER(CLines), _ip, _ip, _ip, C.POINTER(CStats)]
    lib.fso_maglineop.argtypes = [vp, C.POINTER(CLines), C.c_int64, _dp, C.c_int32, _fp, _ip, _ip,
                                  _ip, C.POINTER(CStats)]
    lib.fso_radtran.argtypes = [vp, _dp, _dp]
    lib.fso_magradtran.argtypes = [vp, _dp, C.c_int32, C.c_int32, _dp, _dp]
    lib.fso_flux.argtypes = [vp, _dp, _dp]
    lib.fso_get_lineop.argtypes = [vp, _dp]
    lib.fso_set_lineop.argtypes = [vp, _dp]
    lib.fso_get_maglineop.argtypes = [vp, _dp]
    lib.fso_hvoigt.argtypes = [C.c_double, C.c_double]
    lib.fso_hvoigt.restype = C.c_double
    lib.fso_hvoigt_exact.argtypes = [C.c_double, C.c_double]
    lib.fso_hvoigt_exact.restype = C.c_double
    lib.fso_hvoigt_line.argtypes = [C.c_double, C.c_double]
    lib.fso_hvoigt_line.restype = C.c_double
    lib.fso_fvoigt.argtypes = [C.c_double, C.c_double, _dp, _dp]
    lib.fso_wvoigt_exact.argtypes = [C.c_double, C.c_double, _dp, _dp]
    lib.fso_voigt_table.argtypes = [C.c_int, C.c_int]
    lib.fso_voigt_table.restype = C.c_float
    lib.fso_air2vac.argtypes = [C.c_double]
    lib.fso_air2vac.restype = C.c_double
    lib.fso_wave_idx.argtypes = [vp, C.c_double]
    lib.fso_wave_idx.restype = C.c_int
    lib.fso_xline.argtypes = [vp, C.c_int32, C.POINTER(CLines), C.c_int64, _dp, _dp, _dp]
    lib.fso_locprfpp.argtypes = [_dp, _dp, C.c_int, _dp, _dp, _dp, C.c_int]
    lib.fso_solve4x4.argtypes = [_dp, _dp, _dp]
    lib.fso_stokes_prf_pp.argtypes = [vp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.fso_field_geometry.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp]
    lib.fso_zeeman_pattern.argtypes = [C.c_float] * 4 + [_ip, _fp, _fp, _ip, _fp, _fp]
    lib.fso_gauss_mu.argtypes = [C.c_int, C.c_int, _dp, _dp]
    lib.fso_mu_quadrature.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp]
    lib.fso_sphere_geometry.argtypes = [vp, _ip, _ip, _ip, _dp, _dp]
    lib.fso_locprfsph.argtypes = [vp, _dp, _dp, _dp]
    lib.fso_sph_field_geometry.argtypes = [vp, _dp, C.c_int, C.c_int, _dp]
    lib.fso_stokes_prf_sph.argtypes = [vp, C.c_int, _dp, _dp, _dp]
    lib.fso_mu_quadrature_r.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
    lib.fso_semi_area.argtypes = [C.c_double, C.c_double]
    lib.fso_semi_area.restype = C.c_double
    lib.fso_convolgau.argtypes = [C.c_int, _dp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double]
    lib.fso_integ_convol.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]
    lib.fso_get_spectra.argtypes = [vp, _dp, _dp, _dp]
    lib.fso_set_spectra.argtypes = [vp, _dp, _dp, _dp]
    lib.fso_lin_str.argtypes = [vp, C.c_double, _dp, _dp]
    lib.fso_lin_str.restype = C.c_float
    lib.fso_last_error.restype = C.c_char_p
    return lib


_ORACLE = None


def oracle_lib() -> C.CDLL:
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = _load_oracle()
    return _ORACLE


class Oracle:
    """FP64 CPU restatement with the same call sequence as the GPU library."""

    def __init__(self, setup: Setup):
        self.lib = oracle_lib()
        self.setup = setup
        self._cs = setup.cstruct()
        self.ctx = C.c_void_p()
        rc = self.lib.fso_init(C.byref(self._cs), C.byref(self.ctx))
        if rc != 0:
            self.ctx = C.c_void_p()
            raise RuntimeError(self.lib.fso_last_error().decode())

    def close(self):
        if self.ctx:
            self.lib.fso_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lineop(self, lines: LineList, nlines_lim: int | None = None, linstr: bool = False):
        n = lines.n
        nlim = n if nlines_lim is None else nlines_lim
        cl = lines.cstruct()
        flags = np.zeros(n, np.int32)
        il = np.zeros(n, np.int32)
        ir = np.zeros(n, np.int32)
        ls = np.zeros(n, np.float32) if linstr else None
        st = CStats()
        rc = self.lib.fso_lineop(self.ctx, lines.list_type, C.byref(cl), nlim, ptr(ls, C.c_float),
                                 ptr(flags, C.c_int32), ptr(il, C.c_int32), ptr(ir, C.c_int32),
                                 C.byref(st))
        if rc != 0:
            raise RuntimeError(self.lib.fso_last_error().decode())
        return dict(flags=flags, ileft=il, iright=ir, linstr=ls, stats=st.as_dict())

    def pass1(self, lines: LineList):
        n = lines.n
        cl = lines.cstruct()
        flags = np.zeros(n, np.int32)
        il = np.zeros(n, np.int32)
        ir = np.zeros(n, np.int32)
        st = CStats()
        rc = self.lib.fso_lineop_pass1(self.ctx, lines.list_type, C.byref(cl), ptr(flags, C.c_int32),
                                       ptr(il, C.c_int32), ptr(ir, C.c_int32), C.byref(st))
        if rc != 0:
            raise RuntimeError(self.lib.fso_last_error().decode())
        return dict(flags=flags, ileft=il, iright=ir, stats=st.as_dict())

    def maglineop(self, lines: LineList, bfield, bframe: int, nlines_lim: int | None = None,
                  linstr: bool = False):
        n = lines.n
        nlim = n if nlines_lim is None else nlines_lim
        cl = lines.cstruct()
        b = np.ascontiguousarray(bfield, np.float64)
        flags = np.zeros(n, np.int32)
        il = np.zeros(n, np.int32)
        ir = np.zeros(n, np.int32)
        ls = np.zeros(n, np.float32) if linstr else None
        st = CStats()
        rc = self.lib.fso_maglineop(self.ctx, C.byref(cl), nlim, ptr(b, C.c_double), bframe,
                                    ptr(ls, C.c_float), ptr(flags, C.c_int32), ptr(il, C.c_int32),
                                    ptr(ir, C.c_int32), C.byref(st))
        if rc != 0:
            raise RuntimeError(self.lib.fso_last_error().decode())
        return dict(flags=flags, ileft=il, iright=ir, linstr=ls, stats=st.as_dict())

    def get_lineop(self) -> np.ndarray:
        out = np.zeros((self.setup.nwl, self.setup.ntau))
        self.lib.fso_get_lineop(self.ctx, ptr(out, C.c_double))
        return out

    def set_lineop(self, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float64)
        self.lib.fso_set_lineop(self.ctx, ptr(a, C.c_double))

    def get_maglineop(self) -> np.ndarray:
        out = np.zeros((self.setup.nwl, self.setup.ntau, 6))
        self.lib.fso_get_maglineop(self.ctx, ptr(out, C.c_double))
        return out

    def radtran(self):
        s = self.setup
        spl = np.zeros((s.nwl, s.nmu))
        spc = np.zeros((s.nwl, s.nmu))
        self.lib.fso_radtran(self.ctx, ptr(spl, C.c_double), ptr(spc, C.c_double))
        return spl, spc

    def magradtran(self, bfield, bframe: int):
        s = self.setup
        b = np.ascontiguousarray(bfield, np.float64)
        spz = np.zeros((s.nwl, 4, s.nmu))
        spc = np.zeros((s.nwl, s.nmu))
        rc = self.lib.fso_magradtran(self.ctx, ptr(b, C.c_double), b.shape[0], bframe,
                                     ptr(spz, C.c_double), ptr(spc, C.c_double))
        assert rc == 0
        return spz, spc

    def flux(self):
        s = self.setup
        sfl = np.zeros(s.nwl)
        sfc = np.zeros(s.nwl)
        self.lib.fso_flux(self.ctx, ptr(sfl, C.c_double), ptr(sfc, C.c_double))
        return sfl, sfc

    def integ_convol(self, cimode: int, vsini: float, vmacro: float, vgau: float, quad: int):
        """IntegConvol: mode 2 returns (SfL, SfC) of DiskInteg; mode 1 broadens the held
        intensity spectra in place (read them with spectra())."""
        n = self.setup.nwl
        sfl, sfc = np.zeros(n), np.zeros(n)
        rc = self.lib.fso_integ_convol(self.ctx, cimode, vsini, vmacro, vgau, quad, ptr(sfl, C.c_double),
                                       ptr(sfc, C.c_double))
        if rc != 0:
            raise RuntimeError(self.lib.fso_last_error().decode())
        return sfl, sfc

    def set_spectra(self, spl, spc, spz=None):
        """Replaces the held intensities: spl or spz (a polarised solution), and spc."""
        c = np.ascontiguousarray(spc, np.float64)
        if spz is not None:
            z = np.ascontiguousarray(spz, np.float64)
            self.lib.fso_set_spectra(self.ctx, None, ptr(c, C.c_double), ptr(z, C.c_double))
        else:
            l = np.ascontiguousarray(spl, np.float64)
            self.lib.fso_set_spectra(self.ctx, ptr(l, C.c_double), ptr(c, C.c_double), None)

    def spectra(self, magnetic: bool = False):
        s = self.setup
        spc = np.zeros((s.nwl, s.nmu))
        if magnetic:
            spz = np.zeros((s.nwl, 4, s.nmu))
            self.lib.fso_get_spectra(self.ctx, None, ptr(spc, C.c_double), ptr(spz, C.c_double))
            return spz, spc
        spl = np.zeros((s.nwl, s.nmu))
        self.lib.fso_get_spectra(self.ctx, ptr(spl, C.c_double), ptr(spc, C.c_double), None)
        return spl, spc

    def sphere_geometry(self) -> dict:
        """SphereInit's rays: ntau_mu, graz [nmu]; map (1-based layers), rhox_mu, muloc
        [nmu][2 ntau]."""
        m, L = self.setup.nmu, 2 * self.setup.ntau
        nt, gz = np.zeros(m, np.int32), np.zeros(m, np.int32)
        mp = np.zeros((m, L), np.int32)
        rm, ml = np.zeros((m, L)), np.zeros((m, L))
        rc = self.lib.fso_sphere_geometry(self.ctx, ptr(nt, C.c_int32), ptr(gz, C.c_int32), ptr(mp, C.c_int32),
                                          ptr(rm, C.c_double), ptr(ml, C.c_double))
        assert rc == 0
        return dict(ntau_mu=nt, graz=gz, map=mp, rhox_mu=rm, muloc=ml)

    def locprfsph(self, xki, src) -> np.ndarray:
        out = np.zeros(self.setup.nmu)
        a = np.ascontiguousarray(xki, np.float64)
        b = np.ascontiguousarray(src, np.float64)
        self.lib.fso_locprfsph(self.ctx, ptr(out, C.c_double), ptr(a, C.c_double), ptr(b, C.c_double))
        return out

    def sph_field_geometry(self, bfield, bframe: int) -> np.ndarray:
        """[nmu][2 ntau][5]: SPhi2, CPhi2, CPhi, S2Psi, C2Psi at each ray point."""
        b = np.ascontiguousarray(bfield, np.float64)
        geo = np.zeros((self.setup.nmu, 2 * self.setup.ntau, 5))
        rc = self.lib.fso_sph_field_geometry(self.ctx, ptr(b, C.c_double), b.shape[0], bframe,
                                             ptr(geo, C.c_double))
        if rc != 0:
            raise RuntimeError(self.lib.fso_last_error().decode())
        return geo

    def xline(self, lines: LineList, i: int):
        nt = self.setup.ntau
        cop, vv, av = np.zeros(nt), np.zeros(nt), np.zeros(nt)
        cl = lines.cstruct()
        self.lib.fso_xline(self.ctx, lines.list_type, C.byref(cl), i, ptr(cop, C.c_double),
                           ptr(vv, C.c_double), ptr(av, C.c_double))
        return cop, vv, av


# --------------------------------------------------------------------------------------
# reference kernels
# --------------------------------------------------------------------------------------

def ref_available() -> bool:
    return REF_SO.exists()


def build_ref() -> bool:
    """Builds oracle/_ref from the reference sources when the reference tree is present."""
    if REF_SO.exists():
        return True
    if not Path("/root/reference/src/voigt.f90").exists():
        return False
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], capture_output=True)
    return r.returncode == 0 and REF_SO.exists()


_REF = None


def ref_lib() -> C.CDLL:
    global _REF
    if _REF is None:
        lib = C.CDLL(str(REF_SO))
        lib.ref_waveidx.restype = C.c_int32
        lib.ref_waveidx.argtypes = [C.c_double]
        lib.ref_air2vac.restype = C.c_double
        lib.ref_air2vac.argtypes = [C.c_double]
        lib.ref_nthreads.restype = C.c_int32
        lib.ref_semi_area.restype = C.c_double
        lib.ref_semi_area.argtypes = [C.c_double, C.c_double]
        lib.ref_pow10.argtypes = [C.c_float, C.POINTER(C.c_float)]
        lib.ref_vac2air.restype = C.c_double
        lib.ref_vac2air.argtypes = [C.c_double]
        lib.ref_bounds.argtypes = [C.c_int, _dp, C.c_double, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.ref_convolgau.argtypes = [C.c_int, _dp, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double]
        lib.ref_diskinteg.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _dp,
                                      C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        lib.ref_mu_quadrature_r.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
        _REF = lib
    return _REF


class Ref:
    """Reference module globals filled from a Setup (one Ntau per process)."""

    SLOTS = ("h1", "he1", "hmol", "h2", "he2")

    def __init__(self, setup: Setup, slot_h2: int | None = None, slot_he2: int | None = None):
        self.lib = ref_lib()
        self.setup = s = setup
        nsp = int(s.spc_ion.shape[0])
        nslot = int(s.fraction_pf.shape[1])
        # slots of the perturbers: the workload builder puts H I, He I, H2, H II, He II
        # right after the line species
        base = nsp
        self.slot = dict(h1=base + 1, he1=base + 2, hmol=base + 3,
                         h2=slot_h2 or base + 4, he2=slot_he2 or base + 5)
        frac = np.ascontiguousarray(s.fraction_pf.copy())
        frac[:, self.slot["h1"] - 1] = s.frac_h1
        frac[:, self.slot["he1"] - 1] = s.frac_he1
        frac[:, self.slot["hmol"] - 1] = s.frac_hmol
        potion = np.zeros(nslot, np.float32)
        potion[: s.pot_ion.shape[0]] = s.pot_ion
        xam = np.ones(nslot, np.float32)
        xam[: s.xamass.shape[0]] = s.xamass
        cstd = s.contop_std if s.contop_std is not None else np.ones(s.ntau)
        self._keep = (frac, potion, xam, cstd)
        d = C.c_double
        self.lib.ref_setup(
            C.c_int(s.ntau), C.c_int(s.nwl), C.c_int(s.nmu), C.c_int(s.wmode), d(s.wfirst),
            d(s.dstep), d(s.rstep), d(s.wstart), ptr(s.wave, d), C.c_int(s.motype),
            ptr(s.rhox, d), ptr(s.temp, d), ptr(s.xne, d), ptr(s.rho, d), ptr(s.contop, d),
            ptr(np.ascontiguousarray(cstd, np.float64), d), C.c_int(nslot),
            ptr(s.fraction_pf, C.c_float), ptr(frac, C.c_float), ptr(potion, C.c_float),
            ptr(xam, C.c_float), C.c_int(nsp), ptr(s.spc_ion, C.c_int32), ptr(s.spc_slot, C.c_int32),
            C.c_int(self.slot["h1"]), C.c_int(self.slot["he1"]), C.c_int(self.slot["hmol"]),
            C.c_int(self.slot["h2"]), C.c_int(self.slot["he2"]),
            d(s.eps_opac), d(s.vwfudge), d(s.h2brd), d(s.vwmol), d(s.vmicro),
            C.c_int(int(bool(s.molbrd))), ptr(s.xmu, d), ptr(s.wmu, d))

    def hestark(self):
        nt = self.setup.ntau
        n = C.c_int32()
        wl, elo = np.zeros(64), np.zeros(64)
        gam = np.zeros((64, nt))
        shf = np.zeros((64, nt))
        self.lib.ref_hestark(C.byref(n), ptr(wl, C.c_double), ptr(elo, C.c_double),
                             ptr(gam, C.c_double), ptr(shf, C.c_double))
        k = n.value
        # (Ntau, NheTab) column-major == [nhe][ntau] C-order
        g = gam.reshape(-1)[: k * nt].reshape(k, nt)
        sh = shf.reshape(-1)[: k * nt].reshape(k, nt)
        return wl[:k].copy(), elo[:k].copy(), g.copy(), sh.copy()

    def set_lines(self, L: LineList):
        n = L.n
        z = np.zeros(n, np.float32)
        ismol = np.zeros(n, np.int32) if L.is_mol is None else L.is_mol.astype(np.int32)
        if L.pi_off is not None:
            args = (C.c_int(1), ptr(L.pi_off, C.c_int32), ptr(L.spl_pi, C.c_float), ptr(L.str_pi, C.c_float),
                    ptr(L.si_off, C.c_int32), ptr(L.spl_si, C.c_float), ptr(L.str_si, C.c_float))
        else:
            zi = np.zeros(1, np.int32)
            zf = np.zeros(1, np.float32)
            args = (C.c_int(0), ptr(zi, C.c_int32), ptr(zf, C.c_float), ptr(zf, C.c_float),
                    ptr(zi, C.c_int32), ptr(zf, C.c_float), ptr(zf, C.c_float))
        gq = L.gqst if L.gqst is not None else z
        gv = L.gvdw if L.gvdw is not None else z
        self._lines_keep = (z, ismol, gq, gv, L)
        self.lib.ref_set_lines(C.c_int(L.list_type), C.c_int(n), ptr(L.wl, C.c_double), ptr(L.elo, C.c_float),
                               ptr(L.gf, C.c_float), ptr(L.grad, C.c_float), ptr(gq, C.c_float),
                               ptr(gv, C.c_float), ptr(L.spc, C.c_int32), ptr(ismol, C.c_int32), *args)
        self.nlines = n

    def xline(self):
        n, nt = self.nlines, self.setup.ntau
        cop, vv, av = (np.zeros((n, nt)) for _ in range(3))
        flag = np.zeros(n, np.int32)
        self.lib.ref_xline(ptr(cop, C.c_double), ptr(vv, C.c_double), ptr(av, C.c_double),
                           ptr(flag, C.c_int32))
        return cop, vv, av, flag

    def set_windows(self, flags, il, ir):
        f = np.ascontiguousarray(flags, np.int32)
        a = np.ascontiguousarray(il, np.int32)
        b = np.ascontiguousarray(ir, np.int32)
        self.lib.ref_set_windows(ptr(f, C.c_int32), ptr(a, C.c_int32), ptr(b, C.c_int32))

    def set_block(self, cop, vv, av):
        n = cop.shape[0]
        self._blk = tuple(np.ascontiguousarray(x, np.float64) for x in (cop, vv, av))
        self.lib.ref_set_block(C.c_int(n), *(ptr(x, C.c_double) for x in self._blk))

    def lineop1(self, iwl1: int, iwl2: int, l1: int, l2: int) -> np.ndarray:
        out = np.zeros((iwl2 - iwl1 + 1, self.setup.ntau))
        self.lib.ref_lineop1(C.c_int(iwl1), C.c_int(iwl2), C.c_int(l1), C.c_int(l2), ptr(out, C.c_double))
        return out

    def maglineop1(self, iwl1: int, iwl2: int, l1: int, l2: int) -> np.ndarray:
        out = np.zeros((iwl2 - iwl1 + 1, self.setup.ntau, 6))
        self.lib.ref_maglineop1(C.c_int(iwl1), C.c_int(iwl2), C.c_int(l1), C.c_int(l2), ptr(out, C.c_double))
        return out

    def analytic_window(self, i: int, cop, vv, av, contl) -> float:
        dw = C.c_double()
        arrs = [np.ascontiguousarray(x, np.float64) for x in (cop, vv, av, contl)]
        self.lib.ref_analytic_window(C.c_int(i), *(ptr(x, C.c_double) for x in arrs), C.byref(dw))
        return dw.value

    def mag_window(self, i: int, bsplit: float, cop, vv, av, contl):
        il, ir = C.c_int32(), C.c_int32()
        arrs = [np.ascontiguousarray(x, np.float64) for x in (cop, vv, av, contl)]
        self.lib.ref_mag_window(C.c_int(i), C.c_double(bsplit), *(ptr(x, C.c_double) for x in arrs),
                                C.byref(il), C.byref(ir))
        return il.value, ir.value

    def radtran(self, iwl1: int, iwl2: int, lop: np.ndarray):
        n = iwl2 - iwl1 + 1
        lop = np.ascontiguousarray(lop, np.float64)
        spl = np.zeros((n, self.setup.nmu))
        spc = np.zeros((n, self.setup.nmu))
        self.lib.ref_radtran(C.c_int(iwl1), C.c_int(iwl2), ptr(lop, C.c_double), ptr(spl, C.c_double),
                             ptr(spc, C.c_double))
        return spl, spc

    def locprfpp(self, xki, src):
        out = np.zeros(self.setup.nmu)
        a = np.ascontiguousarray(xki, np.float64)
        b = np.ascontiguousarray(src, np.float64)
        self.lib.ref_locprfpp(ptr(out, C.c_double), ptr(a, C.c_double), ptr(b, C.c_double))
        return out

    def sphere(self, radius: float, height) -> dict:
        m, nt = self.setup.nmu, self.setup.ntau
        h = np.ascontiguousarray(height, np.float64)
        ntmu, graz = np.zeros(m, np.int32), np.zeros(m, np.int32)
        mp = np.zeros((m, 2 * nt), np.int32)  # (2 Ntau, Nmu) column-major
        rm, ml = np.zeros((m, 2 * nt)), np.zeros((m, 2 * nt))
        self.lib.ref_sphere(C.c_double(radius), ptr(h, C.c_double), ptr(ntmu, C.c_int32), ptr(graz, C.c_int32),
                            ptr(mp, C.c_int32), ptr(rm, C.c_double), ptr(ml, C.c_double))
        return dict(ntau_mu=ntmu, graz=graz, map=mp, rhox_mu=rm, muloc=ml)

    def locprfsph(self, xki, src):
        out = np.zeros(self.setup.nmu)
        a = np.ascontiguousarray(xki, np.float64)
        b = np.ascontiguousarray(src, np.float64)
        self.lib.ref_locprfsph(ptr(out, C.c_double), ptr(a, C.c_double), ptr(b, C.c_double))
        return out

    def stokesprfsph(self, iwl, mlop, lop, src, geo):
        """geo [nmu][2 ntau][5] as the oracle's sph_field_geometry."""
        out = np.zeros((4, self.setup.nmu))
        g5 = [np.ascontiguousarray(geo[:, :, q]) for q in range(5)]
        arrs = [np.ascontiguousarray(x, np.float64) for x in (mlop, lop, src)] + g5
        self.lib.ref_stokesprfsph(C.c_int(iwl), *(ptr(x, C.c_double) for x in arrs), ptr(out, C.c_double))
        return out

    def stokesprfpp(self, iwl, mlop, lop, src, geo):
        out = np.zeros((4, self.setup.nmu))
        arrs = [np.ascontiguousarray(x, np.float64) for x in (mlop, lop, src, *geo)]
        self.lib.ref_stokesprfpp(C.c_int(iwl), *(ptr(x, C.c_double) for x in arrs), ptr(out, C.c_double))
        return out


def ref_diskinteg(setup: Setup, quad: int, vsini: float, vmacro: float, vgau: float, spl, spc):
    """The reference's DiskInteg on intensities spl, spc [nwl][nmu]: (SfL, SfC)."""
    lib = ref_lib()
    wave = np.asarray(setup.wave, np.float64)
    wlast = setup.wfirst + wave[-1]
    a = np.ascontiguousarray(spl, np.float64)
    b = np.ascontiguousarray(spc, np.float64)
    xmu = np.ascontiguousarray(setup.xmu, np.float64)
    sfl, sfc = np.zeros(setup.nwl), np.zeros(setup.nwl)
    lib.ref_diskinteg(setup.nmu, setup.nwl, setup.wmode, setup.wfirst, wlast, setup.dstep, quad,
                      ptr(xmu, C.c_double), vsini, vmacro, vgau, ptr(a, C.c_double), ptr(b, C.c_double),
                      ptr(sfl, C.c_double), ptr(sfc, C.c_double))
    return sfl, sfc


def ref_convolgau(setup: Setup, s, fwhm: float):
    lib = ref_lib()
    wave = np.asarray(setup.wave, np.float64)
    out = np.array(s, np.float64, copy=True)
    lib.ref_convolgau(out.shape[0], ptr(out, C.c_double), fwhm, setup.wmode, setup.wfirst, setup.wfirst + wave[-1],
                      setup.dstep)
    return out


def ref_xcont(temp, xne, rho, pops: dict, sources, wl_vac, xna=None, hop_nolev=False) -> np.ndarray:
    """The reference's Xcont (src/contop.f90:267-314), compiled from its sources, at vacuum
    wavelengths wl_vac with only the Ifop sources given switched on (hop_nolev: Ifop(1) = -1):
    Copac [n][ntau].  pops holds the EOS rows (fsg_contop_eos's names) that the harness puts
    in its slots 1 H I, 2 H II, 3 H-, 4 He I, 5 H2, 6 He II, 7 He III, 8 C I, 9 Mg I, 10 Al I,
    11 Si I, 12 Fe I, 13 CH, 14 OH, 15 NH (16, the other species, empty), 17 C II, 18 Mg II,
    19 Si II, 20 N I, 21 O I, 22 Ca II and 23-70 HOTOP's stages I-VI of C, N, O, Ne, Mg, Si,
    S, Fe (hot_frac [ntau][8][6]; hot_fpf [ntau][21] the FractionPF of C I-IV, N I-V, O I-VI,
    Ne I-VI); rows not given stay zero."""
    lib = ref_lib()
    n = int(len(wl_vac))
    ntau = int(len(temp))
    fpf = np.zeros((ntau, 70), np.float32)
    frac = np.zeros((ntau, 70), np.float32)
    fpf[:, 0] = pops["fpf_h1"]
    fpf[:, 1] = pops["fpf_h2"]
    frac[:, 2] = pops["frac_hmin"]
    fpf[:, 3] = pops["fpf_he1"]
    frac[:, 4] = pops["frac_hmol"]
    # Fraction and FractionPF of the other rows as the harness reads them: H I, H II and He I
    # with their FractionPF (Fraction of H I and He I, which H2COLLOP reads, when given), H-
    # and H2 with their Fraction
    frac[:, 0], frac[:, 1], frac[:, 3] = fpf[:, 0], fpf[:, 1], fpf[:, 3]
    fpf[:, 2], fpf[:, 4] = frac[:, 2], frac[:, 4]
    if "frac_h1" in pops:
        frac[:, 0] = pops["frac_h1"]
    if "frac_he1" in pops:
        frac[:, 3] = pops["frac_he1"]
    if "fpf_he2" in pops:
        fpf[:, 5], frac[:, 5] = pops["fpf_he2"], pops["frac_he2"]
        fpf[:, 6] = frac[:, 6] = pops["fpf_he3"]
    for slot, key in ((7, "fpf_c1"), (8, "fpf_mg1"), (9, "fpf_al1"), (10, "fpf_si1"), (11, "fpf_fe1"),
                      (12, "fpf_ch"), (13, "fpf_oh")):
        if key in pops:
            fpf[:, slot] = frac[:, slot] = pops[key]
    if "frac_nh" in pops:
        fpf[:, 14] = frac[:, 14] = pops["frac_nh"]
    for slot, key in ((16, "fpf_c2"), (17, "fpf_mg2"), (18, "fpf_si2"), (19, "fpf_n1"), (20, "fpf_o1"),
                      (21, "fpf_ca2")):
        if key in pops:
            fpf[:, slot] = frac[:, slot] = pops[key]
    if "hot_frac" in pops:
        hf = np.asarray(pops["hot_frac"], np.float32).reshape(ntau, 48)
        frac[:, 22:70] = hf
        fpf[:, 22:70] = hf
        # the FractionPF rows the bound-free edges read: C I-IV, N I-V, O I-VI, Ne I-VI
        cols = [22 + 6 * k + i for k, nst in enumerate((4, 5, 6, 6)) for i in range(nst)]
        fpf[:, cols] = np.asarray(pops["hot_fpf"], np.float32).reshape(ntau, 21)
    ifop = np.zeros(20, np.int32)
    for s in sources:
        ifop[int(s) - 1] = 1
    if hop_nolev:
        ifop[0] = -1
    out = np.zeros((n, ntau))
    ar = lambda a: np.ascontiguousarray(a, np.float64)
    t, ne, r = ar(temp), ar(xne), ar(rho)
    na = ar(np.zeros(ntau) if xna is None else xna)
    w = ar(wl_vac)
    lib.ref_xcont(C.c_int(ntau), ptr(t, C.c_double), ptr(ne, C.c_double), ptr(na, C.c_double), ptr(r, C.c_double),
                  ptr(np.ascontiguousarray(fpf), C.c_float), ptr(np.ascontiguousarray(frac), C.c_float),
                  ptr(ifop, C.c_int32), C.c_int(n), ptr(w, C.c_double), ptr(out, C.c_double))
    return out


def ref_air2vac_array(wl) -> np.ndarray:
    lib = ref_lib()
    return np.array([lib.ref_air2vac(float(x)) for x in wl])
