"""ctypes mirror of include/fsgpu.h.

The structures here are the Python view of the C-ABI that the Fortran host binds with
ISO_C_BINDING.  ``Setup`` and ``LineList`` hold numpy arrays in the host's layouts and
build the C structures on demand, keeping the arrays alive for the duration of a call.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

FSG_FP64 = 0
FSG_FP32 = 1
FSG_LIST_ATOMIC = 0
FSG_LIST_MOLECULAR = 1
FSG_LIST_MIXED = 2

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint8)


class CSetup(C.Structure):
    _fields_ = [
        ("ntau", C.c_int32), ("nwl", C.c_int32), ("nmu", C.c_int32),
        ("wmode", C.c_int32),
        ("wfirst", C.c_double), ("dstep", C.c_double), ("rstep", C.c_double),
        ("wstart", C.c_double),
        ("wave", _dp),
        ("motype", C.c_int32),
        ("rhox", _dp), ("temp", _dp), ("xne", _dp), ("rho", _dp),
        ("contop", _dp), ("contop_std", _dp),
        ("pf_ld", C.c_int32), ("nslot", C.c_int32),
        ("fraction_pf", _fp), ("pot_ion", _fp), ("xamass", _fp),
        ("frac_h1", _fp), ("frac_he1", _fp), ("frac_hmol", _fp),
        ("nspc", C.c_int32), ("spc_ion", _ip), ("spc_slot", _ip),
        ("he1_slot", C.c_int32),
        ("nhe", C.c_int32), ("he_wl", _dp), ("he_elo", _dp),
        ("he_wl_tol", C.c_double), ("he_elo_tol", C.c_double),
        ("he_gam", _dp), ("he_shf", _dp),
        ("eps_opac", C.c_double), ("vwfudge", C.c_double), ("h2brd", C.c_double),
        ("vwmol", C.c_double), ("vmicro", C.c_double),
        ("molbrd", C.c_int32),
        ("xmu", _dp), ("wmu", _dp),
        ("precision", C.c_int32), ("batched", C.c_int32), ("device", C.c_int32),
        ("radius", C.c_double), ("height", _dp),
    ]


class CSpecies(C.Structure):
    """fsg_species_entry: (symbol, ion) -> species slot of the host's table."""
    _fields_ = [("el", C.c_char * 8), ("ion", C.c_int32), ("slot", C.c_int32), ("is_mol", C.c_int32)]


class CIsotope(C.Structure):
    """fsg_isotope_entry: (element, mass number) -> isotopic fraction of the host's table."""
    _fields_ = [("el", C.c_char * 8), ("a", C.c_int32), ("frac", C.c_double)]


class CLines(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("wl", _dp),
        ("elo", _fp), ("gf", _fp), ("grad", _fp),
        ("gqst", _fp), ("gvdw", _fp),
        ("spc", _ip),
        ("is_mol", _up),
        ("pi_off", _ip), ("spl_pi", _fp), ("str_pi", _fp),
        ("si_off", _ip), ("spl_si", _fp), ("str_si", _fp),
    ]


class CAtmosphere(C.Structure):
    """fsg_atmosphere: a new EOS / continuum state for fsg_set_atmosphere (NULL: keep)."""
    _fields_ = [
        ("temp", _dp), ("xne", _dp), ("rho", _dp), ("rhox", _dp), ("contop", _dp), ("contop_std", _dp),
        ("fraction_pf", _fp), ("frac_h1", _fp), ("frac_he1", _fp), ("frac_hmol", _fp),
        ("he_gam", _dp),
    ]


CONTOP_HOP_NOLEV = 1 << 20  # FSG_CONTOP_HOP_NOLEV: Ifop(1) = -1, HOP without levels 1-6


class CContopEos(C.Structure):
    """fsg_contop_eos: the EOS rows the device's continuum sources read ([ntau] REAL(4))."""
    _fields_ = [("fpf_h1", _fp), ("fpf_h2", _fp), ("frac_hmin", _fp), ("fpf_he1", _fp), ("frac_hmol", _fp),
                ("frac_he2", _fp), ("fpf_he2", _fp), ("fpf_he3", _fp),
                ("fpf_c1", _fp), ("fpf_mg1", _fp), ("fpf_al1", _fp), ("fpf_si1", _fp), ("fpf_fe1", _fp),
                ("fpf_ch", _fp), ("fpf_oh", _fp), ("frac_nh", _fp), ("frac_h1", _fp), ("frac_he1", _fp),
                ("fpf_c2", _fp), ("fpf_mg2", _fp), ("fpf_si2", _fp), ("fpf_n1", _fp), ("fpf_o1", _fp),
                ("fpf_ca2", _fp), ("hot_frac", _fp), ("hot_fpf", _fp)]


class CStats(C.Structure):
    _fields_ = [
        ("nlines_in", C.c_int64), ("nlines_use", C.c_int64), ("nwide", C.c_int64),
        ("nblocks", C.c_int64), ("nsegments", C.c_int64), ("evals", C.c_int64),
        ("ms_pass1", C.c_double), ("ms_opacity", C.c_double),
        ("evals_core", C.c_int64), ("ms_far", C.c_double), ("ms_core", C.c_double),
        ("core_tile", C.c_int32), ("far_ring", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def ptr(a: Optional[np.ndarray], ctype):
    """Pointer to a C-contiguous array of the right dtype, or NULL."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


@dataclass
class Setup:
    """Host-side module globals that LineopCalc/RadTranCalc read (src/shared.f90:392-548)."""
    wave: np.ndarray              # [nwl] offsets from wfirst
    wmode: int
    wfirst: float
    dstep: float
    rstep: float
    wstart: float
    rhox: np.ndarray
    temp: np.ndarray
    xne: np.ndarray
    rho: np.ndarray
    contop: np.ndarray            # [nwl][ntau]
    fraction_pf: np.ndarray       # [ntau][pf_ld] float32
    pot_ion: np.ndarray           # [nslot] float32
    xamass: np.ndarray            # [nslot] float32
    frac_h1: np.ndarray           # [ntau] float32
    frac_he1: np.ndarray
    frac_hmol: np.ndarray
    spc_ion: np.ndarray           # [nspc] int32
    spc_slot: np.ndarray          # [nspc] int32 (1-based)
    xmu: np.ndarray
    wmu: np.ndarray
    motype: int = 0
    contop_std: Optional[np.ndarray] = None
    he1_slot: int = 0
    he_wl: Optional[np.ndarray] = None
    he_elo: Optional[np.ndarray] = None
    he_gam: Optional[np.ndarray] = None   # [nhe][ntau]
    he_shf: Optional[np.ndarray] = None
    he_wl_tol: float = 1.0
    he_elo_tol: float = 0.01
    eps_opac: float = 1e-4
    vwfudge: float = 1.0
    h2brd: float = 0.0
    vwmol: float = 0.0
    vmicro: float = 0.0
    molbrd: bool = False
    radius: float = -1.0                  # spherical models: Radius [cm] (<= 0: plane-parallel)
    height: Optional[np.ndarray] = None   # [ntau] Atmos%Height [cm]
    # extra host data that is not part of the C structure (kept for writers and tests)
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.wave = _f64(self.wave)
        for k in ("rhox", "temp", "xne", "rho", "xmu", "wmu"):
            setattr(self, k, _f64(getattr(self, k)))
        self.contop = _f64(self.contop)
        self.contop_std = _f64(self.contop_std)
        for k in ("fraction_pf", "pot_ion", "xamass", "frac_h1", "frac_he1", "frac_hmol"):
            setattr(self, k, _f32(getattr(self, k)))
        self.spc_ion = _i32(self.spc_ion)
        self.spc_slot = _i32(self.spc_slot)
        for k in ("he_wl", "he_elo", "he_gam", "he_shf", "height"):
            setattr(self, k, _f64(getattr(self, k)))

    @property
    def ntau(self) -> int:
        return int(self.temp.shape[0])

    @property
    def nwl(self) -> int:
        return int(self.wave.shape[0])

    @property
    def nmu(self) -> int:
        return int(self.xmu.shape[0])

    @property
    def wavelengths(self) -> np.ndarray:
        return self.wave + self.wfirst

    def cstruct(self, precision: int = FSG_FP64, batched: int | None = None,
                device: int = 0) -> CSetup:
        if batched is None:
            batched = 1 if precision == FSG_FP32 else 0
        assert self.contop.shape == (self.nwl, self.ntau)
        assert self.fraction_pf.ndim == 2 and self.fraction_pf.shape[0] == self.ntau
        nhe = 0 if self.he_wl is None else int(self.he_wl.shape[0])
        cs = CSetup(
            ntau=self.ntau, nwl=self.nwl, nmu=self.nmu, wmode=int(self.wmode),
            wfirst=self.wfirst, dstep=self.dstep, rstep=self.rstep, wstart=self.wstart,
            wave=ptr(self.wave, C.c_double), motype=int(self.motype),
            rhox=ptr(self.rhox, C.c_double), temp=ptr(self.temp, C.c_double),
            xne=ptr(self.xne, C.c_double), rho=ptr(self.rho, C.c_double),
            contop=ptr(self.contop, C.c_double), contop_std=ptr(self.contop_std, C.c_double),
            pf_ld=int(self.fraction_pf.shape[1]), nslot=int(self.pot_ion.shape[0]),
            fraction_pf=ptr(self.fraction_pf, C.c_float), pot_ion=ptr(self.pot_ion, C.c_float),
            xamass=ptr(self.xamass, C.c_float), frac_h1=ptr(self.frac_h1, C.c_float),
            frac_he1=ptr(self.frac_he1, C.c_float), frac_hmol=ptr(self.frac_hmol, C.c_float),
            nspc=int(self.spc_ion.shape[0]), spc_ion=ptr(self.spc_ion, C.c_int32),
            spc_slot=ptr(self.spc_slot, C.c_int32), he1_slot=int(self.he1_slot),
            nhe=nhe, he_wl=ptr(self.he_wl, C.c_double), he_elo=ptr(self.he_elo, C.c_double),
            he_wl_tol=self.he_wl_tol, he_elo_tol=self.he_elo_tol,
            he_gam=ptr(self.he_gam, C.c_double), he_shf=ptr(self.he_shf, C.c_double),
            eps_opac=self.eps_opac, vwfudge=self.vwfudge, h2brd=self.h2brd,
            vwmol=self.vwmol, vmicro=self.vmicro, molbrd=int(bool(self.molbrd)),
            xmu=ptr(self.xmu, C.c_double), wmu=ptr(self.wmu, C.c_double),
            precision=int(precision), batched=int(batched), device=int(device),
            radius=float(self.radius), height=ptr(self.height, C.c_double),
        )
        return cs


@dataclass
class LineList:
    """One line list after the host reader's conversions (src/ioutils.f90:1055-1076)."""
    wl: np.ndarray
    elo: np.ndarray
    gf: np.ndarray
    grad: np.ndarray
    spc: np.ndarray
    gqst: Optional[np.ndarray] = None
    gvdw: Optional[np.ndarray] = None
    is_mol: Optional[np.ndarray] = None
    list_type: int = FSG_LIST_ATOMIC
    pi_off: Optional[np.ndarray] = None
    spl_pi: Optional[np.ndarray] = None
    str_pi: Optional[np.ndarray] = None
    si_off: Optional[np.ndarray] = None
    spl_si: Optional[np.ndarray] = None
    str_si: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.wl = _f64(self.wl)
        for k in ("elo", "gf", "grad", "gqst", "gvdw", "spl_pi", "str_pi", "spl_si", "str_si"):
            setattr(self, k, _f32(getattr(self, k)))
        self.spc = _i32(self.spc)
        for k in ("pi_off", "si_off"):
            setattr(self, k, _i32(getattr(self, k)))
        if self.is_mol is not None:
            self.is_mol = np.ascontiguousarray(self.is_mol, dtype=np.uint8)

    @property
    def n(self) -> int:
        return int(self.wl.shape[0])

    def take(self, idx: np.ndarray) -> "LineList":
        """Subset of lines (keeps the Zeeman CSR arrays consistent)."""
        idx = np.asarray(idx)
        kw = dict(list_type=self.list_type, meta=dict(self.meta))
        for k in ("wl", "elo", "gf", "grad", "spc", "gqst", "gvdw", "is_mol"):
            v = getattr(self, k)
            kw[k] = None if v is None else v[idx]
        if self.pi_off is not None:
            for off, spl, st in (("pi_off", "spl_pi", "str_pi"), ("si_off", "spl_si", "str_si")):
                o = getattr(self, off)
                cnt = (o[1:] - o[:-1])[idx]
                new_off = np.zeros(len(idx) + 1, np.int32)
                np.cumsum(cnt, out=new_off[1:])
                sel = np.concatenate([np.arange(o[i], o[i + 1]) for i in idx]) if len(idx) else np.zeros(0, int)
                kw[off] = new_off
                kw[spl] = getattr(self, spl)[sel]
                kw[st] = getattr(self, st)[sel]
        return LineList(**kw)

    def cstruct(self) -> CLines:
        return CLines(
            n=self.n, wl=ptr(self.wl, C.c_double),
            elo=ptr(self.elo, C.c_float), gf=ptr(self.gf, C.c_float),
            grad=ptr(self.grad, C.c_float), gqst=ptr(self.gqst, C.c_float),
            gvdw=ptr(self.gvdw, C.c_float), spc=ptr(self.spc, C.c_int32),
            is_mol=ptr(self.is_mol, C.c_uint8),
            pi_off=ptr(self.pi_off, C.c_int32), spl_pi=ptr(self.spl_pi, C.c_float),
            str_pi=ptr(self.str_pi, C.c_float), si_off=ptr(self.si_off, C.c_int32),
            spl_si=ptr(self.spl_si, C.c_float), str_si=ptr(self.str_si, C.c_float),
        )
