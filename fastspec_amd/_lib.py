"""Loader of libfsgpu.so (the HIP library).  There is no CPU fallback: if the library is
missing or no GPU is visible, every call raises."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from .abi import CAtmosphere, CContopEos, CIsotope, CLines, CSetup, CSpecies, CStats

LIB_PATH = Path(__file__).resolve().parent / "libfsgpu.so"

# every symbol declared in include/fsgpu.h
EXPORTS = (
    "fsg_init", "fsg_finalize", "fsg_last_error", "fsg_lineop", "fsg_maglineop",
    "fsg_get_lineop", "fsg_set_lineop", "fsg_get_maglineop", "fsg_radtran",
    "fsg_magradtran", "fsg_flux", "fsg_lines_upload", "fsg_lines_free", "fsg_lineop_dev",
    "fsg_device_ptr", "fsg_stream", "fsg_reset_lineop", "fsg_maglineop_dev", "fsg_lines_set_gf",
    "fsg_flux_vjp", "fsg_lineop_vjp", "fsg_integ_convol", "fsg_get_spectra",
    "fsg_lines_scan", "fsg_lines_read", "fsg_lines_stream", "fsg_init_batch", "fsg_batch_offsets",
    "fsg_lines_batch", "fsg_integ_convol_noalloc", "fsg_flux_dev", "fsg_flux_vjp_dev", "fsg_lineop_vjp_dev",
    "fsg_set_atmosphere", "fsg_contop",
)

_LIB = None


class FsgError(RuntimeError):
    pass


def _share_torch_runtime() -> None:
    """torch's ROCm wheel carries its own HIP runtime under the soname the library links
    (libamdhip64.so.7).  Loaded after torch, the library binds to that runtime, so library
    and torch tensors share the device; two runtimes in one process do not (torch then sees
    no GPU).  FSG_NO_TORCH=1 leaves torch out (the library then loads the system runtime)."""
    if os.environ.get("FSG_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except Exception:  # torch absent: the library's own runtime
        pass


def load() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    _share_torch_runtime()
    if not LIB_PATH.exists():
        raise FsgError(f"{LIB_PATH} is not built: run `python -m fastspec_amd.build` "
                       "(the HIP library is required; there is no CPU path)")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int32)
    lib.fsg_last_error.restype = C.c_char_p
    lib.fsg_init.argtypes = [C.POINTER(CSetup), C.POINTER(vp)]
    lib.fsg_finalize.argtypes = [vp]
    lib.fsg_lineop.argtypes = [vp, C.c_int32, C.POINTER(CLines), C.c_int64, fp, ip, ip, ip, C.POINTER(CStats)]
    lib.fsg_maglineop.argtypes = [vp, C.POINTER(CLines), C.c_int64, dp, C.c_int32, fp, ip, ip, ip,
                                  C.POINTER(CStats)]
    lib.fsg_get_lineop.argtypes = [vp, dp]
    lib.fsg_set_lineop.argtypes = [vp, dp]
    lib.fsg_get_maglineop.argtypes = [vp, dp]
    lib.fsg_radtran.argtypes = [vp, dp, dp]
    lib.fsg_magradtran.argtypes = [vp, dp, C.c_int32, C.c_int32, dp, dp]
    lib.fsg_flux.argtypes = [vp, dp, dp]
    lib.fsg_integ_convol.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_double]
    lib.fsg_integ_convol_noalloc.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_double]
    lib.fsg_get_spectra.argtypes = [vp, dp, dp, dp]
    lib.fsg_lines_upload.argtypes = [vp, C.c_int32, C.POINTER(CLines), C.POINTER(vp)]
    lib.fsg_lines_free.argtypes = [vp]
    lib.fsg_lineop_dev.argtypes = [vp, vp, C.c_int64, C.POINTER(CStats)]
    lib.fsg_maglineop_dev.argtypes = [vp, vp, C.c_int64, dp, C.c_int32, C.POINTER(CStats)]
    lib.fsg_lines_set_gf.argtypes = [vp, vp, C.c_int32]
    lib.fsg_flux_vjp.argtypes = [vp, dp, dp, dp, dp, dp]
    lib.fsg_lineop_vjp.argtypes = [vp, vp, C.c_int64, dp, dp, dp]
    lib.fsg_flux_dev.argtypes = [vp, vp, vp]
    lib.fsg_flux_vjp_dev.argtypes = [vp, vp, vp, vp, dp]
    lib.fsg_lineop_vjp_dev.argtypes = [vp, vp, C.c_int64, vp, dp, dp]
    lib.fsg_device_ptr.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    lib.fsg_stream.argtypes = [vp, C.POINTER(vp)]
    lib.fsg_reset_lineop.argtypes = [vp]
    i64p = C.POINTER(C.c_int64)
    sp = C.POINTER(CSpecies)
    lib.fsg_lines_scan.argtypes = [C.c_char_p, C.c_double, C.c_double, ip, i64p, i64p]
    isp = C.POINTER(CIsotope)
    lib.fsg_lines_read.argtypes = [C.c_char_p, C.c_int64, C.c_int64, sp, C.c_int32, isp, C.c_int32, C.c_int32,
                                   C.c_int32, dp, fp, fp, fp, fp, fp, ip, C.POINTER(C.c_uint8), i64p]
    lib.fsg_lines_stream.argtypes = [vp, C.c_char_p, C.c_double, C.c_double, sp, C.c_int32, isp, C.c_int32,
                                     C.c_int32, C.c_int32, C.POINTER(vp), i64p]
    lib.fsg_init_batch.argtypes = [C.POINTER(CSetup), C.c_int32, C.POINTER(vp)]
    lib.fsg_batch_offsets.argtypes = [vp, i64p]
    lib.fsg_lines_batch.argtypes = [vp, vp, i64p, ip, C.c_int64, C.POINTER(vp)]
    lib.fsg_set_atmosphere.argtypes = [vp, C.c_int32, C.POINTER(CAtmosphere)]
    lib.fsg_contop.argtypes = [vp, C.POINTER(CContopEos), C.c_int32, C.c_int32, dp]
    _LIB = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        raise FsgError(load().fsg_last_error().decode())
