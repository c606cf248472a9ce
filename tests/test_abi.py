"""The C-ABI library loads and exports every entry point declared in include/fsgpu.h.
No compute call is made here (this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from fastspec_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "fsgpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fsg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_loader_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_last_error_is_empty_string_initially():
    lib = _lib.load()
    assert isinstance(lib.fsg_last_error(), bytes)


def test_init_rejects_bad_setup_without_gpu_work():
    from fastspec_amd.abi import CSetup
    lib = _lib.load()
    cs = CSetup()
    cs.ntau = 1   # invalid before any device call
    ctx = ctypes.c_void_p()
    rc = lib.fsg_init(ctypes.byref(cs), ctypes.byref(ctx))
    assert rc != 0
    assert b"Ntau" in lib.fsg_last_error()


def test_init_batch_rejects_mismatched_setups_without_gpu_work():
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP64, CSetup
    lib = _lib.load()
    a = W.build_setup(W.grey_atmosphere(5000.0, 4.0), 5000.0, 5001.0, 0.01, 0)
    b = W.build_setup(W.grey_atmosphere(5000.0, 4.0, ntau=48), 5000.0, 5001.0, 0.01, 0)
    c = W.build_setup(W.grey_atmosphere(5000.0, 4.0), 5000.0, 5001.0, 0.01, 0, motype=1)
    ctx = ctypes.c_void_p()
    for other, msg in ((b, b"depth count"), (c, b"Motype 0")):
        keep = [a.cstruct(precision=FSG_FP64), other.cstruct(precision=FSG_FP64)]
        arr = (CSetup * 2)(*keep)
        assert lib.fsg_init_batch(arr, 2, ctypes.byref(ctx)) != 0
        assert msg in lib.fsg_last_error()
    assert lib.fsg_init_batch(arr, 0, ctypes.byref(ctx)) != 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libfsgpu.so")
    with pytest.raises(_lib.FsgError):
        _lib.load()
