"""Builds libfsgpu.so in-tree with hipcc for gfx950 (no JIT, no torch extension).

    python -m fastspec_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libfsgpu.so"
CALIB = PKG / "hbm_copy_cal"
SOURCES = ["fsg_kernels.hip", "fsg_magnetic.hip", "fsg_adjoint.hip", "fsg_convol.hip", "fsg_contop.hip",
           "fsg_ingest.hip", "fsg_hdf5.hip", "fsg_api.hip"]
ARCH = os.environ.get("FSG_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.  MFMA accumulators are kept in
# ordinary VGPRs (no AGPR copies around the compensated sums of the FP32 opacity kernel).
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-mllvm", "-amdgpu-mfma-vgpr-form",
            "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libfsgpu.so needs the ROCm toolchain")


def build(verbose: bool = False) -> Path:
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs, procs = [], []
    # the HBM counter calibration program (a 1 GiB copy) builds beside the library
    cal = [hipcc(), *HIPFLAGS, str(CSRC / "hbm_copy_cal.hip"), "-o", str(CALIB)]
    procs.append(("hbm_copy_cal.hip", subprocess.Popen(cal, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                                       text=True)))
    for src in SOURCES:  # the translation units compile concurrently
        obj = objdir / (Path(src).stem + ".o")
        cmd = [hipcc(), *HIPFLAGS, "-I", str(CSRC), "-c", str(CSRC / src), "-o", str(obj)]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(str(obj))
    failed = []
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(src)
        elif verbose:
            sys.stderr.write(err)
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = LIB.with_suffix(".so.tmp")
    # zlib inflates the HDF5 lists' chunks; libhdf5 itself is loaded at run time (dlopen)
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lz", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
