"""Benchmark of the line-opacity + formal-solution hot path on one MI355X.

Workload: the large atomic configuration -- 2,000,000 seeded lines, lambda log-uniform in
3000-10000 A, log-lambda grid at R = 10^6 (1,203,973 points), grey Eddington atmosphere
(Teff 5777 K, log g 4.44, 72 depths), 7 mu angles.  Synthetic inputs (see
fastspec_amd/workloads.py); lines and host tables are uploaded to HBM before timing.

One step = one spectrum: LineOpac reset, pass 1 (rejection + windows), width binning,
coefficient fill, tiled opacity accumulation (FP32: far-wing series on the matrix cores,
core zones with the Voigt table on the vector units, blocked Kahan sums), FP64 DELO-Bezier
for continuum and line+continuum at 7 mu, flux sums.  The FP64 plain GPU path (table Voigt,
FP64 sums, one bin) is timed on the same input and the FP32 flux is compared with it.

Further legs beside the headline line's workload (each reported in its own key): the plain
configuration (C1, 20,000 lines), the Stokes configuration (C4, 5,000 Zeeman lines), the
cool-dwarf molecular configuration (C3, 200,000,000 diatomic lines streamed from a sidecar
in blocks of 10^7), the differentiable batch (C5), LinStrCalc, ingest, convolution, the C2
synthesis from host buffers to host buffers, and the reference's CPU kernels.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "line·depth·λ evals/s + spectra/s; FP32-batched vs FP64 GPU at ≤1e-4 flux err"
PEAK_FP32_TFLOPS = 157.3     # published FP32 vector / dense matrix peak (vector: packed FP32)
PEAK_FP32_VALU_TFLOPS = 78.65  # FP32 vector without packed instructions (the kernels avoid them)
PEAK_FP64_TFLOPS = 78.6
FLOP_PER_MAG_EVAL = 24       # Zeeman component evaluation (two table lerps, strength, sums)
FLOP_PER_STOKES_POINT = 740  # Stokes DELO-Bezier per (lambda, mu, depth)
FLOP_PER_RT_LAMBDA = 27052   # scalar DELO for line and continuum, 72 depths, 7 mu
FLOP_PER_EVAL = 14          # U_op cost of the table-lerp evaluation (fixed definition)
MFMA_FLOP_PER_EVAL = 6      # far-wing series T (k0 + T (k1 + T k2)): three FMA per evaluation
TRAFFIC_MD = ROOT / "profiles" / "opacity_c2_hbm_traffic.md"


def measured_traffic():
    """HBM bytes per launch of the opacity stage from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel), or None."""
    try:
        for line in TRAFFIC_MD.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


C5_THREADS = 8  # host threads driving the differentiable batch's contexts

# Vsini, Vmacro (km/s) and the instrumental profile (R = 100,000) of the convolution leg
CONVOL = (10.0, 3.0, 1e5)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_path(setup, lines, precision, batched, steps, warmup):
    from fastspec_amd.synth import FastSpecGPU

    g = FastSpecGPU(setup, precision=precision, batched=batched)
    dl = g.upload(lines)
    st = None
    for _ in range(warmup):
        g.reset_lineop()
        st = g.lineop_dev(dl)
        g.RadTranCalc(intensities=False)
    t0 = time.perf_counter()
    kms, p1ms = [], []
    for _ in range(steps):
        g.reset_lineop()
        st = g.lineop_dev(dl)
        g.RadTranCalc(intensities=False)   # synchronises the library stream
        kms.append(st["ms_opacity"])
        p1ms.append(st["ms_pass1"])
    t1 = time.perf_counter()
    sfl, sfc = g.flux()
    # IntegConvol(2) on the same intensities: rotation, macroturbulence and the instrumental
    # profile; each call synchronises and copies SfL, SfC back, so wall-clock covers it whole
    g.IntegConvol(2, *CONVOL)
    reps = 5
    t2 = time.perf_counter()
    for _ in range(reps):
        g.IntegConvol(2, *CONVOL)
    conv_ms = (time.perf_counter() - t2) * 1e3 / reps
    dl.free()
    g.close()
    return dict(ms_per_step=(t1 - t0) * 1e3 / steps, stats=st, opacity_ms=float(np.mean(kms)),
                pass1_ms=float(np.mean(p1ms)), fnorm=sfl / sfc, conv_ms=conv_ms)


def run_linstr(setup, lines):
    """LinStrCalc for every retained line (SAVE flag 1: two DELO solves per mu at the line
    centre) as the difference of two FP32 LineopCalc calls with and without it; both
    upload the list and copy flags and windows back."""
    from fastspec_amd.abi import FSG_FP32
    from fastspec_amd.synth import FastSpecGPU

    g = FastSpecGPU(setup, precision=FSG_FP32, batched=1)
    g.LineopCalc(lines, linstr=True)
    t = {}
    for flag in (False, True, False, True):
        g.reset_lineop()
        t0 = time.perf_counter()
        r = g.LineopCalc(lines, linstr=flag)
        t.setdefault(flag, []).append(time.perf_counter() - t0)
    g.close()
    nret = int(r["stats"]["nlines_use"])
    dt = (min(t[True]) - min(t[False])) * 1e3
    return {"ms": dt, "lines": nret, "delo_solves": 2 * setup.nmu * nret,
            "lines_per_s": nret / (dt * 1e-3) if dt > 0 else None,
            "note": "difference of LineopCalc with and without LinStr (min of two calls each)"}


def run_ingest(setup, nlines: int):
    """The C2 list as a raw-column sidecar (the /Lines columns of its HDF5 form) streamed to
    HBM by fsg_lines_stream: ReadHDF's selection and conversions by 16 host threads into
    pinned chunks while the previous chunk copies.  The files sit in the page cache (just
    written), so this is the reader's own rate, not the disk's."""
    import tempfile

    from fastspec_amd import linelist as LL
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP32
    from fastspec_amd.synth import FastSpecGPU

    cols = W.atomic_columns(nlines, 3000.0, 10000.0, W.SEED_BASE + 2, log_uniform=True)
    nbytes = sum(np.asarray(v).nbytes for v in cols.values())
    with tempfile.TemporaryDirectory() as tmp:
        path = LL.write_sidecar(Path(tmp) / "c2.fsl", cols)
        g = FastSpecGPU(setup, precision=FSG_FP32, batched=1)
        w0, w1 = setup.wfirst - 5.0, setup.wfirst + float(setup.wave[-1]) + 5.0
        LL.StreamedLines(g, path, w0, w1, species=W.ATOMIC_SPECIES_TABLE, nthreads=16).free()
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            sl = LL.StreamedLines(g, path, w0, w1, species=W.ATOMIC_SPECIES_TABLE, nthreads=16)
            t.append(time.perf_counter() - t0)
            n = sl.n
            sl.free()
        g.close()
        h5 = _ingest_hdf5(g_setup=setup, cols=cols, tmp=Path(tmp), w0=w0, w1=w1)
    dt = min(t)
    return {"s": dt, "lines": n, "lines_per_s": n / dt, "file_bytes": nbytes, "gb_per_s": nbytes / dt / 1e9,
            "threads": 16, "note": "fsg_lines_stream from a sidecar in the page cache (min of 3)",
            "hdf5": h5}


PY_H5 = "/opt/conda/bin/python3.9"  # the image's Python with h5py (writes the HDF5 list only)


def _ingest_hdf5(g_setup, cols, tmp: Path, w0: float, w1: float):
    """The same list in the reference's HDF5 layout (W contiguous, the other columns in
    262,144-row chunks with shuffle + gzip 4), written by h5py in a child process, then
    streamed to HBM by fsg_lines_stream: libhdf5 reads the raw chunks, 16 host threads
    inflate and convert them."""
    import subprocess

    from fastspec_amd import linelist as LL
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP32
    from fastspec_amd.synth import FastSpecGPU

    if not Path(PY_H5).exists():
        return {"skipped": f"{PY_H5} (h5py) not present"}
    npz, h5 = tmp / "cols.npz", tmp / "c2.h5"
    np.savez(npz, **cols)
    code = ("import sys, numpy as np; sys.path.insert(0, sys.argv[1]); from fastspec_amd.linelist import write_hdf5; "
            "d = np.load(sys.argv[2]); write_hdf5(sys.argv[3], {k: d[k] for k in d.files})")
    t0 = time.perf_counter()
    r = subprocess.run([PY_H5, "-c", code, str(ROOT), str(npz), str(h5)], capture_output=True, text=True)
    if r.returncode != 0:
        return {"skipped": "h5py writer failed: " + r.stderr[-300:]}
    write_s = time.perf_counter() - t0
    g = FastSpecGPU(g_setup, precision=FSG_FP32, batched=1)
    LL.StreamedLines(g, h5, w0, w1, species=W.ATOMIC_SPECIES_TABLE, nthreads=16).free()
    t = []
    for _ in range(3):
        t0 = time.perf_counter()
        sl = LL.StreamedLines(g, h5, w0, w1, species=W.ATOMIC_SPECIES_TABLE, nthreads=16)
        t.append(time.perf_counter() - t0)
        n = sl.n
        sl.free()
    g.close()
    dt = min(t)
    return {"s": dt, "lines": n, "lines_per_s": n / dt, "file_bytes": h5.stat().st_size, "write_s": write_s,
            "threads": 16, "note": "fsg_lines_stream from the HDF5 file in the page cache (min of 3)"}


def run_c5(steps: int, warmup: int, contexts: bool = False):
    """The differentiable batch: 256 grey atmospheres with their own windows over one shared
    list of 50,000 lines (FP32 opacity, FP64 solver).  A step synthesises all 256 spectra
    (gf of the shared list set on the device, LineopCalc, RadTranCalc, flux); the adjoint
    step adds the reverse mode of L = sum a F/Fc and gathers dL/d log gf of the shared
    lines.  Default: one batch context, every stage launched once for all atmospheres over
    their concatenated grids.  contexts=True: one context per atmosphere, driven from
    C5_THREADS host threads (the earlier design, for comparison).  Contexts and uploads are
    made before timing."""
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP32
    from fastspec_amd.synth import FastSpecBatch, FastSpecGPU

    lines, models = W.config_c5()
    loggf = np.log10(lines.gf.astype(np.float64))
    rng = np.random.default_rng(11)
    weights = [rng.normal(size=setup.nwl) for setup, _ in models]
    if contexts:
        ctxs = []
        for (setup, idx), a in zip(models, weights):
            g = FastSpecGPU(setup, precision=FSG_FP32)
            ctxs.append((g, g.upload(lines.take(idx)), idx, a))

        def one(item, gf, adjoint):
            g, dl, idx, a = item
            g.set_gf(dl, gf[idx])
            g.reset_lineop()
            ev = g.lineop_dev(dl)["evals"]
            g.RadTranCalc(intensities=False)
            sfl, sfc = g.flux()
            gg = None
            if adjoint:
                g.flux_vjp(a / sfc, -a * sfl / sfc**2)
                gg, _, _ = g.lineop_vjp(dl)
            return ev, gg

        from concurrent.futures import ThreadPoolExecutor
        pool = ThreadPoolExecutor(max_workers=C5_THREADS)

        def step(adjoint):
            gf = np.power(np.float32(10.0), loggf.astype(np.float32))
            grad = np.zeros(lines.n)
            evals = 0
            res = list(pool.map(lambda it: one(it, gf, adjoint), ctxs))
            for (ev, gg), (_, _, idx, _) in zip(res, ctxs):
                evals += ev
                if adjoint:
                    grad[idx] += gg
            return evals, grad
    else:
        b = FastSpecBatch([setup for setup, _ in models], precision=FSG_FP32)
        shared = b.upload(lines)
        bl = b.batch_lines(shared, [idx for _, idx in models])
        # the weights over the concatenated grid (zero at the padding points between the
        # atmospheres): the loss gradient in three array operations instead of 256 x 5
        aw = b.join(weights)
        live = aw != 0.0

        def step(adjoint):
            gf = np.power(np.float32(10.0), loggf.astype(np.float32))
            b.set_gf(bl, gf)
            b.reset_lineop()
            evals = b.lineop_dev(bl)["evals"]
            b.RadTranCalc(intensities=False)
            sfl, sfc = b.flux(joined=True)
            grad = None
            if adjoint:
                q = aw / np.where(live, sfc, 1.0)  # dL/dSfL = a / Fc
                b.flux_vjp(q, -q * sfl / np.where(live, sfc, 1.0))  # dL/dSfC = -a F / Fc^2
                gg, _, _ = b.lineop_vjp(bl)
                grad = bl.to_shared(gg)
            return evals, grad

    out = {}
    for name, adj in (("forward", False), ("forward_adjoint", True)):
        for _ in range(warmup):
            step(adj)
        t0 = time.perf_counter()
        for _ in range(steps):
            evals, grad = step(adj)
        dt = (time.perf_counter() - t0) / steps
        out[name] = {"s_per_batch": dt, "spectra_per_s": len(models) / dt, "evals_per_s": evals / dt}
        if adj:
            out[name]["grad_loggf_nonzero"] = int(np.count_nonzero(grad))
    if contexts:
        pool.shutdown()
        for g, dl, _, _ in ctxs:
            dl.free()
            g.close()
        how = f"one context per atmosphere, driven from {C5_THREADS} host threads"
    else:
        bl.free()
        shared.free()
        b.close()
        how = "one batch context: every stage launched once for all atmospheres over their concatenated grids"
    out["workload"] = (f"{len(models)} grey atmospheres (Teff 4500-6500 K, log g 3.5-4.8), windows 5-50 A "
                       f"in 5000-5500 A at 0.01 A ({sum(s.nwl for s, _ in models)} points), shared list "
                       f"of {lines.n} lines; FP32 opacity, FP64 solver; adjoint of L = sum a F/Fc; {how}")
    out["evals_per_batch"] = evals
    return out


def run_host_to_host(setup, lines, steps: int):
    """The C2 synthesis timed as the hot path's seams see it, from lines and ContOpac in host
    memory to the spectra in host memory: each step uploads ContOpac (fsg_set_atmosphere,
    693 MB, transposed on the device), then fsg_lineop uploads the list and returns flags and
    windows, fsg_radtran and fsg_flux copy SpL, SpC, SfL, SfC back.  The context (grid,
    tables) is made beforehand; its set-up time is reported beside it."""
    from fastspec_amd.abi import FSG_FP32
    from fastspec_amd.synth import FastSpecGPU

    t0 = time.perf_counter()
    g = FastSpecGPU(setup, precision=FSG_FP32, batched=1)
    init_s = time.perf_counter() - t0
    g.set_atmosphere(contop=setup.contop)
    g.LineopCalc(lines)
    g.RadTranCalc()
    g.flux()
    t, tc = [], []
    for _ in range(steps):
        g.reset_lineop()
        t0 = time.perf_counter()
        g.set_atmosphere(contop=setup.contop)
        t1 = time.perf_counter()
        g.LineopCalc(lines)
        g.RadTranCalc()
        g.flux()
        t.append(time.perf_counter() - t0)
        tc.append(t1 - t0)
    g.close()
    ms = float(np.median(t)) * 1e3
    return {"ms_per_spectrum": ms, "spectra_per_s": 1e3 / ms, "init_ms": init_s * 1e3,
            "contop_upload_ms": float(np.median(tc)) * 1e3,
            "note": "ContOpac upload (fsg_set_atmosphere) + fsg_lineop from host lines + fsg_radtran with "
                    "SpL/SpC to host + fsg_flux to host (median of the steps); init_ms: fsg_init"}


def _device_steps(g, dl, steps, warmup, nlim=None):
    for _ in range(warmup):
        g.reset_lineop()
        g.lineop_dev(dl, nlim)
        g.RadTranCalc(intensities=False)
    t0 = time.perf_counter()
    st = None
    for _ in range(steps):
        g.reset_lineop()
        st = g.lineop_dev(dl, nlim)
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
    dt = (time.perf_counter() - t0) / steps
    return dt, st, sfl / sfc


def run_c1(steps: int, warmup: int):
    """The plain configuration in full (20,000 lines, 5000-5100 A, 0.01 A, 72 depths, 7 mu):
    FP32 batched and FP64 plain, lines resident in HBM."""
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP32, FSG_FP64
    from fastspec_amd.synth import FastSpecGPU

    setup, lines = W.config_c1()
    out = {"workload": f"C1 plain: {lines.n} lines 5000-5100 A, {setup.nwl} points, 72 depths, 7 mu"}
    fn = {}
    for name, prec, batched in (("fp32_batched", FSG_FP32, 1), ("fp64_plain", FSG_FP64, 0)):
        g = FastSpecGPU(setup, precision=prec, batched=batched)
        dl = g.upload(lines)
        dt, st, fn[name] = _device_steps(g, dl, steps, warmup)
        dl.free()
        g.close()
        out[name] = {"ms_per_spectrum": dt * 1e3, "spectra_per_s": 1.0 / dt, "evals_per_s": st["evals"] / dt,
                     "opacity_ms": st["ms_opacity"], "pass1_ms": st["ms_pass1"]}
    out["evals_per_spectrum"] = st["evals"]
    out["max_flux_err_fp32_vs_fp64"] = float(np.max(np.abs(fn["fp32_batched"] - fn["fp64_plain"])))
    return out


def run_c4(steps: int, warmup: int):
    """The Stokes configuration in full (5,000 Zeeman lines, 6000-6100 A, 0.005 A, B = 2 kG at
    45/30 deg): MagLineopCalc (six Faraday-Voigt profile sums per Zeeman component) and the
    Stokes DELO-Bezier solution, FP32 profiles and FP64 profiles; U_mag = (component,
    depth, lambda) evaluations inside the windows."""
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP32, FSG_FP64
    from fastspec_amd.synth import FastSpecGPU

    setup, lines, bfield, bframe = W.config_c4()
    ncomp = np.diff(lines.pi_off) + np.diff(lines.si_off)
    out = {"workload": f"C4 Stokes: {lines.n} Zeeman lines 6000-6100 A, {setup.nwl} points, 72 depths, "
                       f"7 mu, B = 2 kG at 45/30 deg"}
    for name, prec in (("fp32_profiles", FSG_FP32), ("fp64_profiles", FSG_FP64)):
        g = FastSpecGPU(setup, precision=prec)
        r = g.MagLineopCalc(lines, bfield, bframe)
        keep = (r["flags"] & 1) == 0
        w = (r["iright"] - r["ileft"] + 1).astype(np.int64)
        umag = int(np.sum(w[keep] * ncomp[keep])) * setup.ntau
        dl = g.upload(lines)
        for _ in range(warmup):
            g.reset_lineop()
            g.maglineop_dev(dl, bfield, bframe)
            g.MagRadTranCalc(bfield, bframe, intensities=False)
        t_op, t_rt, t_all, st = [], [], [], None
        for _ in range(steps):
            g.reset_lineop()
            t0 = time.perf_counter()
            st = g.maglineop_dev(dl, bfield, bframe)
            t1 = time.perf_counter()
            g.MagRadTranCalc(bfield, bframe, intensities=False)
            g.flux()
            t2 = time.perf_counter()
            t_op.append(st["ms_opacity"])
            t_rt.append((t2 - t1) * 1e3)
            t_all.append((t2 - t0) * 1e3)
        dl.free()
        g.close()
        ms_op, ms_rt, ms = float(np.median(t_op)), float(np.median(t_rt)), float(np.median(t_all))
        stokes_flop = FLOP_PER_STOKES_POINT * setup.nwl * setup.nmu * setup.ntau
        out[name] = {"ms_per_spectrum": ms, "spectra_per_s": 1e3 / ms, "u_mag": umag,
                     "u_mag_per_s": umag / (ms_op * 1e-3), "opacity_ms": ms_op,
                     "opacity_frac_of_fp32_peak": umag * FLOP_PER_MAG_EVAL / (ms_op * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                     "line_depth_lambda_evals": int(st["evals"]),
                     "stokes_rt_ms": ms_rt,
                     "stokes_rt_frac_of_fp64_peak": stokes_flop / (ms_rt * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                     "note": "stokes_rt_ms: fsg_magradtran + fsg_flux wall time (Stokes DELO, "
                             f"{FLOP_PER_STOKES_POINT} flop per (lambda, mu, depth))"}
    return out


def run_c3(steps: int, fp64: bool):
    """The cool-dwarf molecular configuration: 200,000,000 seeded diatomic (FeH) lines
    log-uniform in 4000-25000 A on an R = 3x10^5 log grid (549,775 points), Teff 3200 K, log g
    5.0, 72 depths, 7 mu.  The list is written as a sidecar of the HDF5 columns (W, E, LGGF,
    GRAD; 4 GB) and streamed to HBM by fsg_lines_stream (ReadHDF's selection and conversions
    on 16 host threads into pinned chunks); the synthesis then accumulates it in blocks of
    NlinesLim = 10^7 lines."""
    import shutil
    import tempfile

    from fastspec_amd import linelist as LL
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP32, FSG_FP64
    from fastspec_amd.synth import FastSpecGPU

    need = 4.2e9
    tmp_root = None
    for cand in (tempfile.gettempdir(), "/dev/shm", str(ROOT / "gpurun_out")):
        try:
            if shutil.disk_usage(cand).free > 2 * need:
                tmp_root = cand
                break
        except OSError:
            pass
    if tmp_root is None:
        return {"skipped": "no file system with 8.4 GB free for the 4 GB sidecar"}
    setup, spc = W.config_c3_setup()
    out = {"workload": f"C3 molecular: {W.C3_NLINES} FeH lines 4000-25000 A, R=3e5 log grid ({setup.nwl} "
                       f"points), 72 depths, 7 mu, grey Teff 3200 logg 5.0, NlinesLim {W.C3_NLINES_LIM}"}
    with tempfile.TemporaryDirectory(dir=tmp_root) as tmp:
        t0 = time.perf_counter()
        nbytes = W.write_diatomic_sidecar(Path(tmp) / "feh")
        out["sidecar"] = {"bytes": nbytes, "write_s": time.perf_counter() - t0}
        log(f"C3 sidecar written: {nbytes / 1e9:.2f} GB in {out['sidecar']['write_s']:.1f} s")
        for name, prec, batched in (("fp32_batched", FSG_FP32, 1), ("fp64_plain", FSG_FP64, 0)):
            if name == "fp64_plain" and not fp64:
                continue
            g = FastSpecGPU(setup, precision=prec, batched=batched)
            t0 = time.perf_counter()
            sl = LL.StreamedLines(g, Path(tmp) / "feh", *W.C3_RANGE, mol_slot=spc, nthreads=16)
            ingest = time.perf_counter() - t0
            nst = steps if name == "fp32_batched" else 1
            dt, st, fn = _device_steps(g, sl, nst, 1 if name == "fp32_batched" else 0, W.C3_NLINES_LIM)
            sl.free()
            g.close()
            out[name] = {"ms_per_spectrum": dt * 1e3, "spectra_per_s": 1.0 / dt, "evals_per_s": st["evals"] / dt,
                         "evals": st["evals"], "lines": sl.n, "retained": st["nlines_use"],
                         "blocks": st["nblocks"], "pass1_ms": st["ms_pass1"], "opacity_ms": st["ms_opacity"],
                         "ingest_s": ingest, "ingest_lines_per_s": sl.n / ingest,
                         "roofline_frac_14flop": st["evals"] * FLOP_PER_EVAL / (st["ms_opacity"] * 1e-3) / 1e12 /
                         PEAK_FP32_TFLOPS}
            out.setdefault("_fn", {})[name] = fn
            log(f"C3 {name}: {json.dumps(out[name])}")
    fns = out.pop("_fn", {})
    if len(fns) == 2:
        out["max_flux_err_fp32_vs_fp64"] = float(np.max(np.abs(fns["fp32_batched"] - fns["fp64_plain"])))
    return out


def run_contop(setup):
    """fsg_contop over the C2 grid: all thirteen continuum sources of Xcont (H I, H2+, H-,
    He I, He II, He-, COOLOP's metals and molecules, LUKEOP's and HOTOP's ions, Rayleigh by
    H I, He I and H2, Thomson), one lane per wavelength, 72 depths, FP64, ContOpac left on
    the device.  The reference's Xcont (oracle/_ref, -O3, one thread: its harness loops
    serially; ContopCalc is an OpenMP loop over wavelengths) is timed on 20,000 of the grid's
    wavelengths for the same sources."""
    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP64
    from fastspec_amd.synth import FastSpecGPU

    pops = W.continuum_populations(setup)
    g = FastSpecGPU(setup, precision=FSG_FP64)
    g.ContopCalc(**pops, copy_out=False)
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        g.ContopCalc(**pops, copy_out=False)
    ms = (time.perf_counter() - t0) * 1e3 / reps
    g.close()
    n, nt = setup.nwl, setup.ntau
    bytes_alg = n * (8 * nt + 8)  # ContOpac written, the grid read
    res = {"ms": ms, "nwl": n, "ntau": nt, "sources": list(FastSpecGPU.CONTOP_DEVICE_SOURCES),
           "points_per_s": n * nt / (ms * 1e-3),
           "hbm_gbs_algorithmic": bytes_alg / (ms * 1e-3) / 1e9,
           "note": "wall clock of fsg_contop (host set-up of the depth state and tables, one launch, "
                   "synchronised); ContOpac stays in HBM"}
    try:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_bridge
        o3 = ROOT / "oracle" / "_ref" / "o3" / "libfsref.so"
        if o3.exists():
            oracle_bridge.REF_SO = o3
            oracle_bridge._REF = None
            wl = oracle_bridge.ref_air2vac_array(setup.wavelengths[::max(1, n // 20000)][:20000])
            cpu = {}
            for name, srcs in (("all_13_sources", FastSpecGPU.CONTOP_DEVICE_SOURCES),):
                t1 = time.perf_counter()
                oracle_bridge.ref_xcont(setup.temp, setup.xne, setup.rho, pops, srcs, wl)
                us = (time.perf_counter() - t1) * 1e6 / len(wl)
                cpu[name] = {"us_per_wavelength_1core": us, "full_grid_s_1core": us * n * 1e-6}
            res["cpu_reference_xcont"] = cpu
            res["speedup_vs_1core"] = cpu["all_13_sources"]["full_grid_s_1core"] / (ms * 1e-3)
    except Exception as e:  # the reference build is optional on the GPU host
        res["cpu_reference_xcont"] = f"unavailable: {e}"
    return res


def host_cpu():
    """Core count and model of the host CPU."""
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "model": model}


def cpu_baseline():
    """The reference's own OpenMP kernels (oracle/_ref, built from the reference sources)
    on a bounded sample of the same workload: the C2 grid and line density restricted to
    a wavelength range around 5000 A.  Times LineOpacity1 over the sample grid and the
    RadTranCalc loop (LocPrfPP); counts evaluations from the windows."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_bridge
    from oracle_bridge import Oracle, Ref

    from fastspec_amd import workloads as W

    o3 = ROOT / "oracle" / "_ref" / "o3" / "libfsref.so"
    if not o3.exists():
        return None
    oracle_bridge.REF_SO = o3  # the -O3 build (the reference Makefile's level) for timing
    width = 500.0  # about 20 s of the reference kernels on 16 host cores (1000 A took 47 s)
    setup, lines = W.config_c2(wl_lo=5000.0, wl_hi=5000.0 + width)
    orc = Oracle(setup)
    ro = orc.pass1(lines)
    ref = Ref(setup)
    ref.set_lines(lines)
    nthreads = int(ref.lib.ref_nthreads())
    evals = int(ro["stats"]["evals"])
    t0 = time.perf_counter()
    cop, vv, av, _ = ref.xline()  # XlineSetup + XlineDepth of every line (pass 1's coefficients)
    t1 = time.perf_counter()
    ref.set_windows(ro["flags"] & ~8, ro["ileft"], ro["iright"])
    ref.set_block(cop, vv, av)
    lop = ref.lineop1(1, setup.nwl, 1, lines.n)
    t2 = time.perf_counter()
    ref.radtran(1, setup.nwl, lop)
    t3 = time.perf_counter()
    dt = t3 - t0
    cpu = host_cpu()
    return {"value": evals / dt, "unit": "line·depth·λ evals/s", "cores": nthreads,
            "kind": "reference",
            "sample": f"C2 restricted to {5000:.0f}-{5000 + width:.0f} A ({setup.nwl} grid points, "
                      f"{lines.n} lines, {evals:.3e} evals); reference XlineSetup/XlineDepth + LineOpacity1 + "
                      f"LocPrfPP (amdflang -O3 -fopenmp, {nthreads} threads; windows from the oracle's pass 1) "
                      f"in {dt:.2f} s",
            "seconds": dt, "stages_s": {"xline": t1 - t0, "lineopacity1": t2 - t1, "locprfpp": t3 - t2},
            "host": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--nlines", type=int, default=2_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-linstr", action="store_true")
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-contop", action="store_true")
    ap.add_argument("--c3-fp64", action="store_true", help="also time the FP64 plain path on C3")
    ap.add_argument("--no-host", action="store_true", help="skip the host-to-host C2 timing")
    ap.add_argument("--only-c5", action="store_true", help="run only the differentiable batch (profiling)")
    ap.add_argument("--only", choices=["c1", "c3", "c4", "c5", "contop"],
                    help="run one leg alone and print its result (profiling; not the benchmark line)")
    ap.add_argument("--c5-contexts", action="store_true",
                    help="differentiable batch with one context per atmosphere (the earlier design)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X "
              "(multi-GPU lambda sharding is described in DESIGN.md, not implemented).", file=sys.stderr)
        sys.exit(2)

    from fastspec_amd import workloads as W
    from fastspec_amd.abi import FSG_FP32, FSG_FP64

    if args.only_c5:  # the differentiable batch alone (for profiling); not the benchmark line
        args.only = "c5"
    if args.only:  # one leg alone (for profiling); not the benchmark line
        legs = {"c1": lambda: run_c1(args.steps, args.warmup), "c4": lambda: run_c4(args.steps, args.warmup),
                "c3": lambda: run_c3(max(1, min(args.steps, 2)), args.c3_fp64),
                "c5": lambda: run_c5(args.steps, args.warmup, args.c5_contexts),
                "contop": lambda: run_contop(W.config_c2(nlines=10)[0])}
        print(json.dumps({args.only: legs[args.only]()}), flush=True)
        return

    log("building the C2 workload")
    setup, lines = W.config_c2(nlines=args.nlines)
    log(f"nwl={setup.nwl} ntau={setup.ntau} nmu={setup.nmu} nlines={lines.n}")
    r32 = run_path(setup, lines, FSG_FP32, 1, args.steps, args.warmup)
    st = r32["stats"]
    log(f"fp32 batched: {r32['ms_per_step']:.2f} ms/step, opacity kernel {r32['opacity_ms']:.2f} ms, "
        f"evals {st['evals']:.3e}, retained {st['nlines_use']}")
    evals = float(st["evals"])
    value = evals / (r32["ms_per_step"] * 1e-3)
    achieved = evals * FLOP_PER_EVAL / (r32["opacity_ms"] * 1e-3) / 1e12
    # the two kernels of the stage: far wings on the matrix cores (depths padded to blocks
    # of 16), core zones on the vector units
    far = evals - float(st["evals_core"])
    pad = 16 * ((setup.ntau + 15) // 16) / setup.ntau
    far_tf = far * MFMA_FLOP_PER_EVAL / (st["ms_far"] * 1e-3) / 1e12 if st["ms_far"] > 0 else None
    kernels = {
        "k_opacity32_mfma": {"ms": st["ms_far"], "evals": far, "mfma_flop_per_eval": MFMA_FLOP_PER_EVAL,
                             "lds_ring_slots": st["far_ring"],
                             "depth_padding": pad, "tflops": far_tf,
                             "frac_of_fp32_mfma_peak": far_tf / PEAK_FP32_TFLOPS if far_tf else None,
                             "issued_frac": far_tf * pad / PEAK_FP32_TFLOPS if far_tf else None},
        "k_opacity32_core": {"ms": st["ms_core"], "evals": float(st["evals_core"]), "tile_points": st["core_tile"],
                             "evals_per_s": st["evals_core"] / (st["ms_core"] * 1e-3) if st["ms_core"] > 0 else None},
    }
    core = kernels["k_opacity32_core"]
    if core["evals_per_s"]:
        # the fixed 14 flop per evaluation against the published FP32 peak and against the
        # vector units' rate without packed FP32 instructions (which the kernels do not use)
        core["frac_of_fp32_peak_14flop"] = core["evals_per_s"] * FLOP_PER_EVAL / 1e12 / PEAK_FP32_TFLOPS
        core["frac_of_fp32_valu_unpacked_14flop"] = core["evals_per_s"] * FLOP_PER_EVAL / 1e12 / PEAK_FP32_VALU_TFLOPS
    out = {
        "metric": METRIC, "value": value, "unit": "line·depth·λ evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 opacity (Kahan), f64 formal solution", "data": "synthetic",
        "config": {"workload": "C2 large atomic: 2,000,000 lines log-uniform 3000-10000 A, R=1e6 "
                               "log grid (1,203,973 points), 72 depths, 7 mu, grey Teff 5777 logg 4.44",
                   "nlines": lines.n, "nwl": setup.nwl, "ntau": setup.ntau, "nmu": setup.nmu,
                   "parallelism": "single GPU"},
        "spectra_per_s": 1e3 / r32["ms_per_step"],
        "spectra_per_s_device_resident": 1e3 / r32["ms_per_step"],
        "integ_convol": {"ms": r32["conv_ms"], "cimode": 2, "vsini": CONVOL[0], "vmacro": CONVOL[1],
                         "resolution": CONVOL[2], "note": "fsg_integ_convol incl. host copy of SfL, SfC"},
        "evals_per_step": evals,
        "retained_lines": st["nlines_use"],
        # compute-bound: the far-wing series is a dense contraction over lines (FP32 MFMA,
        # 157.3 TFLOP/s dense); achieved counts the fixed 14 flop per evaluation over the
        # whole opacity stage (both kernels and the final add), kernels[] the per-kernel view
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic(),
                     "compute": "fp32 MFMA (far wings) + fp32 VALU (core zones)",
                     "kernel": "opacity stage: k_opacity32_mfma + k_opacity32_core (concurrent streams)",
                     "kernel_ms": r32["opacity_ms"], "flop_per_eval": FLOP_PER_EVAL, "kernels": kernels},
    }
    if not args.no_fp64:
        r64 = run_path(setup, lines, FSG_FP64, 0, args.fp64_steps, 0)
        err = float(np.max(np.abs(r32["fnorm"] - r64["fnorm"])))
        log(f"fp64 plain: {r64['ms_per_step']:.2f} ms/step; max |dFn| fp32 vs fp64 = {err:.3e}")
        out["fp64_plain"] = {"ms_per_step": r64["ms_per_step"], "spectra_per_s": 1e3 / r64["ms_per_step"],
                             "evals_per_s": evals / (r64["ms_per_step"] * 1e-3),
                             "opacity_kernel_ms": r64["opacity_ms"]}
        out["speedup_vs_fp64_plain"] = r64["ms_per_step"] / r32["ms_per_step"]
        out["max_flux_err_vs_fp64"] = err
    if not args.no_linstr:
        log("LinStrCalc for the retained C2 lines")
        out["linstr"] = run_linstr(setup, lines)
        log(json.dumps(out["linstr"]))
    if not args.no_ingest:
        log("line-list ingest from a sidecar")
        out["ingest"] = run_ingest(setup, args.nlines)
        log(json.dumps(out["ingest"]))
    if not args.no_host:
        log("C2 from host buffers to host buffers")
        out["c2_host_to_host"] = run_host_to_host(setup, lines, args.steps)
        log(json.dumps(out["c2_host_to_host"]))
        # spectra/s as the hot path's time: inputs in host memory to outputs in host memory
        out["spectra_per_s"] = out["c2_host_to_host"]["spectra_per_s"]
        out["spectra_per_s_note"] = ("host memory to host memory (ContOpac and lines uploaded, spectra "
                                     "copied back each step); spectra_per_s_device_resident: inputs in HBM")
    if not args.no_contop:
        log("continuum opacity on the C2 grid")
        out["contop"] = run_contop(setup)
        log(json.dumps(out["contop"]))
    if not args.no_c1:
        log("C1 plain configuration")
        out["c1_plain"] = run_c1(args.steps, args.warmup)
        log(json.dumps(out["c1_plain"]))
    if not args.no_c4:
        log("C4 Stokes configuration")
        out["c4_stokes"] = run_c4(args.steps, args.warmup)
        log(json.dumps(out["c4_stokes"]))
    if not args.no_c3:
        log("C3 molecular configuration (200M lines)")
        del lines
        out["c3_molecular"] = run_c3(max(1, min(args.steps, 2)), args.c3_fp64)
    if not args.no_c5:
        log("differentiable batch (256 atmospheres)")
        out["c5_differentiable_batch"] = run_c5(3, 1, args.c5_contexts)
        log(json.dumps(out["c5_differentiable_batch"]))
    if not args.no_cpu_baseline:
        log("timing the reference CPU kernels on a bounded sample")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
