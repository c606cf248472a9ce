"""Batch contexts (fsg_init_batch, fsg_lines_batch): many atmospheres synthesised in one
launch sequence over their concatenated grids give, per atmosphere, the results of an
ordinary context on that atmosphere alone -- opacity, flux, and the reverse mode (dL/dT,
dL/d ln FractionPF, dL/d log gf) -- and the FP64 batch flux agrees with the FP64 oracle."""
import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64
from fastspec_amd.synth import FastSpecBatch, FastSpecGPU
from fastspec_amd._lib import FsgError
from oracle_bridge import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c5_small():
    # differentiable-batch workload, scaled down: 6 atmospheres, windows of 5-50 A
    lines, models = W.config_c5(nmodels=6, nlines=4000)
    rng = np.random.default_rng(3)
    weights = [rng.normal(size=s.nwl) for s, _ in models]
    return lines, models, weights


def _single(setup, lines, prec, a=None):
    with FastSpecGPU(setup, precision=prec) as g:
        dl = g.upload(lines)
        st = g.lineop_dev(dl)
        lop = g.get_lineop()
        g.RadTranCalc(intensities=False)
        sfl, sfc = g.flux()
        out = dict(sfl=sfl, sfc=sfc, lop=lop, evals=st["evals"])
        if a is not None:
            _, _, gt = g.flux_vjp(a / sfc, -a * sfl / sfc**2)
            ggf, gfpf, gtl = g.lineop_vjp(dl)
            out.update(gt=gt + gtl, ggf=ggf, gfpf=gfpf)
        dl.free()
    return out


def _batch(lines, models, prec, weights=None):
    setups = [s for s, _ in models]
    with FastSpecBatch(setups, precision=prec) as b:
        shared = b.upload(lines)
        bl = b.batch_lines(shared, [idx for _, idx in models])
        st = b.lineop_dev(bl)
        lop = b.get_lineop()
        b.RadTranCalc(intensities=False)
        sfl, sfc = b.flux()
        out = dict(sfl=sfl, sfc=sfc, lop=lop, evals=st["evals"], off=b.off.copy())
        if weights is not None:
            gsfl = [a / c for a, c in zip(weights, sfc)]
            gsfc = [-a * l / c**2 for a, l, c in zip(weights, sfl, sfc)]
            _, _, gt = b.flux_vjp(gsfl, gsfc)
            ggf, gfpf, gtl = b.lineop_vjp(bl)
            out.update(gt=gt + gtl, ggf=bl.per_atmosphere(ggf), gfpf=gfpf, ggf_shared=bl.to_shared(ggf))
        bl.free()
        shared.free()
    return out


def test_joined_arrays_match_lists(c5_small):
    """flux(joined=True) is flux()'s lists in the concatenated layout, and flux_vjp takes the
    gradients in either form (the atomic sums of dL/dT may add in another order)."""
    lines, models, weights = c5_small
    setups = [s for s, _ in models[:3]]
    with FastSpecBatch(setups, precision=FSG_FP32) as b:
        shared = b.upload(lines)
        bl = b.batch_lines(shared, [idx for _, idx in models[:3]])
        b.lineop_dev(bl)
        b.RadTranCalc(intensities=False)
        sfl, sfc = b.flux()
        jl, jc = b.flux(joined=True)
        assert jl.shape == (b.nwl_total,) and jc.shape == (b.nwl_total,)
        for k, su in enumerate(setups):
            assert np.array_equal(jl[b.off[k]:b.off[k] + su.nwl], sfl[k])
            assert np.array_equal(jc[b.off[k]:b.off[k] + su.nwl], sfc[k])
        gl = [a / c for a, c in zip(weights[:3], sfc)]
        gc = [-a * l / c**2 for a, l, c in zip(weights[:3], sfl, sfc)]
        _, _, gt1 = b.flux_vjp(gl, gc)
        _, _, gt2 = b.flux_vjp(b.join(gl), b.join(gc))
        assert np.count_nonzero(gt1) > 0
        np.testing.assert_allclose(gt2, gt1, rtol=1e-12, atol=0.0)
        bl.free()
        shared.free()


def test_layout(c5_small):
    lines, models, _ = c5_small
    out = _batch(lines, models[:3], FSG_FP64)
    off = out["off"]
    assert off[0] == 0 and np.all(off % 256 == 0)
    for b, (s, _) in enumerate(models[:3]):
        assert off[b + 1] - off[b] >= s.nwl


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_batch_matches_single_contexts(c5_small, prec):
    lines, models, _ = c5_small
    bo = _batch(lines, models, prec)
    evals = 0
    for b, (setup, idx) in enumerate(models):
        so = _single(setup, lines.take(idx), prec)
        evals += so["evals"]
        if prec == FSG_FP64:
            # the same sums in the same order: identical
            assert np.array_equal(bo["lop"][b], so["lop"])
            assert np.array_equal(bo["sfl"][b], so["sfl"])
            assert np.array_equal(bo["sfc"][b], so["sfc"])
        else:
            # the core-zone tile width (autotuned per grid) may group the compensated sums
            # differently
            scale = so["lop"].max() + 1e-300
            assert np.max(np.abs(bo["lop"][b] - so["lop"])) <= 1e-6 * scale
            assert np.max(np.abs(bo["sfl"][b] / bo["sfc"][b] - so["sfl"] / so["sfc"])) <= 1e-7
            assert np.array_equal(bo["sfc"][b], so["sfc"])
    assert bo["evals"] == evals


def test_batch_flux_matches_oracle(c5_small):
    lines, models, _ = c5_small
    bo = _batch(lines, models[:2], FSG_FP64)
    for b, (setup, idx) in enumerate(models[:2]):
        o = Oracle(setup)
        o.lineop(lines.take(idx))
        o.radtran()
        sfl, sfc = o.flux()
        assert np.max(np.abs(bo["sfl"][b] / bo["sfc"][b] - sfl / sfc)) <= 1e-9


def test_batch_reverse_mode_matches_single_contexts(c5_small):
    lines, models, weights = c5_small
    bo = _batch(lines, models, FSG_FP64, weights)
    shared = np.zeros(lines.n)
    for b, (setup, idx) in enumerate(models):
        so = _single(setup, lines.take(idx), FSG_FP64, weights[b])
        # dL/dT sums use atomic adds (order free): equal to rounding
        assert np.allclose(bo["gt"][b], so["gt"], rtol=1e-10, atol=1e-12 * np.abs(so["gt"]).max())
        assert np.allclose(bo["gfpf"][b], so["gfpf"], rtol=1e-10, atol=1e-12 * np.abs(so["gfpf"]).max())
        assert np.allclose(bo["ggf"][b], so["ggf"], rtol=1e-10, atol=1e-12 * np.abs(so["ggf"]).max())
        np.add.at(shared, idx, so["ggf"])
    assert np.allclose(bo["ggf_shared"], shared, rtol=1e-10, atol=1e-12 * np.abs(shared).max())


def test_batch_refusals(c5_small):
    lines, models, _ = c5_small
    setups = [s for s, _ in models[:2]]
    with FastSpecBatch(setups, precision=FSG_FP64) as b:
        shared = b.upload(lines)
        with pytest.raises(FsgError, match="batch lists"):
            b.lineop_dev(shared)  # a plain list on a batch context
        with pytest.raises(FsgError, match="not sorted"):
            b.batch_lines(shared, [models[0][1][::-1], models[1][1]])
        shared.free()
    with FastSpecGPU(setups[0], precision=FSG_FP64) as g, FastSpecBatch(setups, precision=FSG_FP64) as b:
        shared = b.upload(lines)
        bl = b.batch_lines(shared, [models[0][1], models[1][1]])
        with pytest.raises(FsgError, match="batch lists"):
            g.lineop_dev(bl)  # a batch list on an ordinary context
        bl.free()
        shared.free()


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_batch_in_chunks(c5_small, prec):
    """NlinesLim below the batch list's size: chunks of coefficient records mix the
    atmospheres' lines; the opacity is the same sum in another order."""
    lines, models, _ = c5_small
    setups = [s for s, _ in models]
    with FastSpecBatch(setups, precision=prec) as b:
        shared = b.upload(lines)
        bl = b.batch_lines(shared, [idx for _, idx in models])
        st = b.lineop_dev(bl, NlinesLim=max(bl.n // 5, 1))
        lop = b.get_lineop()
        bl.free()
        shared.free()
    assert st["nblocks"] >= 2
    tol = 1e-12 if prec == FSG_FP64 else 1e-6
    for k, (setup, idx) in enumerate(models):
        so = _single(setup, lines.take(idx), prec)
        assert np.max(np.abs(lop[k] - so["lop"])) <= tol * so["lop"].max()


def test_open_objects_released_at_exit(tmp_path):
    """Contexts and lists left open are released by the wrappers at interpreter exit (device
    lists first) and the process ends cleanly."""
    import subprocess
    import sys
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    child = (
        "import sys; sys.path.insert(0, sys.argv[1])\n"
        "from fastspec_amd import workloads as W\n"
        "from fastspec_amd.abi import FSG_FP32\n"
        "from fastspec_amd.synth import FastSpecBatch, FastSpecGPU\n"
        "lines, models = W.config_c5(nmodels=3, nlines=2000)\n"
        "b = FastSpecBatch([s for s, _ in models], precision=FSG_FP32)\n"
        "shared = b.upload(lines)\n"
        "bl = b.batch_lines(shared, [i for _, i in models])\n"
        "b.lineop_dev(bl)\n"
        "g = FastSpecGPU(models[0][0], precision=FSG_FP32)\n"
        "dl = g.upload(lines.take(models[0][1]))\n"
        "g.lineop_dev(dl)\n"
        "print('done', flush=True)\n"
    )
    r = subprocess.run([sys.executable, "-c", child, str(root)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "done" in r.stdout


def test_torch_autograd_batch(c5_small):
    """The torch.autograd.Function over a batch context: loss and gradients (shared log gf,
    temperature per atmosphere) equal the sums of the per-atmosphere wrapper's."""
    torch = pytest.importorskip("torch")
    from fastspec_amd.autograd import normalised_flux, normalised_flux_batch
    lines, models, weights = c5_small
    models, weights = models[:3], weights[:3]
    loggf0 = np.log10(lines.gf.astype(np.float64))
    with FastSpecBatch([s for s, _ in models], precision=FSG_FP64) as b:
        shared = b.upload(lines)
        bl = b.batch_lines(shared, [idx for _, idx in models])
        loggf = torch.tensor(loggf0, dtype=torch.float64, requires_grad=True)
        temp = torch.tensor(np.stack([s.temp for s, _ in models]), dtype=torch.float64, requires_grad=True)
        f = normalised_flux_batch(b, bl, loggf, temp=temp)
        a = torch.as_tensor(np.concatenate(weights))
        loss = (a * f).sum()
        loss.backward()
        bl.free()
        shared.free()
    total, g_shared = 0.0, np.zeros(lines.n)
    for k, (setup, idx) in enumerate(models):
        lg = torch.tensor(loggf0[idx], dtype=torch.float64, requires_grad=True)
        tk = torch.tensor(setup.temp, dtype=torch.float64, requires_grad=True)
        with FastSpecGPU(setup, precision=FSG_FP64) as g:
            dl = g.upload(lines.take(idx))
            fk = normalised_flux(g, dl, lg, temp=tk)
            lk = (torch.as_tensor(weights[k]) * fk).sum()
            lk.backward()
            dl.free()
        total += lk.item()
        np.add.at(g_shared, idx, lg.grad.numpy())
        assert np.allclose(temp.grad[k].numpy(), tk.grad.numpy(), rtol=1e-9, atol=1e-12 * np.abs(tk.grad.numpy()).max())
    assert abs(loss.item() - total) <= 1e-12 * sum(np.abs(w).sum() for w in weights)
    assert np.allclose(loggf.grad.numpy(), g_shared, rtol=1e-9, atol=1e-12 * np.abs(g_shared).max())
