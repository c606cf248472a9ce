"""The PyTorch side on the device: the library shares torch's HIP runtime (one runtime
mapped, torch keeps its GPU), and the autograd Functions keep CUDA parameters on the device
-- gf in, F/Fc and dL/d log gf out through the device-pointer entry points -- with the
results of the host-tensor path."""
import numpy as np
import pytest

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_one_hip_runtime_shared_with_torch():
    from fastspec_amd import _lib
    _lib.load()
    maps = open("/proc/self/maps").read().splitlines()
    hip = {l.split()[-1] for l in maps if "libamdhip64" in l}
    assert len(hip) == 1, hip
    assert torch.cuda.is_available()
    x = torch.ones(1000, device="cuda", dtype=torch.float64)
    assert float((2 * x).sum()) == 2000.0


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_normalised_flux_on_device_equals_host(prec):
    from fastspec_amd.autograd import normalised_flux
    from fastspec_amd.synth import FastSpecGPU
    atm = W.grey_atmosphere()
    setup = W.build_setup(atm, 5000.0, 5010.0, 0.01, 0)
    lines = W.atomic_lines(400, 5000.0, 5010.0, 808)
    a = np.random.default_rng(1).normal(size=setup.nwl)
    x0 = np.log10(lines.gf.astype(np.float64))
    res = {}
    for dev in ("cpu", "cuda"):
        loggf = torch.tensor(x0, dtype=torch.float64, device=dev, requires_grad=True)
        temp = torch.tensor(setup.temp, dtype=torch.float64, device=dev, requires_grad=True)
        with FastSpecGPU(setup, precision=prec) as g:
            dl = g.upload(lines)
            f = normalised_flux(g, dl, loggf, temp=temp)
            assert f.device.type == dev
            loss = (torch.as_tensor(a, device=dev) * f).sum()
            loss.backward()
            dl.free()
        assert loggf.grad.device.type == dev and temp.grad.device.type == dev
        res[dev] = (float(loss.detach()), loggf.grad.cpu().numpy(), temp.grad.cpu().numpy(), f.detach().cpu().numpy())
    c, d = res["cpu"], res["cuda"]
    # gf = 10**log gf rounded to REAL(4) on either side may differ by one unit in the last place
    assert np.max(np.abs(c[3] - d[3])) < 1e-6
    assert abs(c[0] - d[0]) < 1e-6 * np.sum(np.abs(a))
    assert np.max(np.abs(c[1] - d[1])) <= 1e-5 * np.max(np.abs(c[1]))
    assert np.max(np.abs(c[2] - d[2])) <= 1e-5 * np.max(np.abs(c[2]))


def test_normalised_flux_batch_on_device_equals_host():
    from fastspec_amd.autograd import normalised_flux_batch
    from fastspec_amd.synth import FastSpecBatch
    lines, models = W.config_c5(nmodels=4, nlines=3000)
    x0 = np.log10(lines.gf.astype(np.float64))
    res = {}
    for dev in ("cpu", "cuda"):
        with FastSpecBatch([s for s, _ in models], precision=FSG_FP32) as b:
            shared = b.upload(lines)
            bl = b.batch_lines(shared, [idx for _, idx in models])
            loggf = torch.tensor(x0, dtype=torch.float64, device=dev, requires_grad=True)
            f = normalised_flux_batch(b, bl, loggf)
            w = torch.linspace(-1.0, 1.0, f.numel(), dtype=torch.float64, device=dev)
            (w * f).sum().backward()
            bl.free()
            shared.free()
        res[dev] = (f.detach().cpu().numpy(), loggf.grad.cpu().numpy())
    assert res["cpu"][0].shape == res["cuda"][0].shape
    assert np.max(np.abs(res["cpu"][0] - res["cuda"][0])) < 1e-6
    g = res["cpu"][1]
    assert np.count_nonzero(g) > 100
    assert np.max(np.abs(g - res["cuda"][1])) <= 1e-5 * np.max(np.abs(g))
