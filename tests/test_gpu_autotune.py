"""The autotuner of the FP32 opacity launch geometry (the far-wing kernel's LDS ring depth):
it records its choice per workload signature, a cache file forces a choice, every choice
gives the oracle's opacity, and FP32 results do not depend on it -- two fresh processes
give LineOpac equal bit for bit whatever they time.  Each case runs in a child process so
that the process-wide choice table starts empty."""
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]

CHILD = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[1] + "/tests")
from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32
from fastspec_amd.synth import FastSpecGPU
from oracle_bridge import Oracle
atm = W.grey_atmosphere(5777.0, 4.44, ntau=40)
setup = W.build_setup(atm, 5000.0, 5030.0, 0.01, 0, nmu=3)
lines = W.atomic_lines(3000, 5000.0, 5030.0, 77)
with FastSpecGPU(setup, precision=FSG_FP32, batched=1) as g:
    r = g.LineopCalc(lines)
    lop = g.get_lineop()
if len(sys.argv) > 2:
    np.save(sys.argv[2], lop)
ref = Oracle(setup)
ref.lineop(lines)
lo = ref.get_lineop()
err = float(np.max(np.abs(lop - lo)) / np.max(np.abs(lo)))
st = r["stats"]
print(json.dumps({"core_tile": st["core_tile"], "far_ring": st["far_ring"], "err": err}))
"""


def _run(tmp_path, *extra, **env):
    e = dict(os.environ)
    e.update(env)
    p = subprocess.run([sys.executable, "-c", CHILD, str(ROOT), *extra], capture_output=True, text=True, env=e,
                       timeout=300, cwd=tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    return json.loads(p.stdout.strip().splitlines()[-1])


def test_autotune_records_and_forces(tmp_path):
    cache = tmp_path / "tune.txt"
    a = _run(tmp_path, FSG_AUTOTUNE_CACHE=str(cache))
    assert a["core_tile"] == 64 and a["far_ring"] in (2, 4)
    assert a["err"] <= 1e-5
    rows = cache.read_text().split()
    assert len(rows) == 3 and int(rows[1]) * 64 == a["core_tile"] and int(rows[2]) == a["far_ring"]
    # the other ring depth, forced through the cache file: the same opacity
    cache.write_text(f"{rows[0]} {rows[1]} {6 - int(rows[2])}\n")
    b = _run(tmp_path, FSG_AUTOTUNE_CACHE=str(cache))
    assert b["core_tile"] == 64 and b["far_ring"] == 6 - int(rows[2])
    assert b["err"] <= 1e-5


def test_autotune_off_keeps_defaults(tmp_path):
    r = _run(tmp_path, FSG_AUTOTUNE="0")
    assert r["core_tile"] == 64 and r["far_ring"] == 4
    assert r["err"] <= 1e-5


def test_fp32_bits_do_not_depend_on_the_run(tmp_path):
    """Two fresh processes, each timing its own ring-depth candidates, and a third with the
    tuner off: LineOpac equal bit for bit."""
    outs = []
    for k, env in enumerate(({}, {}, {"FSG_AUTOTUNE": "0"})):
        f = tmp_path / f"lop{k}.npy"
        _run(tmp_path, str(f), **env)
        outs.append(np.load(f))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_wide_core_tile_is_opt_in(tmp_path):
    r = _run(tmp_path, FSG_CORE_TILE="128")
    assert r["core_tile"] == 128 and r["err"] <= 1e-5
