"""A Zeeman list streamed from a sidecar (fsg_lines_stream: host conversion, ZeemanPattern in
the library) gives the MagLineOpac of the same lines uploaded with the patterns of the
Python restatement (fastspec_amd.workloads.zeeman_pattern, pinned bitwise against the
reference's compiled ZeemanPattern in tests/test_oracle_magnetic.py)."""
import sys
from pathlib import Path

import numpy as np
import pytest

from fastspec_amd import linelist as LL
from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_FP32, FSG_FP64, FSG_LIST_MIXED, LineList

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import readhdf as R  # noqa: E402

pytestmark = pytest.mark.gpu
SPECIES = {(e, i): (s, int(m)) for e, i, s, m in W.ATOMIC_SPECIES_TABLE}


def _columns(n, seed):
    cols = W.atomic_columns(n, 6000.0, 6020.0, seed)
    rng = np.random.default_rng(seed)
    jlo = np.empty(n, np.float32)
    jup = np.empty(n, np.float32)
    for i in range(n):
        while True:
            a = 0.5 * rng.integers(0, 13)
            b = a + rng.integers(-1, 2)
            if 0 <= b <= 6 and not (a == 0 and b == 0):
                break
        jlo[i], jup[i] = a, b
    cols.update(JLO=jlo, JUP=jup, GLO=rng.uniform(0.5, 2.5, n).astype(np.float32),
                GUP=rng.uniform(0.5, 2.5, n).astype(np.float32))
    cols["GLO"][::41] = 99.0  # no Lande factor: a line without splitting
    cols["ION"] = cols["ION"].copy()
    cols["ION"][::67] = 6  # dropped
    return cols


def _reference_list(cols, first, count):
    c = R.convert_atomic(cols, first, count, SPECIES)
    keep = np.asarray(cols["ION"][first:first + count]) <= R.MAXION
    pi_off, si_off, sp, rp, ss, rs = [0], [0], [], [], [], []
    for jl, ju, gl, gu in zip(*(np.asarray(cols[k][first:first + count])[keep] for k in ("JLO", "JUP", "GLO", "GUP"))):
        if gl > 90.0 or gu > 90.0:
            sp += [0.0]
            rp += [1.0]
            ss += [0.0, 0.0]
            rs += [1.0, 1.0]
        else:
            P, S = W.zeeman_pattern(float(jl), float(ju), float(gl), float(gu))
            sp += [np.float32(g) for g, _ in P]
            rp += [np.float32(r) for _, r in P]
            ss += [-np.float32(g) for g, _ in S] + [np.float32(g) for g, _ in S]
            rs += [np.float32(r) for _, r in S] * 2
        pi_off.append(len(sp))
        si_off.append(len(ss))
    return LineList(wl=c["wl"], elo=c["elo"], gf=c["gf"], grad=c["grad"], gqst=c["gqst"], gvdw=c["gvdw"],
                    spc=c["spc"], is_mol=c["is_mol"], list_type=FSG_LIST_MIXED,
                    pi_off=np.array(pi_off, np.int32), spl_pi=np.array(sp, np.float32),
                    str_pi=np.array(rp, np.float32), si_off=np.array(si_off, np.int32),
                    spl_si=np.array(ss, np.float32), str_si=np.array(rs, np.float32))


@pytest.mark.parametrize("prec", [FSG_FP64, FSG_FP32])
def test_zeeman_stream_matches_upload(tmp_path, prec):
    from fastspec_amd.synth import FastSpecGPU
    atm = W.grey_atmosphere(5777.0, 4.44, ntau=36)
    setup = W.build_setup(atm, 6002.0, 6018.0, 0.005, 0, nmu=3)
    cols = _columns(1500, 31)
    path = LL.write_sidecar(tmp_path / "z.fsl", cols, name="atom Zeeman")
    win = (6001.0, 6019.0)
    lt, first, count = LL.scan(path, *win)
    ref = _reference_list(cols, first, count)
    bfield = np.array([2000.0, 45.0, 30.0])
    with FastSpecGPU(setup, precision=prec) as g:
        sl = LL.StreamedLines(g, path, *win, species=W.ATOMIC_SPECIES_TABLE)
        assert sl.n == ref.n
        g.reset_lineop()
        g.maglineop_dev(sl, bfield, 1)
        a = g.get_maglineop()
        up = g.upload(ref)
        g.reset_lineop()
        g.maglineop_dev(up, bfield, 1)
        b = g.get_maglineop()
        sl.free()
        up.free()
    assert np.array_equal(a, b)
    assert np.abs(a).max() > 0
