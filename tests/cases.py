"""Shared synthetic cases for the oracle and GPU parity tests."""
import dataclasses

import numpy as np

from fastspec_amd import workloads as W
from fastspec_amd.abi import FSG_LIST_MIXED, FSG_LIST_MOLECULAR, LineList


def with_he_lines(setup, lines, nhe_lines=12, seed=5):
    """Adds a He I species, a small tabulated Stark table and He I lines that match it."""
    rng = np.random.default_rng(seed)
    nt = setup.ntau
    lo, hi = setup.wfirst, setup.wfirst + setup.wave[-1]
    he_wl = np.sort(rng.uniform(lo + 1, hi - 1, 4))
    he_elo = np.array([19.81961, 20.61577, 20.96409, 21.21802])
    ne = setup.xne
    he_gam = np.outer(np.array([2.0, 1.0, 3.0, 1.5]), 1e-6 * ne)
    he_shf = np.outer(np.array([0.05, -0.03, 0.08, 0.02]), (ne / ne.max()) ** 0.5)
    spc_ion = np.concatenate([setup.spc_ion, [1]])
    spc_slot = np.concatenate([setup.spc_slot, [setup.he1_slot]])
    s2 = dataclasses.replace(setup, spc_ion=spc_ion, spc_slot=spc_slot, he_wl=he_wl, he_elo=he_elo,
                             he_gam=he_gam, he_shf=he_shf)
    he_spc = len(spc_ion)
    k = rng.integers(0, 4, nhe_lines)
    wl = he_wl[k] + rng.uniform(-0.5, 0.5, nhe_lines)
    elo = he_elo[k] + rng.uniform(-0.005, 0.005, nhe_lines)
    n = lines.n
    allwl = np.concatenate([lines.wl, wl])
    order = np.argsort(allwl, kind="stable")
    cat = lambda a, b: np.concatenate([a, b])[order]
    he = LineList(
        wl=allwl[order], elo=cat(lines.elo, elo.astype(np.float32)),
        gf=cat(lines.gf, np.full(nhe_lines, 0.5, np.float32)),
        grad=cat(lines.grad, np.full(nhe_lines, 1e8, np.float32)),
        gqst=cat(lines.gqst, np.zeros(nhe_lines, np.float32)),
        gvdw=cat(lines.gvdw, np.zeros(nhe_lines, np.float32)),
        spc=cat(lines.spc, np.full(nhe_lines, he_spc, np.int32)), list_type=lines.list_type)
    return s2, he


def mixed_list(lines, seed=3):
    """The same lines as a mixed (.z-type) list with molecules, ABO and default damping."""
    n = lines.n
    rng = np.random.default_rng(seed)
    gvdw = lines.gvdw.copy()
    abo = rng.random(n) < 0.2
    gvdw[abo] = (rng.integers(200, 400, abo.sum()) + rng.uniform(0.2, 0.3, abo.sum())).astype(np.float32)
    gvdw[(rng.random(n) < 0.2) & ~abo] = 0.0
    gqst = lines.gqst.copy()
    gqst[rng.random(n) < 0.3] = 0.0
    is_mol = (rng.random(n) < 0.15).astype(np.uint8)
    return LineList(wl=lines.wl, elo=lines.elo, gf=lines.gf, grad=lines.grad, spc=lines.spc, gqst=gqst,
                    gvdw=gvdw, is_mol=is_mol, list_type=FSG_LIST_MIXED)


def molecular_list(lines):
    return LineList(wl=lines.wl, elo=lines.elo, gf=lines.gf, grad=lines.grad, spc=lines.spc,
                    list_type=FSG_LIST_MOLECULAR)


def with_zeeman(lines, seed=11):
    """The same lines as a Zeeman (.z-type) list with seeded patterns."""
    z = W.zeeman_patterns(lines.n, seed)
    is_mol = lines.is_mol if lines.is_mol is not None else np.zeros(lines.n, np.uint8)
    return LineList(wl=lines.wl, elo=lines.elo, gf=lines.gf, grad=lines.grad, spc=lines.spc,
                    gqst=lines.gqst, gvdw=lines.gvdw, is_mol=is_mol, list_type=FSG_LIST_MIXED, **z)
