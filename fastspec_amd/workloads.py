"""Seeded synthetic inputs for the hot path.

The reference obtains the atmosphere's thermodynamic state from its EOS and the continuum
opacity from its ATLAS12-style routines; both stay on the host and are outside this
library.  For measurements and parity tests we therefore build those host-side inputs
synthetically, with the same shapes, units and orders of magnitude:

* a grey Eddington atmosphere, T^4 = 3/4 Teff^4 (tau + 2/3), on 72 depths with
  log tau_5000 uniform in [-5, 2];
* hydrostatic equilibrium dP/dtau = g / kappa_std with a smooth H^- like opacity law;
* Saha ionisation with constant partition functions for H, He and the line species,
  giving N/U per cm^3 (FractionPF), electron and perturber densities;
* a continuum opacity kappa_c(lambda, depth) = kappa_std(depth) * phi(lambda), phi smooth;
* line lists drawn from the distributions of the benchmark configurations, already in the
  form the reference's HDF5 reader produces (gf = 10**log gf in REAL(4), damping
  constants as linear REAL(4) values).

Everything is deterministic for a given seed (numpy PCG64).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .abi import FSG_LIST_ATOMIC, FSG_LIST_MIXED, FSG_LIST_MOLECULAR, LineList, Setup

KB = 1.380649e-16
MH = 1.673532499e-24
EV = 1.602176634e-12
VC = 299792.458
SEED_BASE = 20261200

# species of the atomic configurations: (name, Z, ion, ionisation potential [eV] of this
# stage, atomic mass, log10 N/N_H, U(this stage), U(next stage))
ATOMIC_SPECIES = [
    ("Fe1", 26, 1, 7.902, 55.845, -4.50, 25.0, 40.0),
    ("Fe2", 26, 2, 16.199, 55.845, -4.50, 40.0, 10.0),
    ("Ti1", 22, 1, 6.828, 47.867, -7.05, 30.0, 55.0),
    ("Ca1", 20, 1, 6.113, 40.078, -5.66, 1.2, 2.2),
    ("Cr1", 24, 1, 6.767, 51.996, -6.36, 10.0, 7.0),
]


def eddington_temperature(tau: np.ndarray, teff: float) -> np.ndarray:
    return (0.75 * teff**4 * (tau + 2.0 / 3.0)) ** 0.25


def _kappa_std(p: float, t: float) -> float:
    """Smooth H^- like Rosseland-ish opacity law [cm^2/g]."""
    return 0.35 * (p / 1e5) ** 0.55 * (t / 5800.0) ** 7.0 + 2e-4


def saha_ratio(t: np.ndarray, ne: np.ndarray, chi_ev: float, u0: float, u1: float) -> np.ndarray:
    """N(i+1)/N(i) from the Saha equation."""
    return 2.0 * u1 / u0 * 2.4147e15 * t**1.5 * np.exp(-chi_ev * EV / (KB * t)) / ne


@dataclass
class Atmosphere:
    tau: np.ndarray
    temp: np.ndarray
    pgas: np.ndarray
    rhox: np.ndarray
    rho: np.ndarray
    xne: np.ndarray
    n_h: np.ndarray
    kappa_std: np.ndarray
    teff: float
    logg: float


def grey_atmosphere(teff: float = 5777.0, logg: float = 4.44, ntau: int = 72,
                    logtau=(-5.0, 2.0)) -> Atmosphere:
    """Grey Eddington model in hydrostatic equilibrium on a column-mass scale."""
    g = 10.0**logg
    tau = np.logspace(logtau[0], logtau[1], ntau)
    temp = eddington_temperature(tau, teff)
    # integrate dP/dln tau = tau g / kappa(P, T) with RK4 in ln tau
    lnt = np.log(tau)
    p = np.empty(ntau)
    p0 = tau[0] * g / _kappa_std(1.0, temp[0])
    for _ in range(20):  # P = tau g / kappa(P) at the top layer
        p0 = tau[0] * g / _kappa_std(p0, temp[0])
    p[0] = p0

    def rhs(lt, pp):
        tt = float(np.interp(lt, lnt, temp))
        return math.exp(lt) * g / _kappa_std(pp, tt)

    for i in range(ntau - 1):
        h = lnt[i + 1] - lnt[i]
        k1 = rhs(lnt[i], p[i])
        k2 = rhs(lnt[i] + 0.5 * h, p[i] + 0.5 * h * k1)
        k3 = rhs(lnt[i] + 0.5 * h, p[i] + 0.5 * h * k2)
        k4 = rhs(lnt[i] + h, p[i] + h * k3)
        p[i + 1] = p[i] + h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    kstd = np.array([_kappa_std(pp, tt) for pp, tt in zip(p, temp)])
    rhox = p / g
    # electron density: H plus a lumped low-potential metal donor, fixed-point iteration
    ntot = p / (KB * temp)
    n_h = ntot / 1.1
    ne = np.maximum(1e-4 * n_h, 1.0)
    for _ in range(60):
        fh = saha_ratio(temp, ne, 13.598, 2.0, 1.0)
        fm = saha_ratio(temp, ne, 7.5, 20.0, 30.0)
        ne_new = n_h * (fh / (1.0 + fh)) + 1e-4 * n_h * (fm / (1.0 + fm))
        ne = np.sqrt(ne * np.maximum(ne_new, 1.0))
    rho = (ntot - ne) * 1.3 * MH
    return Atmosphere(tau=tau, temp=temp, pgas=p, rhox=rhox, rho=rho, xne=ne, n_h=n_h,
                      kappa_std=kstd, teff=teff, logg=logg)


def continuum_shape(lam: np.ndarray) -> np.ndarray:
    """phi(lambda): smooth H^- bound-free + free-free like wavelength dependence."""
    x = lam / 5000.0
    bf = np.where(lam < 16400.0, x**1.3 * (1.0 - 0.6 * (lam / 16400.0) ** 2), 0.0)
    ff = 0.15 * x**2
    scat = 0.05 * x**-4
    return 0.55 + bf + ff + scat


def gauss_mu(n: int):
    """Gauss-Legendre limb angles on [0,1] (descending) and flux weights 2 w mu."""
    z, w = np.polynomial.legendre.leggauss(n)
    z = z[::-1].copy()
    w = w[::-1].copy()
    x = 0.5 * (1.0 + z)
    wg = 0.5 * w
    return x, 2.0 * wg * x


def wave_grid(wstart: float, wend: float, step: float, wmode: int):
    """The reference's grid (src/synth.f90:56-74): offsets from Wfirst."""
    eps = 1e-12
    if wmode == 0:
        nwl = int(math.floor((wend - wstart) / step + eps)) + 1
        wave = np.arange(nwl, dtype=np.float64) * step
        return wave, wstart, step, 0.0
    rstep = step / VC
    nwl = int(math.floor(math.log(wend / wstart) / rstep + eps)) + 1
    wave = wstart * np.exp(np.arange(nwl, dtype=np.float64) * rstep) - wstart
    return wave, wstart, step, rstep


# diatomic species of the molecular configuration: (name, mass [amu], log10 N/N_H at and
# below the reference temperature t0, dissociation energy [eV], partition function, t0 [K]).
# The number density falls with temperature as exp(D0/kT - D0/kT0) above t0.
MOLECULAR_SPECIES = [("FeH", 56.853, -8.0, 1.63, 3000.0, 3200.0)]


def build_setup(atm: Atmosphere, wstart: float, wend: float, step: float, wmode: int,
                species=ATOMIC_SPECIES, nmu: int = 7, vmicro: float = 1.0,
                motype: int = 0, eps_opac: float = 1e-4, extra_slots: int = 4,
                radius: float = -1.0, molecules=()) -> Setup:
    """Host-side inputs (EOS tables, continuum) on the given grid.

    EOS slots: 1..len(species) are the line species, then H I, He I, H2, H II, He II, then the
    molecules (MOLECULAR_SPECIES entries), species slots after the atomic ones.
    radius > 0 makes a spherical model: layer heights from dh = -dm / rho, zero at
    tau_5000 = 1."""
    ntau = atm.temp.shape[0]
    wave, wfirst, dstep, rstep = wave_grid(wstart, wend, step, wmode)
    lam = wave + wfirst
    t, ne, nh = atm.temp, atm.xne, atm.n_h
    nsp = len(species)
    nmol = len(molecules)
    nslot = nsp + 5 + nmol
    pf_ld = nslot + extra_slots
    fraction_pf = np.zeros((ntau, pf_ld), np.float32)
    pot_ion = np.zeros(nslot, np.float32)
    xamass = np.zeros(nslot, np.float32)
    for j, (_, _z, ion, chi, mass, logab, u0, u1) in enumerate(species):
        if ion == 1:
            r = saha_ratio(t, ne, chi, u0, u1)
            n_stage = 10.0**logab * nh / (1.0 + r)
        else:  # first ion: use the neutral stage's potential of the same element
            chi0 = [s for s in species if s[1] == _z and s[2] == 1][0][3]
            r = saha_ratio(t, ne, chi0, 25.0, u0)
            n_stage = 10.0**logab * nh * r / (1.0 + r)
        fraction_pf[:, j] = (n_stage / u0).astype(np.float32)
        pot_ion[j] = chi
        xamass[j] = mass
    fh = saha_ratio(t, ne, 13.598, 2.0, 1.0)
    n_h1 = nh / (1.0 + fh)
    n_hp = nh - n_h1
    n_he = 0.085 * nh
    fhe = saha_ratio(t, ne, 24.587, 1.0, 2.0)
    n_he1 = n_he / (1.0 + fhe)
    n_he2 = n_he - n_he1
    n_h2 = 1e-6 * nh * np.clip(4000.0 / t, 0.0, 2.0) ** 8
    for j, (n_s, chi, mass, u) in enumerate([(n_h1, 13.598, 1.008, 2.0), (n_he1, 24.587, 4.0026, 1.0),
                                             (n_h2, 15.43, 2.016, 1.0), (n_hp, 0.0, 1.008, 1.0),
                                             (n_he2, 54.418, 4.0026, 2.0)]):
        fraction_pf[:, nsp + j] = (n_s / u).astype(np.float32)
        pot_ion[nsp + j] = chi
        xamass[nsp + j] = mass
    for j, (_, mass, logab, d0, u, t0) in enumerate(molecules):
        frac = 10.0**logab * np.minimum(1.0, np.exp(d0 * EV / KB * (1.0 / t - 1.0 / t0)))
        fraction_pf[:, nsp + 5 + j] = (frac * nh / u).astype(np.float32)
        pot_ion[nsp + 5 + j] = d0
        xamass[nsp + 5 + j] = mass
    contop = (atm.kappa_std[None, :] * continuum_shape(lam)[:, None]).astype(np.float64)
    contop_std = None
    rhox = atm.rhox
    if motype == 1:
        contop_std = atm.kappa_std * continuum_shape(np.array([5000.0]))[0]
        contop = contop / contop_std[None, :]
        rhox = atm.tau.copy()
    xmu, wmu = gauss_mu(nmu)
    height = None
    if radius > 0:
        m, r = atm.rhox, atm.rho
        h = np.zeros(ntau)
        for k in range(1, ntau):
            h[k] = h[k - 1] - (m[k] - m[k - 1]) / (0.5 * (r[k] + r[k - 1]))
        height = h - np.interp(0.0, np.log10(atm.tau), h)
    return Setup(
        wave=wave, wmode=wmode, wfirst=wfirst, dstep=dstep, rstep=rstep, wstart=wstart,
        rhox=rhox, temp=t, xne=ne, rho=atm.rho, contop=contop, contop_std=contop_std,
        fraction_pf=fraction_pf, pot_ion=pot_ion, xamass=xamass,
        frac_h1=n_h1.astype(np.float32), frac_he1=n_he1.astype(np.float32),
        frac_hmol=n_h2.astype(np.float32),
        spc_ion=np.array([s[2] for s in species] + [1] * nmol, np.int32),
        spc_slot=np.concatenate([np.arange(1, nsp + 1), nsp + 5 + np.arange(1, nmol + 1)]).astype(np.int32),
        he1_slot=nsp + 2, xmu=xmu, wmu=wmu, motype=motype, eps_opac=eps_opac, vmicro=vmicro,
        radius=radius, height=height,
        meta=dict(atmosphere=atm, species=[s[0] for s in species] + [m[0] for m in molecules], n_atomic=nsp),
    )


# (name, ionisation energy [eV], log abundance, U0, U1) of the atoms COOLOP reads
COOLOP_ATOMS = [("c1", 11.260, -3.57, 9.3, 5.9), ("mg1", 7.646, -4.40, 1.0, 2.0), ("al1", 5.986, -5.55, 5.9, 1.0),
                ("si1", 8.152, -4.49, 9.4, 5.7), ("fe1", 7.902, -4.50, 25.0, 40.0)]
# (name, log abundance at full association, dissociation energy [eV], T0 [K], U) of its
# molecules; NH's row is a Fraction (U = 1)
COOLOP_MOLECULES = [("ch", -7.5, 3.47, 3000.0, 200.0), ("oh", -7.0, 4.39, 3000.0, 150.0),
                    ("nh", -8.0, 3.47, 3000.0, 1.0)]


# (name, ionisation energies of stages I-V [eV], log abundance) of HOTOP's elements in its
# order (C, N, O, Ne, Mg, Si, S, Fe), and Ca for LUKEOP's Ca II
HOT_ELEMENTS = [("c", (11.260, 24.383, 47.888, 64.494, 392.09), -3.57),
                ("n", (14.534, 29.601, 47.449, 77.473, 97.890), -4.17),
                ("o", (13.618, 35.121, 54.936, 77.414, 113.90), -3.31),
                ("ne", (21.565, 40.963, 63.45, 97.12, 126.21), -4.07),
                ("mg", (7.646, 15.035, 80.144, 109.27, 141.27), -4.40),
                ("si", (8.152, 16.346, 33.493, 45.142, 166.77), -4.49),
                ("s", (10.360, 23.338, 34.79, 47.222, 72.594), -4.88),
                ("fe", (7.902, 16.199, 30.652, 54.91, 75.0), -4.50),
                ("ca", (6.113, 11.872, 50.913, 67.27, 84.5), -5.66)]
HOT_PF = np.array([6.0, 5.0, 4.0, 3.0, 2.0, 1.0])


def continuum_populations(setup: Setup) -> dict:
    """The EOS rows the device's continuum sources read, for a set-up made by build_setup:
    FractionPF of H I, H II, He I and He II (their slots), Fraction of H2 and He II (twice
    its FractionPF, partition function 2), Fraction of H- from the Saha equation against H I
    (binding 0.754 eV, statistical weights 1 and 2), FractionPF of He III from the Saha
    equation against He II (54.418 eV, partition functions 2 and 1), COOLOP's rows
    (COOLOP_ATOMS, COOLOP_MOLECULES, Fraction of H I and He I), and LUKEOP's and HOTOP's
    (HOT_ELEMENTS' ionisation stages)."""
    nsp = int(setup.meta["n_atomic"])  # H I, He I, H2, H II, He II follow the atomic slots
    ev_k = EV / KB
    t = setup.temp
    n_h1 = setup.frac_h1.astype(np.float64)
    c_saha = 4.1412e-16  # (h^2 / (2 pi m_e k))^(3/2) [cm^3 K^(3/2)]
    n_hmin = n_h1 * setup.xne * c_saha * t**-1.5 / 4.0 * np.exp(0.754 * ev_k / t)
    fpf_he2 = setup.fraction_pf[:, nsp + 4].copy()
    n_he2 = 2.0 * fpf_he2.astype(np.float64)
    n_he3 = n_he2 * saha_ratio(t, setup.xne, 54.418, 2.0, 1.0)
    pops = dict(fpf_h1=setup.fraction_pf[:, nsp].copy(), fpf_h2=setup.fraction_pf[:, nsp + 3].copy(),
                frac_hmin=n_hmin.astype(np.float32), fpf_he1=setup.fraction_pf[:, nsp + 1].copy(),
                frac_hmol=setup.frac_hmol.copy(), frac_he2=n_he2.astype(np.float32), fpf_he2=fpf_he2,
                fpf_he3=n_he3.astype(np.float32))
    # COOLOP's rows: neutral C, Mg, Al, Si and Fe from the Saha equation (solar abundances,
    # constant partition functions U0, U1), CH, OH and NH from the dissociation balance of the
    # molecular slots, and the number densities of H I and He I
    n_h = n_h1 + setup.fraction_pf[:, nsp + 3].astype(np.float64)
    for name, chi, logab, u0, u1 in COOLOP_ATOMS:
        n0 = 10.0**logab * n_h / (1.0 + saha_ratio(t, setup.xne, chi, u0, u1))
        pops["fpf_" + name] = (n0 / u0).astype(np.float32)
    for name, logab, d0, t0, u in COOLOP_MOLECULES:
        n_m = 10.0**logab * n_h * np.minimum(1.0, np.exp(d0 * ev_k * (1.0 / t - 1.0 / t0)))
        pops[("frac_" if name == "nh" else "fpf_") + name] = (n_m / u).astype(np.float32)
    pops["frac_h1"] = setup.frac_h1.astype(np.float32)
    pops["frac_he1"] = setup.frac_he1.astype(np.float32)
    # LUKEOP's and HOTOP's rows: stages I-VI of each element from a Saha chain with constant
    # partition functions (HOT_PF)
    stages = {name: ion_stages(t, setup.xne, n_h, chis, logab) for name, chis, logab in HOT_ELEMENTS}
    hot = np.stack([stages[name] for name, _, _ in HOT_ELEMENTS if name != "ca"], axis=1)  # [ntau][8][6]
    pops["hot_frac"] = hot.astype(np.float32)
    nst = (4, 5, 6, 6)  # C I-IV, N I-V, O I-VI, Ne I-VI
    pops["hot_fpf"] = np.concatenate([hot[:, k, :m] / HOT_PF[:m] for k, m in enumerate(nst)], axis=1).astype(np.float32)
    for key, (name, st) in dict(fpf_c2=("c", 1), fpf_mg2=("mg", 1), fpf_si2=("si", 1), fpf_n1=("n", 0),
                                fpf_o1=("o", 0), fpf_ca2=("ca", 1)).items():
        pops[key] = (stages[name][:, st] / HOT_PF[st]).astype(np.float32)
    return pops


def ion_stages(t, ne, n_h, chis, logab):
    """Number densities of stages I-VI [ntau][6] of an element of abundance logab from the
    Saha equation with the ionisation energies chis (eV) and partition functions HOT_PF."""
    n = np.zeros((t.shape[0], 6))
    n[:, 0] = 1.0
    for i, chi in enumerate(chis[:5]):
        n[:, i + 1] = n[:, i] * saha_ratio(t, ne, chi, HOT_PF[i], HOT_PF[i + 1])
    n /= n.sum(axis=1, keepdims=True)
    return n * (10.0**logab * n_h)[:, None]


def _f32pow10(x):
    return np.power(np.float32(10.0), np.asarray(x, np.float32)).astype(np.float32)


def atomic_lines(n: int, wl_lo: float, wl_hi: float, seed: int, log_uniform: bool = False,
                 nspecies: int = len(ATOMIC_SPECIES)) -> LineList:
    """Seeded atomic list with the benchmark distributions (sorted by wavelength)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if log_uniform:
        wl = np.exp(rng.uniform(math.log(wl_lo), math.log(wl_hi), n))
    else:
        wl = rng.uniform(wl_lo, wl_hi, n)
    wl.sort()
    lggf = rng.uniform(-5.0, 0.0, n)
    elo = rng.uniform(0.0, 6.0, n)
    spc = rng.integers(1, nspecies + 1, n)
    grad = rng.normal(8.0, 0.3, n)
    gqst = rng.normal(-6.0, 0.5, n)
    gvdw = rng.normal(-7.5, 0.3, n)
    return LineList(
        wl=wl, elo=elo.astype(np.float32), gf=_f32pow10(lggf), grad=_f32pow10(grad),
        gqst=_f32pow10(gqst), gvdw=_f32pow10(gvdw), spc=spc.astype(np.int32),
        list_type=FSG_LIST_ATOMIC,
    )


# (symbol, ion, species slot, molecule) of ATOMIC_SPECIES, for sidecar line lists
ATOMIC_SPECIES_TABLE = [(name[:-1], ion, k + 1, False) for k, (name, _, ion, *_r) in enumerate(ATOMIC_SPECIES)]


def atomic_columns(n: int, wl_lo: float, wl_hi: float, seed: int, log_uniform: bool = False,
                   nspecies: int = len(ATOMIC_SPECIES)) -> dict:
    """The /Lines columns (EL, ION, W, E, LGGF, GRAD, GQST, GVDW; log10 values as stored in an
    HDF5 list) of atomic_lines(n, wl_lo, wl_hi, seed, ...): the same draws, before the
    reader's conversions."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if log_uniform:
        wl = np.exp(rng.uniform(math.log(wl_lo), math.log(wl_hi), n))
    else:
        wl = rng.uniform(wl_lo, wl_hi, n)
    wl.sort()
    lggf = rng.uniform(-5.0, 0.0, n)
    elo = rng.uniform(0.0, 6.0, n)
    spc = rng.integers(1, nspecies + 1, n)
    grad = rng.normal(8.0, 0.3, n)
    gqst = rng.normal(-6.0, 0.5, n)
    gvdw = rng.normal(-7.5, 0.3, n)
    el = np.array([ATOMIC_SPECIES_TABLE[k - 1][0].encode() for k in range(1, nspecies + 1)], "S3")[spc - 1]
    ion = np.array([ATOMIC_SPECIES_TABLE[k - 1][1] for k in range(1, nspecies + 1)], np.int16)[spc - 1]
    return {"W": wl, "EL": el, "ION": ion, "E": elo.astype(np.float32), "LGGF": lggf.astype(np.float32),
            "GRAD": grad.astype(np.float32), "GQST": gqst.astype(np.float32), "GVDW": gvdw.astype(np.float32)}


def molecular_lines(n: int, wl_lo: float, wl_hi: float, seed: int, spc: int = 1) -> LineList:
    """Seeded single-species diatomic list: log gf ~ N(-4, 1.5), E_low ~ Exp(1 eV)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    wl = np.sort(np.exp(rng.uniform(math.log(wl_lo), math.log(wl_hi), n)))
    lggf = rng.normal(-4.0, 1.5, n)
    elo = rng.exponential(1.0, n)
    grad = rng.normal(8.0, 0.3, n)
    return LineList(wl=wl, elo=elo.astype(np.float32), gf=_f32pow10(lggf), grad=_f32pow10(grad),
                    spc=np.full(n, spc, np.int32), list_type=FSG_LIST_MOLECULAR)


def _diatomic_plan(n: int, wl_lo: float, wl_hi: float, seed: int, nbins: int):
    """Bin counts (multinomial over nbins equal bins of ln lambda) and bin edges."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cnt = rng.multinomial(n, np.full(nbins, 1.0 / nbins))
    a, b = math.log(wl_lo), math.log(wl_hi)
    return cnt, a + (b - a) * np.arange(nbins + 1) / nbins


def _diatomic_bin(seed: int, i: int, lo: float, hi: float, m: int) -> dict:
    # every bin has its own stream, so a slice of the list needs only the bins it touches
    r = np.random.Generator(np.random.PCG64([seed, i]))
    w = np.exp(np.sort(r.uniform(lo, hi, m)))
    lggf = r.normal(-4.0, 1.5, m).astype(np.float32)
    elo = r.exponential(1.0, m).astype(np.float32)
    grad = r.normal(8.0, 0.3, m).astype(np.float32)
    return {"W": w, "E": elo, "LGGF": lggf, "GRAD": grad}


def iter_diatomic_bins(n: int, wl_lo: float, wl_hi: float, seed: int, nbins: int = 0):
    """The seeded diatomic list of the molecular configuration bin by bin, in wavelength
    order (columns as diatomic_columns)."""
    if nbins <= 0:
        nbins = max(1, n // 1_000_000)
    cnt, edges = _diatomic_plan(n, wl_lo, wl_hi, seed, nbins)
    for i in range(nbins):
        yield _diatomic_bin(seed, i, edges[i], edges[i + 1], int(cnt[i]))


def diatomic_columns(n: int, wl_lo: float, wl_hi: float, seed: int, first: int = 0, count: int = -1,
                     nbins: int = 0) -> dict:
    """Lines [first, first + count) of the seeded diatomic list of the molecular
    configuration as the /Lines columns of its HDF5 form (W, E, LGGF, GRAD; log values):
    lambda log-uniform in [wl_lo, wl_hi], log gf ~ N(-4, 1.5), E_low ~ Exp(1 eV), log Grad ~
    N(8, 0.3).  The list is generated bin by bin (equal bins of ln lambda with multinomial
    counts, each bin's draws sorted: the order statistics of n independent draws), so any
    slice of it comes out without the whole list in memory."""
    if count < 0:
        count = n - first
    if nbins <= 0:
        nbins = max(1, n // 1_000_000)
    cnt, edges = _diatomic_plan(n, wl_lo, wl_hi, seed, nbins)
    starts = np.concatenate([[0], np.cumsum(cnt)])
    cols = {k: [] for k in ("W", "E", "LGGF", "GRAD")}
    for i in range(nbins):
        s0, s1 = starts[i], starts[i + 1]
        if s1 <= first or s0 >= first + count:
            continue
        c = _diatomic_bin(seed, i, edges[i], edges[i + 1], int(cnt[i]))
        lo, hi = max(first, s0) - s0, min(first + count, s1) - s0
        for k in cols:
            cols[k].append(c[k][lo:hi])
    return {k: (np.concatenate(v) if v else np.zeros(0, np.float64 if k == "W" else np.float32))
            for k, v in cols.items()}


def write_diatomic_sidecar(path, n: int = 0, wl_lo: float = 0.0, wl_hi: float = 0.0, seed: int = 0) -> int:
    """The molecular configuration's list (default: all 200,000,000 lines) as a sidecar
    directory (W f8, E, LGGF, GRAD f4, header), written bin by bin; returns the bytes
    written."""
    from pathlib import Path
    n = n or C3_NLINES
    wl_lo, wl_hi = (wl_lo, wl_hi) if wl_lo > 0 else C3_RANGE
    seed = seed or SEED_BASE + 3
    path = Path(path)
    path.mkdir(parents=True, exist_ok=True)
    files = {k: open(path / f"{k}.bin", "wb") for k in ("W", "E", "LGGF", "GRAD")}
    nb, wmin, wmax, k = 0, None, None, 0
    try:
        for c in iter_diatomic_bins(n, wl_lo, wl_hi, seed):
            for key, f in files.items():
                a = np.ascontiguousarray(c[key], "<f8" if key == "W" else "<f4")
                f.write(a.tobytes())
                nb += a.nbytes
            if len(c["W"]):
                wmin = c["W"][0] if wmin is None else wmin
                wmax = c["W"][-1]
            k += len(c["W"])
    finally:
        for f in files.values():
            f.close()
    hdr = ["format=fastspec-lines-1", "name=FeH", "wl_vac=0", "hflag=0", f"wmin={float(wmin)!r}",
           f"wmax={float(wmax)!r}", f"nl={k}", "columns=W:f8,E:f4,LGGF:f4,GRAD:f4"]
    (path / "header.txt").write_text("\n".join(hdr) + "\n")
    return nb


def diatomic_lines(cols: dict, spc: int) -> LineList:
    """The reader's conversions of diatomic columns (gf = 10**log gf, Grad = 10**log Grad in
    REAL(4), src/ioutils.f90:1060-1066): a molecular list of species spc."""
    return LineList(wl=np.asarray(cols["W"], np.float64), elo=np.asarray(cols["E"], np.float32),
                    gf=_f32pow10(cols["LGGF"]), grad=_f32pow10(cols["GRAD"]),
                    spc=np.full(len(cols["W"]), spc, np.int32), list_type=FSG_LIST_MOLECULAR)


def zeeman_pattern(jlo: float, jup: float, glo: float, gup: float):
    """Anomalous Zeeman pattern in the reference's convention (src/utils.f90:901-972):
    pi shifts/strengths and one sigma branch, strengths normalised per branch."""
    tol = 1e-5
    nm = int(round(2 * jup)) + 1
    gp, rp, gs, rs = [], [], [], []
    for i in range(nm):
        m = jup - i
        gp.append(m * (gup - glo))
        gs.append(m * gup - (m + 1.0) * glo)
        if jlo == jup:
            rp.append(m * m)
            rs.append(0.25 * (jup - m) * (jup + m + 1.0))
        elif jlo > jup:
            rp.append((jup + 1.0) ** 2 - m * m)
            rs.append(0.25 * (jup + m + 1.0) * (jup + m + 2.0))
        else:
            rp.append(jup**2 - m * m)
            rs.append(0.25 * (jup - m) * (jup - m - 1.0))
    gp = [0.0 if abs(x) < tol else x for x in gp]
    gs = [0.0 if abs(x) < tol else x for x in gs]
    P = [(g, r) for g, r in zip(gp, rp) if abs(r) > tol]
    S = [(-g, r) for g, r in zip(gs, rs) if abs(r) > tol]

    def merge(comp):
        out = []
        for g, r in comp:
            for k, (gm, rm) in enumerate(out):
                if abs(g - gm) < tol:
                    out[k] = (gm, rm + r)
                    break
            else:
                out.append((g, r))
        return out

    P, S = merge(P), merge(S)
    sp = sum(r for _, r in P)
    ss = sum(r for _, r in S)
    P = [(g, r / sp if sp > 0 else r) for g, r in P]
    S = [(g, r / ss if ss > 0 else r) for g, r in S]
    return P, S


def zeeman_level_columns(n: int, seed: int) -> dict:
    """Seeded JLO, JUP, GLO, GUP columns of a Zeeman list (J <= 6 with dJ in {0, +-1}, not
    0 -> 0; Lande g ~ U(0.5, 2.5)); every 41st line has no lower Lande factor (99, a line
    without splitting, as the reference's converter marks it)."""
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
    glo = rng.uniform(0.5, 2.5, n).astype(np.float32)
    gup = rng.uniform(0.5, 2.5, n).astype(np.float32)
    glo[::41] = 99.0
    return {"JLO": jlo, "JUP": jup, "GLO": glo, "GUP": gup}


def zeeman_patterns(n: int, seed: int):
    """Seeded Zeeman patterns for n lines in CSR form: J <= 6, Lande g ~ U(0.5, 2.5); sigma
    components stored blue (negated shifts) then red, as the reference's reader does
    (src/ioutils.f90:1012-1031)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pi_off, si_off = [0], [0]
    spl_pi, str_pi, spl_si, str_si = [], [], [], []
    for _ in range(n):
        while True:
            jlo = 0.5 * rng.integers(0, 13)
            dj = rng.integers(-1, 2)
            jup = jlo + dj
            if jup < 0 or jup > 6 or (jlo == 0 and jup == 0):
                continue
            break
        glo, gup = rng.uniform(0.5, 2.5, 2)
        P, S = zeeman_pattern(float(jlo), float(jup), float(np.float32(glo)), float(np.float32(gup)))
        for g, r in P:
            spl_pi.append(g)
            str_pi.append(r)
        for g, r in S:
            spl_si.append(-g)
            str_si.append(r)
        for g, r in S:
            spl_si.append(g)
            str_si.append(r)
        pi_off.append(len(spl_pi))
        si_off.append(len(spl_si))
    return dict(pi_off=np.array(pi_off, np.int32), spl_pi=np.array(spl_pi, np.float32),
                str_pi=np.array(str_pi, np.float32), si_off=np.array(si_off, np.int32),
                spl_si=np.array(spl_si, np.float32), str_si=np.array(str_si, np.float32))


def magnetic_lines(n: int, wl_lo: float, wl_hi: float, seed: int) -> LineList:
    """Seeded Zeeman list (mixed .z type) with the patterns of zeeman_patterns."""
    base = atomic_lines(n, wl_lo, wl_hi, seed)
    return LineList(
        wl=base.wl, elo=base.elo, gf=base.gf, grad=base.grad, gqst=base.gqst, gvdw=base.gvdw,
        spc=base.spc, is_mol=np.zeros(n, np.uint8), list_type=FSG_LIST_MIXED,
        **zeeman_patterns(n, seed + 7),
    )


# --- the benchmark configurations -------------------------------------------------------

def config_c1(nlines: int = 20000):
    """Plain FP64 scalar synthesis: 20,000 lines in 5000-5100 A, 0.01 A step, 7 mu."""
    atm = grey_atmosphere(5777.0, 4.44)
    setup = build_setup(atm, 5000.0, 5100.0, 0.01, 0)
    lines = atomic_lines(nlines, 5000.0, 5100.0, SEED_BASE + 1)
    return setup, lines


def config_c2(nlines: int = 2_000_000, wl_lo: float = 3000.0, wl_hi: float = 10000.0):
    """Large atomic run: 2,000,000 lines log-uniform in 3000-10000 A on an R=10^6 log grid."""
    atm = grey_atmosphere(5777.0, 4.44)
    setup = build_setup(atm, wl_lo, wl_hi, 0.299792458, 1)
    # line density of the full configuration kept for sub-ranges
    frac = math.log(wl_hi / wl_lo) / math.log(10000.0 / 3000.0)
    n = int(round(nlines * frac)) if (wl_lo, wl_hi) != (3000.0, 10000.0) else nlines
    lines = atomic_lines(n, wl_lo, wl_hi, SEED_BASE + 2, log_uniform=True)
    return setup, lines


C3_NLINES = 200_000_000
C3_RANGE = (4000.0, 25000.0)


def config_c3_setup(wl_lo: float = C3_RANGE[0], wl_hi: float = C3_RANGE[1]):
    """Cool-dwarf molecular run: grey atmosphere Teff 3200 K, log g 5.0, 72 depths, R = 3x10^5
    log grid (step 0.999308193 km/s; 549,775 points over 4000-25000 A), the diatomic FeH as
    the line species.  Returns (setup, the FeH species slot)."""
    atm = grey_atmosphere(3200.0, 5.0)
    setup = build_setup(atm, wl_lo, wl_hi, 0.999308193, 1, molecules=MOLECULAR_SPECIES)
    return setup, len(ATOMIC_SPECIES) + 1


def config_c3(nlines: int = C3_NLINES, wl_lo: float = C3_RANGE[0], wl_hi: float = C3_RANGE[1]):
    """The molecular configuration (or a wavelength slice of it with the full list's line
    density): 200,000,000 seeded diatomic lines log-uniform in 4000-25000 A, log gf ~
    N(-4, 1.5), E_low ~ Exp(1 eV); blocks of NlinesLim = 10^7 lines (C3_NLINES_LIM).  The
    slice's lines are those of the full list that fall into it."""
    setup, spc = config_c3_setup(wl_lo, wl_hi)
    if (wl_lo, wl_hi) == C3_RANGE:
        cols = diatomic_columns(nlines, *C3_RANGE, SEED_BASE + 3)
    else:
        cols = diatomic_slice(nlines, wl_lo, wl_hi)
    return setup, diatomic_lines(cols, spc)


C3_NLINES_LIM = 10_000_000


def diatomic_slice(nlines: int, wl_lo: float, wl_hi: float) -> dict:
    """Columns of the full molecular list's lines inside [wl_lo, wl_hi] (only the bins that
    reach the slice are generated)."""
    nbins = max(1, nlines // 1_000_000)
    a, b = math.log(C3_RANGE[0]), math.log(C3_RANGE[1])
    i0 = max(0, int(math.floor((math.log(wl_lo) - a) / (b - a) * nbins)))
    i1 = min(nbins - 1, int(math.floor((math.log(wl_hi) - a) / (b - a) * nbins)))
    cnt, _ = _diatomic_plan(nlines, *C3_RANGE, SEED_BASE + 3, nbins)
    starts = np.concatenate([[0], np.cumsum(cnt)])
    cols = diatomic_columns(nlines, *C3_RANGE, SEED_BASE + 3, first=int(starts[i0]),
                            count=int(starts[i1 + 1] - starts[i0]), nbins=nbins)
    keep = (cols["W"] >= wl_lo) & (cols["W"] <= wl_hi)
    return {k: v[keep] for k, v in cols.items()}


def config_c4(nlines: int = 5000):
    """Stokes run: 5,000 Zeeman lines in 6000-6100 A, 0.005 A step, B = 2 kG at 45/30 deg."""
    atm = grey_atmosphere(5777.0, 4.44)
    setup = build_setup(atm, 6000.0, 6100.0, 0.005, 0)
    lines = magnetic_lines(nlines, 6000.0, 6100.0, SEED_BASE + 4)
    bfield = np.array([2000.0, 45.0, 30.0])
    return setup, lines, bfield, 1


def config_c5(nmodels: int = 256, nlines: int = 50_000, step: float = 0.01, margin: float = 2.5):
    """Differentiable batch: nmodels grey atmospheres (Teff ~ U(4500, 6500) K, log g ~
    U(3.5, 4.8)), each with its own window of width ~ U(5, 50) A inside 5000-5500 A on a
    0.01 A grid, and one shared seeded list of 50,000 atomic lines lambda ~ U(5000, 5500) A
    (distributions as the plain configuration).  Returns the shared list and per model
    (setup, indices of the shared lines within margin of its window)."""
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 5))
    lines = atomic_lines(nlines, 5000.0, 5500.0, SEED_BASE + 50)
    teff = rng.uniform(4500.0, 6500.0, nmodels)
    logg = rng.uniform(3.5, 4.8, nmodels)
    width = rng.uniform(5.0, 50.0, nmodels)
    start = 5000.0 + rng.uniform(0.0, 1.0, nmodels) * (500.0 - width)
    models = []
    for m in range(nmodels):
        lo = float(np.round(start[m], 2))
        hi = float(np.round(lo + width[m], 2))
        atm = grey_atmosphere(float(teff[m]), float(logg[m]))
        setup = build_setup(atm, lo, hi, step, 0)
        idx = np.nonzero((lines.wl >= lo - margin) & (lines.wl <= hi + margin))[0]
        models.append((setup, idx))
    return lines, models
