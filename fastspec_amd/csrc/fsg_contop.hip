// fsg_contop.hip -- continuum opacity on the device: ContopCalc's per-wavelength loop
// (src/synth.f90:663-669) for the sources of Xcont (src/contop.f90:267-314) that depend on
// the hydrogen, helium and electron populations alone:
//
//   HOP     H I bound-free and free-free        src/contop.f90:1079-1100
//   H2PLOP  H2+ bound-free and free-free        src/contop.f90:1102-1113
//   HMINOP  H- bound-free and free-free         src/contop.f90:1128-1181
//   HRAYOP  Rayleigh scattering by H I          src/contop.f90:1115-1124
//   HEMIOP  He- free-free                       src/contop.f90:1518-1527
//   HE1OP   He I bound-free and free-free       src/contop.f90:1187-1305
//   HE2OP   He II bound-free and free-free      src/contop.f90:1488-1516
//   HERAOP  Rayleigh scattering by He I         src/contop.f90:1529-1539
//   ELECOP  Thomson scattering                  src/contop.f90:2764-2768
//   H2RAOP  Rayleigh scattering by H2           src/contop.f90:2770-2780
//   Linter, Map1 (HMINOP's interpolations)      src/contop.f90:500-601
//   XKARZAS, COULFF (hydrogenic bound-free and   src/contop.f90:625-658, 695-1077
//     free-free)
//   CROSSHE, HE12S1S, HE12S3S, HE12P1P, HE12P3P  src/contop.f90:1309-1486
//     (He I cross-sections)
//
// One lane per wavelength of the grid: the wavelength's own quantities (frequency, the H-
// free-free table interpolated in wavelength for each of its 11 temperatures, the H-
// bound-free cross-section, the Rayleigh cross-sections) are formed once, then the lane runs
// down the depths, writing ContOpac [ntau][nwl] (wavelength fastest: a wavefront's stores are
// contiguous).  Per-depth state (T, hc/kT, kT in eV, 5040/T, Xne, Rho and the five
// populations) comes from the host, formed there as ContopSetup forms it.  All arithmetic is
// FP64 in the reference's order without contraction (the pragma below also covers the
// air-to-vacuum conversion of the shared header), so that each source matches the compiled
// reference to rounding.
#pragma clang fp contract(off)

#include <cmath>
#include <vector>

#include "fsg_contop_tables.h"
#include "fsg_coolop_tables.h"
#include "fsg_ionop_tables.h"
#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

constexpr double kClightA = kVc * 1e13;   // speed of light [A/s]
constexpr double kClightCm = kVc * 1e5;   // speed of light [cm/s]
constexpr int KL_NP = 29;                 // tabulated points per level of the Karzas & Latter tables

// XKARZAS (src/contop.f90:695-1077) for levels n <= 15 (the callers use n <= 9): the
// hydrogenic bound-free cross-section of level n (and orbital l, for l < n <= 6) at
// frequency freq for an effective charge squared zeff2, interpolated linearly in
// log10(cross-section) against log10(freq / zeff2) in the table of the level; 0 below the
// threshold.  The bracketing point is the first tabulated frequency below the argument
// (the tables descend); above the table's top it is the top interval.
__host__ __device__ double xkarzas(double freq, double zeff2, int n, int l) {
  const double flg = log10(freq / zeff2);
  const double *fn = kKlFreqn + (n - 1) * KL_NP;
  const double *x = (l >= n || n > 6) ? kKlXn + (n - 1) * KL_NP : kKlXl + (l * 6 + (n - 1)) * KL_NP;
  if (flg < fn[KL_NP - 1]) return 0.0;
  int i = 2;  // 1-based, as the reference's loop index
  while (i <= KL_NP && !(flg > fn[i - 2])) i++;
  i = i < 3 ? 3 : i;
  const double X = (flg - fn[i - 2]) / (fn[i - 3] - fn[i - 2]) * (x[i - 3] - x[i - 2]) + x[i - 2];
  return exp(X * 2.30258509299405) / zeff2;
}

// COULFF (src/contop.f90:625-658): the hydrogenic free-free Gaunt factor for charge nz (1 or
// 2) at depth temperature ln T = tlog and ln(frequency) = freqlg, bilinear in the table of
// log(gamma^2) and log(h nu / kT)
__device__ double coulff(double tlog, double freqlg, int nz) {
  const double z4 [6] = {0.0, 1.20412, 1.90849, 2.40824, 2.79588, 3.11261};  // log10 of Z^4
  const double z4log = z4[nz - 1];
  const double gamlog = 10.39638 - tlog / 1.15129 + z4log;
  const int igam = max(min((int)(gamlog + 7.0), 10), 1);
  const double hvktlg = (freqlg - tlog) / 1.15129 - 20.63764;
  const int ihvkt = max(min((int)(hvktlg + 9.0), 11), 1);
  const double P = gamlog - (igam - 7), Q = hvktlg - (ihvkt - 9);
  auto A = [](int i, int j) { return kCoulffA[(i - 1) * 11 + (j - 1)]; };
  return (1.0 - P) * ((1.0 - Q) * A(ihvkt, igam) + Q * A(ihvkt + 1, igam)) +
         P * ((1.0 - Q) * A(ihvkt, igam + 1) + Q * A(ihvkt + 1, igam + 1));
}

// CROSSHE (src/contop.f90:1309-1355): He I 1s2 photoionization cross-section, linear in
// wavelength in four tables (every 5, 2, 1 and 0.5 A)
__device__ double crosshe(double freq0) {
  if (freq0 < 5.945209e15) return 0.0;
  const double w = kClightA / freq0;
  int i;
  if (w > 50.0) {
    i = min(92, max(2, (int)(93.0 - (w - 50.0) / 5.0)));
    return ((w - (double)(92 - i) * 5.0 - 50.0) / 5.0 * (kHeX505[i - 2] - kHeX505[i - 1]) + kHeX505[i - 1]) * 1.e-18;
  }
  if (w > 20.0) {
    i = min(16, max(2, (int)(17.0 - (w - 20.0) / 2.0)));
    return ((w - (double)(16 - i) * 2.0 - 20.0) / 2.0 * (kHeX50[i - 2] - kHeX50[i - 1]) + kHeX50[i - 1]) * 1.e-18;
  }
  if (w > 10.0) {
    i = min(11, max(2, (int)(12.0 - (w - 10.0))));
    return ((w - (double)(11 - i) - 10.0) * (kHeX20[i - 2] - kHeX20[i - 1]) + kHeX20[i - 1]) * 1.e-18;
  }
  i = min(21, max(2, (int)(22.0 - w / 0.5)));
  return ((w - (double)(21 - i) * 0.5) / 0.5 * (kHeX10[i - 2] - kHeX10[i - 1]) + kHeX10[i - 1]) * 1.e-18;
}

// The Mathisen (1984) tables of HE12S1S, HE12S3S, HE12P1P, HE12P3P: log10 cross-section
// linear in log10(frequency) over 16 descending points
__device__ double he_table16(double freq0, const double *fq, const double *xs) {
  const double flg = log10(freq0);
  int i = 2;
  while (i <= 16 && !(flg > fq[i - 1])) i++;
  i = min(i, 16);
  const double X = (flg - fq[i - 1]) / (fq[i - 2] - fq[i - 1]) * (xs[i - 2] - xs[i - 1]) + xs[i - 1];
  return pow(10.0, X);
}

// HE12S1S, HE12S3S, HE12P1P, HE12P3P (src/contop.f90:1358-1486): the 1s2s and 1s2p
// photoionization cross-sections (threshold wavenumber e0 [1/cm]); above 2.4 Ry the 2s and
// 2p1P ones follow the resonance fits of the reference
__device__ double he12(int which, double freq0) {
  const double ryd = 109722.267;
  switch (which) {
    case 0: {  // 1s2s 1S
      if (freq0 < 32033.214 * kClightCm) return 0.0;
      if (freq0 > 2.4 * ryd * kClightCm) {
        const double wn = freq0 / kClightCm, ek = (wn - 32033.214) / ryd;
        const double eps = 2.0 * (ek - 2.612316) / 0.00322;
        return 0.008175 * pow(484940.0 / wn, 2.71) * 8.067e-18 * ((eps + 76.21) * (eps + 76.21)) / (1.0 + eps * eps);
      }
      return he_table16(freq0, kHeFreq1s, kHeX1s);
    }
    case 1: {  // 1s2s 3S
      if (freq0 < 38454.691 * kClightCm) return 0.0;
      if (freq0 > 2.4 * ryd * kClightCm) {
        const double wn = freq0 / kClightCm, ek = (wn - 38454.691) / ryd;
        const double eps = 2.0 * (ek - 2.47898) / 0.000780;
        return 0.01521 * pow(470310.0 / wn, 3.12) * 8.067e-18 * ((eps - 122.4) * (eps - 122.4)) / (1.0 + eps * eps);
      }
      return he_table16(freq0, kHeFreq3s, kHeX3s);
    }
    case 2: {  // 1s2p 1P
      if (freq0 < 27175.76 * kClightCm) return 0.0;
      if (freq0 > 2.4 * ryd * kClightCm) {
        const double wn = freq0 / kClightCm, ek = (wn - 27175.76) / ryd;
        const double e1s = 2.0 * (ek - 2.446534) / 0.01037, e1d = 2.0 * (ek - 2.59427) / 0.00538;
        return 0.0009487 * pow(466750.0 / wn, 3.69) * 8.067e-18 *
               ((e1s - 29.30) * (e1s - 29.30) / (1.0 + e1s * e1s) + (e1d + 172.4) * (e1d + 172.4) / (1.0 + e1d * e1d));
      }
      return he_table16(freq0, kHeFreq1p, kHeX1p);
    }
    default:  // 1s2p 3P
      if (freq0 < 29223.753 * kClightCm) return 0.0;
      return he_table16(freq0, kHeFreq3p, kHeX3p);
  }
}

// Linter (src/contop.f90:500-516) at one abscissa: linear interpolation between the first
// tabulated point above x and the one before it (extrapolation past either end); the
// ordinates yo[0], yo[ys], yo[2 ys], ...
__device__ __forceinline__ double linter1(const double *xo, const double *yo, int ys, int n, double x) {
  int i = 1;  // 0-based index of the upper point
  while (x >= xo[i] && i < n - 1) i++;
  const double y0 = yo[(i - 1) * ys], y1 = yo[i * ys];
  return y0 + (y1 - y0) / (xo[i] - xo[i - 1]) * (x - xo[i - 1]);
}

// Map1 (src/contop.f90:518-601) at one abscissa: the parabolas through the neighbouring
// points on either side of x, blended by the size of their curvature; straight lines in the
// first two intervals and past the last point.
__device__ double map1_one(const double *xo, const double *fo, int n, double x) {
  int l = 2;  // 1-based as in the reference
  while (true) {
    if (x < xo[l - 1]) break;
    l++;
    if (l > n) break;
  }
  auto X = [xo](int k) { return xo[k - 1]; };
  auto F = [fo](int k) { return fo[k - 1]; };
  double a, b, c;
  if (l > n || l == 2 || l == 3) {
    l = l > n ? n : l;
    c = 0.0;
    b = (F(l) - F(l - 1)) / (X(l) - X(l - 1));
    a = F(l) - X(l) * b;
  } else {
    const int l1 = l - 1, l2 = l - 2;
    double d = (F(l1) - F(l2)) / (X(l1) - X(l2));
    const double cbac = F(l) / ((X(l) - X(l1)) * (X(l) - X(l2))) +
                        (F(l2) / (X(l) - X(l2)) - F(l1) / (X(l) - X(l1))) / (X(l1) - X(l2));
    const double bbac = d - (X(l1) + X(l2)) * cbac;
    const double abac = F(l2) - X(l2) * d + X(l1) * X(l2) * cbac;
    if (l == n) {
      a = abac;
      b = bbac;
      c = cbac;
    } else {
      d = (F(l) - F(l1)) / (X(l) - X(l1));
      const double cfor = F(l + 1) / ((X(l + 1) - X(l)) * (X(l + 1) - X(l1))) +
                          (F(l1) / (X(l + 1) - X(l1)) - F(l) / (X(l + 1) - X(l))) / (X(l) - X(l1));
      const double bfor = d - (X(l) + X(l1)) * cfor;
      const double afor = F(l1) - X(l1) * d + X(l) * X(l1) * cfor;
      double wt = 0.0;
      if (fabs(cfor) != 0.0) wt = fabs(cfor) / (fabs(cfor) + fabs(cbac));
      a = afor + wt * (abac - afor);
      b = bfor + wt * (bbac - bfor);
      c = cfor + wt * (cbac - cfor);
    }
  }
  return a + (b + c * x) * x;
}

__global__ void __launch_bounds__(256) k_contop(int nwl, int ntau, const double *__restrict__ wave, double wfirst,
                                                const double *__restrict__ dep, const double *__restrict__ tab,
                                                const double *__restrict__ cstd, int motype, unsigned mask, int add,
                                                double *__restrict__ contT) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nwl) return;
  auto row = [dep, ntau](int r) { return dep + (size_t)r * ntau; };
  const double *T = row(CR_T), *Hkt = row(CR_HKT), *Tkev = row(CR_TKEV), *Theta = row(CR_THETA);
  const double *Xne = row(CR_XNE), *Rho = row(CR_RHO);
  const double *Fh1 = row(CR_FH1), *Fh2 = row(CR_FH2), *Fhmin = row(CR_FHMIN);
  const double *Fhe1 = row(CR_FHE1), *Fhmol = row(CR_FHMOL), *Tlog = row(CR_TLOG);
  const double *hmwlog = tab, *hmfflog = tab + CONT_HM_NW;  // [22], [22][11]
  const double *thetaff = tab + CONT_HM_NW * (1 + CONT_HM_NT);  // [11]
  const double *wbf = thetaff + CONT_HM_NT, *bf = wbf + CONT_HMBF_N;  // [85], [85]

  const double clight = kVc * 1e13;
  const double Freq = clight / air2vac(wave[p] + wfirst);
  const double Freqlg = log(Freq);
  const bool on_h2p = (mask >> 1) & 1, on_hmin = (mask >> 2) & 1, on_hray = (mask >> 3) & 1;
  const bool on_hemi = (mask >> 6) & 1, on_hera = (mask >> 7) & 1, on_el = (mask >> 11) & 1;
  const bool on_h2ra = (mask >> 12) & 1;
  const bool on_hop = mask & 1, on_he1 = (mask >> 4) & 1, on_he2 = (mask >> 5) & 1;
  const bool hop_lev = !(mask & CONT_HOP_NOLEV);

  // HOP: the bound-free cross-sections of levels 1-8 (Karzas & Latter, Z = 1)
  double hcont[8], hop_c = 0.0, hop_cf = 0.0;
  if (on_hop) {
    for (int n = 1; n <= 8; n++) hcont[n - 1] = xkarzas(Freq, 1.0, n, n);
    const double f3 = Freq * Freq * Freq;
    hop_cf = 3.6919e8 / f3;
    hop_c = 2.815e29 / f3;
  }
  // HE2OP: levels 1-9 of He II (Z^2 = 4)
  double he2c[9], he2_c = 0.0, he2_cf = 0.0;
  if (on_he2) {
    for (int n = 1; n <= 9; n++) he2c[n - 1] = xkarzas(Freq, 4.0, n, n);
    const double f5 = Freq * 1.0e-05, f3 = f5 * f5 * f5;
    he2_cf = 3.6919e-07 / f3 * 4.0;
    he2_c = 2.815e14 * 2.0 * 2.0 / f3;
  }
  // HE1OP: the ten explicit levels from the lowest one ionizable at this frequency, the
  // photoionizations leaving He II in n = 2 and n = 3 added to their levels
  double he1t[10], he1_c = 0.0, he1_cf = 0.0;
  int he1min = 0;
  if (on_he1) {
    const double f10 = Freq * 1.0e-10, f3 = f10 * f10 * f10;
    he1_cf = 3.6919e8 / f3;
    he1_c = 2.815e-1 / f3;
    const double hefreq[10] = {5.9452090e15, 1.1528440e15, 0.9603331e15, 0.8761076e15, 0.8147104e15,
                               0.4519048e15, 0.4030971e15, 0.3821191e15, 0.3660215e15, 0.3627891e15};
    for (int n = 1; n <= 10 && !he1min; n++)
      if (hefreq[n - 1] <= Freq) he1min = n;
    for (int n = 0; n < 10; n++) he1t[n] = 0.0;
    if (he1min >= 1) {
      const double zeff3[5] = {1.236439, 1.102898, 1.045499, 1.001427, 0.9926};
      const int l3[5] = {0, 0, 1, 2, 1};
      for (int n = he1min; n <= 10; n++) {
        if (n == 1) he1t[0] = crosshe(Freq);
        else if (n == 2) he1t[1] = he12(1, Freq);
        else if (n == 3) he1t[2] = he12(0, Freq);
        else if (n == 4) he1t[3] = he12(3, Freq);
        else if (n == 5) he1t[4] = he12(2, Freq);
        else he1t[n - 1] = xkarzas(Freq, zeff3[n - 6], 3, l3[n - 6]);
      }
      const double ryd = 109722.273 * 2.99792458e10;
      // n = 2 levels, He II left in n = 2: 1s2p 1P, 1s2p 3P, 1s2s 1S, 1s2s 3S
      const double e2 = 527490.06, lev2[4] = {171135.000, 169087.0, 166277.546, 159856.069};
      const int to2[4] = {4, 3, 2, 1};
      for (int q = 0; q < 4; q++) {
        const double fh = (e2 - lev2[q]) * kClightCm;
        if (!(Freq >= fh)) break;
        he1t[to2[q]] = he1t[to2[q]] + xkarzas(Freq, fh / ryd, 1, 0);
      }
      // n = 3 levels, He II left in n = 3: 1s3p 1P, 1s3d, 1s3p 3P, 1s3s 1S, 1s3s 3S
      const double e3 = 588451.59, lev3[5] = {186209.471, 186101.0, 185564.0, 184864.0, 183236.0};
      const int to3[5] = {9, 8, 7, 6, 5};
      for (int q = 0; q < 5; q++) {
        const double fh = (e3 - lev3[q]) * kClightCm;
        if (!(Freq >= fh)) break;
        he1t[to3[q]] = he1t[to3[q]] + xkarzas(Freq, fh / ryd, 1, 0);
      }
    }
  }
  // hydrogenic levels n = 4-27 of He I, from 1.25408e16 Hz on (below 240 A): evaluated in
  // the depth loop, where the rare frequencies that need them pay for them
  const bool he1n = on_he1 && he1min >= 1 && Freq >= 1.25408e16 && Freq >= (588451.59 - 183236.0) * kClightCm;

  // H2PLOP
  const bool h2p = on_h2p && !(Freq > 3.28805e15);
  double Fr = 0.0, Es = 0.0;
  if (h2p) {
    Fr = -3.0233e3 + (3.7797e2 + (-1.82496e1 + (3.9207e-1 - 3.1672e-3 * Freqlg) * Freqlg) * Freqlg) * Freqlg;
    const double F15 = Freq * 1.e-15;
    Es = -7.342e-3 + (-2.409 + (1.028 + (-0.4230 + (0.1224 - 0.01351 * F15) * F15) * F15) * F15) * F15;
  }
  // HMINOP: the free-free table at this wavelength for each tabulated temperature, and the
  // bound-free cross-section
  double fftt[CONT_HM_NT];
  double hminbf = 0.0;
  if (on_hmin) {
    const double wl = clight / Freq * 0.1;
    const double wlog = log(wl);
    for (int it = 0; it < CONT_HM_NT; it++)
      fftt[it] = exp(linter1(hmwlog, hmfflog + it, CONT_HM_NT, CONT_HM_NW, wlog)) / thetaff[it] * 5040.0 * 1.380658e-16;
    if (Freq > 1.82365e14) hminbf = map1_one(wbf, bf, CONT_HMBF_N, wl);
    hminbf = fmax(hminbf, 0.0);
  }
  // HRAYOP, HEMIOP, HERAOP, H2RAOP: their wavelength factors
  double sig_h = 0.0, sig_he = 0.0, sig_h2 = 0.0, A = 0.0, B = 0.0, C = 0.0;
  if (on_hray) {
    const double w = clight / fmin(Freq, 2.463e15), w2 = w * w;
    sig_h = (5.799e-13 + 1.422e-6 / w2 + 2.784 / (w2 * w2)) / (w2 * w2);
  }
  if (on_hemi) {
    A = 3.397e-26 + (-5.216e-11 + 7.039e05 / Freq) / Freq;
    B = -4.116e-22 + (1.067e-06 + 8.135e09 / Freq) / Freq;
    C = 5.081e-17 + (-8.724e-03 - 5.659e12 / Freq) / Freq;
  }
  if (on_hera) {
    const double w = kVc / 1e2 / fmin(Freq * 1.0e-15, 5.15), w2 = w * w;
    const double s1 = 1.0 + (2.44e5 + 5.94e10 / (w2 - 2.90e5)) / w2;
    sig_he = 5.484e-14 / w2 / w2 * s1 * s1;
  }
  if (on_h2ra) {
    const double w = clight / fmin(Freq, 2.922e15), w2 = w * w;
    sig_h2 = (8.14e-13 + 1.28e-6 / w2 + 1.61 / (w2 * w2)) / (w2 * w2);
  }

  for (int k = 0; k < ntau; k++) {
    const double ehvkt = exp(-Freq * Hkt[k]);
    const double stim = 1.0 - ehvkt;
    double ahyd = 0.0, ahe1 = 0.0, ahe2 = 0.0;
    double ah2p = 0.0, ahmin = 0.0, ahemin = 0.0, sh = 0.0, she = 0.0, sel = 0.0, sh2 = 0.0;
    if (on_hop) {
      const double exlim = row(CR_EXLIMH)[k];
      double ex = row(CR_BOLTEXH)[k];
      if (Freq < 4.05933e13) ex = exlim / ehvkt;
      double h = ((ex - exlim) * hop_c + hcont[6] * row(CR_BOLTH + 6)[k] + hcont[7] * row(CR_BOLTH + 7)[k] +
                  coulff(Tlog[k], Freqlg, 1) * row(CR_FREETH)[k] * hop_cf) * stim;
      if (hop_lev) {
        double sm = 0.0;
        for (int n = 0; n < 6; n++) sm = sm + hcont[n] * row(CR_BOLTH + n)[k];
        h = h + sm * stim;
      }
      ahyd = h;
    }
    if (on_he1) {
      const double exlim = row(CR_EXLIMHE1)[k];
      double ex = row(CR_BOLTEXHE1)[k];
      if (Freq < 2.055e14) ex = exlim / ehvkt;
      double he1 = (ex - exlim) * he1_c;
      if (he1min >= 1) {
        for (int n = he1min; n <= 10; n++) he1 = he1 + he1t[n - 1] * row(CR_BOLTHE1 + n - 1)[k];
        if (Freq >= 1.25408e16)
          for (int n = 4; n <= 27; n++)
            he1 = he1 + (he1n ? xkarzas(Freq, 4.0 - 3.0 / (double)(n * n), 1, 0) : 0.0) * row(CR_BOLTNHE1 + n - 4)[k];
      }
      ahe1 = (he1 + coulff(Tlog[k], Freqlg, 1) * row(CR_FREETHE1)[k] * he1_cf) * stim;
    }
    if (on_he2) {
      const double exlim = row(CR_EXLIMHE2)[k];
      double ex = row(CR_BOLTEXHE2)[k];
      if (Freq < 1.31522e14) ex = exlim / ehvkt;
      double he2 = (ex - exlim) * he2_c;
      for (int n = 0; n < 9; n++) he2 = he2 + he2c[n] * row(CR_BOLTHE2 + n)[k];
      he2 = (he2 + coulff(Tlog[k], Freqlg, 2) * he2_cf * row(CR_FREETHE2)[k]) * stim;
      ahe2 = he2 < 1.0e-30 ? 0.0 : he2;
    }
    if (h2p) ah2p = exp(-Es / Tkev[k] + Fr) * Fh1[k] * 2.0 * Fh2[k] / Rho[k] * stim;
    if (on_hmin) {
      const double fft = linter1(thetaff, fftt, 1, CONT_HM_NT, Theta[k]);
      const double hminff = fft * Fh1[k] * 2.0 * Xne[k] / Rho[k];
      const double h = hminbf * 1.e-18 * (1.0 - ehvkt) * Fhmin[k] / Rho[k];
      ahmin = h + hminff;
    }
    if (on_hemi) ahemin = (A * T[k] + B + C / T[k]) * Xne[k] * Fhe1[k] / Rho[k] * 1.0e-20;
    if (on_hray) sh = sig_h * Fh1[k] * 2.0 / Rho[k];
    if (on_hera) she = sig_he * Fhe1[k] / Rho[k];
    if (on_el) sel = 0.6653e-24 * Xne[k] / Rho[k];
    if (on_h2ra) sh2 = sig_h2 * Fhmol[k] / Rho[k];
    // Xcont's sum (the absorption sources, then the scattering ones)
    double copac = ahyd + ah2p + ahmin + ahe1 + ahe2 + ahemin;
    copac = copac + sh + she + sel + sh2;
    if (motype == 1) copac = copac / cstd[k];
    const size_t o = (size_t)k * nwl + p;
    contT[o] = add ? contT[o] + copac : copac;
  }
}

void launch_contop(int nwl, int ntau, const double *wave, double wfirst, const double *dep, const double *tab,
                   const double *cstd, int motype, unsigned mask, int add, double *contT, hipStream_t st) {
  k_contop<<<(nwl + 255) / 256, 256, 0, st>>>(nwl, ntau, wave, wfirst, dep, tab, cstd, motype, mask, add, contT);
}

// ------------------------------------------------------------------------------------
// COOLOP (src/contop.f90:1544-1569): C I, Mg I, Si I, Al I and Fe I bound-free, CH and OH
// photodissociation, H2-H2 / H2-He / H2-H collision-induced absorption, H2- free-free and
// NH photodissociation, nothing above the H I Lyman edge.
// ------------------------------------------------------------------------------------
constexpr double kCoolRyd = 109732.298;
// series limits and edges of the hydrogenic pseudo-continua of C I, Mg I, Si I, C II (two)
// and Mg II (src/contop.f90:72-77)
constexpr double kPsElim[6] = {90820.42, 61671.02, 65747.55, 196664.7, 196664.7 + 52367.06, 121267.61};
constexpr double kPsEdge[6] = {kCoolRyd / 16.0, kCoolRyd / 25.0, kCoolRyd / 25.0,
                               kCoolRyd * 4.0 / 36.0, kCoolRyd * 4.0 / 16.0, kCoolRyd * 4.0 / 64.0};
constexpr int COOL_NX = COOL_NC1 + COOL_NMG1 + COOL_NSI1;  // the three atoms' cross-sections per lane

__device__ __forceinline__ double pow10d(double x) { return pow(10.0, x); }
__device__ __forceinline__ double exp10d(double x) { return exp(x * 2.302585092994046); }

// C1OP's cross-sections at this frequency (src/contop.f90:1575-1652), X[i * ld]
__device__ void c1op_x(double Freq, double Waveno, double *X, int ld) {
  const int lryd[14] = {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 0, 0};
  for (int i = 0; i < COOL_NC1; i++) X[i * ld] = 0.0;
  // 2s2 2p3d and 2p3p / 2p3s Rydberg levels, to C II 2P average
  double Elim = 90862.70;
  for (int i = 0; i < 14; i++) {
    if (Waveno < Elim - kC1Elev[i]) break;
    X[i * ld] = xkarzas(Freq, 9.0 / kCoolRyd * (Elim - kC1Elev[i]), 3, lryd[i]);
  }
  // 2s2 2p2 levels to C II 2P1/2 (weight 1/3) and to 2P3/2 (weight 2/3): Luo & Pradhan
  // fits with the resonances of Burke & Taylor
  for (int pass = 0; pass < 2; pass++) {
    Elim = pass == 0 ? 90820.42 : 90820.42 + 63.42;
    if (Waveno < Elim - kC1Elev[14]) continue;
    const double xs0 = pow10d(-16.80 - (Waveno - Elim + kC1Elev[14]) / 3.00 / kCoolRyd);
    double eps = (Waveno - 97700.0) * 2.0 / 2743.0;
    const double xs1 = (68.e-18 * eps + 118.e-18) / (eps * eps + 1.0);
    X[14 * ld] = pass == 0 ? (xs0 + xs1) / 3.0 : X[14 * ld] + (xs0 + xs1) * 2.0 / 3.0;
    if (Waveno < Elim - kC1Elev[15]) continue;
    const double xd0 = pow10d(-16.80 - (Waveno - Elim + kC1Elev[15]) / 3.00 / kCoolRyd);
    eps = (Waveno - 93917.0) * 2.0 / 9230.0;
    const double xd1 = (22.e-18 * eps + 26.e-18) / (eps * eps + 1.0);
    eps = (Waveno - 111130.0) * 2.0 / 2743.0;
    const double xd2 = (-10.5e-18 * eps + 46.e-18) / (eps * eps + 1.0);
    X[15 * ld] = pass == 0 ? (xd0 + xd1 + xd2) / 3.0 : X[15 * ld] + (xd0 + xd1 + xd2) * 2.0 / 3.0;
    for (int q = 16; q <= 18; q++) {  // 3P2, 3P1, 3P0
      if (Waveno < Elim - kC1Elev[q]) break;
      const double x = pow10d(-16.80 - (Waveno - Elim + kC1Elev[q]) / 3.00 / kCoolRyd);
      X[q * ld] = pass == 0 ? x / 3.0 : X[q * ld] + x * 2.0 / 3.0;
    }
  }
  // 2s2p3 levels to C II 2s2p2 4P
  Elim = 90820.42 + 43003.3;
  for (int i = 19; i < COOL_NC1; i++) {
    if (Waveno < Elim - kC1Elev[i]) break;
    X[i * ld] = xkarzas(Freq, 4.0 / kCoolRyd * (Elim - kC1Elev[i]), 2, 1) * 3.0;
  }
}

// MG1OP's cross-sections (src/contop.f90:1672-1727): Karzas & Latter for the 3s4l levels,
// power-law fits for the lower ones
__device__ void mg1op_x(double Freq, double Waveno, double *X, int ld) {
  const double Elim = 61671.02;
  for (int i = 0; i < COOL_NMG1; i++) X[i * ld] = 0.0;
  const int lk[5] = {3, 3, 2, 2, 1};
  for (int i = 0; i < 15; i++) {
    if (Waveno < Elim - kMg1Elev[i]) break;
    double x;
    if (i < 5) x = xkarzas(Freq, 16.0 / kCoolRyd * (Elim - kMg1Elev[i]), 4, lk[i]);
    else if (i == 5) x = 25.e-18 * pow(13713.986 / Waveno, 2.7);
    else if (i == 6) x = 33.8e-18 * pow(13823.223 / Waveno, 2.8);
    else if (i == 7) x = 45.e-18 * pow(15267.955 / Waveno, 2.7);
    else if (i == 8) x = 0.43e-18 * pow(18167.687 / Waveno, 2.6);
    else if (i == 9) x = 2.1e-18 * pow(20473.617 / Waveno, 2.6);
    else if (i == 10) x = 16.e-18 * pow(26619.756 / Waveno, 2.1) - 7.8e-18 * pow(26619.756 / Waveno, 9.5);
    else if (i <= 13) x = fmax(20.e-18 * pow(39759.842 / Waveno, 2.7), 40.e-18 * pow(39759.842 / Waveno, 14.0));
    else x = 1.1e-18 * pow((Elim - kMg1Elev[14]) / Waveno, 10.0);
    X[i * ld] = x;
  }
}

// SI1OP's cross-sections (src/contop.f90:1763-1866): Karzas & Latter for the Rydberg and
// 3s3p3 levels, fits to Nahar & Pradhan for the 3s2 3p2 ones
__device__ void si1op_x(double Freq, double Waveno, double *X, int ld) {
  const int nryd[22] = {4, 4, 4, 4, 4, 4, 4, 4, 3, 3, 3, 4, 3, 4, 3, 4, 4, 4, 3, 4, 4, 4};
  const int lryd[22] = {2, 3, 2, 2, 2, 3, 2, 2, 2, 2, 2, 1, 2, 1, 2, 1, 1, 1, 2, 1, 0, 0};
  for (int i = 0; i < COOL_NSI1; i++) X[i * ld] = 0.0;
  double Elim = 65939.18;
  for (int i = 0; i < 22; i++) {
    if (Waveno < Elim - kSi1Elev[i]) break;
    const double zeff2 = (double)(nryd[i] * nryd[i]) / kCoolRyd * (Elim - kSi1Elev[i]);
    X[i * ld] = xkarzas(Freq, zeff2, nryd[i], lryd[i]);
  }
  // 3s2 3p2 levels to Si II 2P1/2 (weight 1/3) and 2P3/2 (weight 2/3)
  for (int pass = 0; pass < 2; pass++) {
    Elim = pass == 0 ? 65747.55 : 65747.55 + 287.45;
    if (Waveno < Elim - kSi1Elev[22]) continue;
    double eps = (Waveno - 70000.0) * 2.0 / 6500.0;
    double reson1 = (97.e-18 * eps + 94.e-18) / (eps * eps + 1.0);
    double x = 37.e-18 * pow(50353.180 / Waveno, 2.40) + reson1;
    X[22 * ld] = pass == 0 ? x / 3.0 : X[22 * ld] + x * 2.0 / 3.0;
    if (Waveno < Elim - kSi1Elev[23]) continue;
    eps = (Waveno - 78600.0) * 2.0 / 13000.0;
    reson1 = (-10.e-18 * eps + 77.e-18) / (eps * eps + 1.0);
    x = 24.5e-18 * pow(59448.700 / Waveno, 1.85) + reson1;
    X[23 * ld] = pass == 0 ? x / 3.0 : X[23 * ld] + x * 2.0 / 3.0;
    for (int q = 24; q <= 26; q++) {  // 3P2, 3P1, 3P0
      if (Waveno < Elim - kSi1Elev[q]) break;
      const double y = Waveno <= 74000.0 ? 72.e-18 * pow(65524.393 / Waveno, 1.90)
                                         : 93.e-18 * pow(65524.393 / Waveno, 4.00);
      X[q * ld] = pass == 0 ? y / 3.0 : X[q * ld] + y * 2.0 / 3.0;
    }
  }
  // 3s3p3 levels to Si II 3s3p2 4P
  Elim = 65747.5 + 42824.35;
  for (int i = 27; i < COOL_NSI1; i++) {
    if (Waveno < Elim - kSi1Elev[i]) break;
    X[i * ld] = xkarzas(Freq, 9.0 / kCoolRyd * (Elim - kSi1Elev[i]), 3, 1) * 3.0;
  }
}

// One lane per wavelength as in k_contop; the lane's 73 cross-sections of C I, Mg I and
// Si I go to scratch xs [73][nwl] (a wavefront's columns stay in L1 while it runs down the
// depths) and the molecular tables are interpolated per depth as the reference
// interpolates them (powers of ten as exp(x ln 10), within 1e-14 of the reference's pow at
// the tables' exponents).  Without LDS, three wavefronts share a SIMD.  A level no lane of the wavefront can ionize adds
// exact zeros to the reference's sums: the depth loop takes only the levels of the
// wavefront's masks (of the 73, and of Fe I's 48, whose cross-sections it forms as it goes).
// Acool is added to contT (divided by ContOpacStd for Motype 1).
__global__ void __launch_bounds__(64, 3) k_coolop(int nwl, int ntau, const double *__restrict__ wave, double wfirst,
                                                  const double *__restrict__ dep, const double *__restrict__ tab,
                                                  const double *__restrict__ cstd, int motype, double *__restrict__ xs,
                                                  double *__restrict__ contT) {
  const int lane = threadIdx.x;
  const int p = blockIdx.x * blockDim.x + lane;
  if (p >= nwl) return;
  auto row = [dep, ntau](int r) { return dep + (size_t)r * ntau; };
  const double clight = kVc * 1e13;
  const double Freq = clight / air2vac(wave[p] + wfirst);
  if (Freq > 3.28805e15) return;  // no cool-star opacity above the H I Lyman edge
  const double Waveno = Freq / kClightCm;
  const double Freq3 = 2.815e29 / Freq / Freq / Freq;
  double *X = xs + p;  // X[i * nwl]
  c1op_x(Freq, Waveno, X, nwl);
  mg1op_x(Freq, Waveno, X + (size_t)COOL_NC1 * nwl, nwl);
  si1op_x(Freq, Waveno, X + (size_t)(COOL_NC1 + COOL_NMG1) * nwl, nwl);
  // the levels some lane of the wavefront can ionize (bit i of mx: X level i; of mfe: Fe I
  // level i); the lanes that returned above take no part
  uint64_t mx[2] = {0, 0}, mfe = 0;
  for (int i = 0; i < COOL_NX; i++)
    if (__builtin_amdgcn_ballot_w64(X[(size_t)i * nwl] != 0.0)) mx[i >> 6] |= 1ull << (i & 63);
  for (int i = 0; i < COOL_NFE1; i++)
    if (__builtin_amdgcn_ballot_w64(!(Waveno < 21000.0) && kFe1Wno[i] < Waveno)) mfe |= 1ull << i;
  bool blue[3];
  for (int n = 0; n < 3; n++) blue[n] = kPsElim[n] - kPsEdge[n] >= kPsElim[n] - Waveno;
  // AL1OP (src/contop.f90:1747-1758): the 3p 2P3/2 and 2P1/2 components
  double xal1 = 0.0;
  {
    const double Elim = 48278.37;
    if (!(Waveno < Elim - 112.061)) {
      xal1 = 6.5e-17 * pow((Elim - 112.061) / Waveno, 5.0) * 4.0;
      if (!(Waveno < Elim)) xal1 = xal1 + 6.5e-17 * pow(Elim / Waveno, 5.0) * 2.0;
    }
  }
  // FE1OP (src/contop.f90:1879-1904): nothing below 21000 cm^-1
  const bool fe1 = !(Waveno < 21000.0);
  // CHOP and OHOP (src/contop.f90:1909-2099, 2104-2331): the energy interval of the tables
  const double Evolt = Waveno / 8065.479;
  const int nch = (int)(Evolt * 10.0);
  const bool ch = !(nch < 20 || nch >= 105);
  const double ench = (double)nch * 0.1;
  const int noh = (int)(Evolt * 10.0 - 20.0);
  const bool oh = !(noh <= 0 || noh >= 130);
  const double enoh = (double)noh * 0.1 + 2.0;
  // H2COLLOP (src/contop.f90:2341-2372): the wavenumber interval
  const bool cia = !(Waveno < 50.0 || Waveno > 50.0 + 50.0 * (double)(400 - 1));
  const int nu = cia ? min((int)((Waveno - 50.0) / 50.0) + 1, 400 - 1) : 1;
  const double delnu = (Waveno - 50.0) / 50.0 - (double)(nu - 1);
  const bool h2h = Waveno <= 10000.0;
  // H2MINOP (src/contop.f90:2379-2396): the interval of the Bell table, ends extended
  const double *h2mwlog = tab + CONT_TAB_H2M;
  const double wlog = log(clight / Freq);
  int ih = 1;
  while (ih < 17 && wlog > h2mwlog[ih]) ih++;
  const double dwh = (wlog - h2mwlog[ih - 1]) / (h2mwlog[ih] - h2mwlog[ih - 1]);
  // NHOP (src/contop.f90:2403-2418): the wavelength interval of the NH table
  const double wl = clight / Freq;
  const bool nh = !(wl < 910.0 || wl > 910.0 + 2.0 * (double)(2146 - 1));
  const int inh = nh ? min((int)((wl - 910.0) / 2.0) + 1, 2146 - 1) : 1;
  const double dwn = (wl - 910.0) / 2.0 - (double)(inh - 1);

  const double *T = row(CR_T), *Hkt = row(CR_HKT), *Hckt = row(CR_HCKT), *Tk = row(CR_TK);
  const double *Xne = row(CR_XNE), *Rho = row(CR_RHO);
  for (int j = 0; j < ntau; j++) {
    const double stim = 1.0 - exp(-Freq * Hkt[j]);
    const double hckt = Hckt[j], rho = Rho[j], t = T[j];
    // C1OP, MG1OP, SI1OP: the pseudo-continuum, then the explicit levels
    double aatom[3];
    const int base[3] = {0, COOL_NC1, COOL_NC1 + COOL_NMG1}, nlev[3] = {COOL_NC1, COOL_NMG1, COOL_NSI1};
    const int brow[3] = {CR_BOLTC1, CR_BOLTMG1, CR_BOLTSI1}, frow[3] = {CR_FC1, CR_FMG1, CR_FSI1};
    const double gps[3] = {6.0, 2.0, 6.0};
    for (int s = 0; s < 3; s++) {
      double e1 = row(CR_PSEXP1 + s)[j];
      if (!blue[s]) e1 = exp(-(kPsElim[s] - Waveno) * hckt);
      double h = Freq3 * gps[s] / (kCoolRyd * hckt) * (e1 - row(CR_PSEXP0 + s)[j]);
      for (int i = 0; i < nlev[s]; i++) {
        const int b = base[s] + i;
        if ((mx[b >> 6] >> (b & 63)) & 1ull) h = h + X[(size_t)b * nwl] * row(brow[s] + i)[j];
      }
      aatom[s] = h * row(frow[s])[j] * stim / rho;
    }
    // CHOP, OHOP
    double ach = 0.0, aoh = 0.0;
    if ((ch || oh) && !(t >= 9000.0)) {
      const int ip = max((int)((t - 1000.0) / 200.0 + 1.0), 1);
      const double tp = (double)ip * 200.0 + 800.0;
      const int it = max((int)((t - 2000.0) / 500.0 + 1.0), 1);
      const double tn = (double)it * 500.0 + 1500.0;
      if (ch) {
        const double part = kPartCH[ip - 1] + (kPartCH[ip] - kPartCH[ip - 1]) * (t - tp) / 200.0;
        const double *c0 = kCrossCH + (size_t)(nch - 1) * 15, *c1 = c0 + 15;
        const double x0 = c0[it - 1] + (c1[it - 1] - c0[it - 1]) * (Evolt - ench) / 0.1;
        const double x1 = c0[it] + (c1[it] - c0[it]) * (Evolt - ench) / 0.1;
        const double chop0 = exp((x0 + (x1 - x0) * (t - tn) / 500.0) * 2.30258509299405) * part;
        ach = chop0 * row(CR_FCH)[j] * stim / rho;
      }
      if (oh) {
        const double part = kPartOH[ip - 1] + (kPartOH[ip] - kPartOH[ip - 1]) * (t - tp) / 200.0;
        const double *c0 = kCrossOH + (size_t)(noh - 1) * 15, *c1 = c0 + 15;
        const double x0 = c0[it - 1] + (c1[it - 1] - c0[it - 1]) * (Evolt - enoh) / 0.1;
        const double x1 = c0[it] + (c1[it] - c0[it]) * (Evolt - enoh) / 0.1;
        const double ohop0 = exp((x0 + (x1 - x0) * (t - tn) / 500.0) * 2.30258509299405) * part;
        aoh = ohop0 * row(CR_FOH)[j] * stim / rho;
      }
    }
    // H2COLLOP: the tables interpolated in wavenumber at the two temperatures around T
    double ah2coll = 0.0;
    if (cia) {
      const int it = max(1, min(13 - 1, (int)((t - 1000.0) / 500.0) + 1));
      const double delt = fmax(0.0, fmin(1.0, (t - 1000.0) / 500.0 - (double)(it - 1)));
      auto at = [nu, delnu](const double *tb, int i) {
        return tb[(size_t)(nu - 1) * 13 + (i - 1)] * (1.0 - delnu) + tb[(size_t)nu * 13 + (i - 1)] * delnu;
      };
      const double xh2h2 = at(kCiaH2h2, it) * (1.0 - delt) + at(kCiaH2h2, it + 1) * delt;
      const double xh2he = at(kCiaH2he, it) * (1.0 - delt) + at(kCiaH2he, it + 1) * delt;
      const double xnh2 = row(CR_FHMOL)[j];
      ah2coll = exp10d(xh2h2) * xnh2 + exp10d(xh2he) * row(CR_FRHE1)[j];
      if (h2h) {
        const double xh2h = at(kCiaH2h, it) * (1.0 - delt) + at(kCiaH2h, it + 1) * delt;
        ah2coll = ah2coll + exp10d(xh2h) * row(CR_FRH1)[j];
      }
      ah2coll = ah2coll * xnh2 / rho;
    }
    // H2MINOP
    const double ah2min = exp(row(CR_H2MLK + ih - 1)[j] * (1.0 - dwh) + row(CR_H2MLK + ih)[j] * dwh) * Xne[j] *
                          Tk[j] * row(CR_FHMOL)[j] / rho;
    // NHOP: the cross-sections at this depth's temperature at the two table wavelengths
    double anh = 0.0;
    const double nhnrho = row(CR_NHNRHO)[j];
    if (nh && nhnrho != 0.0) {
      const int n = (int)row(CR_NHN)[j];
      const double w = row(CR_NHW)[j];
      const double *l0 = kNhLog + (size_t)(inh - 1) * 15, *l1 = l0 + 15;
      const double s0 = exp10d(l0[n - 1] + (l0[n] - l0[n - 1]) * w);
      const double s1 = exp10d(l1[n - 1] + (l1[n] - l1[n - 1]) * w);
      anh = (s0 + (s1 - s0) * dwn) * nhnrho * stim;
    }
    // FE1OP: SUM(Xsect * BoltFe1(:, j)) over the levels of the mask
    double afe1 = 0.0;
    for (uint64_t m = mfe; m; m &= m - 1) {
      const int i = __builtin_ctzll(m);
      const double x = (kFe1Wno[i] + 3000.0 - Waveno) / kFe1Wno[i] / 0.1, x2 = x * x;
      const double xsect = (fe1 && kFe1Wno[i] < Waveno) ? 3.e-18 / (1.0 + x2 * x2) : 0.0;
      afe1 = afe1 + xsect * row(CR_BOLTFE1 + i)[j];
    }
    double acool = aatom[0] + aatom[1] + aatom[2] + ach + aoh + ah2coll + ah2min + anh +
                   (xal1 * row(CR_FAL1)[j] + afe1 * row(CR_FFE1)[j]) * stim / rho;
    if (motype == 1) acool = acool / cstd[j];
    const size_t o = (size_t)j * nwl + p;
    contT[o] = contT[o] + acool;
  }
}

void launch_coolop(int nwl, int ntau, const double *wave, double wfirst, const double *dep, const double *tab,
                   const double *cstd, int motype, double *xs, double *contT, hipStream_t st) {
  k_coolop<<<(nwl + 63) / 64, 64, 0, st>>>(nwl, ntau, wave, wfirst, dep, tab, cstd, motype, xs, contT);
}

// ContopInit's factors of COOLOP's members (src/contop.f90:337-420) and the populations
// COOLOP reads, into the rows of dep, in ContopInit's order of operations with the host's
// libm.  T, Hkt, Rho from the rows already filled; pops: FractionPF of C I, Mg I, Al I,
// Si I, Fe I, CH and OH, then Fraction of NH, H I and He I.
void coolop_rows(int ntau, const float *const pops[10], double *dep) {
  const size_t N = (size_t)ntau;
  auto col = [dep, N](int r) { return dep + (size_t)r * N; };
  const double *T = col(CR_T), *Hkt = col(CR_HKT), *Rho = col(CR_RHO);
  for (size_t j = 0; j < N; j++) {
    const double hckt = kClightCm * Hkt[j];
    col(CR_HCKT)[j] = hckt;
    col(CR_TK)[j] = kKbolt * T[j];
    for (int i = 0; i < COOL_NC1; i++) col(CR_BOLTC1 + i)[j] = kC1Glev[i] * std::exp(-kC1Elev[i] * hckt);
    for (int i = 0; i < COOL_NMG1; i++) col(CR_BOLTMG1 + i)[j] = kMg1Glev[i] * std::exp(-kMg1Elev[i] * hckt);
    for (int i = 0; i < COOL_NSI1; i++) col(CR_BOLTSI1 + i)[j] = kSi1Glev[i] * std::exp(-kSi1Elev[i] * hckt);
    for (int i = 0; i < COOL_NFE1; i++) col(CR_BOLTFE1 + i)[j] = kFe1G[i] * std::exp(-kFe1E[i] * kClightCm * Hkt[j]);
    for (int n = 0; n < 3; n++) {
      col(CR_PSEXP0 + n)[j] = std::exp(-kPsElim[n] * hckt);
      col(CR_PSEXP1 + n)[j] = std::exp(-(kPsElim[n] - kPsEdge[n]) * hckt);
    }
    // ln K of the Bell H2- table at this depth: linear in ln theta, theta clamped to the table
    const double thlog = std::log(std::fmin(std::fmax(5040.4 / T[j], kH2mTheta[0]), kH2mTheta[7]));
    int n = 1;
    while (n < 7 && thlog > std::log(kH2mTheta[n])) n++;
    const double dth = (thlog - std::log(kH2mTheta[n - 1])) / (std::log(kH2mTheta[n]) - std::log(kH2mTheta[n - 1]));
    for (int i = 0; i < CONT_H2M_NW; i++)
      col(CR_H2MLK + i)[j] = std::log(kH2mFf[i * 8 + n - 1] * 1.e-26) * (1.0 - dth) +
                             std::log(kH2mFf[i * 8 + n] * 1.e-26) * dth;
    // the NH table's temperature interval (1-based) and weight in 1/T; none outside it
    col(CR_NHN)[j] = 1.0;
    col(CR_NHW)[j] = 0.0;
    col(CR_NHNRHO)[j] = 0.0;
    if (!(T[j] < kNhT[0] || T[j] > kNhT[14])) {
      int m = 1;
      while (m < 15 - 1 && T[j] > kNhT[m]) m++;
      col(CR_NHN)[j] = (double)m;
      col(CR_NHW)[j] = (1.0 / T[j] - 1.0 / kNhT[m - 1]) / (1.0 / kNhT[m] - 1.0 / kNhT[m - 1]);
      col(CR_NHNRHO)[j] = (double)pops[7][j] / Rho[j];
    }
    const int prow[7] = {CR_FC1, CR_FMG1, CR_FAL1, CR_FSI1, CR_FFE1, CR_FCH, CR_FOH};
    for (int q = 0; q < 7; q++) col(prow[q])[j] = (double)pops[q][j];
    col(CR_FRH1)[j] = (double)pops[8][j];
    col(CR_FRHE1)[j] = (double)pops[9][j];
  }
}

// ------------------------------------------------------------------------------------
// LUKEOP (src/contop.f90:2422-2440): C II, Mg II and Si II bound-free, N I, O I and Ca II by
// Seaton's fits; HOTOP (src/contop.f90:2729-2763): free-free of the ions of C, N, O, Ne, Mg,
// Si, S and Fe and the bound-free edges of C II-IV, N II-V, O II-VI and Ne I-VI.
// ------------------------------------------------------------------------------------
// SEATON (src/contop.f90:494-498)
__device__ __forceinline__ double seaton(double freq, double freq0, double xsect, double power, double a) {
  const double r = freq0 / freq;
  return xsect * (a + (1.0 - a) * r) * sqrt(pow(r, (double)(int)(2.0 * power + 0.01)));
}

// C2OP's cross-sections (src/contop.f90:2447-2484), X[i * ld] for the 34 levels
__device__ void c2op_x(double Freq, double Waveno, double *X, size_t ld) {
  for (int i = 0; i < LUKE_NC2; i++) X[i * ld] = 0.0;
  double Elim = 196664.7;  // 2s2 nl (n = 5, 4, 3) to C III 2s2 1S
  for (int i = 0; i < 12; i++) {
    if (Waveno < Elim - kC2Elev[i]) break;
    const double zeff2 = (double)(kC2Nlv1[i] * kC2Nlv1[i]) / kCoolRyd * (Elim - kC2Elev[i]);
    X[i * ld] = xkarzas(Freq, zeff2, kC2Nlv1[i], kC2Llv1[i]);
  }
  Elim = 196664.7 + 52367.06;  // 2s2p 3l to C III 2s2p 3P
  for (int i = 13; i < 27; i++) {
    if (Waveno < Elim - kC2Elev[i]) break;
    X[i * ld] = xkarzas(Freq, 9.0 / kCoolRyd * (Elim - kC2Elev[i]), 3, kC2Llv2[i - 13]);
  }
  Elim = 196664.7 + 137425.70;  // 2p3 to C III 2p2 3P
  for (int i = 31; i < 34; i++) {
    if (Waveno < Elim - kC2Elev[i]) break;
    X[i * ld] = xkarzas(Freq, 4.0 / kCoolRyd * (Elim - kC2Elev[i]), 2, 1) * 3.0;
  }
}

// MG2OP's cross-sections (src/contop.f90:2531-2553), X[i * ld] for the 14 levels
__device__ void mg2op_x(double Freq, double Waveno, double *X, size_t ld) {
  const double Elim = 121267.61;
  for (int i = 0; i < 13; i++) {
    X[i * ld] = 0.0;
    if (Waveno < Elim - kMg2Elev[i]) continue;
    const double zeff2 = (double)(kMg2Nlv[i] * kMg2Nlv[i]) / kCoolRyd * (Elim - kMg2Elev[i]);
    X[i * ld] = xkarzas(Freq, zeff2, kMg2Nlv[i], kMg2Llv[i]);
  }
  X[13 * ld] = 0.0;
  if (Waveno >= Elim - kMg2Elev[13]) {
    const double r = (Elim - kMg2Elev[13]) / Waveno;
    X[13 * ld] = 0.14e-18 * (6.700 * pow(r, 4.0) - 5.700 * pow(r, 5.0));
  }
}

// One lane per wavelength: LUKEOP's and HOTOP's sums added into contT, in Xcont's order
// (Aluke, then Ahot).  C II's and Mg II's 48 cross-sections go to scratch xs [48][nwl] with a
// ballot per level as in k_coolop; the Si II table comes from tab; HOTOP takes its edges in
// the reference's order, each one's addition depending on the sum before it.
__global__ void __launch_bounds__(64, 3) k_lukehot(int nwl, int ntau, const double *__restrict__ wave, double wfirst,
                                                   const double *__restrict__ dep, const double *__restrict__ tab,
                                                   const double *__restrict__ cstd, int motype, unsigned mask,
                                                   double *__restrict__ xs, double *__restrict__ contT) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nwl) return;
  auto row = [dep, ntau](int r) { return dep + (size_t)r * ntau; };
  const bool luke = (mask >> 9) & 1, hot = (mask >> 10) & 1;
  const double clight = kVc * 1e13;
  const double Freq = clight / air2vac(wave[p] + wfirst);
  const double Freqlg = log(Freq);
  const double Waveno = Freq / kClightCm;
  constexpr int NL = LUKE_NC2 + LUKE_NMG2;
  double *X = xs + p;  // X[i * nwl]
  uint64_t mx = 0;
  double x853 = 0.0, x1020 = 0.0, x1130 = 0.0, xo1 = 0.0, x1044 = 0.0, x1218 = 0.0, x1420 = 0.0;
  bool blue[3] = {false, false, false}, si2tab = false;
  int iw = 1;
  double wfrac = 0.0, freq3 = 0.0;
  if (luke) {
    c2op_x(Freq, Waveno, X, (size_t)nwl);
    mg2op_x(Freq, Waveno, X + (size_t)LUKE_NC2 * nwl, (size_t)nwl);
    for (int i = 0; i < NL; i++)
      if (__builtin_amdgcn_ballot_w64(X[(size_t)i * nwl] != 0.0)) mx |= 1ull << i;
    for (int n = 0; n < 3; n++) blue[n] = kPsElim[3 + n] - kPsEdge[3 + n] >= kPsElim[3 + n] - Waveno;
    freq3 = 2.815e29 / Freq / Freq / Freq * 16.0;
    // N1OP, O1OP, CA2OP (src/contop.f90:2505-2529, 2708-2727)
    if (Freq >= 3.517915e15) x853 = seaton(Freq, 3.517915e15, 1.142e-17, 2.0, 4.29);
    if (Freq >= 2.941534e15) x1020 = seaton(Freq, 2.941534e15, 4.410e-18, 1.5, 3.85);
    if (Freq >= 2.653317e15) x1130 = seaton(Freq, 2.653317e15, 4.200e-18, 1.5, 4.34);
    if (Freq >= 3.28805e15) xo1 = 9.0 * seaton(Freq, 3.28805e15, 2.94e-18, 1.0, 2.66);
    if (Freq >= 2.870454e15) {
      double x = 2.870454e15 / Freq;
      x = x * x * x;
      x1044 = 1.08e-19 * x;
    }
    if (Freq >= 2.460127e15) x1218 = 1.64e-17 * sqrt(2.460127e15 / Freq);
    if (Freq >= 2.110779e15) x1420 = seaton(Freq, 2.110779e15, 4.13e-18, 3.0, 0.69);
    // SI2OP (src/contop.f90:2572-2596): the table's wavenumber interval
    si2tab = Waveno >= 12192.48;
    if (si2tab) {
      iw = max(min((int)(Waveno * 0.001), CONT_SI2_NW - 1), 1);
      wfrac = (Waveno - (double)iw * 1000.0) / 1000.0;
    }
  }
  const double *si2 = tab + CONT_TAB_SI2;  // [nu][k]
  const double *T = row(CR_T), *Hkt = row(CR_HKT), *Hckt = row(CR_HCKT), *Tlog = row(CR_TLOG);
  const double *Xne = row(CR_XNE), *Rho = row(CR_RHO);
  for (int j = 0; j < ntau; j++) {
    const double ehvkt = exp(-Freq * Hkt[j]);
    const double stim = 1.0 - ehvkt;
    const double rho = Rho[j];
    double add = 0.0;
    if (luke) {
      const double hckt = Hckt[j];
      // C2OP: the 2s2 nl (n >= 6) and 2s2p nl (n >= 4) pseudo-continua, then the levels
      double e1 = row(CR_PSEXP1 + 3)[j];
      if (!blue[0]) e1 = exp(-(kPsElim[3] - Waveno) * hckt);
      double h = freq3 * 1.0 / (kCoolRyd * 4.0 * hckt) * (e1 - row(CR_PSEXP0 + 3)[j]);
      e1 = row(CR_PSEXP1 + 4)[j];
      if (!blue[1]) e1 = exp(-(kPsElim[4] - Waveno) * hckt);
      h = h + freq3 * 9.0 / (kCoolRyd * 4.0 * hckt) * (e1 - row(CR_PSEXP0 + 4)[j]);
      for (int i = 0; i < LUKE_NC2; i++)
        if ((mx >> i) & 1ull) h = h + X[(size_t)i * nwl] * row(CR_BOLTC2 + i)[j];
      const double ac2 = h * row(CR_FC2)[j] * stim / rho;
      // MG2OP: the n >= 8 pseudo-continuum, then the levels
      e1 = row(CR_PSEXP1 + 5)[j];
      if (!blue[2]) e1 = exp(-(kPsElim[5] - Waveno) * hckt);
      h = freq3 / (kCoolRyd * 4.0 * hckt) * (e1 - row(CR_PSEXP0 + 5)[j]);
      for (int i = 0; i < LUKE_NMG2; i++)
        if ((mx >> (LUKE_NC2 + i)) & 1ull) h = h + X[(size_t)(LUKE_NC2 + i) * nwl] * row(CR_BOLTMG2 + i)[j];
      const double amg2 = h * row(CR_FMG2)[j] * stim / rho;
      // SI2OP: the table, or below it the high levels' free-bound limit
      double asi2;
      if (si2tab) {
        const int it = (int)row(CR_SI2IT)[j];
        const double tf = row(CR_SI2TF)[j];
        const double *a = si2 + (size_t)(iw - 1) * CONT_SI2_NT, *b = a + CONT_SI2_NT;
        const double hs = (a[it - 1] * (1.0 - tf) + a[it] * tf) * (1.0 - wfrac) + (b[it - 1] * (1.0 - tf) + b[it] * tf) * wfrac;
        asi2 = exp(hs) * row(CR_FSI2)[j] * stim / rho;
      } else {
        const double hs = freq3 * (1.0 / ehvkt - 1.0) * row(CR_SI2BN)[j];
        asi2 = hs * row(CR_FSI2)[j] * stim / rho;
      }
      const double an1 = x853 * 4.0 + x1020 * row(CR_BOLTN1 + 1)[j] + x1130 * row(CR_BOLTN1)[j];
      const double aca2 = x1044 + x1218 * row(CR_BOLTCA2)[j] + x1420 * row(CR_BOLTCA2 + 1)[j];
      add = ac2 + amg2 + asi2 + (an1 * row(CR_FN1)[j] + xo1 * row(CR_FO1)[j] + aca2 * row(CR_FCA2)[j]) * stim / rho;
    }
    if (hot) {
      const double tlog = Tlog[j];
      double free = 0.0;
      for (int z = 1; z <= 5; z++) {
        const double t = coulff(tlog, Freqlg, z) * (double)(z * z) * row(CR_HOTFF + z - 1)[j];
        free = z == 1 ? t : free + t;
      }
      double ahot = free * 3.6919e8 / Freq / Freq / Freq * Xne[j] / sqrt(T[j]);
      for (int i = 0; i < HOT_NEDGE; i++) {
        const double *a = kHotA + 7 * i;
        if (Freq < a[0]) continue;
        const double r = a[0] / Freq;
        const double xsect = a[1] * (a[2] + r - a[2] * r) * sqrt(pow(r, (double)(int)a[3]));
        const double xx = xsect * row(CR_XNFP + (int)a[6] - 1)[j] * a[4];
        if (xx > ahot / 100.0) ahot = ahot + xx / row(CR_HOTEXP + i)[j];
      }
      ahot = ahot * stim / rho;
      add = add + ahot;
    }
    if (motype == 1) add = add / cstd[j];
    const size_t o = (size_t)j * nwl + p;
    contT[o] = contT[o] + add;
  }
}

void launch_lukehot(int nwl, int ntau, const double *wave, double wfirst, const double *dep, const double *tab,
                    const double *cstd, int motype, unsigned mask, double *xs, double *contT, hipStream_t st) {
  k_lukehot<<<(nwl + 63) / 64, 64, 0, st>>>(nwl, ntau, wave, wfirst, dep, tab, cstd, motype, mask, xs, contT);
}

// ContopInit's factors of LUKEOP's members (src/contop.f90:337-366): C II and Mg II level
// Boltzmann factors, N I and Ca II ones, the pseudo-continuum exponentials of C II and Mg II,
// and the Si II table's temperature interval, weight and long-wavelength factor
void lukeop_rows(int ntau, const float *const pops[6], double *dep) {
  const size_t N = (size_t)ntau;
  auto col = [dep, N](int r) { return dep + (size_t)r * N; };
  const double *Hkt = col(CR_HKT), *Tkev = col(CR_TKEV), *Tlog = col(CR_TLOG);
  for (size_t j = 0; j < N; j++) {
    const double hckt = kClightCm * Hkt[j];
    col(CR_HCKT)[j] = hckt;
    for (int i = 0; i < LUKE_NC2; i++) col(CR_BOLTC2 + i)[j] = kC2Glev[i] * std::exp(-kC2Elev[i] * hckt);
    for (int i = 0; i < LUKE_NMG2; i++) col(CR_BOLTMG2 + i)[j] = kMg2Glev[i] * std::exp(-kMg2Elev[i] * hckt);
    col(CR_BOLTN1)[j] = 6.0 * std::exp(-3.575 / Tkev[j]);
    col(CR_BOLTN1 + 1)[j] = 10.0 * std::exp(-2.384 / Tkev[j]);
    col(CR_BOLTCA2)[j] = 10.0 * std::exp(-1.697 / Tkev[j]);
    col(CR_BOLTCA2 + 1)[j] = 6.0 * std::exp(-3.142 / Tkev[j]);
    for (int n = 3; n < 6; n++) {
      col(CR_PSEXP0 + n)[j] = std::exp(-kPsElim[n] * hckt);
      col(CR_PSEXP1 + n)[j] = std::exp(-(kPsElim[n] - kPsEdge[n]) * hckt);
    }
    const double tlog10 = Tlog[j] / 2.30258509299405;
    const int it = std::max(std::min((int)((tlog10 - 3.48) / 0.02), 50), 1);
    col(CR_SI2IT)[j] = (double)it;
    col(CR_SI2TF)[j] = (tlog10 - 3.48 - (double)it * 0.02) / 0.02;
    col(CR_SI2BN)[j] = (1.0 * std::exp(-131838.4 * hckt) + 9.0 * std::exp(-184563.09 * hckt)) / (109732.298 * 4.0 * hckt);
    const int prow[6] = {CR_FC2, CR_FMG2, CR_FSI2, CR_FN1, CR_FO1, CR_FCA2};
    for (int q = 0; q < 6; q++) col(prow[q])[j] = (double)pops[q][j];
  }
}

// ContopInit's factors of HOTOP (src/contop.f90:462-490): exp(E / kT) of the edges, the
// free-free sums of the ions of charge 1-5 (HOTOP's sums over the elements, in its order),
// FractionPF of the edges' species
void hotop_rows(int ntau, const float *frac, const float *fpf, double *dep) {
  const size_t N = (size_t)ntau;
  auto col = [dep, N](int r) { return dep + (size_t)r * N; };
  const double *Tkev = col(CR_TKEV);
  for (size_t j = 0; j < N; j++) {
    for (int i = 0; i < HOT_NEDGE; i++) col(CR_HOTEXP + i)[j] = std::exp(kHotA[7 * i + 5] / Tkev[j]);
    const float *f = frac + j * 48;  // [element][stage]
    for (int st = 2; st <= 6; st++) {
      const int nel = st == 6 ? 7 : 8;  // no Fe VI
      double sm = (double)f[st - 1];
      for (int k = 1; k < nel; k++) sm = sm + (double)f[k * 6 + st - 1];
      col(CR_HOTFF + st - 2)[j] = sm;
    }
    for (int q = 0; q < HOT_NFPF; q++) col(CR_XNFP + q)[j] = (double)fpf[j * HOT_NFPF + q];
  }
}

// SI2OPINIT (src/contop.f90:2606-2706) on the host, as the reference builds it: ln of the
// Si II opacity at 200 wavenumbers (1000 cm^-1 steps) and 51 temperatures (log T 3.50-4.50),
// [nu][k]
static void si2_table(double *out) {
  const double ryd = 109732.298;
  double zeff2[44], hck[51], b3s2[51], b3s3p[51], bolt[51][44];
  for (int i = 0; i < 44; i++) zeff2[i] = (double)(kSi2Nlev[i] * kSi2Nlev[i]) / ryd * kSi2Tlev[i];
  for (int k = 1; k <= 51; k++) {
    const double ttab = std::pow(10.0, 3.48 + (double)k * 0.02);
    hck[k - 1] = 6.6256e-27 * 2.99792458e10 / 1.38054e-16 / ttab;
    b3s2[k - 1] = std::exp(-kSi2Elim[0] * hck[k - 1]);
    b3s3p[k - 1] = std::exp(-kSi2Elim[11] * hck[k - 1]);
    for (int i = 0; i < 44; i++) bolt[k - 1][i] = kSi2Glev[i] * std::exp(-kSi2Elev[i] * hck[k - 1]);
  }
  for (int nu = 1; nu <= CONT_SI2_NW; nu++) {
    const double wno = (double)nu * 1000.0, fr = wno * 2.99792458e10;
    const double f3 = 2.815e29 / fr / fr / fr * 16.0;
    double x[44] = {0.0};
    for (int i = 0; i < 11; i++)
      if (!(wno < kSi2Tlev[i])) x[i] = xkarzas(fr, zeff2[i], kSi2Nlev[i], kSi2Llev[i]);
    for (int i = 11; i < 37; i++) {
      if (wno < kSi2Tlev[i]) break;
      x[i] = xkarzas(fr, zeff2[i], kSi2Nlev[i], kSi2Llev[i]);
    }
    for (int i = 37; i < 41; i++) {
      if (wno < kSi2Tlev[i]) break;
      x[i] = xkarzas(fr, zeff2[i], kSi2Nlev[i], kSi2Llev[i]) * 2.0;
    }
    for (int i = 41; i < 44; i++) {
      if (wno < kSi2Tlev[i]) break;
      x[i] = xkarzas(fr, zeff2[i], kSi2Nlev[i], kSi2Llev[i]) * 3.0;
    }
    for (int k = 0; k < 51; k++) {
      double h = f3 * kSi2Glev[44] / (ryd * 4.0 * hck[k]) *
                 (std::exp(-std::fmax(kSi2Elev[44], kSi2Elim[44] - wno) * hck[k]) - b3s2[k]);
      h = h + f3 * kSi2Glev[45] / (ryd * 4.0 * hck[k]) *
                  (std::exp(-std::fmax(kSi2Elev[45], kSi2Elim[45] - wno) * hck[k]) - b3s3p[k]);
      for (int i = 0; i < 44; i++) h = h + x[i] * bolt[k][i];
      out[(size_t)(nu - 1) * CONT_SI2_NT + k] = std::log(std::fmax(h, 1.e-300));
    }
  }
}

// ------------------------------------------------------------------------------------
// host side: HMINOP's tables (src/contop.f90:183-208 and 1133-1162; Bell & Berrington 1987
// free-free absorption, 1e-26 cm^4/dyn per H atom and unit electron pressure, at 22
// wavelengths [um] x 11 temperatures theta = 5040/T; Wishart 1979 bound-free cross-section,
// 1e-18 cm^2, at 85 wavelengths [nm]) and the logarithms ContopInit takes of them (host
// libm, as the reference's host computes them)
// ------------------------------------------------------------------------------------
namespace {
const double kHmWavek[22] = {
    0.5, 0.4, 0.35, 0.3, 0.25, 0.2, 0.18, 0.16,
    0.14, 0.12, 0.1, 0.09, 0.08, 0.07, 0.06, 0.05,
    0.04, 0.03, 0.02, 0.01, 0.008, 0.006,
};
const double kHmFf[242] = {
    0.0178, 0.0222, 0.0308, 0.0402, 0.0498, 0.0596, 0.0695, 0.0795, 0.0896, 0.131, 0.172,
    0.0228, 0.028, 0.0388, 0.0499, 0.0614, 0.0732, 0.0851, 0.0972, 0.11, 0.16, 0.211,
    0.0277, 0.0342, 0.0476, 0.0615, 0.076, 0.0908, 0.105, 0.121, 0.136, 0.199, 0.262,
    0.0364, 0.0447, 0.0616, 0.0789, 0.0966, 0.114, 0.132, 0.15, 0.169, 0.243, 0.318,
    0.052, 0.0633, 0.0859, 0.108, 0.131, 0.154, 0.178, 0.201, 0.225, 0.321, 0.418,
    0.0791, 0.0959, 0.129, 0.161, 0.194, 0.227, 0.26, 0.293, 0.327, 0.463, 0.602,
    0.0965, 0.117, 0.157, 0.195, 0.234, 0.272, 0.311, 0.351, 0.39, 0.549, 0.711,
    0.121, 0.146, 0.195, 0.241, 0.288, 0.334, 0.381, 0.428, 0.475, 0.667, 0.861,
    0.154, 0.188, 0.249, 0.309, 0.367, 0.424, 0.482, 0.539, 0.597, 0.83, 1.07,
    0.208, 0.25, 0.332, 0.409, 0.484, 0.557, 0.63, 0.702, 0.774, 1.06, 1.36,
    0.293, 0.354, 0.468, 0.576, 0.677, 0.777, 0.874, 0.969, 1.06, 1.45, 1.83,
    0.358, 0.432, 0.572, 0.702, 0.825, 0.943, 1.06, 1.17, 1.28, 1.73, 2.17,
    0.448, 0.539, 0.711, 0.871, 1.02, 1.16, 1.29, 1.43, 1.57, 2.09, 2.6,
    0.579, 0.699, 0.924, 1.13, 1.33, 1.51, 1.69, 1.86, 2.02, 2.67, 3.31,
    0.781, 0.94, 1.24, 1.52, 1.78, 2.02, 2.26, 2.48, 2.69, 3.52, 4.31,
    1.11, 1.34, 1.77, 2.17, 2.53, 2.87, 3.2, 3.51, 3.8, 4.92, 5.97,
    1.73, 2.08, 2.74, 3.37, 3.9, 4.5, 5.01, 5.5, 5.95, 7.59, 9.06,
    3.04, 3.65, 4.8, 5.86, 6.86, 7.79, 8.67, 9.5, 10.3, 13.2, 15.6,
    6.79, 8.16, 10.7, 13.1, 15.3, 17.4, 19.4, 21.2, 23.0, 29.5, 35.0,
    27.0, 32.4, 42.6, 51.9, 60.7, 68.9, 76.8, 84.2, 91.4, 117.0, 140.0,
    42.3, 50.6, 66.4, 80.8, 94.5, 107.0, 120.0, 131.0, 142.0, 183.0, 219.0,
    75.1, 90.0, 118.0, 144.0, 168.0, 191.0, 212.0, 234.0, 253.0, 325.0, 388.0,
};
const double kHmTheta[11] = {
    0.5, 0.6, 0.8, 1.0, 1.2, 1.4, 1.6, 1.8,
    2.0, 2.8, 3.6,
};
const double kHmBfWave[85] = {
    18.0, 19.6, 21.4, 23.6, 26.4, 29.8, 34.3, 40.4,
    49.1, 62.6, 111.3, 112.1, 112.67, 112.95, 113.05, 113.1,
    113.2, 113.23, 113.5, 114.4, 121.0, 139.0, 164.0, 175.0,
    200.0, 225.0, 250.0, 275.0, 300.0, 325.0, 350.0, 375.0,
    400.0, 425.0, 450.0, 475.0, 500.0, 525.0, 550.0, 575.0,
    600.0, 625.0, 650.0, 675.0, 700.0, 725.0, 750.0, 775.0,
    800.0, 825.0, 850.0, 875.0, 900.0, 925.0, 950.0, 975.0,
    1000.0, 1025.0, 1050.0, 1075.0, 1100.0, 1125.0, 1150.0, 1175.0,
    1200.0, 1225.0, 1250.0, 1275.0, 1300.0, 1325.0, 1350.0, 1375.0,
    1400.0, 1425.0, 1450.0, 1475.0, 1500.0, 1525.0, 1550.0, 1575.0,
    1600.0, 1610.0, 1620.0, 1630.0, 1643.91,
};
const double kHmBfSigma[85] = {
    0.067, 0.088, 0.117, 0.155, 0.206, 0.283, 0.414, 0.703,
    1.24, 2.33, 11.6, 13.9, 24.3, 66.7, 95.0, 56.6,
    20.0, 14.6, 8.5, 7.1, 5.43, 5.91, 7.29, 7.918,
    9.453, 11.08, 12.75, 14.46, 16.19, 17.92, 19.65, 21.35,
    23.02, 24.65, 26.24, 27.77, 29.23, 30.62, 31.94, 33.17,
    34.32, 35.37, 36.32, 37.17, 37.91, 38.54, 39.07, 39.48,
    39.77, 39.95, 40.01, 39.95, 39.77, 39.48, 39.06, 38.53,
    37.89, 37.13, 36.25, 35.28, 34.19, 33.01, 31.72, 30.34,
    28.87, 27.33, 25.71, 24.02, 22.26, 20.46, 18.62, 16.74,
    14.85, 12.95, 11.07, 9.211, 7.407, 5.677, 4.052, 2.575,
    1.302, 0.8697, 0.4974, 0.1989, 0.0,
};
}  // namespace

void contop_tables(std::vector<double> &tab) {
  tab.assign((size_t)CONT_TAB_SI2 + CONT_SI2_NW * CONT_SI2_NT, 0.0);
  for (int i = 0; i < CONT_H2M_NW; i++) tab[CONT_TAB_H2M + i] = std::log(kH2mWave[i]);
  static std::vector<double> si2;  // atmosphere-independent: built once per process
  if (si2.empty()) {
    si2.resize((size_t)CONT_SI2_NW * CONT_SI2_NT);
    si2_table(si2.data());
  }
  std::copy(si2.begin(), si2.end(), tab.begin() + CONT_TAB_SI2);
  double *w = tab.data(), *ff = w + CONT_HM_NW, *th = ff + CONT_HM_NW * CONT_HM_NT;
  double *wb = th + CONT_HM_NT, *sb = wb + CONT_HMBF_N;
  for (int i = 0; i < CONT_HM_NW; i++) w[i] = std::log(91.134 / kHmWavek[i]);
  for (int i = 0; i < CONT_HM_NW * CONT_HM_NT; i++) ff[i] = std::log(kHmFf[i] * 1.e-26);
  for (int i = 0; i < CONT_HM_NT; i++) th[i] = kHmTheta[i];
  for (int i = 0; i < CONT_HMBF_N; i++) {
    wb[i] = kHmBfWave[i];
    sb[i] = kHmBfSigma[i];
  }
}

}  // namespace fsg
