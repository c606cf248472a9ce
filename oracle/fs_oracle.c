/*
 * fs_oracle.c -- FP64 CPU restatement of FastSpec's line-opacity and formal-solution path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library is checked
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, and only as the checker.  The product path (fastspec_amd, libfsgpu.so) never calls
 * into it.
 *
 * Each function restates the reference routine named in its comment, step for step and in
 * the reference's evaluation order, in plain scalar C (one thread).  Array indices are
 * kept 1-based where the reference's control flow depends on them (line and wavelength
 * indices, Ileft/Iright, Lstart/Lstop); storage is C-ordered with the same memory layout
 * as the Fortran arrays (depth fastest).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   Voigt/Faraday-Voigt tables and lookups  scipy.special.wofz, closed forms, oracle/_ref
 *   XlineSetup/XlineDepth, LineOpacity1/2    oracle/_ref (reference opacity.f90 compiled)
 *   AnalyticWindow, WaveIdx, RefineIdx       oracle/_ref; threshold inversion
 *   LineopCalc driver (pass 1, PACK, wide lines, segmentation)
 *                                            block-size invariance; brute-force window sum
 *   LocPrfPP/TauStep/CentDer, StokesPrfPP    oracle/_ref; I=S for constant S;
 *                                            Eddington-Barbier for linear S; B=0 reduction
 *   ZeemanPattern, GaussMu                   oracle/_ref; numpy leggauss; strength sums
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/fsgpu.h"

#define EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------------------------------
 * Constants (src/shared.f90:19-33)
 * ------------------------------------------------------------------------------------- */
static const double kVc = 299792.458;
static const double kHc = 6.62607015e-27;
static const double kKbolt = 1.380649e-16;
static const double kA0 = 5.29177210903e-11;
static const double kM0 = 1.66053906660e-27;
static const double kCe = 4.803204712e-10;
static const double kMe = 9.1093837139e-28;
static const double kMh = 1.673532499e-24;
static const double kHie = 13.598434005;
static const double kKz = 4.668644778e-13;
static const double kDWstart = 0.2, kDWstep = 2.0, kDwide = 5e3; /* src/shared.f90:430 */

static double kPi, kErg2ev, kEv2Ang;
/* opacity kernel constants (src/opacity.f90:34-38) */
static double XlnC, XlnC4Pi, XlnZ, XlnKM0, XlnDop;

static void init_constants(void) {
  kPi = acos(-1.0);
  kErg2ev = 1e-7 / 1.602176634e-19;
  kEv2Ang = 8065.544005e-8;
  XlnC = kVc * 1e13;
  XlnC4Pi = 4.0 * XlnC * kPi;
  XlnZ = sqrt(kPi) * (kCe * kCe) / (kMe * (kVc * kVc)) * 1e-18;
  XlnKM0 = kKbolt * 1e-7 / kM0;
  XlnDop = 2.0 * kKbolt / (kM0 * 1e3) * 1e16;
}

/* ---------------------------------------------------------------------------------------
 * Voigt module (src/voigt.f90:14-208)
 * ------------------------------------------------------------------------------------- */
#define NVV 1001
#define NVA 401
static const double VgtVmax = 8.0, VgtAmax = 8.0;
static const double SqrtPiInv = 0.56418958354775628695;
static float VgtTab[NVA + 1][NVV + 1]; /* VgtTab(iv, ia): iv fastest */
static double VgtANode[NVA + 1], VgtAIDa[NVA + 1];
static double VgtIDv, VgtIDt, VgtA1, VgtIA1;
static int VgtReady = 0;

/* HVoigtExact (src/voigt.f90:158-187): Chiarella-Reichel trapezoidal sum, h=0.5, N=16 */
EXPORT double fso_hvoigt_exact(double A, double V) {
  const double HH = 0.5, Pi = 3.14159265358979323846;
  double av = fabs(V);
  if (A <= 0.0) return exp(-fmin(av * av, 700.0));
  double complex z = av + A * I;
  double complex z2 = z * z;
  double complex s = 1.0 / z;
  for (int n = 1; n <= 16; n++) {
    double nh = n * HH;
    s = s + 2.0 * z * exp(-nh * nh) / (z2 - nh * nh);
  }
  s = (HH / Pi) * I * s;
  if (A < 6.0) {
    double p = 2.0 * Pi * A / HH;
    double th = 2.0 * Pi * (2.0 * av - round(2.0 * av));
    double sh = sin(0.5 * th);
    double redd = -2.0 * exp(0.5 * p) * sinh(0.5 * p) + exp(p) * 2.0 * sh * sh;
    double imdd = exp(p) * sin(th);
    s = s + 2.0 * cexp((A * A - av * av) + (-2.0 * A * av) * I) / (redd + imdd * I);
  }
  return creal(s);
}

/* VoigtInit (src/voigt.f90:42-72) */
EXPORT void fso_voigt_init(void) {
  if (VgtReady) return;
  double dv = VgtVmax / (double)(NVV - 1);
  double dt = sqrt(VgtAmax) / (double)(NVA - 1);
  VgtIDv = 1.0 / dv;
  VgtIDt = 1.0 / dt;
  for (int ia = 1; ia <= NVA + 1; ia++) {
    double x = (double)(ia - 1) * dt;
    VgtANode[ia - 1] = x * x;
  }
  for (int ia = 1; ia <= NVA; ia++) VgtAIDa[ia - 1] = 1.0 / (VgtANode[ia] - VgtANode[ia - 1]);
  VgtAIDa[NVA] = 0.0;
  VgtA1 = VgtANode[1];
  VgtIA1 = VgtAIDa[0];
  for (int ia = 1; ia <= NVA + 1; ia++) {
    double aa = VgtANode[ia - 1];
    for (int iv = 1; iv <= NVV + 1; iv++) {
      double vv = (double)(iv - 1) * dv;
      VgtTab[ia - 1][iv - 1] = (float)fso_hvoigt_exact(aa, vv);
    }
  }
  VgtReady = 1;
}

/* HVoigt (src/voigt.f90:76-110): table box bilinear lookup, continued fractions outside */
EXPORT double fso_hvoigt(double A, double V) {
  double av = fabs(V);
  if (A < VgtAmax && av < VgtVmax) {
    double fv = av * VgtIDv;
    int iv = (int)fv;
    double tv = fv - (double)iv;
    int ia;
    double ua;
    if (A < VgtA1) {
      ia = 1;
      ua = A * VgtIA1;
    } else {
      ia = (int)(sqrt(A) * VgtIDt) + 1;
      ua = (A - VgtANode[ia - 1]) * VgtAIDa[ia - 1];
    }
    const float *c0p = VgtTab[ia - 1], *c1p = VgtTab[ia];
    double c0 = (double)c0p[iv] + tv * ((double)c0p[iv + 1] - (double)c0p[iv]);
    double c1 = (double)c1p[iv] + tv * ((double)c1p[iv + 1] - (double)c1p[iv]);
    return c0 + ua * (c1 - c0);
  }
  double a2 = A * A, v2 = av * av, s = v2 + a2;
  if (s >= 196.0) return SqrtPiInv * A * (s + 0.5) / (s * s - v2 + a2 + 0.25);
  double p = v2 - a2 - 1.5, q = 2.0 * av * A;
  double dr = av * p - A * q, di = av * q + A * p;
  return SqrtPiInv * ((p + 0.5) * di - q * dr) / (dr * dr + di * di);
}

/* Raw table access for tests (VgtTab(iv, ia), 1-based) */
EXPORT float fso_voigt_table(int iv, int ia) {
  fso_voigt_init();
  return VgtTab[ia - 1][iv - 1];
}

/* FaradayVoigt module (src/voigt.f90:197-393) */
static float FvTab[NVA + 1][NVV + 1][2]; /* FvTab(2, iv, ia) */
static double FvANode[NVA + 1], FvAIDa[NVA + 1];
static double FvIDv, FvIDt, FvA1, FvIA1;
static int FvReady = 0;

/* WVoigtExact (src/voigt.f90:356-393) */
EXPORT void fso_wvoigt_exact(double A, double V, double *re, double *im) {
  const double HH = 0.5, Pi = 3.14159265358979323846;
  double av = fabs(V);
  if (A <= 0.0) {
    int m = 2 * (int)lround(2.0 * av);
    double x0 = av - 0.25 * (double)m;
    double p = 0.0;
    for (int n = -31; n <= 31; n += 2) {
      double d = x0 - 0.25 * (double)n;
      p = p + exp(-d * d) / (double)(m + n);
    }
    *re = exp(-fmin(av * av, 700.0));
    *im = (2.0 / Pi) * p;
    if (V < 0.0) *im = -*im;
    return;
  }
  double complex z = av + A * I;
  double complex z2 = z * z;
  double complex s = 1.0 / z;
  for (int n = 1; n <= 16; n++) {
    double nh = n * HH;
    s = s + 2.0 * z * exp(-nh * nh) / (z2 - nh * nh);
  }
  s = (HH / Pi) * I * s;
  if (A < 6.0) {
    double p = 2.0 * Pi * A / HH;
    double th = 2.0 * Pi * (2.0 * av - round(2.0 * av));
    double sh = sin(0.5 * th);
    double redd = -2.0 * exp(0.5 * p) * sinh(0.5 * p) + exp(p) * 2.0 * sh * sh;
    double imdd = exp(p) * sin(th);
    s = s + 2.0 * cexp((A * A - av * av) + (-2.0 * A * av) * I) / (redd + imdd * I);
  }
  if (V < 0.0) s = conj(s);
  *re = creal(s);
  *im = cimag(s);
}

/* FVoigtInit (src/voigt.f90:219-251) */
EXPORT void fso_fvoigt_init(void) {
  if (FvReady) return;
  double dv = VgtVmax / (double)(NVV - 1);
  double dt = sqrt(VgtAmax) / (double)(NVA - 1);
  FvIDv = 1.0 / dv;
  FvIDt = 1.0 / dt;
  for (int ia = 1; ia <= NVA + 1; ia++) {
    double x = (double)(ia - 1) * dt;
    FvANode[ia - 1] = x * x;
  }
  for (int ia = 1; ia <= NVA; ia++) FvAIDa[ia - 1] = 1.0 / (FvANode[ia] - FvANode[ia - 1]);
  FvAIDa[NVA] = 0.0;
  FvA1 = FvANode[1];
  FvIA1 = FvAIDa[0];
  for (int ia = 1; ia <= NVA + 1; ia++)
    for (int iv = 1; iv <= NVV + 1; iv++) {
      double vv = (double)(iv - 1) * dv, re, im;
      fso_wvoigt_exact(FvANode[ia - 1], vv, &re, &im);
      FvTab[ia - 1][iv - 1][0] = (float)re;
      FvTab[ia - 1][iv - 1][1] = (float)im;
    }
  FvReady = 1;
}

/* FvLine a-state (src/voigt.f90:302-312) */
typedef struct {
  double A, UA;
  int IA;
} fv_line;

static fv_line fv_line_init(double A) {
  fv_line L;
  L.A = A;
  if (A < VgtAmax) {
    L.IA = (int)(sqrt(A) * FvIDt) + 1;
    L.UA = (A - FvANode[L.IA - 1]) * FvAIDa[L.IA - 1];
  } else {
    L.IA = 0;
    L.UA = 0.0;
  }
  return L;
}

/* FVoigtLine (src/voigt.f90:315-352) */
static void fvoigt_line(const fv_line *L, double V, double *VGT, double *FVGT) {
  double av = fabs(V);
  if (L->IA > 0 && av < VgtVmax) {
    double fv = av * FvIDv;
    int iv = (int)fv;
    double tv = fv - (double)iv;
    const float(*t0)[2] = FvTab[L->IA - 1];
    const float(*t1)[2] = FvTab[L->IA];
    double h0 = (double)t0[iv][0] + tv * ((double)t0[iv + 1][0] - (double)t0[iv][0]);
    double h1 = (double)t1[iv][0] + tv * ((double)t1[iv + 1][0] - (double)t1[iv][0]);
    double l0 = (double)t0[iv][1] + tv * ((double)t0[iv + 1][1] - (double)t0[iv][1]);
    double l1 = (double)t1[iv][1] + tv * ((double)t1[iv + 1][1] - (double)t1[iv][1]);
    *VGT = h0 + L->UA * (h1 - h0);
    *FVGT = l0 + L->UA * (l1 - l0);
  } else {
    double a2 = L->A * L->A, v2 = av * av, s = v2 + a2;
    if (s >= 196.0) {
      double w = SqrtPiInv / (s * s - v2 + a2 + 0.25);
      *VGT = w * L->A * (s + 0.5);
      *FVGT = w * av * (s - 0.5);
    } else {
      double p = v2 - a2 - 1.5, q = 2.0 * av * L->A;
      double dr = av * p - L->A * q, di = av * q + L->A * p;
      double w = SqrtPiInv / (dr * dr + di * di);
      *VGT = w * ((p + 0.5) * di - q * dr);
      *FVGT = w * ((p + 0.5) * dr + q * di);
    }
  }
  *FVGT = copysign(fabs(*FVGT), V); /* FVGT = SIGN(FVGT, V) */
}

/* FVoigt (src/voigt.f90:256-299): same table, a-state computed per call */
EXPORT void fso_fvoigt(double A, double V, double *VGT, double *FVGT) {
  fso_fvoigt_init();
  fv_line L;
  L.A = A;
  if (A < VgtAmax && fabs(V) < VgtVmax) {
    if (A < FvA1) {
      L.IA = 1;
      L.UA = A * FvIA1;
    } else {
      L.IA = (int)(sqrt(A) * FvIDt) + 1;
      L.UA = (A - FvANode[L.IA - 1]) * FvAIDa[L.IA - 1];
    }
  } else {
    L.IA = 0;
    L.UA = 0.0;
  }
  fvoigt_line(&L, V, VGT, FVGT);
}

/* VoigtLine a-state and HVoigtLine (src/voigt.f90:113-154) */
typedef struct {
  double A, UA;
  int IA;
} vg_line;

static vg_line voigt_line_init(double A) {
  vg_line L;
  L.A = A;
  if (A < VgtAmax) {
    L.IA = (int)(sqrt(A) * VgtIDt) + 1;
    L.UA = (A - VgtANode[L.IA - 1]) * VgtAIDa[L.IA - 1];
  } else {
    L.IA = 0;
    L.UA = 0.0;
  }
  return L;
}

static double hvoigt_line(const vg_line *L, double V) {
  double av = fabs(V);
  if (L->IA > 0 && av < VgtVmax) {
    double fv = av * VgtIDv;
    int iv = (int)fv;
    double tv = fv - (double)iv;
    const float *c0p = VgtTab[L->IA - 1], *c1p = VgtTab[L->IA];
    double c0 = (double)c0p[iv] + tv * ((double)c0p[iv + 1] - (double)c0p[iv]);
    double c1 = (double)c1p[iv] + tv * ((double)c1p[iv + 1] - (double)c1p[iv]);
    return c0 + L->UA * (c1 - c0);
  }
  double a2 = L->A * L->A, v2 = av * av, s = v2 + a2;
  if (s >= 196.0) return SqrtPiInv * L->A * (s + 0.5) / (s * s - v2 + a2 + 0.25);
  double p = v2 - a2 - 1.5, q = 2.0 * av * L->A;
  double dr = av * p - L->A * q, di = av * q + L->A * p;
  return SqrtPiInv * ((p + 0.5) * di - q * dr) / (dr * dr + di * di);
}

EXPORT double fso_hvoigt_line(double A, double V) {
  fso_voigt_init();
  vg_line L = voigt_line_init(A);
  return hvoigt_line(&L, V);
}

/* ---------------------------------------------------------------------------------------
 * Utilities (src/utils.f90)
 * ------------------------------------------------------------------------------------- */
/* Air2Vac (src/utils.f90:1020-1032) */
EXPORT double fso_air2vac(double LambdaAir) {
  double S = 1e4 / LambdaAir, S2 = S * S;
  double N = 1.0 + 8.336624212083e-5 + 2.408926869968e-2 / (1.301065924522e2 - S2) +
             1.599740894897e-4 / (3.892568793293e1 - S2);
  return (LambdaAir > 1999.352) ? LambdaAir * N : LambdaAir;
}

/* LegendreP (src/utils.f90:1416-1431) */
static void legendre_p(int N, double Z, double *Pn, double *Pn1, double *Pn2) {
  double p1 = 0.0, p = 1.0, p2 = 0.0;
  for (int j = 1; j <= N; j++) {
    p2 = p1;
    p1 = p;
    p = ((2.0 * j - 1.0) * Z * p1 - (j - 1.0) * p2) / j;
  }
  *Pn = p;
  *Pn1 = p1;
  *Pn2 = p2;
}

/* GaussMu (src/utils.f90:1358-1413); X descending on [0,1] */
EXPORT void fso_gauss_mu(int quad, int N, double *X, double *W) {
  if (kPi == 0.0) init_constants();
  double Pn, Pn1, Pn2, Z, Dz, Df, F;
  if (quad == 1) {
    for (int i = 1; i <= (N + 1) / 2; i++) {
      Z = cos(kPi * (i - 0.25) / (N + 0.5));
      for (int it = 1; it <= 100; it++) {
        legendre_p(N, Z, &Pn, &Pn1, &Pn2);
        Df = N * (Z * Pn - Pn1) / (Z * Z - 1.0);
        Dz = Pn / Df;
        Z = Z - Dz;
        if (fabs(Dz) < 1e-15) break;
      }
      legendre_p(N, Z, &Pn, &Pn1, &Pn2);
      Df = N * (Z * Pn - Pn1) / (Z * Z - 1.0);
      X[i - 1] = 0.5 * (1.0 + Z);
      X[N - i] = 0.5 * (1.0 - Z);
      W[i - 1] = 1.0 / ((1.0 - Z * Z) * Df * Df);
      W[N - i] = W[i - 1];
    }
  } else {
    X[0] = 1.0;
    W[0] = 1.0 / (double)(N * N);
    for (int i = 1; i <= N - 1; i++) {
      Z = cos(2.0 * kPi * i / (2.0 * N - 1.0));
      for (int it = 1; it <= 100; it++) {
        legendre_p(N, Z, &Pn, &Pn1, &Pn2);
        F = Pn1 - Pn;
        Df = ((N - 1) * (Z * Pn1 - Pn2) - N * (Z * Pn - Pn1)) / (Z * Z - 1.0);
        Dz = F / Df;
        Z = Z - Dz;
        if (fabs(Dz) < 1e-15) break;
      }
      legendre_p(N, Z, &Pn, &Pn1, &Pn2);
      X[i] = 0.5 * (1.0 + Z);
      W[i] = 0.5 * (1.0 + Z) / ((double)(N * N) * Pn1 * Pn1);
    }
  }
}

/* MuQuadrature (src/utils.f90:1273-1333): flux weights W, nodes X optionally set, and the
 * projected radii R[N+1] of the annulus boundaries (R may be NULL) */
EXPORT void fso_mu_quadrature_r(int quad, int N, int set_nodes, double *X, double *W, double *Rout) {
  if (N == 1) {
    if (set_nodes) X[0] = 1.0;
    W[0] = 1.0;
    if (Rout) {
      Rout[0] = 0.0;
      Rout[1] = 1.0;
    }
    return;
  }
  if (quad == 0) {
    if (set_nodes)
      for (int i = 1; i <= N; i++) X[i - 1] = sqrt(fmax(0.0, 1.0 - ((double)i - 0.5) / (double)N));
    double *Rmu = malloc(sizeof(double) * N), *R = malloc(sizeof(double) * (N + 1));
    for (int i = 0; i < N; i++) Rmu[i] = sqrt(1.0 - X[i] * X[i]);
    R[0] = 0.0;
    for (int i = 2; i <= N; i++) R[i - 1] = sqrt(0.5 * (Rmu[i - 2] * Rmu[i - 2] + Rmu[i - 1] * Rmu[i - 1]));
    R[N] = 1.0;
    for (int i = 1; i <= N; i++) W[i - 1] = R[i] * R[i] - R[i - 1] * R[i - 1];
    if (Rout) memcpy(Rout, R, sizeof(double) * (N + 1));
    free(Rmu);
    free(R);
    return;
  }
  double *Xg = malloc(sizeof(double) * N), *Wg = malloc(sizeof(double) * N);
  fso_gauss_mu(quad, N, Xg, Wg);
  for (int i = 0; i < N; i++) {
    if (set_nodes) X[i] = Xg[i];
    W[i] = 2.0 * Wg[i] * Xg[i];
  }
  if (Rout) {  /* each annulus encloses the projected area of its weight */
    Rout[0] = 0.0;
    for (int i = 1; i <= N - 1; i++) Rout[i] = sqrt(Rout[i - 1] * Rout[i - 1] + W[i - 1]);
    Rout[N] = 1.0;
  }
  free(Xg);
  free(Wg);
}

EXPORT void fso_mu_quadrature(int quad, int N, int set_nodes, double *X, double *W) {
  fso_mu_quadrature_r(quad, N, set_nodes, X, W, NULL);
}

/* MergeShifts (src/utils.f90:976-1004) */
static void merge_shifts(double *G, double *R, int *N, double Tol) {
  double *Gm = malloc(sizeof(double) * (*N + 1)), *Rm = malloc(sizeof(double) * (*N + 1));
  int Nm = 0;
  for (int i = 0; i < *N; i++) {
    int found = 0;
    for (int k = 0; k < Nm; k++)
      if (fabs(G[i] - Gm[k]) < Tol) {
        Rm[k] = Rm[k] + R[i];
        found = 1;
        break;
      }
    if (!found) {
      Gm[Nm] = G[i];
      Rm[Nm] = R[i];
      Nm++;
    }
  }
  for (int k = 0; k < Nm; k++) {
    G[k] = Gm[k];
    R[k] = Rm[k];
  }
  *N = Nm;
  free(Gm);
  free(Rm);
}

/* ZeemanPattern (src/utils.f90:901-972).  Output arrays must hold 2*Jup+1 entries. */
EXPORT void fso_zeeman_pattern(float Jlo, float Jup, float Glo, float Gup, int *Npi, float *Spi,
                               float *Rpi, int *Nsig, float *Ssig, float *Rsig) {
  const double TolZ = 1e-5;
  double Jl = Jlo, Ju = Jup, Gl = Glo, Gu = Gup;
  int Nm = (int)lround(2.0 * Ju) + 1;
  double *gp = malloc(sizeof(double) * Nm), *rp = malloc(sizeof(double) * Nm);
  double *gs = malloc(sizeof(double) * Nm), *rs = malloc(sizeof(double) * Nm);
  for (int i = 1; i <= Nm; i++) {
    double m = Ju - (double)(i - 1);
    gp[i - 1] = m * (Gu - Gl);
    gs[i - 1] = m * Gu - (m + 1.0) * Gl;
    if (Jl == Ju) {
      rp[i - 1] = m * m;
      rs[i - 1] = 0.25 * (Ju - m) * (Ju + m + 1.0);
    } else if (Jl > Ju) {
      rp[i - 1] = (Ju + 1.0) * (Ju + 1.0) - m * m;
      rs[i - 1] = 0.25 * (Ju + m + 1.0) * (Ju + m + 2.0);
    } else {
      rp[i - 1] = Ju * Ju - m * m;
      rs[i - 1] = 0.25 * (Ju - m) * (Ju - m - 1.0);
    }
  }
  for (int i = 0; i < Nm; i++) {
    if (fabs(gp[i]) < TolZ) gp[i] = 0.0;
    if (fabs(gs[i]) < TolZ) gs[i] = 0.0;
  }
  int Np = 0;
  for (int i = 0; i < Nm; i++)
    if (fabs(rp[i]) > TolZ) {
      gp[Np] = gp[i];
      rp[Np] = rp[i];
      Np++;
    }
  int Ns = 0;
  for (int i = 0; i < Nm; i++)
    if (fabs(rs[i]) > TolZ) {
      gs[Ns] = -gs[i];
      rs[Ns] = rs[i];
      Ns++;
    }
  merge_shifts(gp, rp, &Np, TolZ);
  merge_shifts(gs, rs, &Ns, TolZ);
  double s = 0.0;
  for (int i = 0; i < Np; i++) s += rp[i];
  if (s > 0.0)
    for (int i = 0; i < Np; i++) rp[i] = rp[i] / s;
  s = 0.0;
  for (int i = 0; i < Ns; i++) s += rs[i];
  if (s > 0.0)
    for (int i = 0; i < Ns; i++) rs[i] = rs[i] / s;
  *Npi = Np;
  *Nsig = Ns;
  for (int i = 0; i < Np; i++) {
    Spi[i] = (float)gp[i];
    Rpi[i] = (float)rp[i];
  }
  for (int i = 0; i < Ns; i++) {
    Ssig[i] = (float)gs[i];
    Rsig[i] = (float)rs[i];
  }
  free(gp);
  free(rp);
  free(gs);
  free(rs);
}

/* ---------------------------------------------------------------------------------------
 * Context: copies of the host-side module globals
 * ------------------------------------------------------------------------------------- */
typedef struct fso_ctx {
  int ntau, nwl, nmu, wmode, motype, pf_ld, nslot, nspc, he1_slot, nhe, molbrd;
  double wfirst, dstep, rstep, wstart;
  double *wave;                     /* [nwl+1], 1-based */
  double *rhox, *temp, *xne, *rho;  /* [ntau] */
  double *contop;                   /* [nwl][ntau] */
  double *contop_std;               /* [ntau] */
  float *fraction_pf;               /* [ntau][pf_ld] */
  float *pot_ion, *xamass;          /* [nslot] */
  float *frac_h1, *frac_he1, *frac_hmol;
  int *spc_ion, *spc_slot;          /* [nspc] */
  double *he_wl, *he_elo, *he_gam, *he_shf;
  double he_wl_tol, he_elo_tol;
  double eps_opac, vwfudge, h2brd, vwmol, vmicro;
  double *xmu, *wmu;
  /* XlineInit depth constants (src/opacity.f90:29-31) */
  double *XlnNprtb, *XlnGvwprt, *XlnTemp6, *XlnTemp3, *XlnKbT, *XlnKbeT;
  double XlnVturb2;
  /* accumulated opacities and spectra */
  double *lineop;                   /* [nwl][ntau] LineOpac */
  double *maglineop;                /* [nwl][ntau][6] MagLineOpac */
  double *spl, *spc, *spz;          /* [nwl][nmu], [nwl][nmu], [nwl][4][nmu] */
  int magnetic;
  /* spherical ray geometry (SphereInit): per mu, points 1..ntau_mu along the ray */
  double radius;
  int *ntau_mu, *graz_mu;           /* [nmu] */
  int *map_mu;                      /* [nmu][2 ntau] layer index (1-based) of each point */
  double *rhox_mu, *muloc;          /* [nmu][2 ntau] */
} fso_ctx;

static char g_err[512];
EXPORT const char *fso_last_error(void) { return g_err; }

static double *dup_d(const double *p, size_t n) {
  double *q = malloc(sizeof(double) * (n ? n : 1));
  if (p) memcpy(q, p, sizeof(double) * n);
  else memset(q, 0, sizeof(double) * (n ? n : 1));
  return q;
}
static float *dup_f(const float *p, size_t n) {
  float *q = malloc(sizeof(float) * (n ? n : 1));
  if (p) memcpy(q, p, sizeof(float) * n);
  else memset(q, 0, sizeof(float) * (n ? n : 1));
  return q;
}
static int *dup_i(const int32_t *p, size_t n) {
  int *q = malloc(sizeof(int) * (n ? n : 1));
  for (size_t i = 0; i < n; i++) q[i] = p ? p[i] : 0;
  return q;
}

#define WAVE(i) (c->wave[(i)])                          /* 1-based Wave(i)            */
#define CONT(it, iw) (c->contop[(size_t)((iw)-1) * c->ntau + ((it)-1)]) /* ContOpac(it,iw) */
#define LOP(it, iw) (c->lineop[(size_t)((iw)-1) * c->ntau + ((it)-1)])  /* LineOpac(it,iw) */
#define FPF(slot, it) (c->fraction_pf[(size_t)((it)-1) * c->pf_ld + ((slot)-1)])

/* XlineInit (src/opacity.f90:65-93) */
static void xline_init(fso_ctx *c) {
  int N = c->ntau;
  c->XlnNprtb = malloc(sizeof(double) * N);
  c->XlnGvwprt = malloc(sizeof(double) * N);
  c->XlnTemp6 = malloc(sizeof(double) * N);
  c->XlnTemp3 = malloc(sizeof(double) * N);
  c->XlnKbT = malloc(sizeof(double) * N);
  c->XlnKbeT = malloc(sizeof(double) * N);
  c->XlnVturb2 = 1e26 / XlnC / XlnC * (c->vmicro * c->vmicro);
  for (int k = 0; k < N; k++) {
    double Temp = c->temp[k], Xne = c->xne[k];
    double Vh = sqrt(2.0 * kKbolt * Temp / kMh);
    c->XlnNprtb[k] = (double)c->frac_h1[k] + 0.412265 * (double)c->frac_he1[k];
    if (c->h2brd > 0.0) c->XlnNprtb[k] = c->XlnNprtb[k] + c->h2brd * 0.874307 * (double)c->frac_hmol[k];
    c->XlnGvwprt[k] = 17.0 * pow(Vh, 0.6) * c->XlnNprtb[k];
    c->XlnTemp6[k] = pow(Temp / 1e4, 1.0 / 6.0) * Xne;
    c->XlnTemp3[k] = pow(Temp / 1e4, 0.3) * c->XlnNprtb[k];
    c->XlnKbT[k] = kKbolt * Temp;
    c->XlnKbeT[k] = kKbolt * kErg2ev * Temp;
  }
}

/* SphereInit (src/solver.f90:45-120): rays with impact parameter p = R_top sqrt(1 - mu^2);
 * the column-mass steps are scaled by (r_k + r_k-1) / (z_k + z_k-1), z = sqrt(r^2 - p^2);
 * grazing rays (p above the deepest layer) cross the tangent shell and revisit the crossed
 * layers in reverse order */
static int sphere_init(fso_ctx *c, const double *height) {
  const int N = c->ntau, M = c->nmu, L = 2 * N;
  double *Rk = malloc(sizeof(double) * N);
  for (int k = 0; k < N; k++) Rk[k] = c->radius + height[k];
  c->ntau_mu = calloc(M, sizeof(int));
  c->graz_mu = calloc(M, sizeof(int));
  c->map_mu = calloc((size_t)M * L, sizeof(int));
  c->rhox_mu = calloc((size_t)M * L, sizeof(double));
  c->muloc = calloc((size_t)M * L, sizeof(double));
  int rc = 0;
  for (int im = 0; im < M && !rc; im++) {
    double xmu = c->xmu[im];
    if (xmu <= 0.0) {
      snprintf(g_err, sizeof(g_err), "SphereInit: mu = 0 rays are unsupported for spherical models");
      rc = 1;
      break;
    }
    double P = Rk[0] * sqrt((1.0 - xmu) * (1.0 + xmu));
    int graz = P > Rk[N - 1];
    int n = N;
    if (graz) {
      int kt;
      for (kt = 2; kt <= N; kt++)
        if (P >= Rk[kt - 1]) break;
      n = kt - 1;
      if (n < 2) {
        snprintf(g_err, sizeof(g_err),
                 "SphereInit: grazing ray tangent in the top model layer; remove extreme mu values close to 0");
        rc = 1;
        break;
      }
    }
    c->graz_mu[im] = graz;
    int *map = c->map_mu + (size_t)im * L;
    double *rm = c->rhox_mu + (size_t)im * L, *ml = c->muloc + (size_t)im * L;
    rm[0] = c->rhox[0] / xmu;
    double Zk1 = sqrt((Rk[0] - P) * (Rk[0] + P));
    ml[0] = Zk1 / Rk[0];
    map[0] = 1;
    for (int k = 2; k <= n; k++) {
      double Zk = sqrt((Rk[k - 1] - P) * (Rk[k - 1] + P));
      rm[k - 1] = rm[k - 2] + (c->rhox[k - 1] - c->rhox[k - 2]) * (Rk[k - 1] + Rk[k - 2]) / (Zk + Zk1);
      ml[k - 1] = Zk / Rk[k - 1];
      map[k - 1] = k;
      Zk1 = Zk;
    }
    if (!graz) {
      c->ntau_mu[im] = N;
      continue;
    }
    double Chord = 2.0 * sqrt((Rk[n - 1] - P) * (Rk[n - 1] + P));
    rm[n] = rm[n - 1] + Chord * (c->rhox[n] - c->rhox[n - 1]) / (Rk[n - 1] - Rk[n]);
    map[n] = n;
    ml[n] = -ml[n - 1];
    for (int k = n + 2; k <= 2 * n; k++) {
      rm[k - 1] = rm[k - 2] + (rm[2 * n - k + 1] - rm[2 * n - k]);
      map[k - 1] = 2 * n + 1 - k;
      ml[k - 1] = -ml[2 * n - k];
    }
    c->ntau_mu[im] = 2 * n;
  }
  free(Rk);
  return rc;
}

/* ray geometry for tests: ntau_mu/graz [nmu], map/rhox_mu/muloc [nmu][2 ntau] */
EXPORT int fso_sphere_geometry(fso_ctx *c, int32_t *ntau_mu, int32_t *graz, int32_t *map, double *rhox_mu,
                               double *muloc) {
  if (c->radius <= 0.0) return 1;
  size_t L = (size_t)2 * c->ntau;
  for (int im = 0; im < c->nmu; im++) {
    ntau_mu[im] = c->ntau_mu[im];
    graz[im] = c->graz_mu[im];
  }
  for (size_t i = 0; i < (size_t)c->nmu * L; i++) {
    map[i] = c->map_mu[i];
    rhox_mu[i] = c->rhox_mu[i];
    muloc[i] = c->muloc[i];
  }
  return 0;
}

EXPORT int fso_finalize(fso_ctx *c);

EXPORT int fso_init(const fsg_setup *s, fso_ctx **out) {
  init_constants();
  fso_voigt_init();
  fso_ctx *c = calloc(1, sizeof(fso_ctx));
  c->ntau = s->ntau;
  c->nwl = s->nwl;
  c->nmu = s->nmu;
  c->wmode = s->wmode;
  c->motype = s->motype;
  c->pf_ld = s->pf_ld;
  c->nslot = s->nslot;
  c->nspc = s->nspc;
  c->he1_slot = s->he1_slot;
  c->nhe = s->nhe;
  c->molbrd = s->molbrd;
  c->wfirst = s->wfirst;
  c->dstep = s->dstep;
  c->rstep = s->rstep;
  c->wstart = s->wstart;
  c->wave = malloc(sizeof(double) * (s->nwl + 1));
  c->wave[0] = 0.0;
  memcpy(c->wave + 1, s->wave, sizeof(double) * s->nwl);
  c->rhox = dup_d(s->rhox, s->ntau);
  c->temp = dup_d(s->temp, s->ntau);
  c->xne = dup_d(s->xne, s->ntau);
  c->rho = dup_d(s->rho, s->ntau);
  c->contop = dup_d(s->contop, (size_t)s->nwl * s->ntau);
  c->contop_std = dup_d(s->contop_std, s->ntau);
  c->fraction_pf = dup_f(s->fraction_pf, (size_t)s->ntau * s->pf_ld);
  c->pot_ion = dup_f(s->pot_ion, s->nslot);
  c->xamass = dup_f(s->xamass, s->nslot);
  c->frac_h1 = dup_f(s->frac_h1, s->ntau);
  c->frac_he1 = dup_f(s->frac_he1, s->ntau);
  c->frac_hmol = dup_f(s->frac_hmol, s->ntau);
  c->spc_ion = dup_i(s->spc_ion, s->nspc);
  c->spc_slot = dup_i(s->spc_slot, s->nspc);
  c->he_wl = dup_d(s->he_wl, s->nhe);
  c->he_elo = dup_d(s->he_elo, s->nhe);
  c->he_gam = dup_d(s->he_gam, (size_t)s->nhe * s->ntau);
  c->he_shf = dup_d(s->he_shf, (size_t)s->nhe * s->ntau);
  c->he_wl_tol = s->he_wl_tol;
  c->he_elo_tol = s->he_elo_tol;
  c->eps_opac = s->eps_opac;
  c->vwfudge = s->vwfudge;
  c->h2brd = s->h2brd;
  c->vwmol = s->vwmol;
  c->vmicro = s->vmicro;
  c->xmu = dup_d(s->xmu, s->nmu);
  c->wmu = dup_d(s->wmu, s->nmu);
  c->lineop = calloc((size_t)s->nwl * s->ntau, sizeof(double));
  xline_init(c);
  c->radius = s->radius;
  if (c->radius > 0.0 && s->nmu > 0) {
    if (!s->height) {
      snprintf(g_err, sizeof(g_err), "fso_init: a spherical model needs layer heights");
      fso_finalize(c);
      return 1;
    }
    if (sphere_init(c, s->height)) {
      fso_finalize(c);
      return 1;
    }
  }
  *out = c;
  return 0;
}

EXPORT int fso_finalize(fso_ctx *c) {
  if (!c) return 0;
  void *ptrs[] = {c->wave,     c->rhox,      c->temp,      c->xne,       c->rho,       c->contop,
                  c->contop_std, c->fraction_pf, c->pot_ion, c->xamass,  c->frac_h1,   c->frac_he1,
                  c->frac_hmol, c->spc_ion,   c->spc_slot,  c->he_wl,     c->he_elo,    c->he_gam,
                  c->he_shf,    c->xmu,       c->wmu,       c->XlnNprtb,  c->XlnGvwprt, c->XlnTemp6,
                  c->XlnTemp3,  c->XlnKbT,    c->XlnKbeT,   c->lineop,    c->maglineop, c->spl,
                  c->spc,       c->spz,       c->ntau_mu,   c->graz_mu,   c->map_mu,    c->rhox_mu,
                  c->muloc};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  free(c);
  return 0;
}

EXPORT int fso_get_lineop(fso_ctx *c, double *out) {
  memcpy(out, c->lineop, sizeof(double) * (size_t)c->nwl * c->ntau);
  return 0;
}
EXPORT int fso_set_lineop(fso_ctx *c, const double *in) {
  memcpy(c->lineop, in, sizeof(double) * (size_t)c->nwl * c->ntau);
  return 0;
}
EXPORT int fso_get_maglineop(fso_ctx *c, double *out) {
  if (!c->maglineop) memset(out, 0, sizeof(double) * (size_t)c->nwl * c->ntau * 6);
  else memcpy(out, c->maglineop, sizeof(double) * (size_t)c->nwl * c->ntau * 6);
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * Line opacity kernel (src/opacity.f90)
 * ------------------------------------------------------------------------------------- */
typedef struct {
  double WaveL, Hnux, Elo, Gf, Grad, Xmass, Cqst, Qdiv, Cvdw, AboPre, AboVbc, AboExp;
  int IxPF, MdQst, MdVdw, IxHe;
} xline_coef;

/* per-list state shared by the driver routines (LineData module arrays) */
typedef struct {
  int list_type;
  int64_t nlines;
  const fsg_lines *L;
  int64_t *map;     /* packed index -> input index (0-based) */
  int *LinFlag;     /* 1-based */
  double *Wlcent;   /* 1-based */
  int *Ileft, *Iright;
  float *spl_pi, *spl_si; /* scaled Zeeman shifts (magnetic lists), input order */
} list_state;

/* HeMatchIdx (src/hestark.f90:376-389) */
static int he_match_idx(const fso_ctx *c, double WaveL, double Elo) {
  for (int k = 1; k <= c->nhe; k++)
    if (fabs(WaveL - c->he_wl[k - 1]) <= c->he_wl_tol && fabs(Elo - c->he_elo[k - 1]) <= c->he_elo_tol)
      return k;
  return 0;
}

/* XlineSetup (src/opacity.f90:96-235) for line iline (1-based, current order) */
static void xline_setup(const fso_ctx *c, const list_state *ls, int64_t iline, xline_coef *XC) {
  int64_t j = ls->map[iline - 1];
  const fsg_lines *L = ls->L;
  double WaveL = ls->Wlcent[iline] + c->wfirst;
  double Elo = (double)L->elo[j];
  XC->WaveL = WaveL;
  XC->Elo = Elo;
  XC->Gf = (double)L->gf[j];
  double GradL = (double)L->grad[j];
  int ispc = L->spc[j];
  int IonL = c->spc_ion[ispc - 1];
  XC->IxPF = c->spc_slot[ispc - 1];
  double GqstL, GvdwL;
  int IsMol;
  if (ls->list_type == FSG_LIST_ATOMIC) {
    GqstL = (double)L->gqst[j];
    GvdwL = (double)L->gvdw[j];
    IsMol = 0;
  } else if (ls->list_type == FSG_LIST_MOLECULAR) {
    GqstL = 0.0;
    GvdwL = 0.0;
    IsMol = 1;
  } else {
    GqstL = (double)L->gqst[j];
    GvdwL = (double)L->gvdw[j];
    IsMol = (ls->LinFlag[iline] >> 1) & 1;
  }
  XC->Hnux = XlnC * kHc / WaveL;
  XC->Xmass = XlnDop / XlnC / XlnC / (double)c->xamass[XC->IxPF - 1];
  XC->Grad = (GradL > 0.0) ? GradL : 0.22233e16 / (WaveL * WaveL);
  XC->Cqst = 0.0;
  XC->Qdiv = 1.0;
  XC->Cvdw = 0.0;
  XC->AboPre = 0.0;
  XC->AboVbc = 0.0;
  XC->AboExp = 0.0;
  XC->IxHe = 0;
  if (IsMol && !c->molbrd) {
    XC->MdQst = 9;
    XC->MdVdw = 9;
  } else {
    double Wavevac = fso_air2vac(WaveL);
    double Eup = Elo + 1.0 / (Wavevac * kEv2Ang);
    double pot = (double)c->pot_ion[XC->IxPF - 1];
    double t1 = (double)(IonL * IonL) * kHie / (pot - Eup);
    double t2 = (double)(IonL * IonL) * kHie / (pot - Elo);
    double Enu4 = t1 * t1, Enl4 = t2 * t2;
    if (XC->IxPF == c->he1_slot) XC->IxHe = he_match_idx(c, WaveL, Elo);
    if (XC->IxHe > 0) {
      XC->MdQst = 3;
    } else if (GqstL > 0.0) {
      XC->MdQst = 0;
      XC->Cqst = GqstL;
    } else if (IsMol) {
      XC->MdQst = 9;
    } else if (IonL == 1) {
      XC->MdQst = 1;
      XC->Cqst = 2.26e-07 * Enu4;
    } else {
      XC->MdQst = 2;
      XC->Cqst = 5.42e-07 * Enu4;
      XC->Qdiv = (IonL + 1.0) * (IonL + 1.0);
    }
    if ((ls->LinFlag[iline] >> 2) & 1) {
      XC->MdVdw = 2;
      double Sigma = kA0 * kA0 * (double)(int)GvdwL;
      double Alpha = GvdwL - (double)(int)GvdwL;
      double X = 2.0 - Alpha * 0.5;
      double Gx = X - 1.0;
      double GammaF = 1.0 + (-0.5748646 + (0.9512363 + (-0.6998588 + (0.4245549 - 0.1010678 * Gx) * Gx) * Gx) * Gx) * Gx;
      XC->AboPre = 2e6 * (pow(4.0 / kPi, Alpha * 0.5) * GammaF * 1e4 * Sigma);
      XC->AboVbc = 1.0 / 1.00794 + 1.0 / (double)c->xamass[XC->IxPF - 1];
      XC->AboExp = 1.0 - Alpha;
    } else if (GvdwL > 0.0) {
      XC->MdVdw = 0;
      XC->Cvdw = GvdwL;
    } else if (IsMol) {
      XC->MdVdw = 9;
    } else {
      XC->MdVdw = 1;
      double Cw = 1.61e-33 * fabs(Enu4 - Enl4) / (double)(IonL * IonL);
      XC->Cvdw = pow(Cw, 0.4);
    }
  }
  if (IsMol && c->vwmol > 0.0) {
    XC->MdVdw = 0;
    XC->Cvdw = c->vwmol;
  }
}

/* XlineDepth (src/opacity.f90:239-293); itau 1-based */
static void xline_depth(const fso_ctx *c, int itau, const xline_coef *XC, double *LinCopL,
                        double *LinVvgtL, double *LinAvgtL) {
  int k = itau - 1;
  double T = c->temp[k];
  double Xstim = 1.0 - exp(-XC->Hnux / c->XlnKbT[k]);
  double Efract = exp(-XC->Elo / c->XlnKbeT[k]);
  double Dopl = sqrt(T * XC->Xmass + c->XlnVturb2);
  double Dldopl = XC->WaveL * Dopl;
  double Dndopl = Dopl / XC->WaveL;
  double Gqst, Gvdw;
  switch (XC->MdQst) {
    case 0: Gqst = XC->Cqst * c->XlnTemp6[k]; break;
    case 1: Gqst = XC->Cqst * c->xne[k]; break;
    case 2: Gqst = XC->Cqst * c->xne[k] / XC->Qdiv; break;
    case 3: Gqst = c->he_gam[(size_t)(XC->IxHe - 1) * c->ntau + k]; break;
    default: Gqst = 0.0;
  }
  switch (XC->MdVdw) {
    case 0: Gvdw = XC->Cvdw * c->XlnTemp3[k]; break;
    case 1: Gvdw = c->vwfudge * c->XlnGvwprt[k] * XC->Cvdw; break;
    case 2: {
      double Vbar = sqrt(8.0 * XlnKM0 * T / kPi * XC->AboVbc);
      Gvdw = XC->AboPre * pow(Vbar / 1e4, XC->AboExp) * c->XlnNprtb[k];
      break;
    }
    default: Gvdw = 0.0;
  }
  double Gamtot = XC->Grad + Gqst + Gvdw;
  *LinAvgtL = Gamtot / (Dndopl * XlnC4Pi);
  *LinVvgtL = 1.0 / Dldopl;
  double cop = XlnZ * Efract * XC->Gf * Xstim / (c->rho[k] * Dndopl);
  *LinCopL = cop * (double)FPF(XC->IxPF, itau);
}

#define HESHF(itau, ihe) (c->he_shf[(size_t)((ihe)-1) * c->ntau + ((itau)-1)])

/* LineOpacity2 (src/opacity.f90:350-386) */
static void line_opacity2(const fso_ctx *c, const list_state *ls, double Wl, int64_t iline, int iwl,
                          const double *Cop, const double *Vv, const double *Av, double *Linop) {
  int N = c->ntau;
  for (int k = 0; k < N; k++) Linop[k] = 0.0;
  int flag = ls->LinFlag[iline];
  if (!(flag & 1) && iwl >= ls->Ileft[iline] && iwl <= ls->Iright[iline]) {
    int ihe = (flag >> 8) & 0xff;
    double dWl0 = Wl - ls->Wlcent[iline];
    for (int k = 0; k < N; k++) {
      double dWl = dWl0;
      if (ihe > 0) dWl = dWl0 - HESHF(k + 1, ihe);
      double V = dWl * Vv[k];
      Linop[k] = fso_hvoigt(Av[k], V) * Cop[k];
    }
  }
  if (c->motype == 1)
    for (int k = 0; k < N; k++) Linop[k] = Linop[k] / c->contop_std[k];
}

/* WaveIdx (src/opacity.f90:594-608) */
static int wave_idx(const fso_ctx *c, double Wavein) {
  if (c->wmode == 0) {
    double X = fmin(fmax(Wavein / c->dstep, 0.0), (double)(c->nwl - 2));
    return (int)X + 1;
  }
  double X = 1.0 + log(fmax((Wavein / c->wstart) + 1.0, 2.2250738585072014e-308)) / c->rstep;
  return (int)fmin(fmax(X, 1.0), (double)(c->nwl - 1));
}

EXPORT int fso_wave_idx(fso_ctx *c, double w) { return wave_idx(c, w); }

/* AnalyticWindow (src/opacity.f90:678-710) */
static double analytic_window(const fso_ctx *c, const list_state *ls, int64_t iline, const double *Cop,
                              const double *Vv, const double *Av, const double *ContOpacL) {
  const double iSqrtPi = 0.5641895835477563;
  int ihe = (ls->LinFlag[iline] >> 8) & 0xff;
  double dWwin = 0.0;
  for (int k = 0; k < c->ntau; k++) {
    double r = Cop[k] / (c->eps_opac * ContOpacL[k]);
    if (c->motype == 1) r = r / c->contop_std[k];
    if (r <= 1.0) continue;
    double w = sqrt(log(r) + fmax(0.0, Av[k] * (r * iSqrtPi - Av[k]))) / Vv[k];
    if (ihe > 0) w = w + fabs(HESHF(k + 1, ihe));
    dWwin = fmax(dWwin, w);
  }
  return dWwin;
}

/* Planck source function as in RadTranCalc/LinStrCalc (src/synth.f90:1792-1794,1854-1856) */
static void planck(const fso_ctx *c, double Wl, double *Src) {
  const double PlkC1 = kHc * kVc * 1e13 / kKbolt;
  double C1 = PlkC1 / Wl;
  double C2 = exp(50.7649141 - 5.0 * log(Wl));
  for (int k = 0; k < c->ntau; k++) Src[k] = C2 / (exp(C1 / c->temp[k]) - 1.0);
}

/* TauStep (src/solver.f90:591-615); Dtau[0..N-2] set */
static void tau_step(const double *Rhox, const double *Xki, double Xmu, int N, double *Dtau) {
  double Dx1 = (Rhox[1] - Rhox[0]) / Xmu;
  double Yp = (Xki[1] - Xki[0]) / Dx1;
  Dtau[0] = Dx1 * (Xki[1] + 2.0 * Xki[0] + Dx1 / 2.0 * Yp) / 3.0;
  for (int k = 2; k <= N - 1; k++) {
    double Dx = (Rhox[k - 1] - Rhox[k - 2]) / Xmu;
    Dx1 = (Rhox[k] - Rhox[k - 1]) / Xmu;
    double Der = (Xki[k - 1] - Xki[k - 2]) / Dx;
    double Der1 = (Xki[k] - Xki[k - 1]) / Dx1;
    if (Der * Der1 > 0.0) {
      double Lambda = (1.0 + Dx1 / (Dx + Dx1)) / 3.0;
      Yp = (Der / (Lambda * Der1 + (1.0 - Lambda) * Der)) * Der1;
    } else {
      Yp = 0.0;
    }
    Dtau[k - 1] = Dx1 * (Xki[k] + 2.0 * Xki[k - 1] + Dx1 / 2.0 * Yp) / 3.0;
  }
  Dtau[N - 1] = 0.0;
}

/* CentDer (src/solver.f90:618-647): Y[k*M + j] (j fastest, like Y(M,N)) */
static void cent_der(const double *Dxx, const double *Y, int M, int N, double *Yp) {
  for (int k = 2; k <= N - 1; k++) {
    double Dx = Dxx[k - 2], Dx1 = Dxx[k - 1];
    double InvDx = 1.0 / Dx, InvDx1 = 1.0 / Dx1;
    double Lambda = (1.0 + Dx1 / (Dx + Dx1)) / 3.0;
    double Lambda1 = 1.0 - Lambda;
    for (int j = 0; j < M; j++) {
      double Der = (Y[(k - 1) * M + j] - Y[(k - 2) * M + j]) * InvDx;
      double Der1 = (Y[k * M + j] - Y[(k - 1) * M + j]) * InvDx1;
      if (Der * Der1 > 0.0) Yp[(k - 1) * M + j] = (Der / (Lambda * Der1 + Lambda1 * Der)) * Der1;
      else Yp[(k - 1) * M + j] = 0.0;
    }
  }
  for (int j = 0; j < M; j++) {
    Yp[j] = (Y[M + j] - Y[j]) / Dxx[0];
    Yp[(N - 1) * M + j] = (Y[(N - 1) * M + j] - Y[(N - 2) * M + j]) / Dxx[N - 2];
  }
}

/* DELO-Bezier step coefficients (src/solver.f90:154-169) */
static void delo_coef(double Dtau, double *Eps, double *Alpha, double *Beta, double *Gamma) {
  double Dtau2 = Dtau * Dtau;
  double Idtau2 = 1.0 / Dtau2;
  if (Dtau < 5e-3) {
    *Eps = 1.0 - Dtau + 0.5 * Dtau2 - Dtau2 * Dtau / 6.0 + Dtau2 * Dtau2 / 24.0;
    *Alpha = Dtau / 3.0 - Dtau2 / 12.0 + Dtau2 * Dtau / 60.0;
    *Beta = Dtau / 3.0 - Dtau2 * 0.25 + Dtau2 * Dtau * 0.1;
    *Gamma = Dtau / 3.0 - Dtau2 / 6.0 + Dtau2 * Dtau * 0.05;
  } else {
    double E = (Dtau < 100.0) ? exp(-Dtau) : 0.0;
    *Eps = E;
    *Alpha = (Dtau2 - 2.0 * Dtau + 2.0 - 2.0 * E) * Idtau2;
    *Beta = (2.0 - (2.0 + 2.0 * Dtau + Dtau2) * E) * Idtau2;
    *Gamma = (2.0 * Dtau - 4.0 + (2.0 * Dtau + 4.0) * E) * Idtau2;
  }
}

/* LocPrfPP (src/solver.f90:123-179) */
EXPORT void fso_locprfpp(double *Ploc, const double *Xmu, int Nmu, const double *Xki, const double *Src,
                         const double *Rhox, int Ntau) {
  double *Srd = malloc(sizeof(double) * Ntau), *Deltau = malloc(sizeof(double) * Ntau);
  double *U = malloc(sizeof(double) * Ntau);
  tau_step(Rhox, Xki, 1.0, Ntau, Deltau);
  cent_der(Deltau, Src, 1, Ntau, Srd);
  for (int k = 1; k <= Ntau - 1; k++)
    U[k - 1] = 0.5 * (Src[k - 1] + 0.5 * Deltau[k - 1] * Srd[k - 1] + Src[k] - 0.5 * Deltau[k - 1] * Srd[k]);
  for (int im = 0; im < Nmu; im++) {
    double Imu = 1.0 / Xmu[im];
    double Y = Src[Ntau - 1];
    for (int k = Ntau - 1; k >= 1; k--) {
      double Dtau = Deltau[k - 1] * Imu, Eps, Alpha, Beta, Gamma;
      delo_coef(Dtau, &Eps, &Alpha, &Beta, &Gamma);
      Y = Y * Eps + Alpha * Src[k - 1] + Beta * Src[k] + Gamma * U[k - 1];
    }
    Ploc[im] = Y;
  }
  free(Srd);
  free(Deltau);
  free(U);
}

/* DeloScalar (src/solver.f90:219-261): DELO-Bezier integration along one ray from the
 * boundary intensity Y0 */
static double delo_scalar(const double *Deltau, const double *Src, int N, double Y0) {
  double *Srd = malloc(sizeof(double) * N);
  cent_der(Deltau, Src, 1, N, Srd);
  double Y = Y0;
  for (int k = N - 1; k >= 1; k--) {
    double Dtau = Deltau[k - 1], Eps, Alpha, Beta, Gamma;
    delo_coef(Dtau, &Eps, &Alpha, &Beta, &Gamma);
    double U = 0.5 * (Src[k - 1] + 0.5 * Dtau * Srd[k - 1] + Src[k] - 0.5 * Dtau * Srd[k]);
    Y = Y * Eps + Alpha * Src[k - 1] + Beta * Src[k] + Gamma * U;
  }
  free(Srd);
  return Y;
}

/* LocPrfSph (src/solver.f90:182-215): per mu, opacity and source gathered along the ray,
 * optical depth steps on the ray's column mass (no division by mu); grazing rays start
 * from zero, the others from S at the deepest point */
EXPORT void fso_locprfsph(fso_ctx *c, double *Ploc, const double *Xki, const double *Src) {
  int L = 2 * c->ntau;
  double *XkiR = malloc(sizeof(double) * L), *SrcR = malloc(sizeof(double) * L);
  double *Deltau = malloc(sizeof(double) * L);
  for (int im = 0; im < c->nmu; im++) {
    int N = c->ntau_mu[im];
    const int *map = c->map_mu + (size_t)im * L;
    for (int k = 0; k < N; k++) {
      XkiR[k] = Xki[map[k] - 1];
      SrcR[k] = Src[map[k] - 1];
    }
    tau_step(c->rhox_mu + (size_t)im * L, XkiR, 1.0, N, Deltau);
    double Y = c->graz_mu[im] ? 0.0 : SrcR[N - 1];
    Ploc[im] = delo_scalar(Deltau, SrcR, N, Y);
  }
  free(XkiR);
  free(SrcR);
  free(Deltau);
}

/* emergent intensities of one opacity column in the model's geometry */
static void locprf(fso_ctx *c, double *Ploc, const double *Xki, const double *Src) {
  if (c->radius > 0.0) fso_locprfsph(c, Ploc, Xki, Src);
  else fso_locprfpp(Ploc, c->xmu, c->nmu, Xki, Src, c->rhox, c->ntau);
}

/* LinStrCalc (src/synth.f90:1774-1822) */
static float lin_str_calc(fso_ctx *c, double Wl, const double *LineOpacL, const double *ContOpacL) {
  int N = c->ntau, M = c->nmu;
  double *Src = malloc(sizeof(double) * N), *X = malloc(sizeof(double) * N);
  double *IntL = malloc(sizeof(double) * M), *IntC = malloc(sizeof(double) * M);
  planck(c, Wl, Src);
  locprf(c, IntC, ContOpacL, Src);
  for (int k = 0; k < N; k++) X[k] = ContOpacL[k] + LineOpacL[k];
  locprf(c, IntL, X, Src);
  float res;
  if (M == 1) {
    res = (float)(1.0 - IntL[0] / IntC[0]);
  } else {
    double FlxL = 0.0, FlxC = 0.0;
    for (int i = 0; i < M; i++) {
      FlxL = FlxL + c->wmu[i] * IntL[i];
      FlxC = FlxC + c->wmu[i] * IntC[i];
    }
    res = (float)(1.0 - FlxL / FlxC);
  }
  free(Src);
  free(X);
  free(IntL);
  free(IntC);
  return res;
}

/* LinStrCalc of one line-centre column (the checker's closed-form tests call it alone) */
EXPORT float fso_lin_str(fso_ctx *c, double Wl, const double *LineOpacL, const double *ContOpacL) {
  return lin_str_calc(c, Wl, LineOpacL, ContOpacL);
}

/* WlcIdx (src/synth.f90:677-706) over the current (packed) Wlcent */
static int64_t wlc_idx(const list_state *ls, double W) {
  if (ls->nlines == 0 || W < ls->Wlcent[1]) return 0;
  int64_t ia = 1, ib = ls->nlines;
  while (ib - ia > 1) {
    int64_t im = (ia + ib) / 2;
    if (ls->Wlcent[im] <= W) ia = im;
    else ib = im;
  }
  return (ls->Wlcent[ib] <= W) ? ib : ia;
}

/* MarkWideLines (src/synth.f90:709-744) */
static int64_t mark_wide_lines(const fso_ctx *c, list_state *ls, int64_t NlinesStep) {
  int64_t Nwide = 0;
  for (int64_t i = 1; i <= ls->nlines; i++) {
    if (ls->LinFlag[i] & 1) continue;
    int64_t Nrec = wlc_idx(ls, WAVE(ls->Iright[i])) - wlc_idx(ls, WAVE(ls->Ileft[i])) + 1;
    double Dvline = kVc * (WAVE(ls->Iright[i]) - WAVE(ls->Ileft[i])) / (ls->Wlcent[i] + c->wfirst);
    if (Nrec > NlinesStep / 2 || Dvline > kDwide) {
      ls->LinFlag[i] |= 8;
      Nwide++;
    }
  }
  return Nwide;
}

/* MagLineOpacity2 (src/opacity.f90:468-520): Stokes I bound used by the window search */
static void mag_line_opacity2(const fso_ctx *c, const list_state *ls, double Wl, int64_t iline, int iwl,
                              const double *Cop, const double *Vv, const double *Av, double *Linop) {
  int N = c->ntau;
  for (int k = 0; k < N; k++) Linop[k] = 0.0;
  int flag = ls->LinFlag[iline];
  if (!(flag & 1) && iwl >= ls->Ileft[iline] && iwl <= ls->Iright[iline]) {
    int ihe = (flag >> 8) & 0xff;
    int64_t j = ls->map[iline - 1];
    const fsg_lines *L = ls->L;
    double dWl0 = Wl - ls->Wlcent[iline];
    for (int k = 0; k < N; k++) {
      double dWl = dWl0;
      if (ihe > 0) dWl = dWl0 - HESHF(k + 1, ihe);
      vg_line VL = voigt_line_init(Av[k]);
      double Pp = 0.0;
      for (int q = L->pi_off[j]; q < L->pi_off[j + 1]; q++)
        Pp = Pp + (double)L->str_pi[q] * hvoigt_line(&VL, (dWl - (double)ls->spl_pi[q]) * Vv[k]);
      double Ps = 0.0;
      for (int q = L->si_off[j]; q < L->si_off[j + 1]; q++)
        Ps = Ps + (double)L->str_si[q] * hvoigt_line(&VL, (dWl - (double)ls->spl_si[q]) * Vv[k]);
      Linop[k] = fmax(0.5 * Ps, 0.5 * Pp + 0.25 * Ps) * Cop[k];
    }
  }
  if (c->motype == 1)
    for (int k = 0; k < N; k++) Linop[k] = Linop[k] / c->contop_std[k];
}

/* RefineIdx (src/opacity.f90:611-675) */
static void refine_idx(const fso_ctx *c, list_state *ls, int64_t iline, int iwl, int idw, const double *Cop,
                       const double *Vv, const double *Av, const double *ContOpacL, double SplMax, int Mag) {
  int N = c->ntau;
  int isw, ikw;
  double *LineOpacL = malloc(sizeof(double) * N);
  /* 0.5 * (Wave(1) + Wave(Nwl)): default REAL 0.5 promoted to REAL(8) exactly */
  if (ls->Wlcent[iline] > 0.5 * (WAVE(1) + WAVE(c->nwl))) {
    isw = -1;
    ikw = 0;
  } else {
    isw = 1;
    ikw = 1;
  }
  double Wl = ls->Wlcent[iline], dW;
  int ia, ib;
  if (idw == 0) {
    ia = wave_idx(c, Wl + isw * SplMax) + ikw;
  } else {
    dW = (kDWstart * pow(kDWstep, idw - 1)) * (ls->Wlcent[iline] + c->wfirst) / 5e3;
    ia = wave_idx(c, Wl + isw * (SplMax + dW)) + ikw;
  }
  dW = (kDWstart * pow(kDWstep, idw)) * (ls->Wlcent[iline] + c->wfirst) / 5e3;
  ib = wave_idx(c, Wl + isw * (SplMax + dW)) + ikw;
  while (abs(ia - ib) > 1) {
    int im = (ia + ib) / 2;
    if (Mag) mag_line_opacity2(c, ls, WAVE(im), iline, iwl, Cop, Vv, Av, LineOpacL);
    else line_opacity2(c, ls, WAVE(im), iline, iwl, Cop, Vv, Av, LineOpacL);
    double MaxOpac = -INFINITY;
    for (int k = 0; k < N; k++) MaxOpac = fmax(MaxOpac, LineOpacL[k] / ContOpacL[k]);
    if (MaxOpac < c->eps_opac) ib = im;
    else ia = im;
  }
  if (isw == 1) {
    ls->Iright[iline] = ia;
    ls->Ileft[iline] = wave_idx(c, 2.0 * Wl - WAVE(ia));
  } else {
    ls->Ileft[iline] = ia;
    ls->Iright[iline] = wave_idx(c, 2.0 * Wl - WAVE(ia)) + 1;
  }
  free(LineOpacL);
}

/* LineOpacity1 (src/opacity.f90:297-346) with block arrays blk*[jline][ntau] */
static void line_opacity1(const fso_ctx *c, const list_state *ls, double Wl, int64_t L1, int64_t L2, int iwl,
                          int64_t Loff, const double *bCop, const double *bVv, const double *bAv, double *Linop) {
  int N = c->ntau;
  for (int k = 0; k < N; k++) Linop[k] = 0.0;
  for (int64_t iline = L1; iline <= L2; iline++) {
    int flag = ls->LinFlag[iline];
    if ((flag & 1) || (flag & 8) || iwl < ls->Ileft[iline] || iwl > ls->Iright[iline]) continue;
    int ihe = (flag >> 8) & 0xff;
    double dWl0 = Wl - ls->Wlcent[iline];
    size_t b = (size_t)(iline - Loff - 1) * N;
    for (int k = 0; k < N; k++) {
      double dWl = dWl0;
      if (ihe > 0) dWl = dWl0 - HESHF(k + 1, ihe);
      double V = dWl * bVv[b + k];
      double A = bAv[b + k];
      double Vgt = fso_hvoigt(A, V) * bCop[b + k];
      Linop[k] = Linop[k] + Vgt;
    }
  }
  if (c->motype == 1)
    for (int k = 0; k < N; k++) Linop[k] = Linop[k] / c->contop_std[k];
}

/* MagLineOpacity1/3 component sums (src/opacity.f90:391-464, 526-591) for one line */
static void mag_line_sum(const fso_ctx *c, const list_state *ls, double Wl, int64_t iline, const double *Cop,
                         const double *Vv, const double *Av, double *M6 /* [ntau][6] */) {
  int N = c->ntau;
  int ihe = (ls->LinFlag[iline] >> 8) & 0xff;
  int64_t j = ls->map[iline - 1];
  const fsg_lines *L = ls->L;
  int np0 = L->pi_off[j], np1 = L->pi_off[j + 1];
  int ns0 = L->si_off[j], ns1 = L->si_off[j + 1];
  int Nsh = (ns1 - ns0) / 2;
  for (int k = 0; k < N; k++) {
    fv_line FL = fv_line_init(Av[k]);
    double Xn = Cop[k], sc = Vv[k];
    double Vbase = (Wl - ls->Wlcent[iline]) * sc;
    if (ihe > 0) Vbase = Vbase - HESHF(k + 1, ihe) * sc;
    double *m = M6 + (size_t)k * 6, Vgt, Fvgt;
    for (int q = np0; q < np1; q++) {
      double V = Vbase - (double)ls->spl_pi[q] * sc;
      fvoigt_line(&FL, V, &Vgt, &Fvgt);
      double Str = Xn * (double)L->str_pi[q];
      m[0] = m[0] + Str * Vgt;
      m[3] = m[3] + Str * Fvgt;
    }
    for (int q = ns0; q < ns0 + Nsh; q++) {
      double V = Vbase - (double)ls->spl_si[q] * sc;
      fvoigt_line(&FL, V, &Vgt, &Fvgt);
      double Str = Xn * (double)L->str_si[q];
      m[1] = m[1] + Str * Vgt;
      m[4] = m[4] + Str * Fvgt;
    }
    for (int q = ns0 + Nsh; q < ns1; q++) {
      double V = Vbase - (double)ls->spl_si[q] * sc;
      fvoigt_line(&FL, V, &Vgt, &Fvgt);
      double Str = Xn * (double)L->str_si[q];
      m[2] = m[2] + Str * Vgt;
      m[5] = m[5] + Str * Fvgt;
    }
  }
}

static void mag_motype_div(const fso_ctx *c, double *M6) {
  if (c->motype != 1) return;
  for (int k = 0; k < c->ntau; k++)
    for (int q = 0; q < 6; q++) M6[(size_t)k * 6 + q] = M6[(size_t)k * 6 + q] / c->contop_std[k];
}

/* ---------------------------------------------------------------------------------------
 * LineopCalc / MagLineopCalc drivers (src/synth.f90:747-1095, 1098-1455)
 * ------------------------------------------------------------------------------------- */
static int lineop_common(fso_ctx *c, int list_type, const fsg_lines *L, int64_t nlines_lim, int magnetic,
                         const double *bfield, int bframe, float *linstr_out, int32_t *flags_out,
                         int32_t *ileft_out, int32_t *iright_out, fsg_lineop_stats *stats, int pass1_only) {
  int N = c->ntau, Nwl = c->nwl;
  int64_t Nlines = L->n;
  fsg_lineop_stats st;
  memset(&st, 0, sizeof(st));
  st.nlines_in = Nlines;
  if (magnetic) {
    fso_fvoigt_init();
    if (!c->maglineop) c->maglineop = calloc((size_t)Nwl * N * 6, sizeof(double));
    c->magnetic = 1;
  }
  list_state ls;
  memset(&ls, 0, sizeof(ls));
  ls.list_type = list_type;
  ls.nlines = Nlines;
  ls.L = L;
  ls.map = malloc(sizeof(int64_t) * (Nlines ? Nlines : 1));
  ls.LinFlag = calloc(Nlines + 2, sizeof(int));
  ls.Wlcent = malloc(sizeof(double) * (Nlines + 2));
  ls.Ileft = malloc(sizeof(int) * (Nlines + 2));
  ls.Iright = malloc(sizeof(int) * (Nlines + 2));
  for (int64_t i = 1; i <= Nlines; i++) {
    ls.map[i - 1] = i - 1;
    ls.Wlcent[i] = L->wl[i - 1] - c->wfirst;
    ls.Ileft[i] = 1;
    ls.Iright[i] = Nwl;
  }
  int64_t NlinesStep = Nlines < nlines_lim ? Nlines : nlines_lim;
  int64_t nstep_alloc = NlinesStep > 0 ? NlinesStep : 1;
  double *bCop = malloc(sizeof(double) * nstep_alloc * N);
  double *bVv = malloc(sizeof(double) * nstep_alloc * N);
  double *bAv = malloc(sizeof(double) * nstep_alloc * N);
  int DoLinStr = (linstr_out != NULL);
  float *LinStr = DoLinStr ? calloc(Nlines + 1, sizeof(float)) : NULL;
  double Bsplit = 0.0;
  if (magnetic) {
    if (bframe == 0)
      Bsplit = kKz * sqrt(bfield[0] * bfield[0] + bfield[1] * bfield[1] + bfield[2] * bfield[2]);
    else Bsplit = kKz * fabs(bfield[0]);
    int npi = L->pi_off[Nlines], nsi = L->si_off[Nlines];
    ls.spl_pi = malloc(sizeof(float) * (npi ? npi : 1));
    ls.spl_si = malloc(sizeof(float) * (nsi ? nsi : 1));
  }
  double *Cop = malloc(sizeof(double) * N), *Vv = malloc(sizeof(double) * N), *Av = malloc(sizeof(double) * N);
  double *LineOpacL = malloc(sizeof(double) * N), *ContOpacL = malloc(sizeof(double) * N);
  xline_coef XC;

  /* pass 1 over lines (src/synth.f90:833-907, 1182-1273) */
  for (int64_t iline = 1; iline <= Nlines; iline++) {
    int64_t j = iline - 1;
    if (list_type == FSG_LIST_ATOMIC) {
      if (L->gvdw[j] > 10.0f) ls.LinFlag[iline] |= 4;
    } else if (list_type == FSG_LIST_MIXED) {
      if (L->is_mol && L->is_mol[j]) ls.LinFlag[iline] |= 2;
      if (L->gvdw[j] > 10.0f) ls.LinFlag[iline] |= 4;
    }
    double SplMax = 0.0;
    if (magnetic) {
      double wl2 = (ls.Wlcent[iline] + c->wfirst) * (ls.Wlcent[iline] + c->wfirst);
      for (int q = L->pi_off[j]; q < L->pi_off[j + 1]; q++) {
        ls.spl_pi[q] = (float)((double)L->spl_pi[q] * Bsplit * wl2);
        SplMax = fmax(SplMax, fabs((double)ls.spl_pi[q]));
      }
      for (int q = L->si_off[j]; q < L->si_off[j + 1]; q++) {
        ls.spl_si[q] = (float)((double)L->spl_si[q] * Bsplit * wl2);
        SplMax = fmax(SplMax, fabs((double)ls.spl_si[q]));
      }
    }
    xline_setup(c, &ls, iline, &XC);
    if (XC.IxHe > 0) ls.LinFlag[iline] |= (XC.IxHe << 8);
    for (int k = 1; k <= N; k++) xline_depth(c, k, &XC, &Cop[k - 1], &Vv[k - 1], &Av[k - 1]);
    if (iline <= NlinesStep) {
      memcpy(bCop + (size_t)(iline - 1) * N, Cop, sizeof(double) * N);
      memcpy(bVv + (size_t)(iline - 1) * N, Vv, sizeof(double) * N);
      memcpy(bAv + (size_t)(iline - 1) * N, Av, sizeof(double) * N);
    }
    int iwl = wave_idx(c, ls.Wlcent[iline]);
    double dW = (ls.Wlcent[iline] - WAVE(iwl)) / (WAVE(iwl + 1) - WAVE(iwl));
    for (int k = 1; k <= N; k++) ContOpacL[k - 1] = CONT(k, iwl) * (1.0 - dW) + CONT(k, iwl + 1) * dW;
    double MaxOpac;
    if (!magnetic) {
      MaxOpac = -INFINITY;
      for (int k = 0; k < N; k++) {
        double r = (c->motype == 1) ? Cop[k] / (ContOpacL[k] * c->contop_std[k]) : Cop[k] / ContOpacL[k];
        MaxOpac = fmax(MaxOpac, r);
      }
      if (MaxOpac < c->eps_opac) {
        ls.LinFlag[iline] |= 1;
        continue;
      }
    }
    line_opacity2(c, &ls, ls.Wlcent[iline], iline, iwl, Cop, Vv, Av, LineOpacL);
    MaxOpac = -INFINITY;
    for (int k = 0; k < N; k++) MaxOpac = fmax(MaxOpac, LineOpacL[k] / ContOpacL[k]);
    if (MaxOpac < c->eps_opac) {
      ls.LinFlag[iline] |= 1;
      continue;
    } else if (DoLinStr) {
      LinStr[iline] = lin_str_calc(c, ls.Wlcent[iline] + c->wfirst, LineOpacL, ContOpacL);
    }
    if (!magnetic) {
      dW = analytic_window(c, &ls, iline, Cop, Vv, Av, ContOpacL);
      ls.Ileft[iline] = wave_idx(c, ls.Wlcent[iline] - dW);
      int ir = wave_idx(c, ls.Wlcent[iline] + dW) + 1;
      ls.Iright[iline] = ir < Nwl ? ir : Nwl;
    } else {
      for (int idw = 0;; idw++) {
        dW = (kDWstart * pow(kDWstep, idw)) * (ls.Wlcent[iline] + c->wfirst) / 5e3;
        double Wl = ls.Wlcent[iline] + SplMax + dW;
        mag_line_opacity2(c, &ls, Wl, iline, iwl, Cop, Vv, Av, LineOpacL);
        MaxOpac = -INFINITY;
        for (int k = 0; k < N; k++) MaxOpac = fmax(MaxOpac, LineOpacL[k] / ContOpacL[k]);
        if (MaxOpac < c->eps_opac) {
          refine_idx(c, &ls, iline, iwl, idw, Cop, Vv, Av, ContOpacL, SplMax, 1);
          break;
        } else if (ls.Wlcent[iline] - SplMax - dW < WAVE(1) && ls.Wlcent[iline] + SplMax + dW > WAVE(Nwl)) {
          break;
        }
      }
    }
  }
  /* outputs in input order, before PACK */
  for (int64_t i = 1; i <= Nlines; i++) {
    if (flags_out) flags_out[i - 1] = ls.LinFlag[i];
    if (ileft_out) ileft_out[i - 1] = ls.Ileft[i];
    if (iright_out) iright_out[i - 1] = ls.Iright[i];
    if (linstr_out) linstr_out[i - 1] = LinStr[i];
  }
  /* PACK (src/synth.f90:915-933, 1280-1293) */
  int do_pack = magnetic ? (NlinesStep == Nlines)
                         : ((list_type == FSG_LIST_ATOMIC || list_type == FSG_LIST_MIXED) && NlinesStep == Nlines);
  if (do_pack) {
    int64_t n = 0;
    for (int64_t i = 1; i <= Nlines; i++) {
      if (ls.LinFlag[i] & 1) continue;
      n++;
      ls.map[n - 1] = ls.map[i - 1];
      ls.LinFlag[n] = ls.LinFlag[i];
      ls.Wlcent[n] = ls.Wlcent[i];
      ls.Ileft[n] = ls.Ileft[i];
      ls.Iright[n] = ls.Iright[i];
      memmove(bCop + (size_t)(n - 1) * N, bCop + (size_t)(i - 1) * N, sizeof(double) * N);
      memmove(bVv + (size_t)(n - 1) * N, bVv + (size_t)(i - 1) * N, sizeof(double) * N);
      memmove(bAv + (size_t)(n - 1) * N, bAv + (size_t)(i - 1) * N, sizeof(double) * N);
    }
    Nlines = n;
    ls.nlines = n;
  }
  if (pass1_only) {
    for (int64_t i = 1; i <= Nlines; i++)
      if (!(ls.LinFlag[i] & 1)) {
        st.nlines_use++;
        st.evals += (int64_t)(ls.Iright[i] - ls.Ileft[i] + 1) * N;
      }
  } else if (Nlines > 0) {
    int64_t Nwide = mark_wide_lines(c, &ls, NlinesStep);
    st.nwide = Nwide;
    int *Lstart = malloc(sizeof(int) * (Nwl + 1)), *Lstop = malloc(sizeof(int) * (Nwl + 1));
    /* Lstart/Lstop hold line indices; Nlines fits int in every supported list */
    for (int i = 1; i <= Nwl; i++) {
      Lstart[i] = (int)Nlines;
      Lstop[i] = 1;
    }
    int64_t NlinesUse = Nlines;
    for (int64_t iline = 1; iline <= Nlines; iline++) {
      if (ls.LinFlag[iline] & 1) {
        NlinesUse--;
        continue;
      }
      st.evals += (int64_t)(ls.Iright[iline] - ls.Ileft[iline] + 1) * N;
      if (ls.LinFlag[iline] & 8) continue;
      for (int i = ls.Ileft[iline]; i <= ls.Iright[iline]; i++) {
        if (iline < Lstart[i]) Lstart[i] = (int)iline;
        if (iline > Lstop[i]) Lstop[i] = (int)iline;
      }
    }
    st.nlines_use = NlinesUse;
    /* wide lines (src/synth.f90:981-1001, 1339-1361) */
    if (Nwide > 0) {
      for (int64_t iline = 1; iline <= Nlines; iline++) {
        if (!(ls.LinFlag[iline] & 8)) continue;
        xline_setup(c, &ls, iline, &XC);
        for (int k = 1; k <= N; k++) xline_depth(c, k, &XC, &Cop[k - 1], &Vv[k - 1], &Av[k - 1]);
        /* one line's window points are independent (the reference's OpenMP loop over
         * the wide line's window, src/synth.f90:990-999) */
#pragma omp parallel
        {
          double *LopW = malloc(sizeof(double) * N);
          double *M6w = malloc(sizeof(double) * 6 * N);
#pragma omp for schedule(static)
          for (int iwl = ls.Ileft[iline]; iwl <= ls.Iright[iline]; iwl++) {
            if (!magnetic) {
              line_opacity2(c, &ls, WAVE(iwl), iline, iwl, Cop, Vv, Av, LopW);
              for (int k = 1; k <= N; k++) LOP(k, iwl) = LOP(k, iwl) + LopW[k - 1];
            } else {
              memset(M6w, 0, sizeof(double) * 6 * N);
              mag_line_sum(c, &ls, WAVE(iwl), iline, Cop, Vv, Av, M6w);
              mag_motype_div(c, M6w);
              double *dst = c->maglineop + (size_t)(iwl - 1) * N * 6;
              for (int q = 0; q < 6 * N; q++) dst[q] = dst[q] + M6w[q];
            }
          }
          free(LopW);
          free(M6w);
        }
      }
    }
    /* segment loop (src/synth.f90:1003-1088, 1366-1447) */
    int64_t iline1 = 1, ilineb = 1, iline2;
    int iwl1 = 1, iwl2;
    for (;;) {
      if (NlinesStep >= Nlines) {
        iwl2 = Nwl;
      } else {
        int64_t ilo = Nlines + 1, ihi = 0;
        iwl2 = iwl1 - 1;
        for (int iwl = iwl1; iwl <= Nwl; iwl++) {
          int64_t ilot = ilo < Lstart[iwl] ? ilo : Lstart[iwl];
          int64_t ihit = ihi > Lstop[iwl] ? ihi : Lstop[iwl];
          if (ihit - ilot + 1 > NlinesStep) break;
          ilo = ilot;
          ihi = ihit;
          iwl2 = iwl;
        }
        if (iwl2 < iwl1) {
          snprintf(g_err, sizeof(g_err),
                   "Line opacity segmentation: lines contributing to wavelength point #%d span more than %lld "
                   "line records. Increase LimitRAM.",
                   iwl1, (long long)NlinesStep);
          return 1;
        }
        iline1 = ilo > 1 ? ilo : 1;
        if (iline1 > Nlines) iline1 = Nlines;
      }
      iline2 = iline1 + NlinesStep - 1 < Nlines ? iline1 + NlinesStep - 1 : Nlines;
      int64_t Loff = iline1 - 1;
      if (iline1 != ilineb) {
        ilineb = iline1;
        st.nblocks++;
        for (int64_t iline = iline1; iline <= iline2; iline++) {
          if (ls.LinFlag[iline] & 1) continue;
          int64_t jl = iline - iline1;
          xline_setup(c, &ls, iline, &XC);
          for (int k = 1; k <= N; k++)
            xline_depth(c, k, &XC, &bCop[jl * N + k - 1], &bVv[jl * N + k - 1], &bAv[jl * N + k - 1]);
        }
      }
      st.nsegments++;
      /* wavelength points are independent, as in the reference's OpenMP loop over iwl
       * (src/synth.f90:1075-1080); each point's sum keeps its line order, so the result
       * does not depend on the thread count */
#pragma omp parallel
      {
        double *LopW = malloc(sizeof(double) * N);
        double *M6w = malloc(sizeof(double) * 6 * N);
#pragma omp for schedule(dynamic, 16)
        for (int iwl = iwl1; iwl <= iwl2; iwl++) {
          if (!magnetic) {
            line_opacity1(c, &ls, WAVE(iwl), Lstart[iwl], Lstop[iwl], iwl, Loff, bCop, bVv, bAv, LopW);
            for (int k = 1; k <= N; k++) LOP(k, iwl) = LOP(k, iwl) + LopW[k - 1];
          } else {
            memset(M6w, 0, sizeof(double) * 6 * N);
            for (int64_t iline = Lstart[iwl]; iline <= Lstop[iwl]; iline++) {
              int flag = ls.LinFlag[iline];
              if ((flag & 1) || (flag & 8) || iwl < ls.Ileft[iline] || iwl > ls.Iright[iline]) continue;
              size_t b = (size_t)(iline - Loff - 1) * N;
              mag_line_sum(c, &ls, WAVE(iwl), iline, bCop + b, bVv + b, bAv + b, M6w);
            }
            mag_motype_div(c, M6w);
            double *dst = c->maglineop + (size_t)(iwl - 1) * N * 6;
            for (int q = 0; q < 6 * N; q++) dst[q] = dst[q] + M6w[q];
          }
        }
        free(LopW);
        free(M6w);
      }
      if (iwl2 == Nwl) break;
      iwl1 = iwl2 + 1;
    }
    free(Lstart);
    free(Lstop);
  }
  if (stats) *stats = st;
  free(ls.map);
  free(ls.LinFlag);
  free(ls.Wlcent);
  free(ls.Ileft);
  free(ls.Iright);
  free(ls.spl_pi);
  free(ls.spl_si);
  free(bCop);
  free(bVv);
  free(bAv);
  free(LinStr);
  free(Cop);
  free(Vv);
  free(Av);
  free(LineOpacL);
  free(ContOpacL);
  return 0;
}

EXPORT int fso_lineop(fso_ctx *c, int32_t list_type, const fsg_lines *L, int64_t nlines_lim, float *linstr_out,
                      int32_t *flags_out, int32_t *ileft_out, int32_t *iright_out, fsg_lineop_stats *stats) {
  return lineop_common(c, list_type, L, nlines_lim, 0, NULL, 0, linstr_out, flags_out, ileft_out, iright_out,
                       stats, 0);
}

/* pass 1 of LineopCalc only: flags, windows and the evaluation count, no accumulation */
EXPORT int fso_lineop_pass1(fso_ctx *c, int32_t list_type, const fsg_lines *L, int32_t *flags_out,
                            int32_t *ileft_out, int32_t *iright_out, fsg_lineop_stats *stats) {
  return lineop_common(c, list_type, L, L->n > 0 ? L->n : 1, 0, NULL, 0, NULL, flags_out, ileft_out, iright_out,
                       stats, 1);
}

EXPORT int fso_maglineop(fso_ctx *c, const fsg_lines *L, int64_t nlines_lim, const double *bfield, int32_t bframe,
                         float *linstr_out, int32_t *flags_out, int32_t *ileft_out, int32_t *iright_out,
                         fsg_lineop_stats *stats) {
  return lineop_common(c, FSG_LIST_MIXED, L, nlines_lim, 1, bfield, bframe, linstr_out, flags_out, ileft_out,
                       iright_out, stats, 0);
}

/* Per-line coefficients for tests: XlineSetup + XlineDepth of line i (0-based) with the
 * flag bits the driver would set; outputs [ntau] each. */
EXPORT int fso_xline(fso_ctx *c, int32_t list_type, const fsg_lines *L, int64_t i, double *cop, double *vv,
                     double *av) {
  list_state ls;
  memset(&ls, 0, sizeof(ls));
  int64_t map0 = i;
  int flag[2] = {0, 0};
  double wlc[2] = {0.0, L->wl[i] - c->wfirst};
  ls.list_type = list_type;
  ls.nlines = 1;
  ls.L = L;
  ls.map = &map0;
  ls.LinFlag = flag;
  ls.Wlcent = wlc;
  if (list_type == FSG_LIST_ATOMIC && L->gvdw[i] > 10.0f) flag[1] |= 4;
  if (list_type == FSG_LIST_MIXED) {
    if (L->is_mol && L->is_mol[i]) flag[1] |= 2;
    if (L->gvdw[i] > 10.0f) flag[1] |= 4;
  }
  xline_coef XC;
  xline_setup(c, &ls, 1, &XC);
  for (int k = 1; k <= c->ntau; k++) xline_depth(c, k, &XC, &cop[k - 1], &vv[k - 1], &av[k - 1]);
  return 0;
}

/* ---------------------------------------------------------------------------------------
 * Formal solution (src/synth.f90:1825-1901, 1904-2111) and flux (src/convol.f90:233-234)
 * ------------------------------------------------------------------------------------- */
EXPORT int fso_radtran(fso_ctx *c, double *spl_out, double *spc_out) {
  int N = c->ntau, M = c->nmu, Nwl = c->nwl;
  free(c->spl);
  free(c->spc);
  c->spl = malloc(sizeof(double) * (size_t)Nwl * M);
  c->spc = malloc(sizeof(double) * (size_t)Nwl * M);
  c->magnetic = 0;
  /* independent wavelength points, as the reference's OpenMP loop (src/synth.f90:1850) */
#pragma omp parallel
  {
    double *Src = malloc(sizeof(double) * N), *Xki = malloc(sizeof(double) * N);
#pragma omp for schedule(dynamic, 64)
    for (int iwl = 1; iwl <= Nwl; iwl++) {
      planck(c, WAVE(iwl) + c->wfirst, Src);
      for (int k = 1; k <= N; k++) Xki[k - 1] = CONT(k, iwl);
      locprf(c, c->spc + (size_t)(iwl - 1) * M, Xki, Src);
      for (int k = 1; k <= N; k++) Xki[k - 1] = CONT(k, iwl) + LOP(k, iwl);
      locprf(c, c->spl + (size_t)(iwl - 1) * M, Xki, Src);
    }
    free(Src);
    free(Xki);
  }
  if (spl_out) memcpy(spl_out, c->spl, sizeof(double) * (size_t)Nwl * M);
  if (spc_out) memcpy(spc_out, c->spc, sizeof(double) * (size_t)Nwl * M);
  return 0;
}

/* Solve4x4 (src/solver.f90:541-588); A[i][j] = A(i+1, j+1) */
EXPORT void fso_solve4x4(const double *Ain, const double *B, double *X) {
  double M[4][4], R[4];
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) M[i][j] = Ain[i * 4 + j];
    R[i] = B[i];
  }
  for (int k = 0; k < 3; k++) {
    int ip = k;
    double p = fabs(M[k][k]);
    for (int i = k + 1; i < 4; i++)
      if (fabs(M[i][k]) > p) {
        p = fabs(M[i][k]);
        ip = i;
      }
    if (ip != k) {
      for (int j = k; j < 4; j++) {
        double s = M[k][j];
        M[k][j] = M[ip][j];
        M[ip][j] = s;
      }
      double s = R[k];
      R[k] = R[ip];
      R[ip] = s;
    }
    p = 1.0 / M[k][k];
    for (int i = k + 1; i < 4; i++) {
      double f = M[i][k] * p;
      for (int j = k + 1; j < 4; j++) M[i][j] = M[i][j] - f * M[k][j];
      R[i] = R[i] - f * R[k];
    }
  }
  X[3] = R[3] / M[3][3];
  X[2] = (R[2] - M[2][3] * X[3]) / M[2][2];
  X[1] = (R[1] - M[1][2] * X[2] - M[1][3] * X[3]) / M[1][1];
  X[0] = (R[0] - M[0][1] * X[1] - M[0][2] * X[2] - M[0][3] * X[3]) / M[0][0];
}

/* 4x4 helpers: m[i][j] stored row-major */
static void mat_mul(const double *A, const double *B, double *C) {
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += A[i * 4 + k] * B[k * 4 + j];
      C[i * 4 + j] = s;
    }
}
static void mat_vec(const double *A, const double *x, double *y) {
  for (int i = 0; i < 4; i++) {
    double s = 0.0;
    for (int k = 0; k < 4; k++) s += A[i * 4 + k] * x[k];
    y[i] = s;
  }
}

/* The polarised DELO-Bezier integration along one ray of N points (the body shared by
 * StokesPrfPP src/solver.f90:265-394 and StokesPrfSph src/solver.f90:402-538).  Point k
 * (0-based) samples layer map[k] (1-based; identity when map is NULL) at column mass
 * rhoxR[k]; the optical depth steps use Xmu (1 on spherical rays, whose column mass is
 * already measured along the ray); geometry g[k * gs + 0..4] = SPhi2, CPhi2, CPhi, S2Psi,
 * C2Psi (gs = 0 for depth-independent angles).  Grazing rays start from a zero Stokes
 * vector, the others from the unpolarised S at the deepest point.  out4 = I, Q, U, V. */
static void stokes_ray(fso_ctx *c, int iwl, int N, const int *map, const double *rhoxR, double Xmu, const double *g,
                       size_t gs, int graz, const double *Src, double *out4) {
  const int Nt = c->ntau;
  double *XK = malloc(sizeof(double) * 16 * N), *XKD = malloc(sizeof(double) * 16 * N);
  double *XKI = malloc(sizeof(double) * N), *Deltau = malloc(sizeof(double) * N);
  double *SRC4 = malloc(sizeof(double) * 4 * N), *SRD = malloc(sizeof(double) * 4 * N);
  /* XK(i,j,k) column-major -> XK[k*16 + (j)*4 + i]; kept in that layout so that CentDer over
   * the 16 components follows the reference's component order */
  for (int it = 0; it < N; it++) {
    const int ia = map ? map[it] - 1 : it;
    const double *gk = g + (size_t)it * gs;
    const double SPhi2 = gk[0], CPhi2 = gk[1], CPhi = gk[2], S2Psi = gk[3], C2Psi = gk[4];
    const double *m = c->maglineop + ((size_t)(iwl - 1) * Nt + ia) * 6;
    double Php = m[0], Phb = m[1], Phr = m[2], Psp = m[3], Psb = m[4], Psr = m[5];
    double PhiI = 0.5 * (Php * SPhi2 + 0.5 * (Phr + Phb) * (1.0 + CPhi2)) + LOP(ia + 1, iwl) + CONT(ia + 1, iwl);
    double InvPhiI = 1.0 / PhiI;
    double Sabs = 0.5 * (Php - 0.5 * (Phr + Phb)) * SPhi2;
    double PhiQ = Sabs * C2Psi * InvPhiI;
    double PhiU = Sabs * S2Psi * InvPhiI;
    double PhiV = 0.5 * (Phr - Phb) * CPhi * InvPhiI;
    double Sdis = 0.5 * (Psp - 0.5 * (Psr + Psb)) * SPhi2;
    double PsiQ = Sdis * C2Psi * InvPhiI;
    double PsiU = Sdis * S2Psi * InvPhiI;
    double PsiV = 0.5 * (Psr - Psb) * CPhi * InvPhiI;
    XKI[it] = PhiI;
    double *x = XK + (size_t)it * 16; /* x[col*4 + row] */
    x[0 * 4 + 0] = 0.0;   x[1 * 4 + 0] = PhiQ;  x[2 * 4 + 0] = PhiU;  x[3 * 4 + 0] = PhiV;
    x[0 * 4 + 1] = PhiQ;  x[1 * 4 + 1] = 0.0;   x[2 * 4 + 1] = PsiV;  x[3 * 4 + 1] = -PsiU;
    x[0 * 4 + 2] = PhiU;  x[1 * 4 + 2] = -PsiV; x[2 * 4 + 2] = 0.0;   x[3 * 4 + 2] = PsiQ;
    x[0 * 4 + 3] = PhiV;  x[1 * 4 + 3] = PsiU;  x[2 * 4 + 3] = -PsiQ; x[3 * 4 + 3] = 0.0;
    SRC4[it * 4 + 0] = Src[ia];
    SRC4[it * 4 + 1] = Src[ia] * PhiQ;
    SRC4[it * 4 + 2] = Src[ia] * PhiU;
    SRC4[it * 4 + 3] = Src[ia] * PhiV;
  }
  tau_step(rhoxR, XKI, Xmu, N, Deltau);
  cent_der(Deltau, SRC4, 4, N, SRD);
  cent_der(Deltau, XK, 16, N, XKD);
  /* row-major copies for the algebra below */
#define RM(src, k, out)                                                  \
  for (int r_ = 0; r_ < 4; r_++)                                         \
    for (int c_ = 0; c_ < 4; c_++) (out)[r_ * 4 + c_] = (src)[(size_t)(k)*16 + c_ * 4 + r_];
  double K[16], K1[16], KD[16], XKA[16], XKB[16], QA[16], QB[16], AA[16], T1[16];
  double XINT1[4], XINT2[4], v1[4], v2[4], v3[4], v4[4];
  XINT2[0] = graz ? 0.0 : SRC4[(N - 1) * 4 + 0];
  XINT2[1] = XINT2[2] = XINT2[3] = 0.0;
  RM(XK, N - 1, K);
  RM(XKD, N - 1, KD);
  mat_mul(K, K, T1);
  for (int q = 0; q < 16; q++) XKB[q] = T1[q] + KD[q];
  for (int k = N - 1; k >= 1; k--) {
    int k1 = k + 1;
    double Dtau = Deltau[k - 1], Dtau05 = 0.5 * Dtau, Eps, Alpha, Beta, Gamma;
    delo_coef(Dtau, &Eps, &Alpha, &Beta, &Gamma);
    Gamma = Gamma * 0.5;
    RM(XK, k - 1, K);
    RM(XK, k1 - 1, K1);
    RM(XKD, k - 1, KD);
    for (int q = 0; q < 16; q++) XKA[q] = XKB[q];
    mat_mul(K, K, T1);
    for (int q = 0; q < 16; q++) XKB[q] = T1[q] + KD[q];
    for (int q = 0; q < 16; q++) QA[q] = Dtau05 * XKA[q] + (Dtau05 - 1.0) * K1[q];
    for (int q = 0; q < 16; q++) QB[q] = -Dtau05 * XKB[q] - (Dtau05 + 1.0) * K[q];
    for (int q = 0; q < 16; q++) AA[q] = ((q % 5) == 0 ? 1.0 : 0.0) + Alpha * K[q] - Gamma * QB[q];
    double GD = Gamma * Dtau05;
    mat_vec(K1, XINT2, v1);
    mat_vec(QA, XINT2, v2);
    mat_vec(K1, SRC4 + (size_t)(k1 - 1) * 4, v3);
    mat_vec(K, SRC4 + (size_t)(k - 1) * 4, v4);
    for (int q = 0; q < 4; q++)
      XINT1[q] = Eps * XINT2[q] - Beta * v1[q] + Gamma * v2[q] + (Beta + Gamma) * SRC4[(k1 - 1) * 4 + q] -
                 GD * v3[q] + (Alpha + Gamma) * SRC4[(k - 1) * 4 + q] + GD * v4[q] +
                 GD * (SRD[(k - 1) * 4 + q] - SRD[(k1 - 1) * 4 + q]);
    fso_solve4x4(AA, XINT1, XINT2);
  }
#undef RM
  for (int q = 0; q < 4; q++) out4[q] = XINT2[q];
  free(XK);
  free(XKD);
  free(XKI);
  free(Deltau);
  free(SRC4);
  free(SRD);
}

/* StokesPrfPP (src/solver.f90:265-394) for one wavelength; Stokes[4][nmu] */
EXPORT void fso_stokes_prf_pp(fso_ctx *c, int iwl, const double *Src, const double *SPhi2, const double *CPhi2,
                              const double *CPhi, const double *S2Psi, const double *C2Psi, double *Stokes) {
  int Nmu = c->nmu;
  for (int im = 0; im < Nmu; im++) {
    double g[5] = {SPhi2[im], CPhi2[im], CPhi[im], S2Psi[im], C2Psi[im]}, out[4];
    stokes_ray(c, iwl, c->ntau, NULL, c->rhox, c->xmu[im], g, 0, 0, Src, out);
    for (int q = 0; q < 4; q++) Stokes[q * Nmu + im] = out[q];
  }
}

/* StokesPrfSph (src/solver.f90:402-538) for one wavelength along the rays of SphereInit;
 * geo[(im * 2 ntau + k) * 5 + 0..4] = SPhi2S, CPhi2S, CPhiS, S2PsiS, C2PsiS at ray point k;
 * Stokes[4][nmu] */
EXPORT void fso_stokes_prf_sph(fso_ctx *c, int iwl, const double *Src, const double *geo, double *Stokes) {
  int Nmu = c->nmu, L = 2 * c->ntau;
  for (int im = 0; im < Nmu; im++) {
    double out[4];
    stokes_ray(c, iwl, c->ntau_mu[im], c->map_mu + (size_t)im * L, c->rhox_mu + (size_t)im * L, 1.0,
               geo + (size_t)im * L * 5, 5, c->graz_mu[im], Src, out);
    for (int q = 0; q < 4; q++) Stokes[q * Nmu + im] = out[q];
  }
}

/* Field geometry per mu (src/synth.f90:1935-1994) */
EXPORT int fso_field_geometry(int nmu, const double *xmu, const double *bfield, int nbfield, int bframe,
                              double *SPhi2, double *CPhi2, double *CPhi, double *S2Psi, double *C2Psi) {
  if (kPi == 0.0) init_constants();
  double d2r = kPi / 180.0;
  if (bframe == 1) {
    double Sphi = sin(bfield[1] * d2r), Spsi = sin(bfield[2] * d2r), Cpsi = cos(bfield[2] * d2r);
    double cg = cos(bfield[1] * d2r);
    for (int im = 0; im < nmu; im++) {
      CPhi[im] = cg;
      SPhi2[im] = Sphi * Sphi;
      CPhi2[im] = cg * cg;
      S2Psi[im] = 2.0 * Spsi * Cpsi;
      C2Psi[im] = Cpsi * Cpsi - Spsi * Spsi;
    }
  } else if (bframe == 2) {
    if (nbfield < 1 + 2 * nmu) {
      snprintf(g_err, sizeof(g_err), "BFRAME 2 needs 1+2*Nmu BFIELD values");
      return 1;
    }
    for (int im = 1; im <= nmu; im++) {
      double Sphi = sin(bfield[2 * im - 1] * d2r);
      double Spsi = sin(bfield[2 * im] * d2r), Cpsi = cos(bfield[2 * im] * d2r);
      CPhi[im - 1] = cos(bfield[2 * im - 1] * d2r);
      SPhi2[im - 1] = Sphi * Sphi;
      CPhi2[im - 1] = CPhi[im - 1] * CPhi[im - 1];
      S2Psi[im - 1] = 2.0 * Spsi * Cpsi;
      C2Psi[im - 1] = Cpsi * Cpsi - Spsi * Spsi;
    }
  } else {
    double Fldmod = sqrt(bfield[0] * bfield[0] + bfield[1] * bfield[1] + bfield[2] * bfield[2]);
    for (int im = 0; im < nmu; im++) {
      double Sphi, Spsi, Cpsi;
      if (Fldmod > 0.0) {
        double Coslon = xmu[im], Sinlon = sqrt(1.0 - Coslon * Coslon);
        double Bx = bfield[0] * Sinlon + bfield[2] * Coslon, By = bfield[1];
        double Bz = bfield[0] * Coslon - bfield[2] * Sinlon;
        double Bt = sqrt(Bx * Bx + By * By);
        Sphi = Bt / Fldmod;
        CPhi[im] = Bz / Fldmod;
        if (Bt > 0.0) {
          Spsi = By / Bt;
          Cpsi = Bx / Bt;
        } else {
          Spsi = 0.0;
          Cpsi = 1.0;
        }
      } else {
        Sphi = 0.0;
        CPhi[im] = 1.0;
        Spsi = 0.0;
        Cpsi = 1.0;
      }
      SPhi2[im] = Sphi * Sphi;
      CPhi2[im] = CPhi[im] * CPhi[im];
      S2Psi[im] = 2.0 * Spsi * Cpsi;
      C2Psi[im] = Cpsi * Cpsi - Spsi * Spsi;
    }
  }
  return 0;
}

/* Field geometry along the spherical rays (src/synth.f90:2017-2090): BFRAME 1 and 2 as in
 * the plane-parallel case; BFRAME 0 projects the stellar-frame field with the local
 * ray-to-radial cosine MuLoc of each ray point.  geo[(im * 2 ntau + k) * 5 + 0..4] =
 * SPhi2, CPhi2, CPhi, S2Psi, C2Psi. */
EXPORT int fso_sph_field_geometry(fso_ctx *c, const double *bfield, int nbfield, int bframe, double *geo) {
  if (kPi == 0.0) init_constants();
  const double d2r = kPi / 180.0;
  const int M = c->nmu, L = 2 * c->ntau;
  if (bframe == 2 && nbfield < 1 + 2 * M) {
    snprintf(g_err, sizeof(g_err), "BFRAME 2 needs 1+2*Nmu BFIELD values");
    return 1;
  }
  for (int im = 0; im < M; im++) {
    int np = (bframe == 1) ? L : c->ntau_mu[im];
    for (int k = 0; k < np; k++) {
      double Sphi, Cgam, Spsi, Cpsi;
      if (bframe == 1) {
        Sphi = sin(bfield[1] * d2r);
        Spsi = sin(bfield[2] * d2r);
        Cpsi = cos(bfield[2] * d2r);
        Cgam = cos(bfield[1] * d2r);
      } else if (bframe == 2) {
        Sphi = sin(bfield[2 * im + 1] * d2r);
        Spsi = sin(bfield[2 * im + 2] * d2r);
        Cpsi = cos(bfield[2 * im + 2] * d2r);
        Cgam = cos(bfield[2 * im + 1] * d2r);
      } else {
        double Fldmod = sqrt(bfield[0] * bfield[0] + bfield[1] * bfield[1] + bfield[2] * bfield[2]);
        if (Fldmod > 0.0) {
          double Coslon = c->muloc[(size_t)im * L + k], Sinlon = sqrt(1.0 - Coslon * Coslon);
          double Bx = bfield[0] * Sinlon + bfield[2] * Coslon, By = bfield[1];
          double Bz = bfield[0] * Coslon - bfield[2] * Sinlon;
          double Bt = sqrt(Bx * Bx + By * By);
          Sphi = Bt / Fldmod;
          Cgam = Bz / Fldmod;
          if (Bt > 0.0) {
            Spsi = By / Bt;
            Cpsi = Bx / Bt;
          } else {
            Spsi = 0.0;
            Cpsi = 1.0;
          }
        } else {
          Sphi = 0.0;
          Cgam = 1.0;
          Spsi = 0.0;
          Cpsi = 1.0;
        }
      }
      double *g = geo + ((size_t)im * L + k) * 5;
      g[0] = Sphi * Sphi;
      g[1] = Cgam * Cgam;
      g[2] = Cgam;
      g[3] = 2.0 * Spsi * Cpsi;
      g[4] = Cpsi * Cpsi - Spsi * Spsi;
    }
  }
  return 0;
}

EXPORT int fso_magradtran(fso_ctx *c, const double *bfield, int32_t nbfield, int32_t bframe, double *spz_out,
                          double *spc_out) {
  int N = c->ntau, M = c->nmu, Nwl = c->nwl;
  if (!c->maglineop) c->maglineop = calloc((size_t)Nwl * N * 6, sizeof(double));
  free(c->spz);
  free(c->spc);
  c->spz = malloc(sizeof(double) * (size_t)Nwl * 4 * M);
  c->spc = malloc(sizeof(double) * (size_t)Nwl * M);
  c->magnetic = 1;
  double *SPhi2 = malloc(sizeof(double) * M), *CPhi2 = malloc(sizeof(double) * M);
  double *CPhi = malloc(sizeof(double) * M), *S2Psi = malloc(sizeof(double) * M);
  double *C2Psi = malloc(sizeof(double) * M);
  double *geo = NULL;
  if (c->radius > 0.0) {
    geo = calloc((size_t)M * 2 * N * 5, sizeof(double));
    if (fso_sph_field_geometry(c, bfield, nbfield, bframe, geo)) return 1;
  } else if (fso_field_geometry(M, c->xmu, bfield, nbfield, bframe, SPhi2, CPhi2, CPhi, S2Psi, C2Psi)) {
    return 1;
  }
#pragma omp parallel
  {
    double *Src = malloc(sizeof(double) * N), *Xki = malloc(sizeof(double) * N);
#pragma omp for schedule(dynamic, 64)
    for (int iwl = 1; iwl <= Nwl; iwl++) {
      planck(c, WAVE(iwl) + c->wfirst, Src);
      for (int k = 1; k <= N; k++) Xki[k - 1] = CONT(k, iwl);
      locprf(c, c->spc + (size_t)(iwl - 1) * M, Xki, Src);
      if (geo) fso_stokes_prf_sph(c, iwl, Src, geo, c->spz + (size_t)(iwl - 1) * 4 * M);
      else fso_stokes_prf_pp(c, iwl, Src, SPhi2, CPhi2, CPhi, S2Psi, C2Psi, c->spz + (size_t)(iwl - 1) * 4 * M);
    }
    free(Src);
    free(Xki);
  }
  free(geo);
  free(SPhi2);
  free(CPhi2);
  free(CPhi);
  free(S2Psi);
  free(C2Psi);
  if (spz_out) memcpy(spz_out, c->spz, sizeof(double) * (size_t)Nwl * 4 * M);
  if (spc_out) memcpy(spc_out, c->spc, sizeof(double) * (size_t)Nwl * M);
  return 0;
}

/* Flux sums of DiskInteg without convolution (src/convol.f90:97-101, 233-234) */
EXPORT int fso_flux(fso_ctx *c, double *sfl, double *sfc) {
  int M = c->nmu, Nwl = c->nwl;
  for (int i = 0; i < Nwl; i++) {
    sfl[i] = 0.0;
    sfc[i] = 0.0;
  }
  for (int im = 0; im < M; im++)
    for (int i = 0; i < Nwl; i++) {
      double IL = c->magnetic ? c->spz[(size_t)i * 4 * M + im] : c->spl[(size_t)i * M + im];
      sfl[i] = sfl[i] + IL * c->wmu[im];
      sfc[i] = sfc[i] + c->spc[(size_t)i * M + im] * c->wmu[im];
    }
  return 0;
}

/* ------------------------------------------------------------------------------------
 * Convolution and disk integration: IntegConvol (src/convol.f90:27-64), DiskInteg (67-246),
 * ConvolGau (249-300), SemiArea (src/utils.f90:1435-1445)
 * ------------------------------------------------------------------------------------ */
EXPORT double fso_semi_area(double R, double V) {
  if (R <= 0.0) return 0.0;
  double X = fmax(-R, fmin(R, V));
  return 0.5 * (X * sqrt(fmax(0.0, (R - X) * (R + X))) + R * R * asin(X / R));
}

/* out[i] = sum_{j=1..nk} in[clamp(i + j - 1 - h)] kern[j], the padded sums of DiskInteg and
 * ConvolGau (edge values repeated), j ascending; 0-based i, kern[0..nk-1] */
static void conv_clamped(int nw, const double *in, int nk, const double *kern, int h, double *out) {
  for (int i = 0; i < nw; i++) {
    double sm = 0.0;
    for (int j = 0; j < nk; j++) {
      int k = i + j - h;
      k = k < 0 ? 0 : (k > nw - 1 ? nw - 1 : k);
      sm = sm + in[k] * kern[j];
    }
    out[i] = sm;
  }
}

/* strided spectrum <-> contiguous copy */
static void gather_s(int nw, const double *base, size_t stride, double *out) {
  for (int i = 0; i < nw; i++) out[i] = base[(size_t)i * stride];
}
static void scatter_s(int nw, const double *in, double *base, size_t stride) {
  for (int i = 0; i < nw; i++) base[(size_t)i * stride] = in[i];
}

/* ConvolGau on a contiguous spectrum S[nw] (FWHM in km/s) */
EXPORT void fso_convolgau(int nw, double *S, double fwhm, int wmode, double wfirst, double wlast, double dstep) {
  double sigma = wmode == 0 ? fwhm * (wfirst + wlast) / kVc / 2.0 / dstep : fwhm / dstep;
  sigma = sigma / (2.0 * sqrt(2.0 * log(2.0)));
  int j = (int)(10.0 * sigma), jm = (int)((nw - 3.0) / 2.0);
  j = j < jm ? j : jm;
  j = j > 3 ? j : 3;
  const int ng = 2 * j + 1;
  double *kern = malloc(sizeof(double) * ng), norm = 0.0;
  for (int i = 1; i <= ng; i++) {
    double arg = fmax(-20.0, -0.5 * pow((double)(i - j - 1) / sigma, 2));
    kern[i - 1] = exp(arg);
    norm = norm + kern[i - 1];
  }
  for (int i = 0; i < ng; i++) kern[i] = kern[i] / norm;
  const int k = (int)((ng - 1.0) / 2.0);
  double *T = malloc(sizeof(double) * nw);
  conv_clamped(nw, S, ng, kern, k, T);
  memcpy(S, T, sizeof(double) * nw);
  free(T);
  free(kern);
}

static double wlast_of(const fso_ctx *c) { return c->wfirst + c->wave[c->nwl]; }

/* DiskInteg: SfL, SfC [nwl] from SpL (Stokes I after a polarised solution) and SpC, with the
 * quadrature quad's weights and annulus radii for the context's mu */
static int disk_integ(fso_ctx *c, double vsini, double vmacro, double vgau, int quad, double *sfl, double *sfc) {
  const int M = c->nmu, Nwl = c->nwl;
  double *Wq = malloc(sizeof(double) * M), *MuRadii = malloc(sizeof(double) * (M + 1));
  fso_mu_quadrature_r(quad, M, 0, c->xmu, Wq, MuRadii);
  const double wlast = wlast_of(c);
  const double Dv = c->wmode == 0 ? 2.0 * c->dstep / (c->wfirst + wlast) * kVc : c->dstep;
  int Nrk = 2 * (int)(vsini * 1.0 / Dv) + 3;
  const double Sigma = vmacro / Dv / sqrt(2.0);
  int Nmk = (int)(10.0 * Sigma), nmm = (int)((Nwl - 3) / 2);
  Nmk = Nmk < nmm ? Nmk : nmm;
  Nmk = Nmk > 3 ? Nmk : 3;
  const int Nm = 2 * Nmk + 1;
  double *Rkern = malloc(sizeof(double) * (Nrk > 0 ? Nrk : 1));
  double *Mkern = malloc(sizeof(double) * Nm), *Mr = malloc(sizeof(double) * Nm), *Mt = malloc(sizeof(double) * Nm);
  double *SiL = malloc(sizeof(double) * Nwl), *SiC = malloc(sizeof(double) * Nwl);
  double *TL = malloc(sizeof(double) * Nwl), *TC = malloc(sizeof(double) * Nwl);
  for (int i = 0; i < Nwl; i++) sfl[i] = sfc[i] = 0.0;
  for (int imu = 1; imu <= M; imu++) {
    if (c->magnetic) gather_s(Nwl, c->spz + (imu - 1), (size_t)4 * M, SiL);
    else gather_s(Nwl, c->spl + (imu - 1), (size_t)M, SiL);
    gather_s(Nwl, c->spc + (imu - 1), (size_t)M, SiC);
    if (vsini > 0.0) {
      const double R1 = vsini * MuRadii[imu - 1], R2 = vsini * MuRadii[imu];
      const double Maxv = R2;
      Nrk = 2 * (int)(Maxv / Dv) + 3;
      double norm = 0.0;
      for (int i = 1; i <= Nrk; i++) {
        const double V = fabs(Dv * (double)(i - 1 - (Nrk - 1) / 2));
        Rkern[i - 1] = fso_semi_area(R2, V + 0.5 * Dv) - fso_semi_area(R2, V - 0.5 * Dv) -
                       fso_semi_area(R1, V + 0.5 * Dv) + fso_semi_area(R1, V - 0.5 * Dv);
        norm = norm + Rkern[i - 1];
      }
      for (int i = 0; i < Nrk; i++) Rkern[i] = Rkern[i] / norm;
      if (R2 > 0.5 * Dv) {
        const int h = (int)((Nrk - 1.0) / 2.0);
        conv_clamped(Nwl, SiL, Nrk, Rkern, h, TL);
        conv_clamped(Nwl, SiC, Nrk, Rkern, h, TC);
        memcpy(SiL, TL, sizeof(double) * Nwl);
        memcpy(SiC, TC, sizeof(double) * Nwl);
      }
    }
    if (vmacro > 0.0) {
      const double Sigr = Sigma * c->xmu[imu - 1], Sigt = Sigma * sqrt(1.0 - c->xmu[imu - 1] * c->xmu[imu - 1]);
      double *kk[2] = {Mr, Mt};
      const double sg[2] = {Sigr, Sigt};
      for (int t = 0; t < 2; t++) {
        if (sg[t] > 0.0) {
          double norm = 0.0;
          for (int i = 1; i <= Nm; i++) {
            const double arg = fmax(-20.0, -0.5 * pow((double)(i - Nmk - 1) / sg[t], 2));
            kk[t][i - 1] = exp(arg);
            norm = norm + kk[t][i - 1];
          }
          for (int i = 0; i < Nm; i++) kk[t][i] = kk[t][i] / norm;
        } else {
          for (int i = 0; i < Nm; i++) kk[t][i] = 0.0;
          kk[t][Nmk] = 1.0;
        }
      }
      for (int i = 0; i < Nm; i++) Mkern[i] = 0.5 * Mr[i] + 0.5 * Mt[i];
      const int h = (int)((Nm - 1.0) / 2.0);
      conv_clamped(Nwl, SiL, Nm, Mkern, h, TL);
      conv_clamped(Nwl, SiC, Nm, Mkern, h, TC);
      memcpy(SiL, TL, sizeof(double) * Nwl);
      memcpy(SiC, TC, sizeof(double) * Nwl);
    }
    for (int i = 0; i < Nwl; i++) {
      sfl[i] = sfl[i] + SiL[i] * Wq[imu - 1];
      sfc[i] = sfc[i] + SiC[i] * Wq[imu - 1];
    }
  }
  if (vgau > 0.0) {
    fso_convolgau(Nwl, sfl, vgau, c->wmode, c->wfirst, wlast, c->dstep);
    fso_convolgau(Nwl, sfc, vgau, c->wmode, c->wfirst, wlast, c->dstep);
  }
  free(Wq);
  free(MuRadii);
  free(Rkern);
  free(Mkern);
  free(Mr);
  free(Mt);
  free(SiL);
  free(SiC);
  free(TL);
  free(TC);
  return 0;
}

/* IntegConvol: cimode 1 Gaussian broadening of the intensity spectra in place (SpL and SpC,
 * or the four Stokes spectra and SpC); cimode 2 disk integration into sfl, sfc */
EXPORT int fso_integ_convol(fso_ctx *c, int cimode, double vsini, double vmacro, double vgau, int quad,
                            double *sfl, double *sfc) {
  const int M = c->nmu, Nwl = c->nwl;
  if (cimode == 2) return disk_integ(c, vsini, vmacro, vgau, quad, sfl, sfc);
  if (cimode != 1) {
    snprintf(g_err, sizeof g_err, "IntegConvol: unknown mode %d", cimode);
    return 1;
  }
  const double wlast = wlast_of(c);
  double *T = malloc(sizeof(double) * Nwl);
  for (int imu = 0; imu < M; imu++) {
    if (c->magnetic) {
      for (int st = 0; st < 4; st++) {
        gather_s(Nwl, c->spz + (size_t)st * M + imu, (size_t)4 * M, T);
        fso_convolgau(Nwl, T, vgau, c->wmode, c->wfirst, wlast, c->dstep);
        scatter_s(Nwl, T, c->spz + (size_t)st * M + imu, (size_t)4 * M);
      }
    } else {
      gather_s(Nwl, c->spl + imu, (size_t)M, T);
      fso_convolgau(Nwl, T, vgau, c->wmode, c->wfirst, wlast, c->dstep);
      scatter_s(Nwl, T, c->spl + imu, (size_t)M);
    }
    gather_s(Nwl, c->spc + imu, (size_t)M, T);
    fso_convolgau(Nwl, T, vgau, c->wmode, c->wfirst, wlast, c->dstep);
    scatter_s(Nwl, T, c->spc + imu, (size_t)M);
  }
  free(T);
  return 0;
}

/* replaces the held intensity spectra (spz non-NULL: a polarised solution) */
EXPORT int fso_set_spectra(fso_ctx *c, const double *spl, const double *spc, const double *spz) {
  const size_t M = c->nmu, Nwl = c->nwl;
  if (spz) {
    if (!c->spz) c->spz = malloc(sizeof(double) * Nwl * 4 * M);
    memcpy(c->spz, spz, sizeof(double) * Nwl * 4 * M);
    c->magnetic = 1;
  } else {
    if (!c->spl) c->spl = malloc(sizeof(double) * Nwl * M);
    memcpy(c->spl, spl, sizeof(double) * Nwl * M);
    c->magnetic = 0;
  }
  if (!c->spc) c->spc = malloc(sizeof(double) * Nwl * M);
  memcpy(c->spc, spc, sizeof(double) * Nwl * M);
  return 0;
}

/* the intensity spectra as held by the oracle: spl/spc [nwl][nmu], spz [nwl][4][nmu] */
EXPORT int fso_get_spectra(fso_ctx *c, double *spl, double *spc, double *spz) {
  const size_t M = c->nmu, Nwl = c->nwl;
  if (spl && c->spl) memcpy(spl, c->spl, sizeof(double) * Nwl * M);
  if (spc && c->spc) memcpy(spc, c->spc, sizeof(double) * Nwl * M);
  if (spz && c->spz) memcpy(spz, c->spz, sizeof(double) * Nwl * 4 * M);
  return 0;
}
