// fsg_device.h -- device helpers shared by the line-opacity and formal-solution kernels:
// XlineSetup/XlineDepth, the HVoigt table lookup, WaveIdx, the XCD-aware tile order,
// compensated FP32 sums and the DELO-Bezier step functions.
#pragma once

#include "fsg_common.h"

namespace fsg {

// ------------------------------------------------------------------------------------
// per-line kernel coefficients
// ------------------------------------------------------------------------------------
struct XC {
  double WaveL, Hnux, Elo, Gf, Grad, Xmass, Cqst, Qdiv, Cvdw, AboPre, AboVbc, AboExp;
  int IxPF, MdQst, MdVdw, IxHe;
};

__device__ __forceinline__ double air2vac(double L) {  // src/utils.f90:1020-1032
  double S = 1e4 / L, S2 = S * S;
  double N = 1.0 + 8.336624212083e-5 + 2.408926869968e-2 / (1.301065924522e2 - S2) +
             1.599740894897e-4 / (3.892568793293e1 - S2);
  return (L > 1999.352) ? L * N : L;
}

// flag bits as LineopCalc pass 1 sets them before XlineSetup (src/synth.f90:839-844)
__device__ __forceinline__ int initial_flag(const DevLines &L, int64_t i, int list_type) {
  int f = 0;
  if (list_type == 0) {
    if (L.gvdw[i] > 10.0f) f |= 4;
  } else if (list_type == 2) {
    if (L.is_mol && L.is_mol[i]) f |= 2;
    if (L.gvdw[i] > 10.0f) f |= 4;
  }
  return f;
}

// Batch contexts: per-line work (pass 1, coefficient fill, line strengths) runs on the
// set-up of the line's own atmosphere, in that atmosphere's grid; per-point work takes the
// temperature, column mass and wavelength origin of the point's atmosphere.
__device__ __forceinline__ DevSetup line_setup(const DevSetup &S, const DevLines &L, int64_t i) {
  if (S.batch) return S.batch[L.atm[i]];
  return S;
}

__device__ __forceinline__ void point_setup(DevSetup &S, int p) {
  if (S.batch) {
    const DevSetup *B = S.batch + S.pmap[p];
    S.temp = B->temp;
    S.rhox = B->rhox;
    S.wfirst = B->wfirst;
  }
}

// XlineSetup (src/opacity.f90:96-235)
__device__ inline XC xline_setup(const DevSetup &S, const DevLines &L, int64_t i, double wlc, int flag,
                          int list_type) {
  XC xc;
  double WaveL = wlc + S.wfirst;
  double Elo = (double)L.elo[i];
  xc.WaveL = WaveL;
  xc.Elo = Elo;
  xc.Gf = (double)L.gf[i];
  double GradL = (double)L.grad[i];
  int ispc = L.spc[i];
  int IonL = S.spc_ion[ispc - 1];
  xc.IxPF = S.spc_slot[ispc - 1];
  double GqstL, GvdwL;
  bool IsMol;
  if (list_type == 1) {
    GqstL = 0.0;
    GvdwL = 0.0;
    IsMol = true;
  } else {
    GqstL = (double)L.gqst[i];
    GvdwL = (double)L.gvdw[i];
    IsMol = (list_type == 2) ? ((flag >> 1) & 1) : false;
  }
  double xam = (double)S.xamass[xc.IxPF - 1];
  xc.Hnux = kXlnC * kHc / WaveL;
  xc.Xmass = kXlnDop / kXlnC / kXlnC / xam;
  xc.Grad = (GradL > 0.0) ? GradL : 0.22233e16 / (WaveL * WaveL);
  xc.Cqst = 0.0;
  xc.Qdiv = 1.0;
  xc.Cvdw = 0.0;
  xc.AboPre = 0.0;
  xc.AboVbc = 0.0;
  xc.AboExp = 0.0;
  xc.IxHe = 0;
  if (IsMol && !S.molbrd) {
    xc.MdQst = 9;
    xc.MdVdw = 9;
  } else {
    double Wavevac = air2vac(WaveL);
    double Eup = Elo + 1.0 / (Wavevac * kEv2Ang);
    double pot = (double)S.pot_ion[xc.IxPF - 1];
    double t1 = (double)(IonL * IonL) * kHie / (pot - Eup);
    double t2 = (double)(IonL * IonL) * kHie / (pot - Elo);
    double Enu4 = t1 * t1, Enl4 = t2 * t2;
    if (xc.IxPF == S.he1_slot) {  // HeMatchIdx (src/hestark.f90:376-389)
      for (int k = 1; k <= S.nhe; k++)
        if (fabs(WaveL - S.he_wl[k - 1]) <= S.he_wl_tol && fabs(Elo - S.he_elo[k - 1]) <= S.he_elo_tol) {
          xc.IxHe = k;
          break;
        }
    }
    if (xc.IxHe > 0) {
      xc.MdQst = 3;
    } else if (GqstL > 0.0) {
      xc.MdQst = 0;
      xc.Cqst = GqstL;
    } else if (IsMol) {
      xc.MdQst = 9;
    } else if (IonL == 1) {
      xc.MdQst = 1;
      xc.Cqst = 2.26e-07 * Enu4;
    } else {
      xc.MdQst = 2;
      xc.Cqst = 5.42e-07 * Enu4;
      xc.Qdiv = (IonL + 1.0) * (IonL + 1.0);
    }
    if ((flag >> 2) & 1) {
      xc.MdVdw = 2;
      double Sigma = kA0 * kA0 * (double)(int)GvdwL;
      double Alpha = GvdwL - (double)(int)GvdwL;
      double X = 2.0 - Alpha * 0.5;
      double Gx = X - 1.0;
      double GammaF = 1.0 + (-0.5748646 + (0.9512363 + (-0.6998588 + (0.4245549 - 0.1010678 * Gx) * Gx) * Gx) * Gx) * Gx;
      xc.AboPre = 2e6 * (pow(4.0 / kPi, Alpha * 0.5) * GammaF * 1e4 * Sigma);
      xc.AboVbc = 1.0 / 1.00794 + 1.0 / xam;
      xc.AboExp = 1.0 - Alpha;
    } else if (GvdwL > 0.0) {
      xc.MdVdw = 0;
      xc.Cvdw = GvdwL;
    } else if (IsMol) {
      xc.MdVdw = 9;
    } else {
      xc.MdVdw = 1;
      double Cw = 1.61e-33 * fabs(Enu4 - Enl4) / (double)(IonL * IonL);
      xc.Cvdw = pow(Cw, 0.4);
    }
  }
  if (IsMol && S.vwmol > 0.0) {
    xc.MdVdw = 0;
    xc.Cvdw = S.vwmol;
  }
  return xc;
}

// XlineDepth (src/opacity.f90:239-293), k 0-based
__device__ __forceinline__ void xline_depth(const DevSetup &S, int k, const XC &xc, double &cop, double &vv,
                                            double &av) {
  const int N = S.ntau;
  double T = S.temp[k];
  double Xstim = 1.0 - exp(-xc.Hnux / S.xln[4 * N + k]);
  double Efract = exp(-xc.Elo / S.xln[5 * N + k]);
  double Dopl = sqrt(T * xc.Xmass + S.vturb2);
  double Dldopl = xc.WaveL * Dopl;
  double Dndopl = Dopl / xc.WaveL;
  double Gqst, Gvdw;
  switch (xc.MdQst) {
    case 0: Gqst = xc.Cqst * S.xln[2 * N + k]; break;
    case 1: Gqst = xc.Cqst * S.xne[k]; break;
    case 2: Gqst = xc.Cqst * S.xne[k] / xc.Qdiv; break;
    case 3: Gqst = S.he_gam[(size_t)(xc.IxHe - 1) * N + k]; break;
    default: Gqst = 0.0;
  }
  switch (xc.MdVdw) {
    case 0: Gvdw = xc.Cvdw * S.xln[3 * N + k]; break;
    case 1: Gvdw = S.vwfudge * S.xln[1 * N + k] * xc.Cvdw; break;
    case 2: {
      double Vbar = sqrt(8.0 * kXlnKM0 * T / kPi * xc.AboVbc);
      Gvdw = xc.AboPre * pow(Vbar / 1e4, xc.AboExp) * S.xln[0 * N + k];
      break;
    }
    default: Gvdw = 0.0;
  }
  double Gamtot = xc.Grad + Gqst + Gvdw;
  av = Gamtot / (Dndopl * kXlnC4Pi);
  vv = 1.0 / Dldopl;
  double c = S.xlnz * Efract * xc.Gf * Xstim / (S.rho[k] * Dndopl);
  cop = c * (double)S.fpf[(size_t)k * S.pf_ld + (xc.IxPF - 1)];
}

// a-state of the table lookup (src/voigt.f90:86-93 and 113-123)
__device__ __forceinline__ void voigt_astate(const DevSetup &S, double A, int &ia, double &ua) {
  if (A < 8.0) {
    if (A < S.a1) {
      ia = 1;
      ua = A * S.ia1;
    } else {
      ia = (int)(sqrt(A) * S.idt) + 1;
      ua = (A - S.anode[ia - 1]) * S.aida[ia - 1];
    }
  } else {
    ia = 0;
    ua = 0.0;
  }
}

// HVoigt in FP64 (src/voigt.f90:76-110) with a precomputed a-state
__device__ __forceinline__ double hvoigt64(const DevSetup &S, double A, int ia, double ua, double V) {
  double av = fabs(V);
  if (ia > 0 && av < 8.0) {
    double fv = av * S.idv;
    int iv = (int)fv;
    double tv = fv - (double)iv;
    const float *t0 = S.vtab + (size_t)(ia - 1) * TAB_LD;
    const float *t1 = t0 + TAB_LD;
    double c0 = (double)t0[iv] + tv * ((double)t0[iv + 1] - (double)t0[iv]);
    double c1 = (double)t1[iv] + tv * ((double)t1[iv + 1] - (double)t1[iv]);
    return c0 + ua * (c1 - c0);
  }
  double a2 = A * A, v2 = av * av, s = v2 + a2;
  if (s >= 196.0) return kSqrtPiInv * A * (s + 0.5) / (s * s - v2 + a2 + 0.25);
  double p = v2 - a2 - 1.5, q = 2.0 * av * A;
  double dr = av * p - A * q, di = av * q + A * p;
  return kSqrtPiInv * ((p + 0.5) * di - q * dr) / (dr * dr + di * di);
}

__device__ __forceinline__ double hvoigt64_full(const DevSetup &S, double A, double V) {
  int ia;
  double ua;
  voigt_astate(S, A, ia, ua);
  return hvoigt64(S, A, ia, ua, V);
}

// WaveIdx (src/opacity.f90:594-608)
__device__ __forceinline__ int wave_idx(const DevSetup &S, double W) {
  if (S.wmode == 0) {
    double X = fmin(fmax(W / S.dstep, 0.0), (double)(S.nwl - 2));
    return (int)X + 1;
  }
  double X = 1.0 + log(fmax((W / S.wstart) + 1.0, 2.2250738585072014e-308)) / S.rstep;
  return (int)fmin(fmax(X, 1.0), (double)(S.nwl - 1));
}

// XCD-aware tile order: consecutive block ids go round-robin over the 8 XCDs.  Each XCD
// takes runs of XCD_RUN consecutive tiles (lines shared by neighbouring tiles stay in its
// L2), and the runs go round-robin over the XCDs, so that work concentrated in part of the
// grid (a block of a bin-sorted list covers a slice of the wavelengths) still spreads over
// all of them.  The grid holds 8 * per blocks; the rows of complete runs map run by run, the
// rest contiguously per XCD: a bijection onto [0, 8 * per), every tile below ntiles visited
// exactly once.  Runs of 4 to 64 tiles take the same time; the longest keeps the most
// lines of a core zone (several 64-point tiles wide) within one L2.
constexpr int XCD_RUN = 64;
__device__ __forceinline__ int xcd_tile(int bid, int nblk) {
  const int NX = 8;
  const int per = nblk / NX;
  const int x = bid % NX, r = bid / NX;
  const int full = (per / XCD_RUN) * XCD_RUN;
  if (r < full) return ((r / XCD_RUN) * NX + x) * XCD_RUN + r % XCD_RUN;
  return full * NX + x * (per - full) + (r - full);
}

__device__ __forceinline__ void kahan_add(float &acc, float &cmp, float val) {
  float y = val - cmp;
  float t = acc + y;
  cmp = (t - acc) - y;
  acc = t;
}

// CentDer (src/solver.f90:618-647) interior element from the two adjacent steps
__device__ __forceinline__ double cent_der_i(double Dx, double Dx1, double ym, double y0, double yp) {
  double InvDx = 1.0 / Dx, InvDx1 = 1.0 / Dx1;
  double Lambda = (1.0 + Dx1 / (Dx + Dx1)) / 3.0;
  double Lambda1 = 1.0 - Lambda;
  double Der = (y0 - ym) * InvDx, Der1 = (yp - y0) * InvDx1;
  return (Der * Der1 > 0.0) ? (Der / (Lambda * Der1 + Lambda1 * Der)) * Der1 : 0.0;
}

// DELO coefficients (src/solver.f90:154-169)
__device__ __forceinline__ void delo_coef(double Dtau, double &Eps, double &Alpha, double &Beta, double &Gamma) {
  double Dtau2 = Dtau * Dtau;
  double Idtau2 = 1.0 / Dtau2;
  if (Dtau < 5e-3) {
    Eps = 1.0 - Dtau + 0.5 * Dtau2 - Dtau2 * Dtau / 6.0 + Dtau2 * Dtau2 / 24.0;
    Alpha = Dtau / 3.0 - Dtau2 / 12.0 + Dtau2 * Dtau / 60.0;
    Beta = Dtau / 3.0 - Dtau2 * 0.25 + Dtau2 * Dtau * 0.1;
    Gamma = Dtau / 3.0 - Dtau2 / 6.0 + Dtau2 * Dtau * 0.05;
  } else {
    double E = (Dtau < 100.0) ? exp(-Dtau) : 0.0;
    Eps = E;
    Alpha = (Dtau2 - 2.0 * Dtau + 2.0 - 2.0 * E) * Idtau2;
    Beta = (2.0 - (2.0 + 2.0 * Dtau + Dtau2) * E) * Idtau2;
    Gamma = (2.0 * Dtau - 4.0 + (2.0 * Dtau + 4.0) * E) * Idtau2;
  }
}

__device__ __forceinline__ double planck_src(double C1, double C2, double T) { return C2 / (exp(C1 / T) - 1.0); }

// TauStep (src/solver.f90:591-615) element i (0-based) on the column-mass scale rhox with
// the steps divided by Xmu (1 on spherical rays); km, k0, kp = opacity at points i-1, i,
// i+1 (km unused for i = 0)
__device__ __forceinline__ double tau_step_mu(const double *rhox, double xmu, int i, double km, double k0,
                                              double kp) {
  const double Dx1 = (rhox[i + 1] - rhox[i]) / xmu;
  double Yp;
  if (i == 0) {
    Yp = (kp - k0) / Dx1;
  } else {
    const double Dx = (rhox[i] - rhox[i - 1]) / xmu;
    const double Der = (k0 - km) / Dx, Der1 = (kp - k0) / Dx1;
    if (Der * Der1 > 0.0) {
      const double Lambda = (1.0 + Dx1 / (Dx + Dx1)) / 3.0;
      Yp = (Der / (Lambda * Der1 + (1.0 - Lambda) * Der)) * Der1;
    } else {
      Yp = 0.0;
    }
  }
  return Dx1 * (kp + 2.0 * k0 + Dx1 / 2.0 * Yp) / 3.0;
}

// coefficient-record metadata of retained line j (bin order): centre offset, window, centre
// pixel and He I index; XlineSetup of the line in xc
__device__ __forceinline__ void line_meta(const DevSetup &S, const DevLines &L, int list_type, int64_t j,
                                          const int *order, const int *flag, const int *il, const int *ir,
                                          const int *ic, XC &xc, double *wlc_s, int4 *meta_s) {
  int64_t i = order[j];
  int f = flag[i];
  double wlc = L.wl[i] - S.wfirst;
  xc = xline_setup(S, L, i, wlc, f & 0x6, list_type);
  wlc_s[j] = wlc;
  // window and centre in the (concatenated) grid the accumulation kernels run on
  meta_s[j] = make_int4(il[i] + S.goff, ir[i] + S.goff, ic[i] + S.goff, (f >> 8) & 0xff);
}

// FP32 batched path: far-wing expansion records (CoefA) and general-path records (CoefB)
constexpr int MAXMU = 16;
// opacity column of one grid point: continuum, or continuum + accumulated line opacity
template <typename LopT>
struct GridCol {
  const double *cont;
  const LopT *lop;
  int nwl, p;
  bool with_line;
  __device__ __forceinline__ double k(int d) const {
    double v = cont[(size_t)d * nwl + p];
    if (with_line) v = v + (double)lop[(size_t)d * nwl + p];
    return v;
  }
};

// Width-bin histogram of pass 1, aggregated over the wavefront: one atomic per distinct bin
// of the wavefront's lines (instead of one per line on a handful of hot addresses), the
// bin's largest half-width h from a wavefront maximum (h < 0: no half-width).  Every lane of
// the wavefront calls it; lanes with valid false take no part.
__device__ __forceinline__ void wave_bin_update(int *bin_count, int *bin_hmax, unsigned key, int h, bool valid) {
  const int lane = (int)(threadIdx.x & 63);
  uint64_t todo = __builtin_amdgcn_ballot_w64(valid);
  while (todo) {
    const int leader = __builtin_ctzll(todo);
    const unsigned k = (unsigned)__builtin_amdgcn_readlane((int)key, leader);
    const bool mine = valid && key == k;
    const uint64_t m = __builtin_amdgcn_ballot_w64(mine);
    int hv = mine ? h : -1;
    for (int off = 32; off > 0; off >>= 1) hv = max(hv, __shfl_xor(hv, off, 64));
    if (lane == leader) {
      atomicAdd(&bin_count[k], (int)__builtin_popcountll(m));
      if (hv >= 0) atomicMax(&bin_hmax[k], hv);
    }
    todo &= ~m;
  }
}

}  // namespace fsg
