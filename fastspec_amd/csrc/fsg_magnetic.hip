// fsg_magnetic.hip -- the polarised path: Faraday-Voigt table, MagLineopCalc pass 1 with
// the Zeeman window search, the six-profile opacity accumulation and the 4x4 Stokes
// DELO-Bezier formal solution.
//
// Reference routines restated here (file:line in the reference tree):
//   FVoigtInit, FvLine, FVoigtLine        src/voigt.f90:219-251, 302-352
//   WVoigtExact                           src/voigt.f90:356-393
//   MagLineopCalc pass 1                  src/synth.f90:1182-1273
//   MagLineOpacity1/2, RefineIdx          src/opacity.f90:391-520, 611-675
//   MagRadTranCalc field geometry         src/synth.f90:1935-2090
//   StokesPrfPP, StokesPrfSph, Solve4x4   src/solver.f90:265-394, 402-538, 541-588
#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

// ------------------------------------------------------------------------------------
// Faraday-Voigt table: H = Re w and L = Im w on the Voigt table's (v, a) grid
// ------------------------------------------------------------------------------------
// WVoigtExact: w(z) for z = |V| + i A by the Chiarella-Reichel sums (step h = 1/2) with
// the pole correction for A < 6; the A = 0 limit uses the Dawson-type sum.  Im w changes
// sign with V.
__device__ void wvoigt_exact(double A, double V, double &re, double &im) {
  const double HH = 0.5;
  double av = fabs(V);
  if (A <= 0.0) {
    int m = 2 * (int)lround(2.0 * av);
    double x0 = av - 0.25 * (double)m;
    double p = 0.0;
    for (int n = -31; n <= 31; n += 2) {
      double d = x0 - 0.25 * (double)n;
      p = p + exp(-d * d) / (double)(m + n);
    }
    re = exp(-fmin(av * av, 700.0));
    im = (2.0 / kPi) * p;
    if (V < 0.0) im = -im;
    return;
  }
  // s = 1/z + sum_n 2 z e^{-(n h)^2} / (z^2 - (n h)^2)
  double zr = av, zi = A;
  double z2r = zr * zr - zi * zi, z2i = 2.0 * zr * zi;
  double den = zr * zr + zi * zi;
  double sr = zr / den, si = -zi / den;
  for (int n = 1; n <= 16; n++) {
    double nh = n * HH;
    double e = 2.0 * exp(-nh * nh);
    double nr = e * zr, ni = e * zi;
    double dr = z2r - nh * nh, di = z2i;
    double dd = dr * dr + di * di;
    sr += (nr * dr + ni * di) / dd;
    si += (ni * dr - nr * di) / dd;
  }
  // s = i (h/pi) s
  double c = HH / kPi;
  double tr = -c * si, ti = c * sr;
  if (A < 6.0) {
    double p = 2.0 * kPi * A / HH;
    double th = 2.0 * kPi * (2.0 * av - round(2.0 * av));
    double sh = sin(0.5 * th);
    double redd = -2.0 * exp(0.5 * p) * sinh(0.5 * p) + exp(p) * 2.0 * sh * sh;
    double imdd = exp(p) * sin(th);
    // + 2 exp(A^2 - av^2 - 2 i A av) / (redd + i imdd)
    double m = 2.0 * exp(A * A - av * av);
    double er = m * cos(2.0 * A * av), ei = -m * sin(2.0 * A * av);
    double dd = redd * redd + imdd * imdd;
    tr += (er * redd + ei * imdd) / dd;
    ti += (ei * redd - er * imdd) / dd;
  }
  if (V < 0.0) ti = -ti;
  re = tr;
  im = ti;
}

// table [NVA+1][TAB_LD] of (H, L) pairs, FP32 like the reference's FvTab(2, iv, ia)
__global__ void k_fvoigt_table(float2 *tab, const double *anode, double dv) {
  int iv = blockIdx.x * blockDim.x + threadIdx.x;  // 0..NVV
  int ia = blockIdx.y;                             // 0..NVA
  if (iv > NVV) return;
  double re, im;
  wvoigt_exact(anode[ia], (double)iv * dv, re, im);
  tab[(size_t)ia * TAB_LD + iv] = make_float2((float)re, (float)im);
}

// a-state of VoigtLine/FvLine (src/voigt.f90:113-123, 302-312): unlike HVoigt there is no
// separate branch below the first a node
__device__ __forceinline__ void line_astate(const DevSetup &S, double A, int &ia, double &ua) {
  if (A < 8.0) {
    ia = (int)(sqrt(A) * S.idt) + 1;
    ua = (A - S.anode[ia - 1]) * S.aida[ia - 1];
  } else {
    ia = 0;
    ua = 0.0;
  }
}

// FVoigtLine (src/voigt.f90:315-352) in FP64
__device__ __forceinline__ void fvoigt64(const float2 *ftab, double idv, double A, int ia, double ua, double V,
                                         double &H, double &L) {
  double av = fabs(V);
  if (ia > 0 && av < 8.0) {
    double fv = av * idv;
    int iv = (int)fv;
    double tv = fv - (double)iv;
    const float2 *t0 = ftab + (size_t)(ia - 1) * TAB_LD;
    const float2 *t1 = t0 + TAB_LD;
    float2 a0 = t0[iv], a1 = t0[iv + 1], b0 = t1[iv], b1 = t1[iv + 1];
    double h0 = (double)a0.x + tv * ((double)a1.x - (double)a0.x);
    double h1 = (double)b0.x + tv * ((double)b1.x - (double)b0.x);
    double l0 = (double)a0.y + tv * ((double)a1.y - (double)a0.y);
    double l1 = (double)b0.y + tv * ((double)b1.y - (double)b0.y);
    H = h0 + ua * (h1 - h0);
    L = l0 + ua * (l1 - l0);
  } else {
    double a2 = A * A, v2 = av * av, s = v2 + a2;
    if (s >= 196.0) {
      double w = kSqrtPiInv / (s * s - v2 + a2 + 0.25);
      H = w * A * (s + 0.5);
      L = w * av * (s - 0.5);
    } else {
      double p = v2 - a2 - 1.5, q = 2.0 * av * A;
      double dr = av * p - A * q, di = av * q + A * p;
      double w = kSqrtPiInv / (dr * dr + di * di);
      H = w * ((p + 0.5) * di - q * dr);
      L = w * ((p + 0.5) * dr + q * di);
    }
  }
  L = copysign(fabs(L), V);
}

// FVoigtLine in FP32.  The two-pole fraction is evaluated from its real and imaginary
// parts rearranged without cancellation (z = v + i a, p = v^2 - a^2 - 3/2, q = 2 v a,
// s = |z|^2):  H = a (p (p + 1/2) + q^2 + v^2) / (s (p^2 + q^2)) / sqrt(pi),
// L = v (p (p + 1/2) + q^2 - a^2) / (s (p^2 + q^2)) / sqrt(pi).
__device__ __forceinline__ void fvoigt32(const float2 *ftab, float idv, float A, int ia, float ua, float V, float &H,
                                         float &L) {
  float av = fabsf(V);
  if (ia > 0 && av < 8.0f) {
    float fv = av * idv;
    int iv = (int)fv;
    float tv = fv - (float)iv;
    const float2 *t0 = ftab + (size_t)(ia - 1) * TAB_LD + iv;
    const float2 *t1 = t0 + TAB_LD;
    float2 a0 = t0[0], a1 = t0[1], b0 = t1[0], b1 = t1[1];
    float h0 = fmaf(tv, a1.x - a0.x, a0.x), h1 = fmaf(tv, b1.x - b0.x, b0.x);
    float l0 = fmaf(tv, a1.y - a0.y, a0.y), l1 = fmaf(tv, b1.y - b0.y, b0.y);
    H = fmaf(ua, h1 - h0, h0);
    L = fmaf(ua, l1 - l0, l0);
  } else {
    const float sp = (float)kSqrtPiInv;
    float a2 = A * A, v2 = av * av, s = v2 + a2;
    if (s >= 196.0f) {
      float w = sp * __builtin_amdgcn_rcpf(fmaf(s, s, a2 + 0.25f - v2));
      H = w * A * (s + 0.5f);
      L = w * av * (s - 0.5f);
    } else {
      float p = v2 - a2 - 1.5f, q = 2.0f * av * A;
      float pp = fmaf(p, p + 0.5f, q * q);
      float w = sp * __builtin_amdgcn_rcpf(s * fmaf(p, p, q * q));
      H = w * A * (pp + v2);
      L = w * av * (pp - a2);
    }
  }
  L = copysignf(fabsf(L), V);
}

// ------------------------------------------------------------------------------------
// MagLineopCalc pass 1 (src/synth.f90:1182-1273): one lane per line
// ------------------------------------------------------------------------------------
// one wavefront per line: the lanes share the line's depths (and its Zeeman components for
// the shift scaling), maxima over depths by wavefront reduction (exact, so the decisions are
// those of a serial depth loop); the search itself is wave-uniform
__device__ __forceinline__ double wave_max64(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// MagLineOpacity2 (src/opacity.f90:468-520): the Stokes I bound max(Ps/2, Pp/2 + Ps/4)
// LinCop / ContOpac at Wl, maximum over the lane's depths.  The component shifts are the
// scaled REAL(4) shifts of this field (shift * Bsplit * lambda^2), formed here as pass 1 stores
// them.
__device__ double mag_bound_lane(const DevSetup &S, const XC &xc, const MagLine &Z, int64_t i, double wlc, int ihe,
                                 int iwl, double dW, double Wl, double bsplit, double wl2, int lane) {
  const int N = S.ntau;
  const int q0 = Z.pi_off[i], q1 = Z.pi_off[i + 1], r0 = Z.si_off[i], r1 = Z.si_off[i + 1];
  double mx = -INFINITY;
  const double dWl0 = Wl - wlc;
  for (int k = lane; k < N; k += 64) {
    double cop, vv, av;
    xline_depth(S, k, xc, cop, vv, av);
    double dWl = dWl0;
    if (ihe > 0) dWl = dWl0 - S.he_shf[(size_t)(ihe - 1) * S.ntp + k];
    int ia;
    double ua;
    line_astate(S, av, ia, ua);
    double Pp = 0.0, Ps = 0.0;
    for (int q = q0; q < q1; q++)
      Pp = Pp + (double)Z.str_pi[q] *
                    hvoigt64(S, av, ia, ua, (dWl - (double)(float)((double)Z.spl_pi[q] * bsplit * wl2)) * vv);
    for (int q = r0; q < r1; q++)
      Ps = Ps + (double)Z.str_si[q] *
                    hvoigt64(S, av, ia, ua, (dWl - (double)(float)((double)Z.spl_si[q] * bsplit * wl2)) * vv);
    double lo = fmax(0.5 * Ps, 0.5 * Pp + 0.25 * Ps) * cop;
    if (S.motype == 1) lo = lo / S.cstd[k];
    const double *ck = S.contT + (size_t)k * S.nwl;
    double kc = ck[iwl - 1] * (1.0 - dW) + ck[iwl] * dW;
    mx = fmax(mx, lo / kc);
  }
  return mx;
}

// MagLineopCalc pass 1: blocks of MAGP1_WAVES wavefronts, one line per wavefront
constexpr int MAGP1_WAVES = 4;

__global__ void __launch_bounds__(64 * MAGP1_WAVES) k_magpass1(DevSetup S, DevLines L, MagLine Z, double bsplit,
                                                               int *flag_out, int *il_out, int *ir_out, int *ic_out,
                                                               unsigned *key_out, int *bin_count, int *bin_hmax,
                                                               int batched) {
  const int lane = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * MAGP1_WAVES + (threadIdx.x >> 6);
  if (i >= L.n) return;  // wave-uniform
  const int N = S.ntau;
  int flag = initial_flag(L, i, 2);
  const double wlc = L.wl[i] - S.wfirst;
  // the Zeeman pattern in Angstrom for this field: shift * Bsplit * lambda^2, stored REAL(4)
  const double wl2 = (wlc + S.wfirst) * (wlc + S.wfirst);
  double SplMax = 0.0;
  for (int q = Z.pi_off[i] + lane; q < Z.pi_off[i + 1]; q += 64) {
    float v = (float)((double)Z.spl_pi[q] * bsplit * wl2);
    Z.sp_pi[q] = v;
    SplMax = fmax(SplMax, fabs((double)v));
  }
  for (int q = Z.si_off[i] + lane; q < Z.si_off[i + 1]; q += 64) {
    float v = (float)((double)Z.spl_si[q] * bsplit * wl2);
    Z.sp_si[q] = v;
    SplMax = fmax(SplMax, fabs((double)v));
  }
  SplMax = wave_max64(SplMax);
  XC xc = xline_setup(S, L, i, wlc, flag, 2);
  if (xc.IxHe > 0) flag |= (xc.IxHe << 8);
  const int ihe = xc.IxHe;
  const int iwl = wave_idx(S, wlc);
  const double dW = (wlc - S.wave[iwl - 1]) / (S.wave[iwl] - S.wave[iwl - 1]);
  // field-free centre test (LineOpacity2 at the centre); there is no bound test here
  double maxc = -INFINITY;
  for (int k = lane; k < N; k += 64) {
    double cop, vv, av;
    xline_depth(S, k, xc, cop, vv, av);
    double dWl = (ihe > 0) ? 0.0 - S.he_shf[(size_t)(ihe - 1) * S.ntp + k] : 0.0;
    double lo = hvoigt64_full(S, av, dWl * vv) * cop;
    if (S.motype == 1) lo = lo / S.cstd[k];
    const double *ck = S.contT + (size_t)k * S.nwl;
    double kc = ck[iwl - 1] * (1.0 - dW) + ck[iwl] * dW;
    maxc = fmax(maxc, lo / kc);
  }
  maxc = wave_max64(maxc);
  auto bound = [&](double Wl) {
    return wave_max64(mag_bound_lane(S, xc, Z, i, wlc, ihe, iwl, dW, Wl, bsplit, wl2, lane));
  };
  int il = 1, ir = S.nwl;
  const bool weak = maxc < S.eps_opac;
  if (weak) {
    flag |= 1;
  } else {
    // doubling search of the half width beyond the outermost component, then RefineIdx:
    // bisection in index space between the last two trial widths
    const double wabs = wlc + S.wfirst;
    for (int idw = 0;; idw++) {
      const double dWo = (kDWstart * pow(kDWstep, (double)idw)) * wabs / 5e3;
      if (bound(wlc + SplMax + dWo) < S.eps_opac) {
        int isw, ikw;
        if (wlc > 0.5 * (S.wave[0] + S.wave[S.nwl - 1])) {
          isw = -1;
          ikw = 0;
        } else {
          isw = 1;
          ikw = 1;
        }
        int ia, ib;
        if (idw == 0) {
          ia = wave_idx(S, wlc + isw * SplMax) + ikw;
        } else {
          const double d1 = (kDWstart * pow(kDWstep, (double)(idw - 1))) * wabs / 5e3;
          ia = wave_idx(S, wlc + isw * (SplMax + d1)) + ikw;
        }
        ib = wave_idx(S, wlc + isw * (SplMax + dWo)) + ikw;
        while (abs(ia - ib) > 1) {
          int im = (ia + ib) / 2;
          if (bound(S.wave[im - 1]) < S.eps_opac) ib = im;
          else ia = im;
        }
        if (isw == 1) {
          ir = ia;
          il = wave_idx(S, 2.0 * wlc - S.wave[ia - 1]);
        } else {
          il = ia;
          ir = wave_idx(S, 2.0 * wlc - S.wave[ia - 1]) + 1;
        }
        break;
      } else if (wlc - SplMax - dWo < S.wave[0] && wlc + SplMax + dWo > S.wave[S.nwl - 1]) {
        break;
      }
    }
  }
  if (lane != 0) return;
  flag_out[i] = flag;
  il_out[i] = il;
  ir_out[i] = ir;
  ic_out[i] = iwl;
  unsigned key;
  if (weak) {
    key = WEAK_KEY;
  } else {
    int h = max(max(iwl - il, ir - iwl), 0);
    key = batched ? (h == 0 ? 0u : (unsigned)(32 - __clz(h))) : 0u;
    atomicMax(&bin_hmax[key], h);
  }
  atomicAdd(&bin_count[key], 1);
  key_out[i] = key;
}

// ------------------------------------------------------------------------------------
// per-(line, depth) records of the polarised opacity kernel, lines in bin order, depths
// padded to ntp with zero records (zero strength)
// ------------------------------------------------------------------------------------
template <typename Real>
struct MagCoef {
  Real sc, a, xn, ua;  // LinVvgt (1/Doppler width), LinAvgt, LinCop, a-interpolation weight
  int ia, pad;         // a cell of the table (0 = outside the table box)
};

template <typename Real>
__global__ void k_magfill(DevSetup S, DevLines L, MagLine Z, int64_t nret, const int *order, const int *flag,
                          const int *il, const int *ir, const int *ic, MagCoef<Real> *coef, double *wlc_s,
                          int4 *meta_s, int4 *comp_s) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nret) return;
  const int64_t i = order[j];
  const int f = flag[i];
  const double wlc = L.wl[i] - S.wfirst;
  XC xc = xline_setup(S, L, i, wlc, f & 0x6, 2);
  wlc_s[j] = wlc;
  meta_s[j] = make_int4(il[i], ir[i], ic[i], (f >> 8) & 0xff);
  comp_s[j] = make_int4(Z.pi_off[i], Z.pi_off[i + 1], Z.si_off[i], Z.si_off[i + 1]);
  MagCoef<Real> *c = coef + (size_t)j * S.ntp;
  for (int k = 0; k < S.ntp; k++) {
    MagCoef<Real> r{};
    if (k < S.ntau) {
      double cop, vv, av, ua;
      int ia;
      xline_depth(S, k, xc, cop, vv, av);
      line_astate(S, av, ia, ua);
      r.sc = (Real)vv;
      r.a = (Real)av;
      r.xn = (Real)cop;
      r.ua = (Real)ua;
      r.ia = ia;
    }
    c[k] = r;
  }
}

// ------------------------------------------------------------------------------------
// six-profile opacity (MagLineOpacity1, src/opacity.f90:391-464).  A block owns MTILE = 64
// wavelength points (one per lane) and all depths, one wavefront per DPW depths.  The
// accumulators are (pi, blue sigma, red sigma) x (absorption H, dispersion L); FP32 sums
// are compensated.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void acc_add(float &s, float &c, float v) { kahan_add(s, c, v); }
__device__ __forceinline__ void acc_add(double &s, double &, double v) { s += v; }

template <typename Real>
__device__ __forceinline__ void fvoigt(const float2 *ftab, double idv, Real A, int ia, Real ua, Real V, Real &H,
                                       Real &L) {
  if constexpr (std::is_same<Real, float>::value) fvoigt32(ftab, (float)idv, A, ia, ua, V, H, L);
  else fvoigt64(ftab, idv, A, ia, ua, V, H, L);
}

template <typename Real>
__global__ void __launch_bounds__(1024) k_magopacity(DevSetup S, const float2 *__restrict__ ftab, int ntiles,
                                                     int nbins, int rstride, const int2 *__restrict__ rng,
                                                     const MagCoef<Real> *__restrict__ coef,
                                                     const double *__restrict__ wlc_s, const int4 *__restrict__ meta_s,
                                                     const int4 *__restrict__ comp_s, const float *__restrict__ sp_pi,
                                                     const float *__restrict__ str_pi, const float *__restrict__ sp_si,
                                                     const float *__restrict__ str_si, Real *__restrict__ mlop) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const int N = S.ntau, Nw = S.nwl;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int p0 = t * MTILE + 1, p1 = min(p0 + MTILE - 1, Nw);
  const int p = p0 + lane;
  const bool valid = p <= Nw;
  const double Wp = valid ? S.wave[p - 1] : 0.0;
  const int d0 = wave * DPW;
  Real acc[6][DPW], cmp[6][DPW];
#pragma unroll
  for (int q = 0; q < 6; q++)
#pragma unroll
    for (int k = 0; k < DPW; k++) {
      acc[q][k] = 0;
      cmp[q][k] = 0;
    }
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * rstride + b];
    for (int j = r.x; j < r.y; j++) {
      const int4 m = meta_s[j];
      if (m.y < p0 || m.x > p1) continue;
      if (!valid || p < m.x || p > m.y) continue;
      const int4 cs = comp_s[j];
      const int nsh = (cs.w - cs.z) / 2;
      const double dl = Wp - wlc_s[j];
      const MagCoef<Real> *c = coef + (size_t)j * S.ntp + d0;
#pragma unroll
      for (int k = 0; k < DPW; k++) {
        const MagCoef<Real> ck = c[k];
        const Real sc = ck.sc;
        Real vb = (Real)dl * sc;
        if (m.w > 0) vb = vb - (Real)S.he_shf[(size_t)(m.w - 1) * S.ntp + d0 + k] * sc;
        for (int q = cs.x; q < cs.y; q++) {
          Real H, Lf;
          fvoigt<Real>(ftab, S.idv, ck.a, ck.ia, ck.ua, vb - (Real)sp_pi[q] * sc, H, Lf);
          const Real st = ck.xn * (Real)str_pi[q];
          acc_add(acc[0][k], cmp[0][k], st * H);
          acc_add(acc[3][k], cmp[3][k], st * Lf);
        }
        for (int q = cs.z; q < cs.z + nsh; q++) {
          Real H, Lf;
          fvoigt<Real>(ftab, S.idv, ck.a, ck.ia, ck.ua, vb - (Real)sp_si[q] * sc, H, Lf);
          const Real st = ck.xn * (Real)str_si[q];
          acc_add(acc[1][k], cmp[1][k], st * H);
          acc_add(acc[4][k], cmp[4][k], st * Lf);
        }
        for (int q = cs.z + nsh; q < cs.w; q++) {
          Real H, Lf;
          fvoigt<Real>(ftab, S.idv, ck.a, ck.ia, ck.ua, vb - (Real)sp_si[q] * sc, H, Lf);
          const Real st = ck.xn * (Real)str_si[q];
          acc_add(acc[2][k], cmp[2][k], st * H);
          acc_add(acc[5][k], cmp[5][k], st * Lf);
        }
      }
    }
  }
  if (!valid) return;
#pragma unroll
  for (int k = 0; k < DPW; k++) {
    const int d = d0 + k;
    if (d >= N) continue;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      Real v = acc[q][k];
      if constexpr (std::is_same<Real, float>::value) v = v - cmp[q][k];
      if (S.motype == 1) v = v / (Real)S.cstd[d];
      const size_t o = ((size_t)q * N + d) * Nw + (p - 1);
      mlop[o] = mlop[o] + v;
    }
  }
}

// ------------------------------------------------------------------------------------
// Stokes DELO-Bezier (StokesPrfPP src/solver.f90:265-394, StokesPrfSph 402-538): one lane
// per (mu, lambda), lambda fastest; FP64, the recursion streamed from the deepest point
// upward over a sliding window of three points.  A ray point samples one model layer: the
// layer itself in plane-parallel geometry, the ray map of SphereInit on spherical rays.
// The normalised propagation matrix K' has six distinct entries (PhiQ, PhiU, PhiV, PsiQ,
// PsiU, PsiV) and zeros on the diagonal; CentDer is odd in its data, so the derivative of
// K' is K' built from the derivatives of the six entries.
// ------------------------------------------------------------------------------------
struct Geo {  // sin^2 gamma, cos^2 gamma, cos gamma, sin 2 chi, cos 2 chi
  double sphi2, cphi2, cphi, s2psi, c2psi;
};

struct SDepth {
  double phi;    // PhiI
  double k6[6];  // PhiQ, PhiU, PhiV, PsiQ, PsiU, PsiV (divided by PhiI)
  double s4[4];  // source vector S (1, PhiQ, PhiU, PhiV)
};

// the opacities of one wavelength point seen along one ray
template <typename LopT, typename MlopT>
struct StokesRay {
  const DevSetup *S;
  const LopT *lop;
  const MlopT *mlop;
  size_t p;
  double C1, C2;
  const int *map;  // ray point -> layer (nullptr: identity)
  const Geo *geo;  // geometry of point j at geo[j * gs]
  int gs;
  __device__ __forceinline__ int layer(int j) const { return map ? map[j] : j; }
  __device__ __forceinline__ double mval(int q, int d) const {
    return (double)mlop[((size_t)q * S->ntau + d) * S->nwl + p];
  }
  __device__ __forceinline__ double phi_i(int d, const Geo &g, double m0, double m1, double m2) const {
    return 0.5 * (m0 * g.sphi2 + 0.5 * (m2 + m1) * (1.0 + g.cphi2)) + (double)lop[(size_t)d * S->nwl + p] +
           S->contT[(size_t)d * S->nwl + p];
  }
  __device__ __forceinline__ double phi(int j) const {
    const int d = layer(j);
    return phi_i(d, geo[(size_t)j * gs], mval(0, d), mval(1, d), mval(2, d));
  }
  __device__ __forceinline__ void at(int j, SDepth &o) const {
    const int d = layer(j);
    const Geo g = geo[(size_t)j * gs];
    const double Php = mval(0, d), Phb = mval(1, d), Phr = mval(2, d);
    const double Psp = mval(3, d), Psb = mval(4, d), Psr = mval(5, d);
    o.phi = phi_i(d, g, Php, Phb, Phr);
    const double InvPhiI = 1.0 / o.phi;
    const double Sabs = 0.5 * (Php - 0.5 * (Phr + Phb)) * g.sphi2;
    o.k6[0] = Sabs * g.c2psi * InvPhiI;
    o.k6[1] = Sabs * g.s2psi * InvPhiI;
    o.k6[2] = 0.5 * (Phr - Phb) * g.cphi * InvPhiI;
    const double Sdis = 0.5 * (Psp - 0.5 * (Psr + Psb)) * g.sphi2;
    o.k6[3] = Sdis * g.c2psi * InvPhiI;
    o.k6[4] = Sdis * g.s2psi * InvPhiI;
    o.k6[5] = 0.5 * (Psr - Psb) * g.cphi * InvPhiI;
    const double Src = planck_src(C1, C2, S->temp[d]);
    o.s4[0] = Src;
    o.s4[1] = Src * o.k6[0];
    o.s4[2] = Src * o.k6[1];
    o.s4[3] = Src * o.k6[2];
  }
};

// K' (row-major) from the six entries
__device__ __forceinline__ void kmat(const double *k6, double *K) {
  K[0] = 0.0;    K[1] = k6[0];  K[2] = k6[1];   K[3] = k6[2];
  K[4] = k6[0];  K[5] = 0.0;    K[6] = k6[5];   K[7] = -k6[4];
  K[8] = k6[1];  K[9] = -k6[5]; K[10] = 0.0;    K[11] = k6[3];
  K[12] = k6[2]; K[13] = k6[4]; K[14] = -k6[3]; K[15] = 0.0;
}

__device__ __forceinline__ void mv4(const double *A, const double *x, double *y) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 4; k++) s += A[i * 4 + k] * x[k];
    y[i] = s;
  }
}

// K'^2 + K'_d, row-major
__device__ __forceinline__ void ksq_plus(const double *k6, const double *kd6, double *X) {
  double K[16], KD[16];
  kmat(k6, K);
  kmat(kd6, KD);
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 4; k++) s += K[i * 4 + k] * K[k * 4 + j];
      X[i * 4 + j] = s + KD[i * 4 + j];
    }
}

// Solve4x4 (src/solver.f90:541-588): Gaussian elimination with partial pivoting.  Row
// exchanges are done by selects so that every array index stays static (registers).
__device__ __forceinline__ void solve4x4(const double *A, const double *B, double *X) {
  double M[4][4], R[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
#pragma unroll
    for (int j = 0; j < 4; j++) M[i][j] = A[i * 4 + j];
    R[i] = B[i];
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    int ip = k;
    double pv = fabs(M[k][k]);
#pragma unroll
    for (int i = k + 1; i < 4; i++)
      if (fabs(M[i][k]) > pv) {
        pv = fabs(M[i][k]);
        ip = i;
      }
#pragma unroll
    for (int i = k + 1; i < 4; i++) {
      const bool sw = (ip == i);
#pragma unroll
      for (int j = k; j < 4; j++) {
        double a = M[k][j], b = M[i][j];
        M[k][j] = sw ? b : a;
        M[i][j] = sw ? a : b;
      }
      double a = R[k], b = R[i];
      R[k] = sw ? b : a;
      R[i] = sw ? a : b;
    }
    const double inv = 1.0 / M[k][k];
#pragma unroll
    for (int i = k + 1; i < 4; i++) {
      const double f = M[i][k] * inv;
#pragma unroll
      for (int j = k + 1; j < 4; j++) M[i][j] = M[i][j] - f * M[k][j];
      R[i] = R[i] - f * R[k];
    }
  }
  X[3] = R[3] / M[3][3];
  X[2] = (R[2] - M[2][3] * X[3]) / M[2][2];
  X[1] = (R[1] - M[1][2] * X[2] - M[1][3] * X[3]) / M[1][1];
  X[0] = (R[0] - M[0][1] * X[1] - M[0][2] * X[2] - M[0][3] * X[3]) / M[0][0];
}

// geo: [nmu] angles (plane-parallel) or [nmu][2 ntau] angles per ray point (spherical)
template <typename LopT, typename MlopT>
__global__ void __launch_bounds__(64) k_stokes(DevSetup S, const LopT *__restrict__ lop,
                                               const MlopT *__restrict__ mlop, const Geo *__restrict__ geo,
                                               double *__restrict__ spz) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int M = S.nmu, Nw = S.nwl, L = 2 * S.ntau;
  if (gid >= (int64_t)Nw * M) return;
  const int im = (int)(gid / Nw), p = (int)(gid - (int64_t)im * Nw);
  const double Wl = S.wave[p] + S.wfirst;
  const bool sph = S.sph != 0;
  // the ray: its points, column mass and the Xmu of the optical depth steps
  const int N = sph ? S.ntau_mu[im] : S.ntau;
  const double *rhox = sph ? S.rhox_mu + (size_t)im * L : S.rhox;
  const double xmu = sph ? 1.0 : S.xmu[im];
  const bool graz = sph && S.graz_mu[im] != 0;
  const StokesRay<LopT, MlopT> col{&S,
                                   lop,
                                   mlop,
                                   (size_t)p,
                                   kPlkC1 / Wl,
                                   exp(50.7649141 - 5.0 * log(Wl)),
                                   sph ? S.map_mu + (size_t)im * L : nullptr,
                                   sph ? geo + (size_t)im * L : geo + im,
                                   sph ? 1 : 0};
  // window: w0 = point i-1, w1 = point i, w2 = point i+1; dt_i = Deltau(i), dt_im1 = Deltau(i-1)
  SDepth w0, w1, w2;
  int i = N - 2;
  col.at(N - 1, w2);
  col.at(N - 2, w1);
  col.at(N - 3, w0);
  double dt_i = tau_step_mu(rhox, xmu, i, w0.phi, w1.phi, w2.phi);
  double dt_im1 = tau_step_mu(rhox, xmu, i - 1, (i >= 2) ? col.phi(i - 2) : 0.0, w0.phi, w1.phi);
  // derivatives at the deepest point (CentDer's one-sided last point)
  double kd_ip1[6], sd_ip1[4];
#pragma unroll
  for (int q = 0; q < 6; q++) kd_ip1[q] = (w2.k6[q] - w1.k6[q]) / dt_i;
#pragma unroll
  for (int q = 0; q < 4; q++) sd_ip1[q] = (w2.s4[q] - w1.s4[q]) / dt_i;
  double XKB[16];
  ksq_plus(w2.k6, kd_ip1, XKB);
  // thermalised and unpolarised at the deepest layer; grazing rays enter with zero
  double X2[4] = {graz ? 0.0 : w2.s4[0], 0.0, 0.0, 0.0};
  for (; i >= 0; i--) {
    double kd_i[6], sd_i[4];
    if (i == 0) {
#pragma unroll
      for (int q = 0; q < 6; q++) kd_i[q] = (w2.k6[q] - w1.k6[q]) / dt_i;
#pragma unroll
      for (int q = 0; q < 4; q++) sd_i[q] = (w2.s4[q] - w1.s4[q]) / dt_i;
    } else {
#pragma unroll
      for (int q = 0; q < 6; q++) kd_i[q] = cent_der_i(dt_im1, dt_i, w0.k6[q], w1.k6[q], w2.k6[q]);
#pragma unroll
      for (int q = 0; q < 4; q++) sd_i[q] = cent_der_i(dt_im1, dt_i, w0.s4[q], w1.s4[q], w2.s4[q]);
    }
    const double Dtau = dt_i, Dtau05 = 0.5 * Dtau;
    double Eps, Alpha, Beta, Gamma;
    delo_coef(Dtau, Eps, Alpha, Beta, Gamma);
    Gamma = Gamma * 0.5;
    double K[16], K1[16], XKA[16], AA[16], QA[16];
    kmat(w1.k6, K);
    kmat(w2.k6, K1);
#pragma unroll
    for (int q = 0; q < 16; q++) XKA[q] = XKB[q];
    ksq_plus(w1.k6, kd_i, XKB);
#pragma unroll
    for (int q = 0; q < 16; q++) {
      QA[q] = Dtau05 * XKA[q] + (Dtau05 - 1.0) * K1[q];
      const double qb = -Dtau05 * XKB[q] - (Dtau05 + 1.0) * K[q];
      AA[q] = ((q % 5) == 0 ? 1.0 : 0.0) + Alpha * K[q] - Gamma * qb;
    }
    const double GD = Gamma * Dtau05;
    double v1[4], v2[4], v3[4], v4[4], X1[4];
    mv4(K1, X2, v1);
    mv4(QA, X2, v2);
    mv4(K1, w2.s4, v3);
    mv4(K, w1.s4, v4);
#pragma unroll
    for (int q = 0; q < 4; q++)
      X1[q] = Eps * X2[q] - Beta * v1[q] + Gamma * v2[q] + (Beta + Gamma) * w2.s4[q] - GD * v3[q] +
              (Alpha + Gamma) * w1.s4[q] + GD * v4[q] + GD * (sd_i[q] - sd_ip1[q]);
    solve4x4(AA, X1, X2);
    // move the window up one point
#pragma unroll
    for (int q = 0; q < 6; q++) kd_ip1[q] = kd_i[q];
#pragma unroll
    for (int q = 0; q < 4; q++) sd_ip1[q] = sd_i[q];
    w2 = w1;
    w1 = w0;
    dt_i = dt_im1;
    if (i >= 2) {
      col.at(i - 2, w0);
      dt_im1 = tau_step_mu(rhox, xmu, i - 2, (i >= 3) ? col.phi(i - 3) : 0.0, w0.phi, w1.phi);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++) spz[((size_t)q * M + im) * Nw + p] = X2[q];
}

// flux sum of Stokes I (src/fastspec.f90:437-440 with src/convol.f90:233-234)
__global__ void k_stokes_flux(DevSetup S, const double *spz, double *sfl) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= S.nwl) return;
  double f = 0.0;
  for (int m = 0; m < S.nmu; m++) f = f + spz[(size_t)m * S.nwl + p] * S.wmu[m];
  sfl[p] = f;
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
static inline unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_fvoigt_table(float2 *tab, const double *anode, double dv, hipStream_t st) {
  dim3 grid(nb(NVV + 1, 256), NVA + 1);
  k_fvoigt_table<<<grid, 256, 0, st>>>(tab, anode, dv);
}

void launch_magpass1(const DevSetup &S, const DevLines &L, const MagLine &Z, double bsplit, int *flag, int *il,
                     int *ir, int *ic, unsigned *key, int *bin_count, int *bin_hmax, int batched, hipStream_t st) {
  if (L.n == 0) return;
  k_magpass1<<<nb(L.n, MAGP1_WAVES), 64 * MAGP1_WAVES, 0, st>>>(S, L, Z, bsplit, flag, il, ir, ic, key, bin_count,
                                                              bin_hmax, batched);
}

size_t magcoef_size(int fp32) { return fp32 ? sizeof(MagCoef<float>) : sizeof(MagCoef<double>); }

void launch_magfill(const DevSetup &S, const DevLines &L, const MagLine &Z, int64_t nret, const int *order,
                    const int *flag, const int *il, const int *ir, const int *ic, void *coef, int fp32, double *wlc_s,
                    int4 *meta_s, int4 *comp_s, hipStream_t st) {
  if (!nret) return;
  if (fp32)
    k_magfill<float><<<nb(nret, 128), 128, 0, st>>>(S, L, Z, nret, order, flag, il, ir, ic, (MagCoef<float> *)coef,
                                                    wlc_s, meta_s, comp_s);
  else
    k_magfill<double><<<nb(nret, 128), 128, 0, st>>>(S, L, Z, nret, order, flag, il, ir, ic,
                                                     (MagCoef<double> *)coef, wlc_s, meta_s, comp_s);
}

void launch_magopacity(const DevSetup &S, const float2 *ftab, int ntiles, int nbins, int rstride, const int2 *rng,
                       const void *coef, const double *wlc_s, const int4 *meta_s, const int4 *comp_s,
                       const MagLine &Z, void *mlop, int fp32, hipStream_t st) {
  if (ntiles <= 0 || nbins <= 0) return;
  const int nw = S.ntp / DPW;
  const int nblk8 = 8 * ((ntiles + 7) / 8);
  if (fp32)
    k_magopacity<float><<<nblk8, 64 * nw, 0, st>>>(S, ftab, ntiles, nbins, rstride, rng, (const MagCoef<float> *)coef,
                                                   wlc_s, meta_s, comp_s, Z.sp_pi, Z.str_pi, Z.sp_si, Z.str_si,
                                                   (float *)mlop);
  else
    k_magopacity<double><<<nblk8, 64 * nw, 0, st>>>(S, ftab, ntiles, nbins, rstride, rng,
                                                    (const MagCoef<double> *)coef, wlc_s, meta_s, comp_s, Z.sp_pi,
                                                    Z.str_pi, Z.sp_si, Z.str_si, (double *)mlop);
}

void launch_stokes(const DevSetup &S, const void *lop, const void *mlop, int fp32, const double *geo, double *spz,
                   double *sfl, hipStream_t st) {
  const int64_t n = (int64_t)S.nwl * S.nmu;
  const Geo *g = (const Geo *)geo;
  if (fp32)
    k_stokes<float, float><<<nb(n, 64), 64, 0, st>>>(S, (const float *)lop, (const float *)mlop, g, spz);
  else
    k_stokes<double, double><<<nb(n, 64), 64, 0, st>>>(S, (const double *)lop, (const double *)mlop, g, spz);
  k_stokes_flux<<<nb(S.nwl, 256), 256, 0, st>>>(S, spz, sfl);
}

}  // namespace fsg
