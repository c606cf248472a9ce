// fsg_adjoint.hip -- reverse mode of the plane-parallel formal solution with the flux sums
// and of the line opacity: the gradients of a loss L(SfL, SfC) with respect to LineOpac,
// ContOpac, the temperature of each depth, log gf of each line and the EOS number
// densities FractionPF, at fixed host inputs (continuum, EOS, electron density).
//
// Forward routines differentiated here (file:line in the reference tree):
//   LocPrfPP, TauStep, CentDer, DELO coefficients   src/solver.f90:123-179, 591-647
//   flux sums of DiskInteg                          src/convol.f90:233-234
//   XlineDepth, HVoigt, LineOpacity1                src/opacity.f90:239-346, src/voigt.f90:76-110
#include <algorithm>

#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

// ------------------------------------------------------------------------------------
// derivatives of the forward pieces
// ------------------------------------------------------------------------------------
// DELO-Bezier coefficients and their derivatives with respect to Dtau (src/solver.f90:154-169)
__device__ __forceinline__ void delo_coef_d(double x, double &E, double &A, double &B, double &G, double &dE,
                                            double &dA, double &dB, double &dG) {
  delo_coef(x, E, A, B, G);
  const double x2 = x * x;
  if (x < 5e-3) {
    dE = -1.0 + x - 0.5 * x2 + x2 * x / 6.0;
    dA = 1.0 / 3.0 - x / 6.0 + x2 / 20.0;
    dB = 1.0 / 3.0 - 0.5 * x + 0.3 * x2;
    dG = 1.0 / 3.0 - x / 3.0 + 0.15 * x2;
  } else {
    dE = (x < 100.0) ? -E : 0.0;
    const double ix2 = 1.0 / x2;
    dA = (2.0 * x - 2.0 - 2.0 * dE) * ix2 - 2.0 * A / x;
    dB = (-(2.0 + 2.0 * x) * E - (2.0 + 2.0 * x + x2) * dE) * ix2 - 2.0 * B / x;
    dG = (2.0 + 2.0 * E + (2.0 * x + 4.0) * dE) * ix2 - 2.0 * G / x;
  }
}

// the monotone harmonic-mean slope Yp = Der Der1 / (l Der1 + (1 - l) Der) (zero unless the
// one-sided slopes agree in sign) and its partial derivatives
struct Slope {
  double yp, d_der, d_der1, d_l;
};
__device__ __forceinline__ Slope mono_slope(double der, double der1, double l) {
  Slope s{0.0, 0.0, 0.0, 0.0};
  if (der * der1 > 0.0) {
    const double Q = l * der1 + (1.0 - l) * der, iq2 = 1.0 / (Q * Q);
    s.yp = (der / Q) * der1;
    s.d_der = l * der1 * der1 * iq2;
    s.d_der1 = (1.0 - l) * der * der * iq2;
    s.d_l = -der * der1 * (der1 - der) * iq2;
  }
  return s;
}

// TauStep element j (Xmu = 1): d Deltau(j) / d kappa at points j-1, j, j+1
__device__ __forceinline__ void tau_step_grad(const double *rhox, int j, double km, double k0, double kp, double &gm,
                                              double &g0, double &gp) {
  const double Dx1 = rhox[j + 1] - rhox[j];
  if (j == 0) {  // Yp = (kp - k0) / Dx1: Deltau = Dx1 (kp + k0) / 2 up to rounding
    gm = 0.0;
    g0 = Dx1 * (2.0 - 0.5) / 3.0;
    gp = Dx1 * (1.0 + 0.5) / 3.0;
    return;
  }
  const double Dx = rhox[j] - rhox[j - 1];
  const double der = (k0 - km) / Dx, der1 = (kp - k0) / Dx1;
  const double l = (1.0 + Dx1 / (Dx + Dx1)) / 3.0;
  const Slope s = mono_slope(der, der1, l);
  const double h = Dx1 * Dx1 / 6.0;  // d Deltau / d Yp
  gm = h * (-s.d_der / Dx);
  g0 = 2.0 * Dx1 / 3.0 + h * (s.d_der / Dx - s.d_der1 / Dx1);
  gp = Dx1 / 3.0 + h * (s.d_der1 / Dx1);
}

// CentDer interior element Sd(i) = f(Dx = Deltau(i-1), Dx1 = Deltau(i); y(i-1), y(i), y(i+1))
// and its partial derivatives
struct CentGrad {
  double dx, dx1, ym, y0, yp;
};
__device__ __forceinline__ CentGrad cent_der_grad(double Dx, double Dx1, double ym, double y0, double yp) {
  const double der = (y0 - ym) / Dx, der1 = (yp - y0) / Dx1;
  const double l = (1.0 + Dx1 / (Dx + Dx1)) / 3.0;
  const Slope s = mono_slope(der, der1, l);
  const double ss = (Dx + Dx1) * (Dx + Dx1);
  CentGrad g;
  g.dx = s.d_der * (-der / Dx) + s.d_l * (-Dx1 / (3.0 * ss));
  g.dx1 = s.d_der1 * (-der1 / Dx1) + s.d_l * (Dx / (3.0 * ss));
  g.ym = -s.d_der / Dx;
  g.y0 = s.d_der / Dx - s.d_der1 / Dx1;
  g.yp = s.d_der1 / Dx1;
  return g;
}

// ------------------------------------------------------------------------------------
// reverse mode of LocPrfPP + flux for one opacity column of one wavelength point
// ------------------------------------------------------------------------------------
// Forward, from the deepest point up (as LocPrfPP): Deltau(i) by TauStep, Sd = CentDer(S),
// U(i) = (S(i) + Deltau Sd(i) / 2 + S(i+1) - Deltau Sd(i+1) / 2) / 2 and per mu
// Y <- Y eps + alpha S(i) + beta S(i+1) + gamma U(i) with the coefficients at Deltau / mu;
// F = sum_mu Wmu Y.  The intensities entering each step are kept in scratch (the
// checkpoint of the reverse sweep).  Reverse, from the top down: adjoints of Y (per mu), U,
// Sd, Deltau, kappa and S, each emitted as soon as no later step can add to it.
// Scratch rows (stride ld, one column per lane): Yin [(N-1) M], Deltau [N], Sd [N], S [N];
// the adjoints of S go to gsrc (rows of stride ld).  MM bounds the angle count (the per-angle
// registers of a lane).
template <int MM, typename Col>
__device__ void rt_vjp_column(const DevSetup &S, const Col &col, double gF, const double *imu, double C1, double C2,
                              double *scr, size_t ld, double *gkap, size_t gstride, double *gsrc) {
  const int N = S.ntau, M = S.nmu;
  double *Yin = scr, *dts = scr + (size_t)(N - 1) * M * ld, *sds = dts + (size_t)N * ld, *srcs = sds + (size_t)N * ld;
  // ---- forward sweep (streamed as locprf_pp), storing the checkpoint
  {
    int i = N - 2;
    double S_ip1 = planck_src(C1, C2, S.temp[N - 1]);
    double S_i = planck_src(C1, C2, S.temp[N - 2]);
    double S_im1 = planck_src(C1, C2, S.temp[N - 3]);
    srcs[(size_t)(N - 1) * ld] = S_ip1;
    double dt_i = tau_step_mu(S.rhox, 1.0, i, col.k(i - 1), col.k(i), col.k(i + 1));
    double dt_im1 = tau_step_mu(S.rhox, 1.0, i - 1, (i >= 2) ? col.k(i - 2) : 0.0, col.k(i - 1), col.k(i));
    double Sd_ip1 = (S_ip1 - S_i) / dt_i;
    sds[(size_t)(N - 1) * ld] = Sd_ip1;
    double Y[MM];
    for (int m = 0; m < M; m++) Y[m] = S_ip1;
    for (; i >= 0; i--) {
      const double Sd_i = (i == 0) ? (S_ip1 - S_i) / dt_i : cent_der_i(dt_im1, dt_i, S_im1, S_i, S_ip1);
      const double U = 0.5 * (S_i + 0.5 * dt_i * Sd_i + S_ip1 - 0.5 * dt_i * Sd_ip1);
      dts[(size_t)i * ld] = dt_i;
      sds[(size_t)i * ld] = Sd_i;
      srcs[(size_t)i * ld] = S_i;
      for (int m = 0; m < M; m++) {
        Yin[((size_t)i * M + m) * ld] = Y[m];
        double Eps, Alpha, Beta, Gamma;
        delo_coef(dt_i * imu[m], Eps, Alpha, Beta, Gamma);
        Y[m] = Y[m] * Eps + Alpha * S_i + Beta * S_ip1 + Gamma * U;
      }
      Sd_ip1 = Sd_i;
      S_ip1 = S_i;
      S_i = S_im1;
      dt_i = dt_im1;
      if (i >= 2) {
        S_im1 = planck_src(C1, C2, S.temp[i - 2]);
        dt_im1 = tau_step_mu(S.rhox, 1.0, i - 2, (i >= 3) ? col.k(i - 3) : 0.0, col.k(i - 2), col.k(i - 1));
      }
    }
  }
  // ---- reverse sweep
  double yb[MM];
  for (int m = 0; m < M; m++) yb[m] = gF * S.wmu[m];
  double sb_prev = 0.0, sb_cur = 0.0, sb_next = 0.0;  // S adjoints at i-1, i, i+1
  double sdb_cur = 0.0, sdb_next = 0.0;               // Sd adjoints at i, i+1
  double dtb_prev = 0.0, dtb_cur = 0.0;               // Deltau adjoints at i-1, i
  double kb0 = 0.0, kb1 = 0.0, kb2 = 0.0;             // kappa adjoints at i-2, i-1, i
  for (int i = 0; i <= N - 2; i++) {
    const double dt = dts[(size_t)i * ld], Sd_i = sds[(size_t)i * ld], Sd_ip1 = sds[(size_t)(i + 1) * ld];
    const double S_i = srcs[(size_t)i * ld], S_ip1 = srcs[(size_t)(i + 1) * ld];
    const double U = 0.5 * (S_i + 0.5 * dt * Sd_i + S_ip1 - 0.5 * dt * Sd_ip1);
    double ub = 0.0;
    for (int m = 0; m < M; m++) {
      double E, A, B, G, dE, dA, dB, dG;
      delo_coef_d(dt * imu[m], E, A, B, G, dE, dA, dB, dG);
      const double y = yb[m];
      dtb_cur += y * (Yin[((size_t)i * M + m) * ld] * dE + dA * S_i + dB * S_ip1 + dG * U) * imu[m];
      sb_cur += y * A;
      sb_next += y * B;
      ub += y * G;
      yb[m] = y * E;
    }
    // U(i)
    sb_cur += 0.5 * ub;
    sb_next += 0.5 * ub;
    sdb_cur += 0.25 * dt * ub;
    sdb_next -= 0.25 * dt * ub;
    dtb_cur += 0.25 * (Sd_i - Sd_ip1) * ub;
    // Sd(i) is complete
    if (i == 0) {
      sb_next += sdb_cur / dt;
      sb_cur -= sdb_cur / dt;
      dtb_cur += sdb_cur * (-Sd_i / dt);
    } else {
      const double dtm = dts[(size_t)(i - 1) * ld], S_im1 = srcs[(size_t)(i - 1) * ld];
      const CentGrad cg = cent_der_grad(dtm, dt, S_im1, S_i, S_ip1);
      dtb_prev += sdb_cur * cg.dx;
      dtb_cur += sdb_cur * cg.dx1;
      sb_prev += sdb_cur * cg.ym;
      sb_cur += sdb_cur * cg.y0;
      sb_next += sdb_cur * cg.yp;
      // Deltau(i-1) is complete: kappa at i-2, i-1, i
      double gm, g0, gp;
      tau_step_grad(S.rhox, i - 1, (i >= 2) ? col.k(i - 2) : 0.0, col.k(i - 1), col.k(i), gm, g0, gp);
      kb0 += dtb_prev * gm;
      kb1 += dtb_prev * g0;
      kb2 += dtb_prev * gp;
      if (i >= 2) gkap[(size_t)(i - 2) * gstride] = kb0;
      gsrc[(size_t)(i - 1) * ld] = sb_prev;
      kb0 = kb1;
      kb1 = kb2;
      kb2 = 0.0;
    }
    if (i == 0) {  // nothing below point 0 to retire: kappa window starts at points -1, 0, 1
      kb0 = 0.0;
      kb1 = 0.0;
      kb2 = 0.0;
    }
    sb_prev = sb_cur;
    sb_cur = sb_next;
    sb_next = 0.0;
    sdb_cur = sdb_next;
    sdb_next = 0.0;
    dtb_prev = dtb_cur;
    dtb_cur = 0.0;
  }
  // deepest point: Sd(N-1) one-sided, the boundary Y = S(N-1), Deltau(N-2) complete
  {
    const double dt = dts[(size_t)(N - 2) * ld], Sd_last = sds[(size_t)(N - 1) * ld];
    sb_cur += sdb_cur / dt;
    sb_prev -= sdb_cur / dt;
    dtb_prev += sdb_cur * (-Sd_last / dt);
    for (int m = 0; m < M; m++) sb_cur += yb[m];
    double gm, g0, gp;
    tau_step_grad(S.rhox, N - 2, (N >= 3) ? col.k(N - 3) : 0.0, col.k(N - 2), col.k(N - 1), gm, g0, gp);
    kb0 += dtb_prev * gm;
    kb1 += dtb_prev * g0;
    kb2 += dtb_prev * gp;
    gkap[(size_t)(N - 3) * gstride] = kb0;
    gkap[(size_t)(N - 2) * gstride] = kb1;
    gkap[(size_t)(N - 1) * gstride] = kb2;
    gsrc[(size_t)(N - 2) * ld] = sb_prev;
    gsrc[(size_t)(N - 1) * ld] = sb_cur;
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Eight sums across the wavefront at once: each exchange level hands half of a lane's values
// to its partner and keeps the other half, so the eight values need 4 + 2 + 1 exchanges and
// then three more within groups of eight lanes.  On return lanes 8k .. 8k + 7 hold the sum of
// value k over the wavefront.
__device__ __forceinline__ double wave_sum8(const double (&s)[8], int lane) {
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
  double t4[4], t2[2];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double keep = h5 ? s[4 + i] : s[i], give = h5 ? s[i] : s[4 + i];
    t4[i] = keep + __shfl_xor(give, 32, 64);
  }
#pragma unroll
  for (int i = 0; i < 2; i++) {
    const double keep = h4 ? t4[2 + i] : t4[i], give = h4 ? t4[i] : t4[2 + i];
    t2[i] = keep + __shfl_xor(give, 16, 64);
  }
  double t = (h3 ? t2[1] : t2[0]) + __shfl_xor(h3 ? t2[0] : t2[1], 8, 64);
  t += __shfl_xor(t, 4, 64);
  t += __shfl_xor(t, 2, 64);
  t += __shfl_xor(t, 1, 64);
  return t;
}

// One lane per wavelength point of [p_lo, p_lo + np): the line column (cont + LineOpac,
// adjoint gsfl) and the continuum column (adjoint gsfc).  Outputs dL/dLineOpac and
// dL/dContOpac [ntau][nwl] and the source-function part of dL/dT [ntau] (accumulated).
template <typename LopT, int MM>
__global__ void __launch_bounds__(64) k_rt_vjp(DevSetup S, const LopT *lop, const double *gsfl, const double *gsfc,
                                               int p_lo, int np, double *scr, double *glop, double *gcont,
                                               double *gtemp) {
  const int lane_p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = lane_p < np;
  const int p = p_lo + (valid ? lane_p : 0);
  const int N = S.ntau, M = S.nmu;
  // batch contexts: the wavefront's points (p_lo is a multiple of 64) belong to one
  // atmosphere, whose dL/dT row receives the wave's sums
  if (S.batch) gtemp += (size_t)S.pmap[p_lo + blockIdx.x * blockDim.x] * N;
  point_setup(S, p);
  const double Wl = S.wave[p] + S.wfirst;
  const double C1 = kPlkC1 / Wl, C2 = exp(50.7649141 - 5.0 * log(Wl));
  double imu[MM];
  for (int m = 0; m < M; m++) imu[m] = 1.0 / S.xmu[m];
  // scratch: the checkpoint rows, then the S adjoints of the line and continuum columns
  const size_t ld = (size_t)np;
  double *scr_l = scr + (valid ? lane_p : 0);
  double *sL = scr_l + ((size_t)(N - 1) * M + 3 * (size_t)N) * ld, *sC = sL + (size_t)N * ld;
  if (valid) {
    // kappa adjoints go straight to the outputs; the line column's are added to the
    // continuum's, since ContOpac enters both columns
    GridCol<LopT> line{S.contT, lop, S.nwl, p, true};
    rt_vjp_column<MM>(S, line, gsfl[p], imu, C1, C2, scr_l, ld, glop + p, (size_t)S.nwl, sL);
    GridCol<LopT> cont{S.contT, lop, S.nwl, p, false};
    rt_vjp_column<MM>(S, cont, gsfc[p], imu, C1, C2, scr_l, ld, gcont + p, (size_t)S.nwl, sC);
    for (int k = 0; k < N; k++) gcont[(size_t)k * S.nwl + p] += glop[(size_t)k * S.nwl + p];
  }
  // dL/dT from the source function: (S-bar_L + S-bar_C) dS/dT, dS/dT = S (C1/T^2)(1 + S/C2)
  for (int k = 0; k < N; k++) {
    double v = 0.0;
    if (valid) {
      const double T = S.temp[k], Sk = planck_src(C1, C2, T);
      v = (sL[(size_t)k * ld] + sC[(size_t)k * ld]) * Sk * (C1 / (T * T)) * (1.0 + Sk / C2);
    }
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) atomicAdd(&gtemp[k], v);
  }
}

// ------------------------------------------------------------------------------------
// reverse mode of the line opacity: per (line, depth) the sum over the line's window of
// dL/dLineOpac x contribution, and the temperature derivative of the contributions
// ------------------------------------------------------------------------------------
// H(a, v) of the table path with its partial derivatives, in FP32 (the reverse mode's
// general path: gradients are needed to 1e-3, and the per-(line, depth) sums stay FP64): in
// the table box the bilinear interpolant's slopes; outside it the continued fractions
// w(z) ~ i z / (sqrt(pi)(z^2 - 1/2)) (one pole) and i (z^2 - 1) / (sqrt(pi) z (z^2 - 3/2))
// (two poles), H = Re w, z = |v| + i a, dH/dv = Re w', dH/da = -Im w'; one reciprocal per
// point.
__device__ __forceinline__ float hvoigt32_d(const DevSetup &S, float A, int ia, float ua, float aida, float V,
                                            float &Hv, float &Ha) {
  const float av = fabsf(V), sg = (V < 0.0f) ? -1.0f : 1.0f;
  const float idv = (float)S.idv;
  if (ia > 0 && av < 8.0f) {
    const float fv = av * idv;
    const int iv = (int)fv;
    const float tv = fv - (float)iv;
    const float *t0 = S.vtab + (size_t)(ia - 1) * TAB_LD + iv;
    const float a0 = t0[0], a1 = t0[1], b0 = t0[TAB_LD], b1 = t0[TAB_LD + 1];
    const float c0 = a0 + tv * (a1 - a0), c1 = b0 + tv * (b1 - b0);
    Hv = sg * idv * ((1.0f - ua) * (a1 - a0) + ua * (b1 - b0));
    Ha = (c1 - c0) * aida;
    return c0 + ua * (c1 - c0);
  }
  const float a2 = A * A, v2 = av * av, s = v2 + a2;
  const float zr = av, zi = A;
  const float z2r = v2 - a2, z2i = 2.0f * av * A;
  const float kp = (float)kSqrtPiInv;
  float H, wr, wi;
  if (s >= 196.0f) {
    const float dr = z2r - 0.5f, di = z2i;
    const float rd = __builtin_amdgcn_rcpf(dr * dr + di * di), rdd = rd * rd;
    H = kp * A * (s + 0.5f) * rd;
    const float nr = z2r + 0.5f, ni = z2i;
    const float d2r = dr * dr - di * di, d2i = 2.0f * dr * di;
    const float qr = (nr * d2r + ni * d2i) * rdd, qi = (ni * d2r - nr * d2i) * rdd;
    wr = kp * qi;
    wi = -kp * qr;
  } else {
    const float p = v2 - a2 - 1.5f, q = 2.0f * av * A;
    const float dr = av * p - A * q, di = av * q + A * p;
    const float rg = __builtin_amdgcn_rcpf(dr * dr + di * di), rgg = rg * rg;
    H = kp * ((p + 0.5f) * di - q * dr) * rg;
    const float gr = dr, gi = di;
    const float t1r = 2.0f * (zr * gr - zi * gi), t1i = 2.0f * (zr * gi + zi * gr);
    const float ar = z2r - 1.0f, ai = z2i, br = 3.0f * z2r - 1.5f, bi = 3.0f * z2i;
    const float t2r = ar * br - ai * bi, t2i = ar * bi + ai * br;
    const float nr = t1r - t2r, ni = t1i - t2i;
    const float g2r = gr * gr - gi * gi, g2i = 2.0f * gr * gi;
    const float fr = (nr * g2r + ni * g2i) * rgg, fi = (ni * g2r - nr * g2i) * rgg;
    wr = -kp * fi;
    wi = kp * fr;
  }
  Hv = sg * wr;
  Ha = -wi;
  return H;
}

// per-(line, depth) records for the reverse mode: the FP64 path's coefficients with
// d ln LinCop / dT, d ln LinVvgt / dT and d LinAvgt / dT in the record's spare words
__global__ void k_fill64_vjp(DevSetup S0, DevLines L, int list_type, int64_t nret, const int *order, const int *flag,
                             const int *il, const int *ir, const int *ic, Coef64 *coef, CoefV *coefv, double *wlc_s,
                             int4 *meta_s, int *slot_s, VjpWin *win, double *dmaxv, CoefR *coefr, int want64) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nret) return;
  const DevSetup S = line_setup(S0, L, order[j]);
  XC xc;
  line_meta(S, L, list_type, j, order, flag, il, ir, ic, xc, wlc_s, meta_s);
  slot_s[j] = xc.IxPF;
  const int N = S.ntau;
  Coef64 *c = coef + (size_t)j * S.ntp;
  CoefV *cv = coefv + (size_t)j * S.ntp;
  CoefR *cr = coefr + (size_t)j * S.ntp;
  for (int k = N; k < S.ntp; k++) {
    if (want64) {
      Coef64 r{};
      r.ia = 1;
      c[k] = r;
    }
    cv[k] = CoefV{};  // padding depths: the series with zero coefficients everywhere
    cr[k] = CoefR{};
  }
  for (int k = 0; k < N; k++) {
    double cop, vv, av, ua;
    int ia;
    xline_depth(S, k, xc, cop, vv, av);
    voigt_astate(S, av, ia, ua);
    Coef64 r{};
    r.vv = vv;
    r.a = av;
    r.cop = cop;
    r.ua = ua;
    r.ia = ia;
    // temperature derivatives of XlineDepth (src/opacity.f90:247-291) at fixed electron
    // and perturber densities
    const double T = S.temp[k];
    const double eh = exp(-xc.Hnux / S.xln[4 * N + k]);
    const double dlnXstim = -eh * (xc.Hnux / (S.xln[4 * N + k] * T)) / (1.0 - eh);
    const double dlnEfract = xc.Elo / (S.xln[5 * N + k] * T);
    const double Dopl2 = T * xc.Xmass + S.vturb2;
    const double dlnDopl = xc.Xmass / (2.0 * Dopl2);
    double dGam = 0.0;
    if (xc.MdQst == 0) dGam += xc.Cqst * S.xln[2 * N + k] / (6.0 * T);
    if (xc.MdVdw == 0) dGam += 0.3 * xc.Cvdw * S.xln[3 * N + k] / T;
    else if (xc.MdVdw == 1) dGam += 0.3 * S.vwfudge * S.xln[1 * N + k] * xc.Cvdw / T;
    else if (xc.MdVdw == 2) {
      const double Vbar = sqrt(8.0 * kXlnKM0 * T / kPi * xc.AboVbc);
      dGam += 0.5 * xc.AboExp * xc.AboPre * pow(Vbar / 1e4, xc.AboExp) * S.xln[0 * N + k] / T;
    }
    const double Dndopl = sqrt(Dopl2) / xc.WaveL;
    r.pad1 = dlnEfract + dlnXstim - dlnDopl;           // d ln LinCop / dT
    r.pad2 = -dlnDopl;                                 // d ln LinVvgt / dT
    r.pad3 = dGam / (Dndopl * kXlnC4Pi) - av * dlnDopl;  // d LinAvgt / dT
    if (want64) c[k] = r;
    // far-wing series H = a/sqrt(pi) (u + f1 u^2 + f2 u^3), u = 1/v^2, of the one-pole
    // fraction (the FP32 path's expansion), valid to 1e-7 relative where v^2 >= 400 (a^2 + 1/2):
    //   c     = LinCop H
    //   dc/dT = c dlnLinCop/dT + LinCop (v dH/dv dlnLinVvgt/dT + dH/da dLinAvgt/dT),
    //   v dH/dv = -2 a/sqrt(pi) (u + 2 f1 u^2 + 3 f2 u^3)
    //   dH/da   = 1/sqrt(pi) (u + f1 u^2 + f2 u^3) + a/sqrt(pi) (f1a u^2 + f2a u^3)
    {
      const double a2 = av * av, be = a2 + 0.5, al = 2.0 * a2 - 1.0;
      const double f1 = be - al, f2 = -al * f1 - be * be;
      const double f1a = -2.0 * av, f2a = -4.0 * av * f1 + 2.0 * av * al - 4.0 * av * be;  // d f/da
      const double A0 = kSqrtPiInv * av, w = 1.0 / (vv * vv);
      const double cA = cop * A0, c1 = cA * r.pad1, c2 = -2.0 * cA * r.pad2, c3 = cop * r.pad3;
      CoefV q;
      q.k1 = cA * w;
      q.k2 = cA * f1 * w * w;
      q.k3 = cA * f2 * w * w * w;
      q.m1 = (c1 + c2 + c3 * kSqrtPiInv) * w;
      q.m2 = (c1 * f1 + 2.0 * c2 * f1 + c3 * (kSqrtPiInv * f1 + A0 * f1a)) * w * w;
      q.m3 = (c1 * f2 + 3.0 * c2 * f2 + c3 * (kSqrtPiInv * f2 + A0 * f2a)) * w * w * w;
      q.dthr = sqrt(fmax(196.0, 400.0 * be)) / vv * (1.0 + 1e-12);
      q.pad = 0.0;
      cv[k] = q;
      CoefR g;
      g.dthr = q.dthr;
      g.vv = vv;
      g.a = (float)av;
      g.cop = (float)cop;
      g.ua = (float)ua;
      g.aida = ia > 0 ? (float)S.aida[ia - 1] : 0.0f;
      g.p1 = (float)r.pad1;
      g.p2 = (float)r.pad2;
      g.p3 = (float)r.pad3;
      g.ia = ia;
      cr[k] = g;
    }
  }
  // the window's end wavelengths (the grid of the line's atmosphere: the same values as the
  // concatenated grid of a batch holds) and each wavefront's largest threshold
  const int4 mm = meta_s[j];
  win[j] = VjpWin{S.wave[mm.x - S.goff - 1], S.wave[mm.y - S.goff - 1]};
  const int nwv = S.ntp / DPW;
  for (int w = 0; w < nwv; w++) {
    double mx = 0.0;
    for (int k = w * DPW; k < (w + 1) * DPW; k++) mx = fmax(mx, cv[k].dthr);
    dmaxv[(size_t)j * nwv + w] = mx;
  }
}

// The far-wing test of the reverse mode for a line and a tile: the distance from the line
// centre to the points it adds to in the tile (window and tile; 0 if the centre lies among
// them, or for a line with a He I Stark shift).  At a depth whose series threshold dthr it
// reaches, every such point is in the far wing.  k_lineop_vjp and k_lineop_vjp_mfma split
// the (line, depth, tile) triples with this value and the comparison dmin >= dthr.
// The window's and the tile's first and last wavelengths give those points: the grid rises
// within an atmosphere, so the later first point and the earlier last point bound them.
__device__ __forceinline__ double tile_dmin(const VjpWin w, int he, double wc, double Wt0, double Wt1) {
  const double Wlo = fmax(w.wlo, Wt0), Whi = fmin(w.whi, Wt1);
  return (he > 0) ? 0.0 : ((wc < Wlo) ? (Wlo - wc) : (wc > Whi ? (wc - Whi) : 0.0));
}

// One block per VJP_TILE-point wavelength tile, one wavefront per DPW depths, VJP_PPL
// points per lane.  For each candidate line the wave reduces sum_p gLop(d, p) c(d, p)
// per depth into sl[line][d]; sum gLop dc/dT stays in registers per (lane, depth).  With
// far_done set the depths in the far wing over the tile (tile_dmin >= dthr) are the
// matrix-core kernel's: a wavefront then takes only the others, and skips a line none of
// whose depths it has.
__global__ void __launch_bounds__(1024) k_lineop_vjp(DevSetup S, int ntiles, int nbins, int rstride,
                                                     const int2 *__restrict__ rng, const Coef64 *__restrict__ coef,
                                                     const CoefV *__restrict__ coefv,
                                                     const double *__restrict__ wlc_s,
                                                     const int4 *__restrict__ meta_s,
                                                     const VjpWin *__restrict__ win, const double *__restrict__ dmaxv,
                                                     const double *__restrict__ glop, double *__restrict__ sl,
                                                     double *__restrict__ gtemp, int far_done) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const int N = S.ntau;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int p0 = t * VJP_TILE + 1, p1 = min(p0 + VJP_TILE - 1, S.nwl);
  if (S.batch) gtemp += (size_t)S.pmap[p0 - 1] * N;  // the tile's atmosphere (batch contexts)
  int pp[VJP_PPL];
  bool valid[VJP_PPL];
  double Wp[VJP_PPL];
#pragma unroll
  for (int q = 0; q < VJP_PPL; q++) {
    pp[q] = p0 + lane + 64 * q;
    valid[q] = pp[q] <= S.nwl;
    Wp[q] = valid[q] ? S.wave[pp[q] - 1] : 0.0;
  }
  const int d0 = wave * DPW;
  double g[VJP_PPL][DPW], tT[DPW];
#pragma unroll
  for (int k = 0; k < DPW; k++) {
    const int d = d0 + k;
    const double inv = (d < N && S.motype == 1) ? 1.0 / S.cstd[d] : 1.0;
    tT[k] = 0.0;
#pragma unroll
    for (int q = 0; q < VJP_PPL; q++) g[q][k] = (valid[q] && d < N) ? glop[(size_t)d * S.nwl + pp[q] - 1] * inv : 0.0;
  }
  // the tile's candidates are shared among gridDim.y workgroups (every output is an atomic
  // sum), so that small grids still fill the device
  const int slice = blockIdx.y, nslice = gridDim.y;
  const double Wt0 = S.wave[p0 - 1], Wt1 = S.wave[p1 - 1];
  const int nwv = S.ntp / DPW;
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * rstride + b];
    if (r.x + slice >= r.y) continue;
    // each candidate's meta, centre, window ends and the wave's largest threshold are
    // loaded one candidate ahead
    int jn = r.x + slice;
    int4 nm = meta_s[jn];
    double nwc = wlc_s[jn], ndx = dmaxv[(size_t)jn * nwv + wave];
    VjpWin nw_ = win[jn];
    for (int j = r.x + slice; j < r.y; j += nslice) {
      const int4 m = nm;
      const double wc = nwc, dmx = ndx;
      const VjpWin ww = nw_;
      jn = min(j + nslice, r.y - 1);
      nm = meta_s[jn];
      nwc = wlc_s[jn];
      ndx = dmaxv[(size_t)jn * nwv + wave];
      nw_ = win[jn];
      if (m.y < p0 || m.x > p1) continue;
      const double dmin = tile_dmin(ww, m.w, wc, Wt0, Wt1);  // wave-uniform
      // nothing left here when every depth of the wave is in the far wing
      if (far_done && dmin >= dmx) continue;
      const Coef64 *c = coef + (size_t)j * S.ntp + d0;
      const CoefV *cv = coefv + (size_t)j * S.ntp + d0;
      const double *hs = (m.w > 0) ? S.he_shf + (size_t)(m.w - 1) * S.ntp + d0 : nullptr;
      bool act[VJP_PPL];
      double dl[VJP_PPL], T1[VJP_PPL], T2[VJP_PPL], T3[VJP_PPL];
#pragma unroll
      for (int q = 0; q < VJP_PPL; q++) {
        act[q] = valid[q] && pp[q] >= m.x && pp[q] <= m.y;
        dl[q] = Wp[q] - wc;
        // T = 1/dlambda^2, zero outside the window: the series adds nothing there (unused
        // when the series is the matrix-core kernel's)
        const double d2 = (act[q] && !far_done) ? dl[q] * dl[q] : 1.0;
        double r = __builtin_amdgcn_rcp(d2);
        r = r * (2.0 - d2 * r);  // Newton steps to full FP64 precision
        r = r * (2.0 - d2 * r);
        T1[q] = act[q] ? r : 0.0;
        T2[q] = T1[q] * T1[q];
        T3[q] = T2[q] * T1[q];
      }
      double sk[DPW];  // per depth: this lane's share of the sum over the tile's points
#pragma unroll
      for (int k = 0; k < DPW; k++) {
        const CoefV q3 = cv[k];
        double s = 0.0;
        if (dmin >= q3.dthr && far_done) {
          // the matrix-core kernel's
        } else if (dmin >= q3.dthr) {
          // far wing at every point: the series in T with the precomputed coefficients
#pragma unroll
          for (int q = 0; q < VJP_PPL; q++) {
            const double x1 = g[q][k] * T1[q], x2 = g[q][k] * T2[q], x3 = g[q][k] * T3[q];
            s += q3.k1 * x1 + q3.k2 * x2 + q3.k3 * x3;
            tT[k] += q3.m1 * x1 + q3.m2 * x2 + q3.m3 * x3;
          }
        } else {
          const Coef64 ck = c[k];
          const float fa = (float)ck.a, fua = (float)ck.ua, fcop = (float)ck.cop;
          const float faida = ck.ia > 0 ? (float)S.aida[ck.ia - 1] : 0.0f;
          const float p1 = (float)ck.pad1, p2 = (float)ck.pad2, p3 = (float)ck.pad3;
#pragma unroll
          for (int q = 0; q < VJP_PPL; q++) {
            if (!act[q]) continue;
            const float v = (float)((hs ? dl[q] - hs[k] : dl[q]) * ck.vv);
            float Hv, Ha;
            const float H = hvoigt32_d(S, fa, ck.ia, fua, faida, v, Hv, Ha);
            const float cval = H * fcop;
            s += g[q][k] * (double)cval;
            // d c / dT = c d ln LinCop/dT + LinCop (Hv v d ln LinVvgt/dT + Ha d LinAvgt/dT)
            tT[k] += g[q][k] * (double)(cval * p1 + fcop * (Hv * v * p2 + Ha * p3));
          }
        }
        sk[k] = s;
      }
      // the wavefront's 9 depths: 8 sums by one transposed reduction, the ninth alone
      static_assert(DPW == 9, "eight depths in one reduction and one more");
      double s8[8];
#pragma unroll
      for (int k = 0; k < 8; k++) s8[k] = sk[k];
      const double t = wave_sum8(s8, lane);
      if ((lane & 7) == 0 && t != 0.0) atomicAdd(&sl[(size_t)j * S.ntp + d0 + (lane >> 3)], t);
      const double t8 = wave_sum(sk[8]);
      if (lane == 0 && t8 != 0.0) atomicAdd(&sl[(size_t)j * S.ntp + d0 + 8], t8);
    }
  }
#pragma unroll
  for (int k = 0; k < DPW; k++) {
    const double v = wave_sum(tT[k]);
    if (lane == 0 && d0 + k < N && v != 0.0) atomicAdd(&gtemp[d0 + k], v);
  }
}

// The general (line, depth, tile) triples of the reverse mode when the far-wing ones are the
// matrix-core kernel's: as k_lineop_vjp, but a wavefront reads the nine CoefR records of its
// depths with one load per lane (lane k: depth k) as soon as it takes a line, finds its
// general depths by one ballot (dmin < dthr, the test of k_lineop_vjp_mfma) and hands each
// general depth's record to all lanes by lane reads.
__global__ void __launch_bounds__(1024) k_lineop_vjp_gen(DevSetup S, int ntiles, int nbins, int rstride,
                                                         const int2 *__restrict__ rng, const CoefR *__restrict__ coefr,
                                                         const double *__restrict__ wlc_s,
                                                         const int4 *__restrict__ meta_s,
                                                         const VjpWin *__restrict__ win,
                                                         const double *__restrict__ dmaxv,
                                                         const double *__restrict__ glop, double *__restrict__ sl,
                                                         double *__restrict__ gtemp) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const int N = S.ntau;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int p0 = t * VJP_TILE + 1, p1 = min(p0 + VJP_TILE - 1, S.nwl);
  if (S.batch) gtemp += (size_t)S.pmap[p0 - 1] * N;  // the tile's atmosphere (batch contexts)
  int pp[VJP_PPL];
  bool valid[VJP_PPL];
  double Wp[VJP_PPL];
#pragma unroll
  for (int q = 0; q < VJP_PPL; q++) {
    pp[q] = p0 + lane + 64 * q;
    valid[q] = pp[q] <= S.nwl;
    Wp[q] = valid[q] ? S.wave[pp[q] - 1] : 0.0;
  }
  const int d0 = wave * DPW;
  double g[VJP_PPL][DPW], tT[DPW];
#pragma unroll
  for (int k = 0; k < DPW; k++) {
    const int d = d0 + k;
    const double inv = (d < N && S.motype == 1) ? 1.0 / S.cstd[d] : 1.0;
    tT[k] = 0.0;
#pragma unroll
    for (int q = 0; q < VJP_PPL; q++) g[q][k] = (valid[q] && d < N) ? glop[(size_t)d * S.nwl + pp[q] - 1] * inv : 0.0;
  }
  const int slice = blockIdx.y, nslice = gridDim.y;
  const double Wt0 = S.wave[p0 - 1], Wt1 = S.wave[p1 - 1];
  const int nwv = S.ntp / DPW;
  const int kl = lane < DPW ? lane : 0;  // the depth whose record this lane reads
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * rstride + b];
    if (r.x + slice >= r.y) continue;
    // each candidate's meta, centre, window ends and the wave's largest threshold are
    // loaded one candidate ahead
    int jn = r.x + slice;
    int4 nm = meta_s[jn];
    double nwc = wlc_s[jn], ndx = dmaxv[(size_t)jn * nwv + wave];
    VjpWin nw_ = win[jn];
    for (int j = r.x + slice; j < r.y; j += nslice) {
      const int4 m = nm;
      const double wc = nwc, dmx = ndx;
      const VjpWin ww = nw_;
      jn = min(j + nslice, r.y - 1);
      nm = meta_s[jn];
      nwc = wlc_s[jn];
      ndx = dmaxv[(size_t)jn * nwv + wave];
      nw_ = win[jn];
      if (m.y < p0 || m.x > p1) continue;
      const double dmin = tile_dmin(ww, m.w, wc, Wt0, Wt1);  // wave-uniform
      if (dmin >= dmx) continue;  // every depth of the wave in the far wing
      const CoefR rec = coefr[(size_t)j * S.ntp + d0 + kl];
      const uint64_t gen = __builtin_amdgcn_ballot_w64(lane < DPW && dmin < rec.dthr);
      const double *hs = (m.w > 0) ? S.he_shf + (size_t)(m.w - 1) * S.ntp + d0 : nullptr;
      bool act[VJP_PPL];
      double dl[VJP_PPL];
#pragma unroll
      for (int q = 0; q < VJP_PPL; q++) {
        act[q] = valid[q] && pp[q] >= m.x && pp[q] <= m.y;
        dl[q] = Wp[q] - wc;
      }
      double sk[DPW];
#pragma unroll
      for (int k = 0; k < DPW; k++) {
        double s = 0.0;
        if ((gen >> k) & 1) {
          // depth k's record from lane k
          const double vv = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(rec.vv), k),
                                             __builtin_amdgcn_readlane(__double2loint(rec.vv), k));
          const float fa = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.a), k));
          const float fcop = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.cop), k));
          const float fua = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.ua), k));
          const float faida = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.aida), k));
          const float p1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.p1), k));
          const float p2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.p2), k));
          const float p3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rec.p3), k));
          const int ia = __builtin_amdgcn_readlane(rec.ia, k);
#pragma unroll
          for (int q = 0; q < VJP_PPL; q++) {
            if (!act[q]) continue;
            const float v = (float)((hs ? dl[q] - hs[k] : dl[q]) * vv);
            float Hv, Ha;
            const float H = hvoigt32_d(S, fa, ia, fua, faida, v, Hv, Ha);
            const float cval = H * fcop;
            s += g[q][k] * (double)cval;
            // d c / dT = c d ln LinCop/dT + LinCop (Hv v d ln LinVvgt/dT + Ha d LinAvgt/dT)
            tT[k] += g[q][k] * (double)(cval * p1 + fcop * (Hv * v * p2 + Ha * p3));
          }
        }
        sk[k] = s;
      }
      static_assert(DPW == 9, "eight depths in one reduction and one more");
      double s8[8];
#pragma unroll
      for (int k = 0; k < 8; k++) s8[k] = sk[k];
      const double tt = wave_sum8(s8, lane);
      if ((lane & 7) == 0 && tt != 0.0) atomicAdd(&sl[(size_t)j * S.ntp + d0 + (lane >> 3)], tt);
      if ((gen >> 8) & 1) {
        const double t8 = wave_sum(sk[8]);
        if (lane == 0 && t8 != 0.0) atomicAdd(&sl[(size_t)j * S.ntp + d0 + 8], t8);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < DPW; k++) {
    const double v = wave_sum(tT[k]);
    if (lane == 0 && d0 + k < N && v != 0.0) atomicAdd(&gtemp[d0 + k], v);
  }
}

typedef float mf4v __attribute__((ext_vector_type(4)));

// Reverse mode of the far-wing series on the matrix cores.  For a tile of VJP_TILE points and a
// group of 16 candidate lines the sums over the tile's points
//     G_n[d][l] = sum_p gLop(d, p) T(l, p)^n,   n = 1..3,  T = 1/dlambda^2 (0 outside the window)
// are a product of the tile's gLop (depths x points, FP32 in LDS) with X_n (points x lines,
// computed per lane): v_mfma_f32_16x16x4f32 with A = 16 depths x 4 points, B = 4 points x 16
// lines, and the tile's points as the contraction in VJP_TILE / 4 steps.  Where the line is in
// its far wing at the depth over the tile (tile_dmin >= dthr, as k_lineop_vjp decides), the
// per-(line, depth) sum is sum_n k_n G_n (added to sl) and the temperature derivative's share
// sum_n m_n G_n (added to the depth's dL/dT); the other (line, depth) pairs are left to
// k_lineop_vjp.  The four wavefronts take the tile's groups in turn; He I lines (Stark
// shifted) are never in the far wing here.  Lane layout of v_mfma_f32_16x16x4f32: A[r][k] in
// lane 16k + r, B[k][c] in lane 16k + c, D[4 (i / 16) + e][i % 16] in element e of lane i.
template <int NQ>
__global__ void __launch_bounds__(256) k_lineop_vjp_mfma(DevSetup S, int ntiles, int nbins, int rstride,
                                                         const int2 *__restrict__ rng, const CoefV *__restrict__ coefv,
                                                         const double *__restrict__ wlc_s,
                                                         const int4 *__restrict__ meta_s,
                                                         const VjpWin *__restrict__ win,
                                                         const double *__restrict__ glop, double *__restrict__ sl,
                                                         double *__restrict__ gtemp) {
  constexpr int LDG = VJP_TILE + 4;  // row stride of the gLop tile: the A reads of a step fall two lanes per bank
  __shared__ float gt[NQ * 16][LDG];
  __shared__ float ofs_s[VJP_TILE];
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;  // whole workgroups leave together (the barrier below)
  const int N = S.ntau, nwl = S.nwl;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int c = lane & 15, kq = lane >> 4;  // line of the group (B column) / point of the step
  const int p0 = t * VJP_TILE + 1, p1 = min(p0 + VJP_TILE - 1, nwl);
  if (S.batch) gtemp += (size_t)S.pmap[p0 - 1] * N;  // the tile's atmosphere (batch contexts)
  const double Wt0 = S.wave[p0 - 1], Wt1 = S.wave[p1 - 1];
  // the tile's dL/dLineOpac (Motype 1: over ContOpacStd, as k_lineop_vjp scales it), zero
  // past the grid and at the padding depths; the points' offsets from the first point
  for (int i = threadIdx.x; i < NQ * 16 * VJP_TILE; i += blockDim.x) {
    const int d = i / VJP_TILE, pi = i - d * VJP_TILE, p = p0 + pi;
    float v = 0.0f;
    if (d < N && p <= nwl) {
      const double inv = (S.motype == 1) ? 1.0 / S.cstd[d] : 1.0;
      v = (float)(glop[(size_t)d * nwl + p - 1] * inv);
    }
    gt[d][pi] = v;
  }
  for (int i = threadIdx.x; i < VJP_TILE; i += blockDim.x)
    ofs_s[i] = (p0 + i <= nwl) ? (float)(S.wave[p0 + i - 1] - Wt0) : 0.0f;
  __syncthreads();
  double tT[NQ][4];
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) tT[q][e] = 0.0;
  int gidx = 0;  // the tile's groups of 16 candidates, in bin order; wavefront w takes w, w + 4, ...
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * rstride + b];
    for (int base = r.x; base < r.y; base += 16, gidx++) {
      if ((gidx & 3) != wave) continue;
      const int j = (base + c < r.y) ? base + c : -1;
      int4 m = make_int4(1, 0, 0, 1);
      double wc = 0.0;
      VjpWin ww{0.0, 0.0};
      if (j >= 0) {
        m = meta_s[j];
        wc = wlc_s[j];
        ww = win[j];
      }
      const bool meets = j >= 0 && m.w == 0 && m.y >= p0 && m.x <= p1;
      if (!__builtin_amdgcn_ballot_w64(meets)) continue;
      const double dmin = meets ? tile_dmin(ww, m.w, wc, Wt0, Wt1) : 0.0;
      // B: T^n at (point, this lane's line); dl = (point - first point) + (first point -
      // centre) in FP32, relative error below 2^-23 (1 + 2 tile width / |dl|), |dl| >= dthr
      // wherever the sums are used
      const float bs = (float)(Wt0 - wc);
      mf4v C[3][NQ];
#pragma unroll
      for (int n = 0; n < 3; n++)
#pragma unroll
        for (int q = 0; q < NQ; q++) C[n][q] = mf4v{0.f, 0.f, 0.f, 0.f};
      for (int st = 0; st < VJP_TILE / 4; st++) {
        const int pi = 4 * st + kq, p = p0 + pi;
        const bool act = meets && p >= m.x && p <= m.y && p <= nwl;
        const float x = ofs_s[pi] + bs;
        const float T1 = act ? __builtin_amdgcn_rcpf(fmaxf(x * x, 1e-12f)) : 0.0f;
        const float T2 = T1 * T1, T3 = T2 * T1;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          const float a = gt[16 * q + c][4 * st + kq];
          C[0][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, T1, C[0][q], 0, 0, 0);
          C[1][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, T2, C[1][q], 0, 0, 0);
          C[2][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, T3, C[2][q], 0, 0, 0);
        }
      }
      // this lane's (line c, depth 16 q + 4 kq + e) sums: the far-wing pairs only
      if (meets) {
        const CoefV *cv = coefv + (size_t)j * S.ntp;
#pragma unroll
        for (int q = 0; q < NQ; q++)
#pragma unroll
          for (int e = 0; e < 4; e++) {
            const int d = 16 * q + 4 * kq + e;
            if (d >= N) continue;
            const CoefV rec = cv[d];
            if (!(dmin >= rec.dthr)) continue;
            const double g1 = C[0][q][e], g2 = C[1][q][e], g3 = C[2][q][e];
            const double sv = rec.k1 * g1 + rec.k2 * g2 + rec.k3 * g3;
            if (sv != 0.0) atomicAdd(&sl[(size_t)j * S.ntp + d], sv);
            tT[q][e] += rec.m1 * g1 + rec.m2 * g2 + rec.m3 * g3;
          }
      }
    }
  }
  // dL/dT: the 16 lanes of a line column hold the same depths
#pragma unroll
  for (int q = 0; q < NQ; q++)
#pragma unroll
    for (int e = 0; e < 4; e++) {
      double v = tT[q][e];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      const int d = 16 * q + 4 * kq + e;
      if (c == 0 && d < N && v != 0.0) atomicAdd(&gtemp[d], v);
    }
}

// per line: dL/d log10 gf = ln 10 sum_d sl[line][d] (to the line's input index), and
// dL/d ln FractionPF[d][slot] += sl[line][d]
__global__ void k_lineop_vjp_scatter(DevSetup S, int64_t nret, const int *order, const int *atm, const int *slot_s,
                                     const double *sl, double *g_loggf, double *g_fpf) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nret) return;
  if (atm) g_fpf += (size_t)atm[order[j]] * S.ntau * S.pf_ld;  // the line's atmosphere (batch lists)
  const double *r = sl + (size_t)j * S.ntp;
  const int slot = slot_s[j];
  double tot = 0.0;
  for (int d = 0; d < S.ntau; d++) {
    tot += r[d];
    if (r[d] != 0.0) atomicAdd(&g_fpf[(size_t)d * S.pf_ld + (slot - 1)], r[d]);
  }
  g_loggf[order[j]] = 2.302585092994045684 * tot;
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
static inline unsigned nb(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

size_t rt_vjp_scratch_per_point(const DevSetup &S) {
  return sizeof(double) * ((size_t)(S.ntau - 1) * S.nmu + 5 * (size_t)S.ntau);
}

void launch_rt_vjp(const DevSetup &S, const void *lop, int fp32, const double *gsfl, const double *gsfc, int p_lo,
                   int np, double *scr, double *glop, double *gcont, double *gtemp, hipStream_t st) {
  if (np <= 0) return;
  // up to 8 angles (the usual disk-integration sets): half the per-angle registers (172
  // VGPRs; forcing 3 wavefronts per SIMD moves the angle arrays to scratch and is slower)
  if (fp32) {
    if (S.nmu <= 8)
      k_rt_vjp<float, 8><<<nb(np, 64), 64, 0, st>>>(S, (const float *)lop, gsfl, gsfc, p_lo, np, scr, glop, gcont,
                                                    gtemp);
    else
      k_rt_vjp<float, MAXMU><<<nb(np, 64), 64, 0, st>>>(S, (const float *)lop, gsfl, gsfc, p_lo, np, scr, glop,
                                                        gcont, gtemp);
  } else {
    if (S.nmu <= 8)
      k_rt_vjp<double, 8><<<nb(np, 64), 64, 0, st>>>(S, (const double *)lop, gsfl, gsfc, p_lo, np, scr, glop,
                                                     gcont, gtemp);
    else
      k_rt_vjp<double, MAXMU><<<nb(np, 64), 64, 0, st>>>(S, (const double *)lop, gsfl, gsfc, p_lo, np, scr, glop,
                                                         gcont, gtemp);
  }
}

void launch_fill64_vjp(const DevSetup &S, const DevLines &L, int list_type, int64_t nret, const int *order,
                       const int *flag, const int *il, const int *ir, const int *ic, Coef64 *coef, CoefV *coefv,
                       double *wlc_s, int4 *meta_s, int *slot_s, VjpWin *win, double *dmaxv, CoefR *coefr,
                       hipStream_t st) {
  // the Coef64 records feed only the kernel that runs past the matrix cores' 128 depths
  const int want64 = (S.ntau + 15) / 16 > MFMA_MAX_NQ;
  if (nret)
    k_fill64_vjp<<<nb(nret, 128), 128, 0, st>>>(S, L, list_type, nret, order, flag, il, ir, ic, coef, coefv, wlc_s,
                                                meta_s, slot_s, win, dmaxv, coefr, want64);
}

void launch_lineop_vjp(const DevSetup &S, int ntiles, int nbins, int rstride, const int2 *rng, const Coef64 *coef,
                       const CoefV *coefv, const CoefR *coefr, const double *wlc_s, const int4 *meta_s,
                       const VjpWin *win, const double *dmaxv, const double *glop, double *sl, double *gtemp,
                       hipStream_t st) {
  if (ntiles <= 0 || nbins <= 0) return;
  const int nw = S.ntp / DPW;
  const int nb = 8 * ((ntiles + 7) / 8);
  // enough workgroups for about 8 wavefronts per SIMD of the 256 CUs (a 5-50 A window at
  // 0.01 A has a few dozen tiles; a 1.2-million-point grid has thousands and keeps one
  // workgroup per tile)
  const int nslice = std::max(1, std::min(64, (8 * 1024) / std::max(1, nb * nw)));
  // the 1024-thread bound keeps 128 registers and 4 wavefronts per SIMD, which outruns the
  // spill-free 138-register build at 3 (256 atmospheres of 2,760 points on average: 86 vs
  // 117 ms)
  // the far-wing (line, depth, tile) triples on the matrix cores (up to 128 depths; the
  // width bins, not the He I group, which is the last of the nbins groups)
  const int nq = (S.ntau + 15) / 16;
  int far_done = 0;
  switch (nq) {
#define FSG_VJP_MFMA(Q)                                                                                   \
  case Q:                                                                                                \
    k_lineop_vjp_mfma<Q><<<nb, 256, 0, st>>>(S, ntiles, nbins - 1, rstride, rng, coefv, wlc_s, meta_s, win, glop, \
                                             sl, gtemp);                                                 \
    far_done = 1;                                                                                        \
    break;
    FSG_VJP_MFMA(1)
    FSG_VJP_MFMA(2)
    FSG_VJP_MFMA(3)
    FSG_VJP_MFMA(4)
    FSG_VJP_MFMA(5)
    FSG_VJP_MFMA(6)
    FSG_VJP_MFMA(7)
    FSG_VJP_MFMA(8)
#undef FSG_VJP_MFMA
    default:
      break;
  }
  if (far_done)
    k_lineop_vjp_gen<<<dim3(nb, nslice), 64 * nw, 0, st>>>(S, ntiles, nbins, rstride, rng, coefr, wlc_s, meta_s, win,
                                                           dmaxv, glop, sl, gtemp);
  else
    k_lineop_vjp<<<dim3(nb, nslice), 64 * nw, 0, st>>>(S, ntiles, nbins, rstride, rng, coef, coefv, wlc_s, meta_s, win,
                                                       dmaxv, glop, sl, gtemp, 0);
}

void launch_lineop_vjp_scatter(const DevSetup &S, int64_t nret, const int *order, const int *atm, const int *slot_s,
                               const double *sl, double *g_loggf, double *g_fpf, hipStream_t st) {
  if (nret) k_lineop_vjp_scatter<<<nb(nret, 128), 128, 0, st>>>(S, nret, order, atm, slot_s, sl, g_loggf, g_fpf);
}

}  // namespace fsg
