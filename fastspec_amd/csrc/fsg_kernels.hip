// fsg_kernels.hip -- HIP kernels for gfx950: Voigt tables, line-opacity pass 1, width
// binning support, coefficient fill, the tiled opacity accumulation and the scalar
// DELO-Bezier formal solution with the flux sum.
//
// Reference routines restated here (file:line in the reference tree):
//   HVoigtExact/VoigtInit      src/voigt.f90:42-72, 158-187
//   HVoigt                     src/voigt.f90:76-110
//   XlineInit/Setup/Depth      src/opacity.f90:65-293
//   LineOpacity1/2             src/opacity.f90:297-386
//   WaveIdx, AnalyticWindow    src/opacity.f90:594-608, 678-710
//   pass 1 of LineopCalc       src/synth.f90:833-907
//   RadTranCalc, LocPrfPP      src/synth.f90:1825-1895, src/solver.f90:123-179
//   LocPrfSph, DeloScalar      src/solver.f90:182-261
//   TauStep, CentDer           src/solver.f90:591-647
//   flux sums of DiskInteg     src/convol.f90:233-234
#include <hipcub/hipcub.hpp>

#include "fsg_kernels.h"
#include "fsg_device.h"

namespace fsg {

// ------------------------------------------------------------------------------------
// Voigt table (VoigtInit): FP64 Chiarella-Reichel sums, stored FP32
// ------------------------------------------------------------------------------------
__device__ double hvoigt_exact(double A, double V) {
  const double HH = 0.5;
  double av = fabs(V);
  if (A <= 0.0) return exp(-fmin(av * av, 700.0));
  // z = av + i A;  s = 1/z + sum 2 z e^{-(nh)^2} / (z^2 - (nh)^2)
  double zr = av, zi = A;
  double z2r = zr * zr - zi * zi, z2i = 2.0 * zr * zi;
  double den = zr * zr + zi * zi;
  double sr = zr / den, si = -zi / den;
  for (int n = 1; n <= 16; n++) {
    double nh = n * HH;
    double e = 2.0 * exp(-nh * nh);
    double nr = e * zr, ni = e * zi;              // 2 z e^{-nh^2}
    double dr = z2r - nh * nh, di = z2i;          // z^2 - nh^2
    double dd = dr * dr + di * di;
    sr += (nr * dr + ni * di) / dd;
    si += (ni * dr - nr * di) / dd;
  }
  // s = i (HH/pi) s
  double c = HH / kPi;
  double tr = -c * si, ti = c * sr;
  if (A < 6.0) {
    double p = 2.0 * kPi * A / HH;
    double th = 2.0 * kPi * (2.0 * av - rint(2.0 * av));
    double sh = sin(0.5 * th);
    double redd = -2.0 * exp(0.5 * p) * sinh(0.5 * p) + exp(p) * 2.0 * sh * sh;
    double imdd = exp(p) * sin(th);
    // 2 exp(A^2 - av^2 - 2 i A av) / (redd + i imdd)
    double m = 2.0 * exp(A * A - av * av);
    double er = m * cos(2.0 * A * av), ei = -m * sin(2.0 * A * av);
    double dd = redd * redd + imdd * imdd;
    tr += (er * redd + ei * imdd) / dd;
    ti += (ei * redd - er * imdd) / dd;
  }
  (void)ti;
  return tr;
}

__global__ void k_voigt_table(float *tab, const double *anode, double dv) {
  int iv = blockIdx.x * blockDim.x + threadIdx.x;  // 0..NVV
  int ia = blockIdx.y;                             // 0..NVA
  if (iv > NVV) return;
  tab[(size_t)ia * TAB_LD + iv] = (float)hvoigt_exact(anode[ia], (double)iv * dv);
}

// ------------------------------------------------------------------------------------
// XlineInit: per-depth constants [6][ntau] (src/opacity.f90:65-93)
// ------------------------------------------------------------------------------------
__global__ void k_xline_init(int ntau, const double *temp, const double *xne, const float *fh1,
                             const float *fhe1, const float *fhmol, double h2brd, double *xln) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ntau) return;
  double T = temp[k], Xne = xne[k];
  double Vh = sqrt(2.0 * kKbolt * T / kMh);
  double np = (double)fh1[k] + 0.412265 * (double)fhe1[k];
  if (h2brd > 0.0) np = np + h2brd * 0.874307 * (double)fhmol[k];
  xln[0 * ntau + k] = np;                                // XlnNprtb
  xln[1 * ntau + k] = 17.0 * pow(Vh, 0.6) * np;          // XlnGvwprt
  xln[2 * ntau + k] = pow(T / 1e4, 1.0 / 6.0) * Xne;     // XlnTemp6
  xln[3 * ntau + k] = pow(T / 1e4, 0.3) * np;            // XlnTemp3
  xln[4 * ntau + k] = kKbolt * T;                        // XlnKbT
  xln[5 * ntau + k] = kKbolt * kErg2ev * T;              // XlnKbeT
}

// ------------------------------------------------------------------------------------
// pass 1 of LineopCalc (src/synth.f90:833-907): one lane per line.  The three depth
// maxima (bound test, centre test, analytic window) are streamed in one depth loop; the
// decisions are taken afterwards in the reference's order.
// ------------------------------------------------------------------------------------
// Three waves per SIMD (168 VGPRs, a few spills) hide more of the FP64 latency chains than
// the two that 195 VGPRs allow: 7 % less time on 2x10^8 lines.
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3))) k_pass1(DevSetup S0, DevLines L, int list_type, int *flag_out, int *il_out, int *ir_out,
                        int *ic_out, unsigned *key_out, int *bin_count, int *bin_hmax, int batched) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < L.n;  // every lane reaches the wavefront's bin update
  unsigned key = 0;
  int h = -1;
  if (valid) {
  const DevSetup S = line_setup(S0, L, i);  // windows in the line's own grid
  const int N = S.ntau;
  int flag = initial_flag(L, i, list_type);
  double wlc = L.wl[i] - S.wfirst;
  XC xc = xline_setup(S, L, i, wlc, flag, list_type);
  if (xc.IxHe > 0) flag |= (xc.IxHe << 8);
  int ihe = xc.IxHe;
  int iwl = wave_idx(S, wlc);
  double w0 = S.wave[iwl - 1], w1 = S.wave[iwl];
  double dW = (wlc - w0) / (w1 - w0);
  const double iSqrtPi = 0.5641895835477563;
  double maxb = -INFINITY, maxc = -INFINITY, win = 0.0;
  for (int k = 0; k < N; k++) {
    double cop, vv, av;
    xline_depth(S, k, xc, cop, vv, av);
    const double *ck = S.contT + (size_t)k * S.nwl;
    double kc = ck[iwl - 1] * (1.0 - dW) + ck[iwl] * dW;
    double cs = S.cstd[k];
    double rb = (S.motype == 1) ? cop / (kc * cs) : cop / kc;
    maxb = fmax(maxb, rb);
    // LineOpacity2 at the line centre: dWl0 = 0, shifted by HeShf for He I lines
    double dWl = 0.0;
    if (ihe > 0) dWl = 0.0 - S.he_shf[(size_t)(ihe - 1) * S.ntp + k];
    double lo = hvoigt64_full(S, av, dWl * vv) * cop;
    if (S.motype == 1) lo = lo / cs;
    maxc = fmax(maxc, lo / kc);
    // AnalyticWindow (src/opacity.f90:678-710)
    double r = cop / (S.eps_opac * kc);
    if (S.motype == 1) r = r / cs;
    if (r > 1.0) {
      double w = sqrt(log(r) + fmax(0.0, av * (r * iSqrtPi - av))) / vv;
      if (ihe > 0) w = w + fabs(S.he_shf[(size_t)(ihe - 1) * S.ntp + k]);
      win = fmax(win, w);
    }
  }
  int il = 1, ir = S.nwl;
  bool weak = (maxb < S.eps_opac) || (maxc < S.eps_opac);
  if (weak) {
    flag |= 1;
  } else {
    il = wave_idx(S, wlc - win);
    int r2 = wave_idx(S, wlc + win) + 1;
    ir = r2 < S.nwl ? r2 : S.nwl;
  }
  flag_out[i] = flag;
  il_out[i] = il;
  ir_out[i] = ir;
  ic_out[i] = iwl;
  if (weak) {
    key = WEAK_KEY;
  } else {
    h = max(iwl - il, ir - iwl);
    if (ihe > 0) key = HE_KEY;
    else key = batched ? (h == 0 ? 0u : (unsigned)(32 - __clz(h))) : 0u;
  }
  key_out[i] = key;
  }  // valid
  wave_bin_update(bin_count, bin_hmax, key, h, valid);
}

// batch lists: line k is line idx[k] of the source list
__global__ void k_gather_lines(DevLines src, const int64_t *idx, int64_t n, double *wl, float *elo, float *gf,
                               float *grad, float *gqst, float *gvdw, int *spc, uint8_t *is_mol) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = idx[k];
  wl[k] = src.wl[i];
  elo[k] = src.elo[i];
  gf[k] = src.gf[i];
  grad[k] = src.grad[i];
  spc[k] = src.spc[i];
  if (gqst) gqst[k] = src.gqst[i];
  if (gvdw) gvdw[k] = src.gvdw[i];
  if (is_mol) is_mol[k] = src.is_mol[i];
}

__global__ void k_iota(int *v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (int)i;
}

// Sum over retained lines of window pixels x Ntau: the metric's evaluation count.  A
// grid-stride loop with one atomic per workgroup (a single counter takes every atomic, so
// one per wavefront over 10^8 lines would serialise on it).
constexpr int COUNT_BLOCKS = 2048;
__global__ void __launch_bounds__(256) k_count_evals(int64_t n, const int *flag, const int *il, const int *ir, int ntau,
                                                     unsigned long long *acc) {
  __shared__ unsigned long long part[4];
  unsigned long long v = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    if (!(flag[i] & 1)) v += (unsigned long long)(ir[i] - il[i] + 1) * (unsigned long long)ntau;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(acc, t);
  }
}

// MarkWideLines criterion (src/synth.f90:709-744) over the (packed or full) list of
// centre offsets wlc_list[0..m): WlcIdx = number of entries <= W.
__device__ __forceinline__ int64_t upper_count(const double *a, int64_t m, double W) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (a[mid] <= W) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_mark_wide(DevSetup S, int64_t n, const double *wl, const int *flag, const int *il,
                            const int *ir, const double *wlc_list, int64_t m, int64_t nstep,
                            unsigned long long *nwide) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int w = 0;
  if (i < n && !(flag[i] & 1)) {
    double wr = S.wave[ir[i] - 1], wL = S.wave[il[i] - 1];
    int64_t nrec = upper_count(wlc_list, m, wr) - upper_count(wlc_list, m, wL) + 1;
    double wlc = wl[i] - S.wfirst;
    double dv = kVc * (wr - wL) / (wlc + S.wfirst);
    if (nrec > nstep / 2 || dv > kDwide) w = 1;
  }
  unsigned long long v = w;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(nwide, v);
}

__global__ void k_wlc_list(int64_t n, const double *wl, double wfirst, const int *flag, const int *pos,
                           double *out, int pack) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!pack) {
    out[i] = wl[i] - wfirst;
  } else if (!(flag[i] & 1)) {
    out[pos[i]] = wl[i] - wfirst;
  }
}

// exclusive positions of retained lines (inclusive scan result shifted)
__global__ void k_retained_flag(int64_t n, const int *flag, int *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (flag[i] & 1) ? 0 : 1;
}

// ------------------------------------------------------------------------------------
// coefficient fill: one wavefront per 64 retained lines in bin-sorted order.  Lanes take the
// lines for XlineSetup and the core-zone search, and the flattened (line, depth) pairs for
// the records, so the records of consecutive lines go out as one contiguous stream.
// ------------------------------------------------------------------------------------
constexpr int FILL_LINES = 64;
__global__ void __launch_bounds__(64) k_fill32(DevSetup S0, DevLines L, int list_type, int64_t nret,
                         const int *order, const int *flag, const int *il, const int *ir, const int *ic, CoefA *ca,
                         CoefB *cb, float *dmaxw, int2 *zone_s, LineRec *lrec, double *wlc_s, int4 *meta_s) {
  __shared__ XC sxc[FILL_LINES];
  __shared__ unsigned sgm[FILL_LINES * 16];  // per depth group (ntau <= 144: at most 16): largest dthr (bits)
  const int lane = threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.x * FILL_LINES;
  const int nl = (int)min((int64_t)FILL_LINES, nret - j0);
  const int ntp = S0.ntp, N = S0.ntau, G = ntp / DPW;  // one depth count for every line of a batch
  for (int i = lane; i < FILL_LINES * G; i += 64) sgm[i] = 0u;
  if (lane < nl) {
    const int64_t j = j0 + lane;
    const DevSetup S = line_setup(S0, L, order[j]);
    line_meta(S, L, list_type, j, order, flag, il, ir, ic, sxc[lane], wlc_s, meta_s);
  }
  __syncthreads();
  CoefA *A = ca + (size_t)j0 * ntp;
  CoefB *B = cb + (size_t)j0 * ntp;
  const int np = nl * ntp;
  for (int p = lane; p < np; p += 64) {
    const int m = p / ntp, k = p - m * ntp;
    CoefA a{};
    CoefB b{};
    if (k < N) {
      const DevSetup S = line_setup(S0, L, order[j0 + m]);
      const XC &xc = sxc[m];
      double cop, vv, av, ua;
      int ia;
      xline_depth(S, k, xc, cop, vv, av);
      voigt_astate(S, av, ia, ua);
      const double a2 = av * av, K = cop * av * kSqrtPiInv;
      const double al = 2.0 * a2 - 1.0, be = a2 + 0.5;
      // series of (1 + be u)/(1 + al u + be^2 u^2): f_n = -al f_{n-1} - be^2 f_{n-2}
      const double f1 = be - al;
      const double f2 = -al * f1 - be * be;
      const double w = 1.0 / (vv * vv);
      a.k0 = (float)(K * w);
      a.k1 = (float)(K * f1 * w * w);
      a.k2 = (float)(K * f2 * w * w * w);
      a.dthr = (float)(sqrt(fmax(196.0, 400.0 * be)) / vv * (1.0 + 1e-4));
      b.vv = (float)vv;
      b.a = (float)av;
      b.a2 = (float)a2;
      b.c1 = (float)(a2 + 0.25);
      b.ka = (float)K;
      b.cop = (float)cop;
      b.ua = (float)ua;
      b.ia = ia;
      // the thresholds are finite and >= 0, so their bit patterns order as the values do
      // (a NaN is skipped, as fmaxf would)
      if (a.dthr > 0.0f) atomicMax(&sgm[p / DPW], __float_as_uint(a.dthr));
    } else {  // padding depths contribute exactly zero on either path
      b.a2 = 1e4f;
      b.c1 = 1e4f + 0.25f;
    }
    A[p] = a;
    B[p] = b;
  }
  __syncthreads();
  // per wavefront depth group: the largest far-wing threshold (padding depths: 0)
  for (int i = lane; i < nl * G; i += 64) dmaxw[(size_t)j0 * G + i] = __uint_as_float(sgm[i]);
  if (lane >= nl) return;
  const int64_t j = j0 + lane;
  const DevSetup S = line_setup(S0, L, order[j]);
  float zmax = 0.0f;
  for (int w = 0; w < G; w++) zmax = fmaxf(zmax, __uint_as_float(sgm[lane * G + w]));
  // core zone of the line: the grid points with |lambda - lambda_c| < max_k dthr (1-based,
  // empty when zl > zr); outside it the far-wing expansion holds at every depth
  const double wc = wlc_s[j], zlo = wc - (double)zmax, zhi = wc + (double)zmax;
  int lo = 0, e = S.nwl;
  while (lo < e) {  // first point above zlo
    const int mid = (lo + e) >> 1;
    if (S.wave[mid] > zlo) e = mid;
    else lo = mid + 1;
  }
  const int zl = lo + 1;
  lo = 0;
  e = S.nwl;
  while (lo < e) {  // first point at or above zhi
    const int mid = (lo + e) >> 1;
    if (S.wave[mid] >= zhi) e = mid;
    else lo = mid + 1;
  }
  zone_s[j] = make_int2(zl + S.goff, lo + S.goff);
  const int4 mm = meta_s[j];
  LineRec r;
  r.il = mm.x;
  r.ir = mm.y;
  r.zl = zl + S.goff;
  r.zr = lo + S.goff;
  r.wc = wc;
  r.pad0 = r.pad1 = 0;
  lrec[j] = r;
}


__global__ void __launch_bounds__(128) k_fill64(DevSetup S0, DevLines L, int list_type, int64_t nret, const int *order, const int *flag,
                         const int *il, const int *ir, const int *ic, Coef64 *coef, double *wlc_s, int4 *meta_s) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nret) return;
  const DevSetup S = line_setup(S0, L, order[j]);
  XC xc;
  line_meta(S, L, list_type, j, order, flag, il, ir, ic, xc, wlc_s, meta_s);
  const int N = S.ntau;
  Coef64 *c = coef + (size_t)j * S.ntp;
  for (int k = N; k < S.ntp; k++) {  // padding depths contribute exactly zero
    Coef64 r{};
    r.ia = 1;
    c[k] = r;
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
    c[k] = r;
  }
}

// candidate ranges: for tile t and bin b, lines of the bin whose centre pixel lies in
// [p0 - hmax_b, p1 + hmax_b] (a contiguous run of the bin's index-ordered lines)
__global__ void k_tile_ranges(int ntiles, int tile, int nbins, int nwl, const int *bin_off, const int *bin_cnt,
                              const int *bin_hmax, const int4 *meta_s, int2 *rng) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (t >= ntiles || b >= nbins) return;
  int p0 = t * tile + 1, p1 = min(p0 + tile - 1, nwl);
  int off = bin_off[b], cnt = bin_cnt[b], h = bin_hmax[b];
  long lo = (long)p0 - h, hi = (long)p1 + h;
  int a = 0, e = cnt;
  while (a < e) {  // first with ic >= lo
    int m = (a + e) >> 1;
    if (meta_s[off + m].z < lo) a = m + 1;
    else e = m;
  }
  int s = a;
  e = cnt;
  while (a < e) {  // first with ic > hi
    int m = (a + e) >> 1;
    if (meta_s[off + m].z <= hi) a = m + 1;
    else e = m;
  }
  rng[(size_t)t * nbins + b] = make_int2(off + s, off + a);
}

// ------------------------------------------------------------------------------------
// line opacity accumulation: one block per 64-point wavelength tile, one wavefront per
// group of DPW depths, one lane per wavelength point.  Per-(line, depth) coefficients
// are wave-uniform and come through scalar loads.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ float contrib32(const DevSetup &S, const Coef32 &c, float v) {
  float av = fabsf(v);
  if (c.ia > 0 && av < 8.0f) {
    // table box: bilinear interpolation of the FP32 table (src/voigt.f90:81-95)
    float fv = av * (float)S.idv;
    int iv = (int)fv;
    float tv = fv - (float)iv;
    const float *t0 = S.vtab + (size_t)(c.ia - 1) * TAB_LD + iv;
    const float *t1 = t0 + TAB_LD;
    float a0 = t0[0], a1 = t0[1], b0 = t1[0], b1 = t1[1];
    float c0 = fmaf(tv, a1 - a0, a0);
    float c1 = fmaf(tv, b1 - b0, b0);
    return fmaf(c.ua, c1 - c0, c0) * c.cop;
  }
  float v2 = av * av;
  float s = v2 + c.a2;
  if (s >= 196.0f) {
    // one-pole continued fraction: a (s + 1/2) / (s^2 - v^2 + a^2 + 1/4) / sqrt(pi)
    float den = fmaf(s, s, c.c1 - v2);
    return c.ka * (s + 0.5f) * __builtin_amdgcn_rcpf(den);
  }
  // two-pole continued fraction, numerator and denominator rearranged without
  // cancellation: a (p (p + 1/2) + q^2 + v^2) / ((v^2 + a^2)(p^2 + q^2)) / sqrt(pi)
  float p = v2 - c.a2 - 1.5f;
  float q = 2.0f * av * c.a;
  float num = fmaf(p, p + 0.5f, fmaf(q, q, v2));
  float den = s * fmaf(p, p, q * q);
  return c.ka * num * __builtin_amdgcn_rcpf(den);
}

// contrib32 for the P points of a lane at once, without branches: both table gathers are
// issued (at a valid cell for the points outside the table box) before the continued
// fractions are evaluated, and each point takes the branch contrib32 would take.  Points
// with act false add nothing.
template <int P>
__device__ __forceinline__ void contrib32_lane(const DevSetup &S, const Coef32 &c, const float (&v)[P],
                                               const bool (&act)[P], float (&out)[P]) {
  float4 e[P];
  float tv[P];
  bool tab[P];
  const size_t row = (size_t)max(c.ia - 1, 0) * TAB_LD;
#pragma unroll
  for (int q = 0; q < P; q++) {
    const float av = fabsf(v[q]);
    tab[q] = c.ia > 0 && av < 8.0f;
    const float fv = av * (float)S.idv;
    const int iv = tab[q] ? (int)fv : 0;
    tv[q] = fv - (float)iv;
    const float *t0 = S.vtab + row + iv, *t1 = t0 + TAB_LD;
    e[q] = make_float4(t0[0], t0[1], t1[0], t1[1]);
  }
#pragma unroll
  for (int q = 0; q < P; q++) {
    const float av = fabsf(v[q]);
    const float v2 = av * av;
    const float s = v2 + c.a2;
    // one-pole continued fraction (s >= 196)
    const float den1 = fmaf(s, s, c.c1 - v2);
    const float r1 = c.ka * (s + 0.5f) * __builtin_amdgcn_rcpf(den1);
    // two-pole continued fraction
    const float p = v2 - c.a2 - 1.5f;
    const float qq = 2.0f * av * c.a;
    const float num = fmaf(p, p + 0.5f, fmaf(qq, qq, v2));
    const float den2 = s * fmaf(p, p, qq * qq);
    const float r2 = c.ka * num * __builtin_amdgcn_rcpf(den2);
    const float c0 = fmaf(tv[q], e[q].y - e[q].x, e[q].x);
    const float c1 = fmaf(tv[q], e[q].w - e[q].z, e[q].z);
    const float rt = fmaf(c.ua, c1 - c0, c0) * c.cop;
    const float r = tab[q] ? rt : (s >= 196.0f ? r1 : r2);
    out[q] = act[q] ? r : 0.0f;
  }
}

__device__ __forceinline__ double contrib64(const DevSetup &S, const Coef64 &c, double v) {
  return hvoigt64(S, c.a, c.ia, c.ua, v) * c.cop;
}


typedef float f4v __attribute__((ext_vector_type(4)));

// Lines are added into plain FP32 partial sums in groups of FOLD lines; each group's partial
// is then added to the running sum with Kahan compensation (blocked compensated summation:
// the uncompensated part is at most FOLD terms long).
constexpr int FOLD = 8;

// FP32 batched path.  A block owns a tile of TILE = 64 x PPL wavelength points (PPL per
// lane, 64 apart) and all depths, one wavefront per DPW depths.  Per line the wave's DPW
// 16-byte CoefA records are wave-uniform and arrive in one batch of scalar loads; when
// every depth is in the far wing for the whole tile (the common case) the depths run as
// straight-line FMA code.  Lines of the tabulated He I Stark table come in a separate
// launch (HE = true) so that the common path carries no per-depth shift.
template <bool HE>
__global__ void __launch_bounds__(1024) k_opacity32(DevSetup S, int ntiles, int nbins, int rstride,
                                                    const int2 *__restrict__ rng, const CoefA *__restrict__ ca,
                                                    const CoefB *__restrict__ cb, const float *__restrict__ dmaxw,
                                                    const double *__restrict__ wlc_s,
                                                    const int4 *__restrict__ meta_s, float *__restrict__ lop) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const int N = S.ntau, nwv = S.ntp / DPW;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int p0 = t * TILE + 1, p1 = min(p0 + TILE - 1, S.nwl);
  const double Wt0 = S.wave[p0 - 1], Wt1 = S.wave[p1 - 1];  // tile edges
  int pp[PPL];
  bool valid[PPL];
  double Wp[PPL];
  float ofs[PPL];  // offset of the point from the tile's first point (sub-angstrom, exact in FP32)
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    pp[q] = p0 + lane + 64 * q;
    valid[q] = pp[q] <= S.nwl;
    Wp[q] = valid[q] ? S.wave[pp[q] - 1] : 0.0;
    ofs[q] = valid[q] ? (float)(Wp[q] - Wt0) : 0.0f;
  }
  const int d0 = wave * DPW;
  float acc[PPL][DPW], cmp[PPL][DPW], part[PPL][DPW];
#pragma unroll
  for (int q = 0; q < PPL; q++)
#pragma unroll
    for (int k = 0; k < DPW; k++) {
      acc[q][k] = 0.f;
      cmp[q][k] = 0.f;
      part[q][k] = 0.f;
    }
  int grp = 0;  // lines in the current partial sums (wave-uniform)
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * rstride + b];
    for (int j = r.x; j < r.y; j++) {
      const int4 m = meta_s[j];
      if (m.y < p0 || m.x > p1) continue;  // window misses the tile (wave-uniform)
      if (grp == FOLD) {
#pragma unroll
        for (int q = 0; q < PPL; q++)
#pragma unroll
          for (int k = 0; k < DPW; k++) {
            kahan_add(acc[q][k], cmp[q][k], part[q][k]);
            part[q][k] = 0.f;
          }
        grp = 0;
      }
      grp++;
      bool act[PPL];
#pragma unroll
      for (int q = 0; q < PPL; q++) act[q] = valid[q] && pp[q] >= m.x && pp[q] <= m.y;
      const double wc = wlc_s[j];
      const f4v *cp = reinterpret_cast<const f4v *>(ca + (size_t)j * S.ntp + d0);
      f4v rr[DPW];
#pragma unroll
      for (int k = 0; k < DPW; k++) rr[k] = cp[k];
      // depths of this wave that leave the far wing somewhere in the tile (bit k): the
      // far-wing expansion applies at depth k over the whole tile when the smallest
      // |lambda - lambda_c| (0 if the centre lies inside the tile) is at least dthr
      unsigned gen = (1u << DPW) - 1;
      if (!HE) {
        const double dmin = (wc < Wt0) ? (Wt0 - wc) : (wc > Wt1 ? (wc - Wt1) : 0.0);
        const unsigned dmb = __builtin_amdgcn_readfirstlane(__float_as_uint((float)dmin));
        gen = 0;
        if (dmb < __float_as_uint(dmaxw[(size_t)j * nwv + wave])) {
#pragma unroll
          for (int k = 0; k < DPW; k++)
            if (dmb < __float_as_uint(rr[k][3])) {
              gen |= 1u << k;
              rr[k][0] = rr[k][1] = rr[k][2] = 0.0f;
            }
        }
        // far wing, straight-line FMA code over all depths (zero coefficients at the general
        // depths).  dl = (point - first point) + (first point - centre), both rounded to FP32:
        // the relative error of dl stays below 2^-23 (1 + tile width / |dl|) where it is used
        const float bs = (float)(Wt0 - wc);
        float T2[PPL];
#pragma unroll
        for (int q = 0; q < PPL; q++) {
          const float x = ofs[q] + bs;
          T2[q] = act[q] ? __builtin_amdgcn_rcpf(fmaxf(x * x, 1e-30f)) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < DPW; k++)
#pragma unroll
          for (int q = 0; q < PPL; q++) {
            const float x = T2[q];
            part[q][k] += x * fmaf(fmaf(rr[k][2], x, rr[k][1]), x, rr[k][0]);
          }
      }
      if (gen) {
        // line core / near wing: the Voigt table or the continued fractions
        double dl[PPL];
#pragma unroll
        for (int q = 0; q < PPL; q++) dl[q] = Wp[q] - wc;
        const CoefB *cbp = cb + (size_t)j * S.ntp + d0;
        const double *hs = HE ? S.he_shf + (size_t)(m.w - 1) * S.ntp + d0 : nullptr;
#pragma unroll
        for (int k = 0; k < DPW; k++) {
          if (!(gen & (1u << k))) continue;
          const CoefB c = cbp[k];
#pragma unroll
          for (int q = 0; q < PPL; q++) {
            if (act[q]) {
              const float x = HE ? (float)(dl[q] - hs[k]) : (float)dl[q];
              part[q][k] += contrib32(S, c, x * c.vv);
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    if (!valid[q]) continue;
#pragma unroll
    for (int k = 0; k < DPW; k++) {
      const int d = d0 + k;
      if (d < N) {
        kahan_add(acc[q][k], cmp[q][k], part[q][k]);
        float v = acc[q][k] - cmp[q][k];
        if (S.motype == 1) v = v / (float)S.cstd[d];
        const size_t o = (size_t)d * S.nwl + (pp[q] - 1);
        lop[o] = lop[o] + v;
      }
    }
  }
}

// FP32 batched path on the matrix cores.  Outside its core zone (|lambda - lambda_c| <
// max_d dthr) the profile of line l at depth d and point p is the series
// T (k0 + T (k1 + T k2)) with T = 1 / (lambda_p - lambda_l)^2, so the far-wing opacity of a
// block of points is a sum of three matrix products,
//     Lop[d, p] += sum_n sum_l K_n[d, l] X_n[l, p],   X_n = T^(n+1)  (0 inside the core zone),
// and the core zones are added by k_opacity32_core.  v_mfma_f32_16x16x4_f32 takes 16 depths x
// 4 lines (A, the CoefA records as loaded) times 4 lines x 16 points (B, computed per lane)
// into a 16 x 16 FP32 block.  A workgroup owns a tile of TILE = 128 points and all depths (NQ
// blocks of 16); wavefront w owns points 32w .. 32w+31 (two 16-point blocks).  The tile's
// lines come from its list 4 per group; the four wavefronts share a ring of group slots in
// LDS whose loads run ahead of the products.  Partial sums hold FOLDG groups of lines and
// are then added to the running sums with Kahan compensation.
// Lane layout of v_mfma_f32_16x16x4_f32: A[r][k] in lane 16k + r, B[k][c] in lane 16k + c,
// D[4 (i / 16) + e][i % 16] in element e of lane i.
typedef float mf4 __attribute__((ext_vector_type(4)));
#ifndef FSG_FOLDG
#define FSG_FOLDG 32
#endif
constexpr int FOLDG = FSG_FOLDG;  // groups of 4 lines per partial sum

__device__ __forceinline__ void kahan_add4(mf4 &acc, mf4 &cmp, mf4 val) {
  const mf4 y = val - cmp;
  const mf4 t = acc + y;
  cmp = (t - acc) - y;
  acc = t;
}

// Reads one ring slot of the matrix-core kernel from LDS: NQ A fragments (lane order, 1 KB
// apart from byte address a) and the two halves of the lane's line record (at byte address
// r and r + 64).  One asm block with its own wait: the compiler does not see these reads, so
// it adds no wait for the direct global-to-LDS loads that fill other slots.
template <int NQ>
struct SlotRead;

template <>
struct SlotRead<1> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %3\n"
                 "ds_read_b128 %1, %4\n"
                 "ds_read_b128 %2, %4 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <>
struct SlotRead<2> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %4\n"
                 "ds_read_b128 %1, %4 offset:1024\n"
                 "ds_read_b128 %2, %5\n"
                 "ds_read_b128 %3, %5 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(A[1]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <>
struct SlotRead<3> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %5\n"
                 "ds_read_b128 %1, %5 offset:1024\n"
                 "ds_read_b128 %2, %5 offset:2048\n"
                 "ds_read_b128 %3, %6\n"
                 "ds_read_b128 %4, %6 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <>
struct SlotRead<4> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %6\n"
                 "ds_read_b128 %1, %6 offset:1024\n"
                 "ds_read_b128 %2, %6 offset:2048\n"
                 "ds_read_b128 %3, %6 offset:3072\n"
                 "ds_read_b128 %4, %7\n"
                 "ds_read_b128 %5, %7 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <>
struct SlotRead<5> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %7\n"
                 "ds_read_b128 %1, %7 offset:1024\n"
                 "ds_read_b128 %2, %7 offset:2048\n"
                 "ds_read_b128 %3, %7 offset:3072\n"
                 "ds_read_b128 %4, %7 offset:4096\n"
                 "ds_read_b128 %5, %8\n"
                 "ds_read_b128 %6, %8 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(A[4]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <>
struct SlotRead<6> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %8\n"
                 "ds_read_b128 %1, %8 offset:1024\n"
                 "ds_read_b128 %2, %8 offset:2048\n"
                 "ds_read_b128 %3, %8 offset:3072\n"
                 "ds_read_b128 %4, %8 offset:4096\n"
                 "ds_read_b128 %5, %8 offset:5120\n"
                 "ds_read_b128 %6, %9\n"
                 "ds_read_b128 %7, %9 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(A[4]), "=&v"(A[5]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <>
struct SlotRead<7> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %9\n"
                 "ds_read_b128 %1, %9 offset:1024\n"
                 "ds_read_b128 %2, %9 offset:2048\n"
                 "ds_read_b128 %3, %9 offset:3072\n"
                 "ds_read_b128 %4, %9 offset:4096\n"
                 "ds_read_b128 %5, %9 offset:5120\n"
                 "ds_read_b128 %6, %9 offset:6144\n"
                 "ds_read_b128 %7, %10\n"
                 "ds_read_b128 %8, %10 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(A[4]), "=&v"(A[5]), "=&v"(A[6]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <>
struct SlotRead<8> {
  __device__ __forceinline__ static void read(unsigned a, unsigned r, mf4 *A, mf4 &h0, mf4 &h1) {
    asm volatile("ds_read_b128 %0, %10\n"
                 "ds_read_b128 %1, %10 offset:1024\n"
                 "ds_read_b128 %2, %10 offset:2048\n"
                 "ds_read_b128 %3, %10 offset:3072\n"
                 "ds_read_b128 %4, %10 offset:4096\n"
                 "ds_read_b128 %5, %10 offset:5120\n"
                 "ds_read_b128 %6, %10 offset:6144\n"
                 "ds_read_b128 %7, %10 offset:7168\n"
                 "ds_read_b128 %8, %11\n"
                 "ds_read_b128 %9, %11 offset:64\n"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&v"(A[0]), "=&v"(A[1]), "=&v"(A[2]), "=&v"(A[3]), "=&v"(A[4]), "=&v"(A[5]), "=&v"(A[6]), "=&v"(A[7]), "=&v"(h0), "=&v"(h1)
                 : "v"(a), "v"(r));
  }
};

template <int NQ, int R>
__global__ void __launch_bounds__(256) k_opacity32_mfma(DevSetup S, int ntiles, const long long *__restrict__ loff,
                                                        const int *__restrict__ list, const CoefA *__restrict__ ca,
                                                        const LineRec *__restrict__ lrec, int64_t mnull,
                                                        float *__restrict__ lop) {
  // A ring of R group slots in LDS shared by the workgroup's 4 wavefronts (they all run the
  // tile's list), filled by direct global-to-LDS loads R - 1 groups ahead: the A fragments of
  // the group (NQ x 64 lanes x 16 B, in lane order; depth block nb loaded by wavefront nb % 4)
  // and the records of its 4 lines (two 16-byte halves each, by wavefront NQ % 4).
  static_assert(R >= 2 && R <= 8, "ring of 2 to 8 slots");
  constexpr int SLOT = NQ * 64 + 8;
  __shared__ mf4 ring[R][SLOT];
  typedef __attribute__((address_space(3))) void lds_void;
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;  // whole workgroups leave together (the barriers below)
  const int N = S.ntau, ntp = S.ntp, nwl = S.nwl;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int col = lane & 15, kq = lane >> 4;  // point / depth within a 16-block; line of the group
  const int p0 = t * TILE + 1;
  const double Wt0 = S.wave[p0 - 1];
  int pp[2];     // this lane's points (B columns)
  float ofs[2];  // their offsets from the tile's first point (sub-angstrom, exact in FP32)
#pragma unroll
  for (int q = 0; q < 2; q++) {
    pp[q] = p0 + 32 * wave + 16 * q + col;
    ofs[q] = pp[q] <= nwl ? (float)(S.wave[pp[q] - 1] - Wt0) : 0.0f;
  }
  mf4 C[2][NQ], acc[2][NQ], cmp[2][NQ];
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int nb = 0; nb < NQ; nb++) {
      C[q][nb] = mf4{0.f, 0.f, 0.f, 0.f};
      acc[q][nb] = mf4{0.f, 0.f, 0.f, 0.f};
      cmp[q][nb] = mf4{0.f, 0.f, 0.f, 0.f};
    }
  // the tile's lines in groups of 4 (k_far_lists): -1 past the end, R groups of -1 after
  // the last line, so the look-ahead of R - 1 groups stays inside the list
  const int4 *lg = reinterpret_cast<const int4 *>(list + loff[t]);
  // this wavefront's share of a group's loads (the same count every group: a past-the-end
  // line loads line 0's coefficients and the empty record at position mnull)
  constexpr int LW = NQ % 4;  // the wavefront that also loads the line records
  const int nla = (NQ - wave + 3) / 4;  // depth blocks wave, wave + 4, ... below NQ
  // per-lane choice among a group's 4 indices as two selects on fixed lane bits
  const bool kq1 = kq & 1, kq2 = kq & 2, ln1 = lane & 1, ln2 = lane & 2;
  auto fetch = [&](int s, const int4 jg) {
    const int j = kq2 ? (kq1 ? jg.w : jg.z) : (kq1 ? jg.y : jg.x);
    const mf4 *src = reinterpret_cast<const mf4 *>(ca + (size_t)(j >= 0 ? j : 0) * ntp + col);
    mf4 *slot = ring[s];
#pragma unroll
    for (int u = 0; u < (NQ + 3) / 4; u++) {
      const int nb = wave + 4 * u;
      if (nb < NQ)
        __builtin_amdgcn_global_load_lds((const void *)(src + 16 * nb), (lds_void *)(slot + 64 * nb), 16, 0, 0);
    }
    if (wave == LW && lane < 8) {  // lane i: half i / 4 of line i % 4
      const int jr = ln2 ? (ln1 ? jg.w : jg.z) : (ln1 ? jg.y : jg.x);
      const mf4 *rs = reinterpret_cast<const mf4 *>(lrec + (jr >= 0 ? (int64_t)jr : mnull)) + (lane >> 2);
      __builtin_amdgcn_global_load_lds((const void *)rs, (lds_void *)(slot + 64 * NQ), 16, 0, 0);
    }
    return jg.x >= 0;
  };
  int ng = 0;  // groups in the partial sums
  mf4 A[NQ], h0, h1;
  // the group in slot s: its A fragments and line records into registers
  auto read = [&](int s) {
    const unsigned sa = static_cast<unsigned>(reinterpret_cast<uintptr_t>((lds_void *)ring[s]));
    SlotRead<NQ>::read(sa + 16 * lane, sa + 1024 * NQ + 16 * kq, A, h0, h1);
  };
  auto run = [&]() {
    if (ng == FOLDG) {
#pragma unroll
      for (int q = 0; q < 2; q++)
#pragma unroll
        for (int nb = 0; nb < NQ; nb++) {
          kahan_add4(acc[q][nb], cmp[q][nb], C[q][nb]);
          C[q][nb] = mf4{0.f, 0.f, 0.f, 0.f};
        }
      ng = 0;
    }
    ng++;
    const int il = __float_as_int(h0[0]), ir = __float_as_int(h0[1]);
    const int zl = __float_as_int(h0[2]), zr = __float_as_int(h0[3]);
    const double wc = __hiloint2double(__float_as_int(h1[1]), __float_as_int(h1[0]));
    // B: T^(n+1) at this lane's (line, point); dl = (point - first point) + (first point -
    // centre) in FP32, relative error below 2^-23 (1 + tile width / |dl|) outside the core zone
    const float bs = (float)(Wt0 - wc);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const float x = ofs[q] + bs;
      const float T = __builtin_amdgcn_rcpf(fmaxf(x * x, 1e-12f));
      const bool act = (pp[q] >= il) & (pp[q] <= ir) & (pp[q] <= nwl) & ((pp[q] < zl) | (pp[q] > zr));
      float X[3];
      X[0] = act ? T : 0.0f;
      X[1] = X[0] * T;
      X[2] = X[1] * T;
#pragma unroll
      for (int n = 0; n < 3; n++)
#pragma unroll
        for (int nb = 0; nb < NQ; nb++)
          C[q][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[nb][n], X[n], C[q][nb], 0, 0, 0);
    }
  };
  // wait until at most n of this wavefront's vector memory operations are outstanding (asm,
  // so that the compiler neither moves nor merges it); n counts the loads of the R - 2 groups
  // issued after the one about to run
#define FSG_VMCNT(n) asm volatile("s_waitcnt vmcnt(%0)" ::"n"(n))
  int nvalid = -1;  // groups in the list, once the look-ahead has met its end
#pragma unroll
  for (int h = 0; h < R - 1; h++)
    if (!fetch(h, lg[h]) && nvalid < 0) nvalid = h;
  int4 jnext = lg[R - 1];  // the indices of the next group to fetch
  for (int g = 0; nvalid < 0 || g < nvalid; g++) {
    // this wavefront's loads of group g are complete; after the barrier every wavefront's
    // are, and every wavefront has finished reading group g - 1, whose slot is refilled next
    const int nl = nla + (wave == LW);
    if (nl == 1) FSG_VMCNT(1 * (R - 2));
    else if (nl == 2) FSG_VMCNT(2 * (R - 2));
    else FSG_VMCNT(3 * (R - 2));
    __builtin_amdgcn_s_barrier();
    const int h = g + R - 1;
    if (!fetch(h % R, jnext) && nvalid < 0) nvalid = h;
    read(g % R);
    jnext = lg[h + 1];  // in flight during the matrix products
    run();
  }
#undef FSG_VMCNT
#pragma unroll
  for (int q = 0; q < 2; q++) {
    if (pp[q] > nwl) continue;
#pragma unroll
    for (int nb = 0; nb < NQ; nb++) {
      kahan_add4(acc[q][nb], cmp[q][nb], C[q][nb]);
      const mf4 sum = acc[q][nb] - cmp[q][nb];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int d = 16 * nb + 4 * kq + e;
        if (d < N) {
          float v = sum[e];
          if (S.motype == 1) v = v / (float)S.cstd[d];
          const size_t o = (size_t)d * nwl + (pp[q] - 1);
          lop[o] = lop[o] + v;
        }
      }
    }
  }
}

// Core zones of the FP32 batched path: every line whose core zone [zl, zr] meets the tile
// inside its window, at the points of zone and window, every depth.  Per (line, depth) the
// far-wing expansion still applies where the whole tile lies beyond dthr (straight-line FMA
// code); the other depths run the Voigt table / continued fractions (contrib32).  Same layout
// as k_opacity32 (a tile of 64 CP points, one wavefront per DPW depths, CP points per lane);
// a line's records of the wavefront's DPW depths come in one vector load (lane k holds depth
// k) and reach all lanes by lane reads.  The candidates come bin by bin from the
// lines sorted by core-zone half width (zrec), NZB - 1 bins per tile in rng.
template <int CP>
__global__ void __launch_bounds__(1024) k_opacity32_core(DevSetup S, int ntiles, const int2 *__restrict__ rng,
                                                         const ZoneRec *__restrict__ zrec, const CoefA *__restrict__ ca,
                                                         const CoefB *__restrict__ cb, const float *__restrict__ dmaxw,
                                                         float *__restrict__ lop) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const int N = S.ntau;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  constexpr int TC = 64 * CP;  // points per tile
  const int p0 = t * TC + 1, p1 = min(p0 + TC - 1, S.nwl);
  const double Wt0 = S.wave[p0 - 1], Wt1 = S.wave[p1 - 1];
  int pp[CP];
  bool valid[CP];
  double Wp[CP];
  float ofs[CP];
#pragma unroll
  for (int q = 0; q < CP; q++) {
    pp[q] = p0 + lane + 64 * q;
    valid[q] = pp[q] <= S.nwl;
    Wp[q] = valid[q] ? S.wave[pp[q] - 1] : 0.0;
    ofs[q] = valid[q] ? (float)(Wp[q] - Wt0) : 0.0f;
  }
  const int d0 = wave * DPW;
  float acc[CP][DPW], cmp[CP][DPW], part[CP][DPW];
#pragma unroll
  for (int q = 0; q < CP; q++)
#pragma unroll
    for (int k = 0; k < DPW; k++) {
      acc[q][k] = 0.f;
      cmp[q][k] = 0.f;
      part[q][k] = 0.f;
    }
  int grp = 0;  // lines in the current partial sums (wave-uniform)
  for (int b = 0; b < NZB - 1; b++) {
    const int2 r = rng[(size_t)t * NZB + b];
    ZoneRec nx = zrec[r.x];  // one record ahead (the array holds one spare record at the end)
    for (int x = r.x; x < r.y; x++) {
      const ZoneRec zr = nx;
      nx = zrec[x + 1];
      const int j = zr.j;
      const int lo = max(zr.lo, p0), hi = min(zr.hi, p1);
      if (lo > hi) continue;  // zone and window miss the tile (wave-uniform)
      const double wc = zr.wc;
      const double dmin = (wc < Wt0) ? (Wt0 - wc) : (wc > Wt1 ? (wc - Wt1) : 0.0);
      const unsigned dmb = __builtin_amdgcn_readfirstlane(__float_as_uint((float)dmin));
      if (grp == FOLD) {
#pragma unroll
        for (int q = 0; q < CP; q++)
#pragma unroll
          for (int k = 0; k < DPW; k++) {
            kahan_add(acc[q][k], cmp[q][k], part[q][k]);
            part[q][k] = 0.f;
          }
        grp = 0;
      }
      grp++;
      bool act[CP];
#pragma unroll
      for (int q = 0; q < CP; q++) act[q] = valid[q] && pp[q] >= lo && pp[q] <= hi;
      // the records of the wavefront's depths with one load per lane (lane k: depth k), the
      // depths that leave the far wing somewhere in the tile by one ballot, and each depth's
      // record handed to all lanes by lane reads
      const int kl = lane < DPW ? lane : 0;
      const f4v ra = reinterpret_cast<const f4v *>(ca + (size_t)j * S.ntp + d0)[kl];  // k0, k1, k2, dthr
      const CoefB rb = cb[(size_t)j * S.ntp + d0 + kl];
      const unsigned gen =
          (unsigned)__builtin_amdgcn_ballot_w64(lane < DPW && dmb < __float_as_uint(ra[3]));  // bit k: depth k
      // far wing: straight-line FMA code over the depths that stay in it for the whole tile,
      // skipped when every depth of the wave takes the general path (wave-uniform);
      // dl = (point - first point) + (first point - centre), both rounded to FP32
      if (gen != (1u << DPW) - 1u) {
        const float bs = (float)(Wt0 - wc);
        float T2[CP];
#pragma unroll
        for (int q = 0; q < CP; q++) {
          const float xx = ofs[q] + bs;
          T2[q] = act[q] ? __builtin_amdgcn_rcpf(fmaxf(xx * xx, 1e-30f)) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < DPW; k++) {
          if ((gen >> k) & 1u) continue;
          const float k0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ra[0]), k));
          const float k1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ra[1]), k));
          const float k2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(ra[2]), k));
#pragma unroll
          for (int q = 0; q < CP; q++) {
            const float xx = T2[q];
            part[q][k] += xx * fmaf(fmaf(k2, xx, k1), xx, k0);
          }
        }
      }
      if (gen) {
        float dl[CP];
#pragma unroll
        for (int q = 0; q < CP; q++) dl[q] = (float)(Wp[q] - wc);
#pragma unroll
        for (int k = 0; k < DPW; k++) {
          if (!((gen >> k) & 1u)) continue;
          CoefB c;
          c.vv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.vv), k));
          c.a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.a), k));
          c.a2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.a2), k));
          c.c1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.c1), k));
          c.ka = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.ka), k));
          c.cop = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.cop), k));
          c.ua = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(rb.ua), k));
          c.ia = __builtin_amdgcn_readlane(rb.ia, k);
          float v[CP], r[CP];
#pragma unroll
          for (int q = 0; q < CP; q++) v[q] = dl[q] * c.vv;
          contrib32_lane(S, c, v, act, r);
#pragma unroll
          for (int q = 0; q < CP; q++) part[q][k] += r[q];
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < CP; q++) {
    if (!valid[q]) continue;
#pragma unroll
    for (int k = 0; k < DPW; k++) {
      const int d = d0 + k;
      if (d < N) {
        kahan_add(acc[q][k], cmp[q][k], part[q][k]);
        float v = acc[q][k] - cmp[q][k];
        if (S.motype == 1) v = v / (float)S.cstd[d];
        const size_t o = (size_t)d * S.nwl + (pp[q] - 1);
        lop[o] = lop[o] + v;
      }
    }
  }
}

// Narrow windows (the lines of the width bins whose half-widths stay below NARROW_HMAX points:
// molecular forests at moderate resolving power, a few points per line).  With one point per
// lane and 64 points per wavefront such a line would keep a few lanes busy; here a workgroup
// owns a tile of NARROW_TILE points and all depths, and each wavefront's lanes are
// NARROW_TILE points x 8 depths (lane = 8 depth + point offset), so a window of a few points
// fills most lanes.  The per-(line, depth) records of eight candidates come through one
// vector load (lane (depth, q) loads candidate q's record at its depth) and reach the lanes
// of each depth by cross-lane permutes; the next eight are loaded while these are
// evaluated.  Every point of a window takes the general path of the core-zone kernel (table
// in the box, one- or two-pole fraction outside: the reference's HVoigt in FP32), and the
// lines of the tile are summed in the order of their bins with the same blocked Kahan
// compensation (groups of FOLD lines).
__global__ void __launch_bounds__(1024) k_opacity32_narrow(DevSetup S, int ntiles, int nbins, const int2 *__restrict__ rng,
                                                           const CoefB *__restrict__ cb,
                                                           const double *__restrict__ wlc_s,
                                                           const int4 *__restrict__ meta_s, float *__restrict__ lop) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;  // workgroup-uniform
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int d = wave * 8 + (lane >> 3);
  const int dc = min(d, S.ntp - 1);
  const int slot = lane & 7;  // the candidate (of eight) whose record this lane loads
  const int p0 = t * NARROW_TILE + 1, p1 = min(p0 + NARROW_TILE - 1, S.nwl);
  const int p = p0 + (lane & 7);
  const bool valid = p <= p1 && d < S.ntau;
  const double Wp = p <= S.nwl ? S.wave[p - 1] : 0.0;
  float acc = 0.f, cmp = 0.f, part = 0.f;
  int grp = 0;  // lines in the partial sum (wave-uniform)
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * nbins + b];
    // candidates 64 at a time: each lane reads one candidate's window and centre, a ballot
    // keeps those whose window meets the tile, and they are taken in candidate order
    for (int base = r.x; base < r.y; base += 64) {
      const int jl = base + lane;
      int4 mm = make_int4(1, 0, 0, 0);
      double wl = 0.0;
      if (jl < r.y) {
        mm = meta_s[jl];
        wl = wlc_s[jl];
      }
      uint64_t rest = __builtin_amdgcn_ballot_w64(jl < r.y && mm.y >= p0 && mm.x <= p1);
      // eight candidates at a time: lane (depth, slot q) loads candidate q's record at its
      // depth, so one load instruction brings eight candidates' records; the next group's
      // loads are issued before the current group is evaluated
      auto take8 = [](uint64_t &rs, int &n) {
        uint64_t g = 0;
        n = 0;
        while (rs && n < 8) {
          g |= rs & (~rs + 1);
          rs &= rs - 1;
          n++;
        }
        return g;
      };
      auto load_slot = [&](uint64_t g, int n) {
        uint64_t u = g;
        for (int q = 0; q < slot && q < n - 1; q++) u &= u - 1;
        return cb[(size_t)(base + __builtin_ctzll(u)) * S.ntp + dc];
      };
      auto eval_group = [&](const CoefB &cur, uint64_t gq, int n) {
        for (int q = 0; q < n; q++) {
          const int k = __builtin_ctzll(gq);
          gq &= gq - 1;
          // candidate q's record at this lane's depth sits in lane (depth, q)
          const int src = (lane & ~7) | q;
          CoefB c;
          c.vv = __shfl(cur.vv, src, 64);
          c.a = __shfl(cur.a, src, 64);
          c.a2 = __shfl(cur.a2, src, 64);
          c.c1 = __shfl(cur.c1, src, 64);
          c.ka = __shfl(cur.ka, src, 64);
          c.cop = __shfl(cur.cop, src, 64);
          c.ua = __shfl(cur.ua, src, 64);
          c.ia = __shfl(cur.ia, src, 64);
          const int il = __builtin_amdgcn_readlane(mm.x, k), ir = __builtin_amdgcn_readlane(mm.y, k);
          const double wck = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(wl), k),
                                              __builtin_amdgcn_readlane(__double2loint(wl), k));
          if (grp == FOLD) {
            kahan_add(acc, cmp, part);
            part = 0.f;
            grp = 0;
          }
          grp++;
          const bool act = valid && p >= il && p <= ir;
          float v[1] = {(float)(Wp - wck) * c.vv}, out[1];
          const bool a1[1] = {act};
          contrib32_lane(S, c, v, a1, out);
          part += out[0];
        }
      };
      int ng = 0;
      uint64_t grpm = rest ? take8(rest, ng) : 0;
      CoefB mine = grpm ? load_slot(grpm, ng) : CoefB{};
      while (grpm) {
        int ngn = 0;
        const uint64_t grpn = rest ? take8(rest, ngn) : 0;
        const CoefB cur = mine;
        if (grpn) mine = load_slot(grpn, ngn);
        eval_group(cur, grpm, ng);
        grpm = grpn;
        ng = ngn;
      }
    }
  }
  if (!valid) return;
  kahan_add(acc, cmp, part);
  float v = acc - cmp;
  if (S.motype == 1) v = v / (float)S.cstd[d];
  const size_t o = (size_t)d * S.nwl + (p - 1);
  lop[o] = lop[o] + v;
}

// The matrix-core kernel's line lists: per tile, the positions of the candidate lines (the
// per-bin ranges of k_tile_ranges) that have a point of the tile inside their window and
// outside their core zone, in candidate order, padded with -1 to a multiple of 4 and followed
// by FAR_END_GROUPS groups of four -1.  cap[t] bounds the list: candidates rounded up to 4,
// plus the end groups.
__global__ void k_far_cap(int ntiles, int nbins, int rstride, const int2 *rng, long long *cap, LineRec *lrec,
                          int64_t mnull) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  if (t == ntiles) {  // the empty record that past-the-end group entries load: no window
    LineRec r;
    r.il = 1;
    r.ir = 0;
    r.zl = 1;
    r.zr = 0;
    r.wc = 0.0;
    r.pad0 = r.pad1 = 0;
    lrec[mnull] = r;
  }
  long long n = 0;
  if (t < ntiles)
    for (int b = 0; b < nbins; b++) {
      const int2 r = rng[(size_t)t * rstride + b];
      n += r.y - r.x;
    }
  cap[t] = t < ntiles ? ((n + 3) & ~3LL) + 4 * FAR_END_GROUPS : 0;  // the entry past the last tile: the total
}

__global__ void __launch_bounds__(64) k_far_lists(int ntiles, int nwl, int nbins, int rstride, const int2 *rng,
                                                  const LineRec *lrec, const long long *loff, int *list) {
  const int t = blockIdx.x;
  if (t >= ntiles) return;
  const int lane = threadIdx.x;
  const int p0 = t * TILE + 1, p1 = min(p0 + TILE - 1, nwl);
  int *out = list + loff[t];
  int n = 0;
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * rstride + b];
    for (int base = r.x; base < r.y; base += 64) {
      const int x = base + lane;
      bool ov = false;
      if (x < r.y) {
        const int4 w = *reinterpret_cast<const int4 *>(lrec + x);  // il, ir, zl, zr
        const int lo = max(w.x, p0), hi = min(w.y, p1);
        ov = (lo <= hi) & !((w.z <= lo) & (w.w >= hi));
      }
      const uint64_t m = __builtin_amdgcn_ballot_w64(ov);
      if (ov) out[n + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = x;
      n += __builtin_popcountll(m);
    }
  }
  const int end = ((n + 3) & ~3) + 4 * FAR_END_GROUPS;
  for (int i = n + lane; i < end; i += 64) out[i] = -1;
}

// core-zone sort keys: (zone bin << 32) | centre point, the bin being the bit length of the
// largest distance from the centre point to the ends of zone ∩ window; lines with no point in
// zone ∩ window go to the last bin (NZB - 1), which no tile reads
__global__ void k_zone_keys(int64_t m, int64_t mnarrow, int64_t mmain, int ntau, const int4 *meta_s,
                            const int2 *zone_s, unsigned long long *key, int *val, unsigned long long *evals_core) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long ev = 0;  // the wavefront's core-zone evaluations, added once
  if (j < m) {
    const int4 mm = meta_s[j];
    const int2 z = zone_s[j];
    const int lo = max(mm.x, z.x), hi = min(mm.y, z.y);
    unsigned b = NZB - 1;
    // positions below mnarrow hold the narrow lines and from mmain on the He I lines (both
    // have kernels of their own)
    if (j >= mnarrow && j < mmain && lo <= hi) {
      const int zh = max(max(mm.z - lo, hi - mm.z), 0);
      b = zh == 0 ? 0u : (unsigned)(32 - __clz(zh));
      ev = (unsigned long long)(hi - lo + 1) * (unsigned long long)ntau;
    }
    key[j] = ((unsigned long long)b << 32) | (unsigned)mm.z;
    val[j] = (int)j;
  }
  for (int off = 32; off > 0; off >>= 1) ev += __shfl_xor(ev, off, 64);
  if ((threadIdx.x & 63) == 0 && ev) atomicAdd(evals_core, ev);
}

// first and one-past-last sorted position of each zone bin (zero for empty bins)
__global__ void k_zone_bounds(int64_t m, const unsigned long long *key, int *zoff, int *zend) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const unsigned b = (unsigned)(key[i] >> 32);
  if (i == 0 || (unsigned)(key[i - 1] >> 32) != b) zoff[b] = (int)i;
  if (i == m - 1 || (unsigned)(key[i + 1] >> 32) != b) zend[b] = (int)(i + 1);
}

// the core-zone kernel's records in core-zone order (one scalar load per candidate)
__global__ void k_zone_records(int64_t m, const int *perm, const int4 *meta_s, const int2 *zone_s,
                               const double *wlc_s, ZoneRec *zrec) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= m) return;
  const int j = perm[x];
  const int4 mm = meta_s[j];
  const int2 z = zone_s[j];
  ZoneRec r;
  r.lo = max(mm.x, z.x);
  r.hi = min(mm.y, z.y);
  r.j = j;
  r.pad0 = 0;
  r.wc = wlc_s[j];
  r.pad1 = 0.0;
  zrec[x] = r;
}

__global__ void k_zero_f32(float *p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// dst += src, and src back to zero for the next use
__global__ void k_add_clear_f32(float *dst, float *src, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    dst[i] = dst[i] + src[i];
    src[i] = 0.0f;
  }
}

__global__ void k_fill_int(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// per tile and zone bin b (NZB - 1 of them), the sorted positions of the lines whose centre
// point lies within 2^b - 1 points of the tile (zh < 2^b in bin b)
__global__ void k_tile_ranges_zone(int ntiles, int tile, int nwl, const int *zoff, const int *zend,
                                   const unsigned long long *key, int2 *rng) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (t >= ntiles || b >= NZB - 1) return;
  const int p0 = t * tile + 1, p1 = min(p0 + tile - 1, nwl);
  const int off = zoff[b], cnt = zend[b] - zoff[b];
  const long h = (1L << b) - 1;
  const long lo = (long)p0 - h, hi = (long)p1 + h;
  int a = 0, e = cnt;
  while (a < e) {  // first with centre >= lo
    const int mid = (a + e) >> 1;
    if ((long)(unsigned)key[off + mid] < lo) a = mid + 1;
    else e = mid;
  }
  const int s = a;
  e = cnt;
  while (a < e) {  // first with centre > hi
    const int mid = (a + e) >> 1;
    if ((long)(unsigned)key[off + mid] <= hi) a = mid + 1;
    else e = mid;
  }
  rng[(size_t)t * NZB + b] = make_int2(off + s, off + a);
}

// FP64 plain path: the reference's arithmetic (table lookup in the box, continued
// fractions outside, plain FP64 sums in line order within a bin).
__global__ void __launch_bounds__(1024) k_opacity64(DevSetup S, int ntiles, int nbins, int rstride,
                                                    const int2 *__restrict__ rng, const Coef64 *__restrict__ coef,
                                                    const double *__restrict__ wlc_s,
                                                    const int4 *__restrict__ meta_s, double *__restrict__ lop) {
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  if (t >= ntiles) return;
  const int N = S.ntau;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int p0 = t * TILE + 1, p1 = min(p0 + TILE - 1, S.nwl);
  int pp[PPL];
  bool valid[PPL];
  double Wp[PPL];
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    pp[q] = p0 + lane + 64 * q;
    valid[q] = pp[q] <= S.nwl;
    Wp[q] = valid[q] ? S.wave[pp[q] - 1] : 0.0;
  }
  const int d0 = wave * DPW;
  double acc[PPL][DPW];
#pragma unroll
  for (int q = 0; q < PPL; q++)
#pragma unroll
    for (int k = 0; k < DPW; k++) acc[q][k] = 0.0;
  for (int b = 0; b < nbins; b++) {
    const int2 r = rng[(size_t)t * rstride + b];
    for (int j = r.x; j < r.y; j++) {
      const int4 m = meta_s[j];
      if (m.y < p0 || m.x > p1) continue;
      const double wc = wlc_s[j];
      const Coef64 *c = coef + (size_t)j * S.ntp + d0;
      const double *hs = (m.w > 0) ? S.he_shf + (size_t)(m.w - 1) * S.ntp + d0 : nullptr;
#pragma unroll
      for (int q = 0; q < PPL; q++) {
        if (!(valid[q] && pp[q] >= m.x && pp[q] <= m.y)) continue;
        const double dl = Wp[q] - wc;
#pragma unroll
        for (int k = 0; k < DPW; k++) {
          const double dld = hs ? dl - hs[k] : dl;
          acc[q][k] += contrib64(S, c[k], dld * c[k].vv);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPL; q++) {
    if (!valid[q]) continue;
#pragma unroll
    for (int k = 0; k < DPW; k++) {
      const int d = d0 + k;
      if (d < N) {
        double v = acc[q][k];
        if (S.motype == 1) v = v / S.cstd[d];
        const size_t o = (size_t)d * S.nwl + (pp[q] - 1);
        lop[o] = lop[o] + v;
      }
    }
  }
}

// ------------------------------------------------------------------------------------
// scalar plane-parallel DELO-Bezier (LocPrfPP) for the continuum and the line+continuum
// opacity of one wavelength per lane, streamed from the bottom of the atmosphere up with
// a sliding window of depth points; FP64 throughout.  Fused flux sums.
// ------------------------------------------------------------------------------------

// TauStep (src/solver.f90:591-615) element i (0-based, i <= N-2), Xmu = 1, from the
// opacities xm, x0, xp at depths i - 1, i, i + 1 (i = 0 uses x0 and xp only)
__device__ __forceinline__ double tau_step_v(const double *rhox, int i, double xm, double x0, double xp) {
  if (i == 0) {
    double Dx1 = (rhox[1] - rhox[0]) / 1.0;
    double Yp = (xp - x0) / Dx1;
    return Dx1 * (xp + 2.0 * x0 + Dx1 / 2.0 * Yp) / 3.0;
  }
  double Dx = (rhox[i] - rhox[i - 1]) / 1.0;
  double Dx1 = (rhox[i + 1] - rhox[i]) / 1.0;
  double Der = (x0 - xm) / Dx, Der1 = (xp - x0) / Dx1, Yp;
  if (Der * Der1 > 0.0) {
    double Lambda = (1.0 + Dx1 / (Dx + Dx1)) / 3.0;
    Yp = (Der / (Lambda * Der1 + (1.0 - Lambda) * Der)) * Der1;
  } else {
    Yp = 0.0;
  }
  return Dx1 * (xp + 2.0 * x0 + Dx1 / 2.0 * Yp) / 3.0;
}

// Emergent intensities at all mu of one opacity column (LocPrfPP, src/solver.f90:123-179),
// streamed from the bottom up; col.k(d) is the opacity at depth d, read once per depth (a
// window of the opacities at the two depths above the last optical-depth step).
template <typename Col>
__device__ __forceinline__ void locprf_pp(const DevSetup &S, const Col &col, double C1, double C2, const double *imu,
                                          double *Y) {
  const int N = S.ntau, M = S.nmu;
  int i = N - 2;
  double S_ip1 = planck_src(C1, C2, S.temp[N - 1]);
  double S_i = planck_src(C1, C2, S.temp[N - 2]);
  double S_im1 = (N >= 3) ? planck_src(C1, C2, S.temp[N - 3]) : 0.0;
  const double kp = col.k(N - 1), k0 = col.k(N - 2);
  double kb = (N >= 3) ? col.k(N - 3) : 0.0;  // k(i - 1)
  double ka = (N >= 4) ? col.k(N - 4) : 0.0;  // k(i - 2)
  double dt_i = tau_step_v(S.rhox, i, kb, k0, kp);
  double dt_im1 = (i >= 1) ? tau_step_v(S.rhox, i - 1, ka, kb, k0) : 0.0;
  double Sd_ip1 = (S_ip1 - S_i) / dt_i;  // Srd(N)
  for (int m = 0; m < M; m++) Y[m] = S_ip1;
  for (; i >= 0; i--) {
    double Sd_i = (i == 0) ? (S_ip1 - S_i) / dt_i : cent_der_i(dt_im1, dt_i, S_im1, S_i, S_ip1);
    double U = 0.5 * (S_i + 0.5 * dt_i * Sd_i + S_ip1 - 0.5 * dt_i * Sd_ip1);
    for (int m = 0; m < M; m++) {
      double Eps, Alpha, Beta, Gamma;
      delo_coef(dt_i * imu[m], Eps, Alpha, Beta, Gamma);
      Y[m] = Y[m] * Eps + Alpha * S_i + Beta * S_ip1 + Gamma * U;
    }
    // shift the window up by one depth point
    Sd_ip1 = Sd_i;
    S_ip1 = S_i;
    S_i = S_im1;
    dt_i = dt_im1;
    if (i >= 2) {
      S_im1 = planck_src(C1, C2, S.temp[i - 2]);
      const double kn = (i >= 3) ? col.k(i - 3) : 0.0;
      dt_im1 = tau_step_v(S.rhox, i - 2, kn, ka, kb);
      kb = ka;
      ka = kn;
    }
  }
}

// DeloScalar along ray im (LocPrfSph, src/solver.f90:182-261) for NC opacity columns that
// share the source function: points gathered through the ray map, optical depth steps on
// the ray's column mass (Xmu = 1), streamed from the deepest point upward.  Grazing rays
// enter with zero intensity, the others with S at the deepest point.
template <int NC, typename Col>
__device__ void delo_ray(const DevSetup &S, const Col *cols, int im, double C1, double C2, double *Y) {
  const int L = 2 * S.ntau, N = S.ntau_mu[im];
  const int *map = S.map_mu + (size_t)im * L;
  const double *rm = S.rhox_mu + (size_t)im * L;
  auto src = [&](int j) { return planck_src(C1, C2, S.temp[map[j]]); };
  auto tau = [&](int c, int j) {  // Deltau(j) of column c
    const double km = (j >= 1) ? cols[c].k(map[j - 1]) : 0.0;
    return tau_step_mu(rm, 1.0, j, km, cols[c].k(map[j]), cols[c].k(map[j + 1]));
  };
  int i = N - 2;
  double S_ip1 = src(N - 1), S_i = src(N - 2), S_im1 = src(N - 3);
  double dt_i[NC], dt_im1[NC], Sd_ip1[NC];
  for (int c = 0; c < NC; c++) {
    dt_i[c] = tau(c, i);
    dt_im1[c] = tau(c, i - 1);
    Sd_ip1[c] = (S_ip1 - S_i) / dt_i[c];
    Y[c] = S.graz_mu[im] ? 0.0 : S_ip1;
  }
  for (; i >= 0; i--) {
    for (int c = 0; c < NC; c++) {
      const double Sd_i = (i == 0) ? (S_ip1 - S_i) / dt_i[c] : cent_der_i(dt_im1[c], dt_i[c], S_im1, S_i, S_ip1);
      double Eps, Alpha, Beta, Gamma;
      delo_coef(dt_i[c], Eps, Alpha, Beta, Gamma);
      const double U = 0.5 * (S_i + 0.5 * dt_i[c] * Sd_i + S_ip1 - 0.5 * dt_i[c] * Sd_ip1[c]);
      Y[c] = Y[c] * Eps + Alpha * S_i + Beta * S_ip1 + Gamma * U;
      Sd_ip1[c] = Sd_i;
      dt_i[c] = dt_im1[c];
    }
    S_ip1 = S_i;
    S_i = S_im1;
    if (i >= 2) {
      S_im1 = src(i - 2);
      for (int c = 0; c < NC; c++) dt_im1[c] = tau(c, i - 2);
    }
  }
}

// spherical RadTranCalc (src/synth.f90:1873-1895): continuum and line+continuum along each
// ray, one lane per wavelength point, flux sums fused
template <typename LopT>
__global__ void __launch_bounds__(128) k_radtran_sph(DevSetup S, const LopT *lop, double *spl, double *spc,
                                                     double *sfl, double *sfc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= S.nwl) return;
  const double Wl = S.wave[p] + S.wfirst;
  const double C1 = kPlkC1 / Wl;
  const double C2 = exp(50.7649141 - 5.0 * log(Wl));
  const GridCol<LopT> cols[2] = {{S.contT, lop, S.nwl, p, false}, {S.contT, lop, S.nwl, p, true}};
  double fc = 0.0, fl = 0.0;
  for (int m = 0; m < S.nmu; m++) {
    double Y[2];
    delo_ray<2>(S, cols, m, C1, C2, Y);
    spc[(size_t)m * S.nwl + p] = Y[0];
    spl[(size_t)m * S.nwl + p] = Y[1];
    fc = fc + Y[0] * S.wmu[m];
    fl = fl + Y[1] * S.wmu[m];
  }
  sfc[p] = fc;
  sfl[p] = fl;
}

template <typename LopT>
__global__ void __launch_bounds__(256) k_radtran(DevSetup S, const LopT *lop, double *spl, double *spc,
                                                 double *sfl, double *sfc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= S.nwl) return;
  point_setup(S, p);
  const int M = S.nmu;
  double Wl = S.wave[p] + S.wfirst;
  double C1 = kPlkC1 / Wl;
  double C2 = exp(50.7649141 - 5.0 * log(Wl));
  double imu[MAXMU], Y[MAXMU];
  for (int m = 0; m < M; m++) imu[m] = 1.0 / S.xmu[m];
  for (int pass = 0; pass < 2; pass++) {  // 0: continuum (SpC), 1: line + continuum (SpL)
    GridCol<LopT> col{S.contT, lop, S.nwl, p, pass == 1};
    locprf_pp(S, col, C1, C2, imu, Y);
    double *out = pass == 1 ? spl : spc;
    double f = 0.0;
    for (int m = 0; m < M; m++) {
      out[(size_t)m * S.nwl + p] = Y[m];
      f = f + Y[m] * S.wmu[m];
    }
    (pass == 1 ? sfl : sfc)[p] = f;
  }
}

// opacity column at a line centre for LinStrCalc: continuum interpolated to the centre and
// the line's own opacity there (LineOpacity2 at the centre, src/synth.f90:867-888)
struct LineCentreCol {
  const DevSetup *S;
  const XC *xc;
  int iwl, ihe;
  double dW;
  bool with_line;
  __device__ __forceinline__ double k(int d) const {
    const double *ck = S->contT + (size_t)d * S->nwl;
    double kc = ck[iwl - 1] * (1.0 - dW) + ck[iwl] * dW;
    if (!with_line) return kc;
    double cop, vv, av;
    xline_depth(*S, d, *xc, cop, vv, av);
    double dWl = (ihe > 0) ? 0.0 - S->he_shf[(size_t)(ihe - 1) * S->ntp + d] : 0.0;
    double lo = hvoigt64_full(*S, av, dWl * vv) * cop;
    if (S->motype == 1) lo = lo / S->cstd[d];
    return kc + lo;
  }
};

// LinStrCalc (src/synth.f90:1774-1822): central depth of each retained line
__global__ void __launch_bounds__(128) k_linstr(DevSetup S0, DevLines L, int list_type, const int *flag,
                                                float *linstr) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  const DevSetup S = line_setup(S0, L, i);
  int f = flag[i];
  if (f & 1) {
    linstr[i] = 0.0f;
    return;
  }
  const int M = S.nmu;
  double wlc = L.wl[i] - S.wfirst;
  XC xc = xline_setup(S, L, i, wlc, f & 0x6, list_type);
  int iwl = wave_idx(S, wlc);
  double dW = (wlc - S.wave[iwl - 1]) / (S.wave[iwl] - S.wave[iwl - 1]);
  double Wl = wlc + S.wfirst;
  double C1 = kPlkC1 / Wl;
  double C2 = exp(50.7649141 - 5.0 * log(Wl));
  double imu[MAXMU], IntC[MAXMU], IntL[MAXMU];
  for (int m = 0; m < M; m++) imu[m] = 1.0 / S.xmu[m];
  const LineCentreCol cols[2] = {{&S, &xc, iwl, xc.IxHe, dW, false}, {&S, &xc, iwl, xc.IxHe, dW, true}};
  if (S.sph) {
    for (int m = 0; m < M; m++) {
      double Y[2];
      delo_ray<2>(S, cols, m, C1, C2, Y);
      IntC[m] = Y[0];
      IntL[m] = Y[1];
    }
  } else {
    locprf_pp(S, cols[0], C1, C2, imu, IntC);
    locprf_pp(S, cols[1], C1, C2, imu, IntL);
  }
  float r;
  if (M == 1) {
    r = (float)(1.0 - IntL[0] / IntC[0]);
  } else {
    double FlxL = 0.0, FlxC = 0.0;
    for (int m = 0; m < M; m++) {
      FlxL = FlxL + S.wmu[m] * IntL[m];
      FlxC = FlxC + S.wmu[m] * IntC[m];
    }
    r = (float)(1.0 - FlxL / FlxC);
  }
  linstr[i] = r;
}

// ------------------------------------------------------------------------------------
// host-side launchers (the only entry points the API translation unit uses)
// ------------------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

void launch_voigt_table(float *tab, const double *anode, double dv, hipStream_t st) {
  dim3 grid(nblk(NVV + 1, 256), NVA + 1);
  k_voigt_table<<<grid, 256, 0, st>>>(tab, anode, dv);
}

void launch_xline_init(int ntau, const double *temp, const double *xne, const float *fh1, const float *fhe1,
                       const float *fhmol, double h2brd, double *xln, hipStream_t st) {
  k_xline_init<<<nblk(ntau, 64), 64, 0, st>>>(ntau, temp, xne, fh1, fhe1, fhmol, h2brd, xln);
}

void launch_pass1(const DevSetup &S, const DevLines &L, int list_type, int *flag, int *il, int *ir, int *ic,
                  unsigned *key, int *bin_count, int *bin_hmax, int batched, hipStream_t st) {
  if (L.n == 0) return;
  k_pass1<<<nblk(L.n, 128), 128, 0, st>>>(S, L, list_type, flag, il, ir, ic, key, bin_count, bin_hmax, batched);
}

void launch_gather_lines(const DevLines &src, const int64_t *idx, int64_t n, double *wl, float *elo, float *gf,
                         float *grad, float *gqst, float *gvdw, int *spc, uint8_t *is_mol, hipStream_t st) {
  if (n > 0) k_gather_lines<<<nblk(n, 256), 256, 0, st>>>(src, idx, n, wl, elo, gf, grad, gqst, gvdw, spc, is_mol);
}

void launch_iota(int *v, int64_t n, hipStream_t st) {
  if (n) k_iota<<<nblk(n, 256), 256, 0, st>>>(v, n);
}

void launch_count_evals(int64_t n, const int *flag, const int *il, const int *ir, int ntau,
                        unsigned long long *acc, hipStream_t st) {
  if (n) k_count_evals<<<(unsigned)std::min<int64_t>(COUNT_BLOCKS, nblk(n, 256)), 256, 0, st>>>(n, flag, il, ir, ntau, acc);
}

void launch_retained_flag(int64_t n, const int *flag, int *out, hipStream_t st) {
  if (n) k_retained_flag<<<nblk(n, 256), 256, 0, st>>>(n, flag, out);
}

void launch_wlc_list(int64_t n, const double *wl, double wfirst, const int *flag, const int *pos, double *out,
                     int pack, hipStream_t st) {
  if (n) k_wlc_list<<<nblk(n, 256), 256, 0, st>>>(n, wl, wfirst, flag, pos, out, pack);
}

void launch_mark_wide(const DevSetup &S, int64_t n, const double *wl, const int *flag, const int *il, const int *ir,
                      const double *wlc_list, int64_t m, int64_t nstep, unsigned long long *nwide, hipStream_t st) {
  if (n) k_mark_wide<<<nblk(n, 256), 256, 0, st>>>(S, n, wl, flag, il, ir, wlc_list, m, nstep, nwide);
}

// ContOpac(Ntau, Nwl) as the host holds it (depth fastest) into the device layout
// [ntau][nwl] (wavelength fastest): 32 x 32 tiles through LDS, both sides coalesced
__global__ void __launch_bounds__(256) k_transpose_cont(int64_t nwl, int ntau, const double *__restrict__ src,
                                                        double *__restrict__ dst) {
  __shared__ double tile[32][33];
  const int64_t w0 = (int64_t)blockIdx.x * 32;
  const int k0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += 8) {
    const int64_t w = w0 + r;
    const int k = k0 + tx;
    if (w < nwl && k < ntau) tile[r][tx] = src[w * ntau + k];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int k = k0 + r;
    const int64_t w = w0 + tx;
    if (w < nwl && k < ntau) dst[(size_t)k * nwl + w] = tile[tx][r];
  }
}

void launch_transpose_cont(int64_t nwl, int ntau, const double *src, double *dst, hipStream_t st) {
  if (nwl <= 0 || ntau <= 0) return;
  const dim3 grid((unsigned)((nwl + 31) / 32), (unsigned)((ntau + 31) / 32));
  k_transpose_cont<<<grid, 256, 0, st>>>(nwl, ntau, src, dst);
}

void launch_fill32(const DevSetup &S, const DevLines &L, int list_type, int64_t nret, const int *order,
                   const int *flag, const int *il, const int *ir, const int *ic, CoefA *ca, CoefB *cb, float *dmaxw,
                   int2 *zone_s, LineRec *lrec, double *wlc_s, int4 *meta_s, hipStream_t st) {
  if (nret)
    k_fill32<<<nblk(nret, FILL_LINES), 64, 0, st>>>(S, L, list_type, nret, order, flag, il, ir, ic, ca, cb, dmaxw, zone_s,
                                              lrec, wlc_s, meta_s);
}


void launch_fill64(const DevSetup &S, const DevLines &L, int list_type, int64_t nret, const int *order,
                   const int *flag, const int *il, const int *ir, const int *ic, Coef64 *coef, double *wlc_s,
                   int4 *meta_s, hipStream_t st) {
  if (nret) k_fill64<<<nblk(nret, 128), 128, 0, st>>>(S, L, list_type, nret, order, flag, il, ir, ic, coef, wlc_s, meta_s);
}

void launch_tile_ranges(int ntiles, int tile, int nbins, int nwl, const int *bin_off, const int *bin_cnt,
                        const int *bin_hmax, const int4 *meta_s, int2 *rng, hipStream_t st) {
  dim3 grid(nblk(ntiles, 256), nbins);
  k_tile_ranges<<<grid, 256, 0, st>>>(ntiles, tile, nbins, nwl, bin_off, bin_cnt, bin_hmax, meta_s, rng);
}

// returns 1 when the matrix-core kernel ran (the core zones then need launch_opacity32_core)
size_t far_scan_tmp_bytes(int ntiles) {
  size_t b = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (long long *)nullptr, (long long *)nullptr, ntiles + 1);
  return b;
}

void launch_far_cap(int ntiles, int nbins, int rstride, const int2 *rng, long long *cap, long long *loff, void *tmp,
                    size_t tmp_bytes, LineRec *lrec, int64_t mnull, hipStream_t st) {
  k_far_cap<<<nblk(ntiles + 1, 256), 256, 0, st>>>(ntiles, nbins, rstride, rng, cap, lrec, mnull);
  (void)hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cap, loff, ntiles + 1, st);
}

void launch_far_lists(int ntiles, int nwl, int nbins, int rstride, const int2 *rng, const LineRec *lrec,
                      const long long *loff, int *list, hipStream_t st) {
  if (ntiles > 0) k_far_lists<<<ntiles, 64, 0, st>>>(ntiles, nwl, nbins, rstride, rng, lrec, loff, list);
}

int launch_opacity32(const DevSetup &S, int ntiles, int nbins, int rstride, const int2 *rng, const CoefA *ca,
                     const CoefB *cb, const float *dmaxw, const LineRec *lrec, const double *wlc_s,
                     const int4 *meta_s, const long long *loff, const int *list, int64_t mnull, float *lineop,
                     int he, int ring, hipStream_t st) {
  int nw = (S.ntau + DPW - 1) / DPW;
  int nb = 8 * ((ntiles + 7) / 8);
  const int nq = (S.ntau + 15) / 16;
  if (he) {
    k_opacity32<true><<<nb, 64 * nw, 0, st>>>(S, ntiles, nbins, rstride, rng, ca, cb, dmaxw, wlc_s, meta_s, lineop);
    return 0;
  }
  static_assert(TILE == 128, "the matrix-core kernel covers 128 points with 4 wavefronts");
  switch (nq) {  // depth blocks of 16: up to 128 depths on the matrix cores
#define FSG_MFMA_CASE(Q)                                                                                  \
  case Q:                                                                                                \
    if (ring == 2) k_opacity32_mfma<Q, 2><<<nb, 256, 0, st>>>(S, ntiles, loff, list, ca, lrec, mnull, lineop); \
    else k_opacity32_mfma<Q, 4><<<nb, 256, 0, st>>>(S, ntiles, loff, list, ca, lrec, mnull, lineop);          \
    return 1;
    FSG_MFMA_CASE(1)
    FSG_MFMA_CASE(2)
    FSG_MFMA_CASE(3)
    FSG_MFMA_CASE(4)
    FSG_MFMA_CASE(5)
    FSG_MFMA_CASE(6)
    FSG_MFMA_CASE(7)
    FSG_MFMA_CASE(8)
#undef FSG_MFMA_CASE
    default:
      k_opacity32<false><<<nb, 64 * nw, 0, st>>>(S, ntiles, nbins, rstride, rng, ca, cb, dmaxw, wlc_s, meta_s, lineop);
      return 0;
  }
}

void launch_zone_order(int64_t m, int64_t mnarrow, int64_t mmain, int ntiles, int nwl, int ntau, const int4 *meta_s,
                       const int2 *zone_s, const double *wlc_s, unsigned long long *key, int *val,
                       unsigned long long *key_sorted, int *perm, ZoneRec *zrec, void *tmp, size_t tmp_bytes,
                       int *zoff, int *zend, unsigned long long *evals_core, hipStream_t st) {
  k_fill_int<<<1, 2 * NZB, 0, st>>>(zoff, 2 * NZB, 0);  // zoff and zend are adjacent
  if (m > 0) {
    k_zone_keys<<<nblk(m, 256), 256, 0, st>>>(m, mnarrow, mmain, ntau, meta_s, zone_s, key, val, evals_core);
    (void)hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key_sorted, val, perm, (int)m, 0, 64, st);
    k_zone_bounds<<<nblk(m, 256), 256, 0, st>>>(m, key_sorted, zoff, zend);
    k_zone_records<<<nblk(m, 256), 256, 0, st>>>(m, perm, meta_s, zone_s, wlc_s, zrec);
  }
}

void launch_core_ranges(int core_pp, int nwl, const int *zoff, const int *zend, const unsigned long long *key_sorted,
                        int2 *rng_core, hipStream_t st) {
  const int tile = 64 * core_pp, ntiles = (nwl + tile - 1) / tile;
  dim3 grid(nblk(ntiles, 256), NZB - 1);
  k_tile_ranges_zone<<<grid, 256, 0, st>>>(ntiles, tile, nwl, zoff, zend, key_sorted, rng_core);
}

void launch_zero_f32(float *p, int64_t n, hipStream_t st) {
  if (n) k_zero_f32<<<nblk(n, 256), 256, 0, st>>>(p, n);
}

void launch_add_clear_f32(float *dst, float *src, int64_t n, hipStream_t st) {
  if (n) k_add_clear_f32<<<nblk(n, 256), 256, 0, st>>>(dst, src, n);
}

size_t zone_sort_tmp_bytes(int64_t m) {
  size_t b = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, b, (unsigned long long *)nullptr, (unsigned long long *)nullptr,
                                           (int *)nullptr, (int *)nullptr, (int)m, 0, 64);
  return b;
}

void launch_opacity32_core(const DevSetup &S, int core_pp, const int2 *rng_core, const ZoneRec *zrec, const CoefA *ca,
                           const CoefB *cb, const float *dmaxw, float *lineop, hipStream_t st) {
  const int nw = (S.ntau + DPW - 1) / DPW;
  const int tile = 64 * core_pp, ntiles = (S.nwl + tile - 1) / tile;
  const int nb = 8 * ((ntiles + 7) / 8);
  if (core_pp == 1) k_opacity32_core<1><<<nb, 64 * nw, 0, st>>>(S, ntiles, rng_core, zrec, ca, cb, dmaxw, lineop);
  else k_opacity32_core<2><<<nb, 64 * nw, 0, st>>>(S, ntiles, rng_core, zrec, ca, cb, dmaxw, lineop);
}

void launch_opacity32_narrow(const DevSetup &S, int nbins, const int2 *rng, const CoefB *cb, const double *wlc_s,
                             const int4 *meta_s, float *lineop, hipStream_t st) {
  const int ntiles = (S.nwl + NARROW_TILE - 1) / NARROW_TILE;
  const int nw = (S.ntau + 7) / 8;
  if (nbins > 0 && ntiles > 0)
    k_opacity32_narrow<<<8 * ((ntiles + 7) / 8), 64 * nw, 0, st>>>(S, ntiles, nbins, rng, cb, wlc_s, meta_s, lineop);
}

void launch_opacity64(const DevSetup &S, int ntiles, int nbins, int rstride, const int2 *rng, const Coef64 *coef,
                      const double *wlc_s, const int4 *meta_s, double *lineop, hipStream_t st) {
  int nw = (S.ntau + DPW - 1) / DPW;
  int nb = 8 * ((ntiles + 7) / 8);
  k_opacity64<<<nb, 64 * nw, 0, st>>>(S, ntiles, nbins, rstride, rng, coef, wlc_s, meta_s, lineop);
}

void launch_linstr(const DevSetup &S, const DevLines &L, int list_type, const int *flag, float *linstr,
                   hipStream_t st) {
  if (L.n) k_linstr<<<nblk(L.n, 128), 128, 0, st>>>(S, L, list_type, flag, linstr);
}

void launch_radtran32(const DevSetup &S, const float *lop, double *spl, double *spc, double *sfl, double *sfc,
                      hipStream_t st) {
  if (S.sph) k_radtran_sph<float><<<nblk(S.nwl, 128), 128, 0, st>>>(S, lop, spl, spc, sfl, sfc);
  else k_radtran<float><<<nblk(S.nwl, 256), 256, 0, st>>>(S, lop, spl, spc, sfl, sfc);
}

void launch_radtran64(const DevSetup &S, const double *lop, double *spl, double *spc, double *sfl, double *sfc,
                      hipStream_t st) {
  if (S.sph) k_radtran_sph<double><<<nblk(S.nwl, 128), 128, 0, st>>>(S, lop, spl, spc, sfl, sfc);
  else k_radtran<double><<<nblk(S.nwl, 256), 256, 0, st>>>(S, lop, spl, spc, sfl, sfc);
}

template __global__ void k_radtran<float>(DevSetup, const float *, double *, double *, double *, double *);
template __global__ void k_radtran<double>(DevSetup, const double *, double *, double *, double *, double *);

}  // namespace fsg
