// fsg_kernels.h -- launchers of the HIP kernels (defined in fsg_kernels.hip).
#pragma once

#include <type_traits>

#include <vector>

#include "fsg_common.h"

namespace fsg {

void launch_voigt_table(float *tab, const double *anode, double dv, hipStream_t st);
void launch_xline_init(int ntau, const double *temp, const double *xne, const float *fh1, const float *fhe1,
                       const float *fhmol, double h2brd, double *xln, hipStream_t st);
void launch_pass1(const DevSetup &S, const DevLines &L, int list_type, int *flag, int *il, int *ir, int *ic,
                  unsigned *key, int *bin_count, int *bin_hmax, int batched, hipStream_t st);
void launch_iota(int *v, int64_t n, hipStream_t st);
void launch_gather_lines(const DevLines &src, const int64_t *idx, int64_t n, double *wl, float *elo, float *gf,
                         float *grad, float *gqst, float *gvdw, int *spc, uint8_t *is_mol, hipStream_t st);
void launch_count_evals(int64_t n, const int *flag, const int *il, const int *ir, int ntau,
                        unsigned long long *acc, hipStream_t st);
void launch_retained_flag(int64_t n, const int *flag, int *out, hipStream_t st);
void launch_wlc_list(int64_t n, const double *wl, double wfirst, const int *flag, const int *pos, double *out,
                     int pack, hipStream_t st);
void launch_mark_wide(const DevSetup &S, int64_t n, const double *wl, const int *flag, const int *il, const int *ir,
                      const double *wlc_list, int64_t m, int64_t nstep, unsigned long long *nwide, hipStream_t st);
void launch_fill32(const DevSetup &S, const DevLines &L, int list_type, int64_t nret, const int *order,
                   const int *flag, const int *il, const int *ir, const int *ic, CoefA *ca, CoefB *cb, float *dmaxw,
                   int2 *zone_s, LineRec *lrec, double *wlc_s, int4 *meta_s, hipStream_t st);
void launch_fill64(const DevSetup &S, const DevLines &L, int list_type, int64_t nret, const int *order,
                   const int *flag, const int *il, const int *ir, const int *ic, Coef64 *coef, double *wlc_s,
                   int4 *meta_s, hipStream_t st);
void launch_tile_ranges(int ntiles, int tile, int nbins, int nwl, const int *bin_off, const int *bin_cnt,
                        const int *bin_hmax, const int4 *meta_s, int2 *rng, hipStream_t st);
int launch_opacity32(const DevSetup &S, int ntiles, int nbins, int rstride, const int2 *rng, const CoefA *ca,
                     const CoefB *cb, const float *dmaxw, const LineRec *lrec, const double *wlc_s,
                     const int4 *meta_s, const long long *loff, const int *list, int64_t mnull, float *lineop,
                     int he, int ring, hipStream_t st);
// the matrix-core kernel's per-tile line lists: capacities and their offsets (ntiles + 1
// entries, the last one the total), then the lists
constexpr int MFMA_MAX_NQ = 8;     // depth blocks of 16 on the matrix cores (up to 128 depths)
constexpr int FAR_END_GROUPS = 8;  // groups of -1 after each list: the deepest look-ahead
size_t far_scan_tmp_bytes(int ntiles);
void launch_far_cap(int ntiles, int nbins, int rstride, const int2 *rng, long long *cap, long long *loff, void *tmp,
                    size_t tmp_bytes, LineRec *lrec, int64_t mnull, hipStream_t st);
void launch_far_lists(int ntiles, int nwl, int nbins, int rstride, const int2 *rng, const LineRec *lrec,
                      const long long *loff, int *list, hipStream_t st);
void launch_zone_order(int64_t m, int64_t mnarrow, int64_t mmain, int ntiles, int nwl, int ntau, const int4 *meta_s,
                       const int2 *zone_s, const double *wlc_s, unsigned long long *key, int *val,
                       unsigned long long *key_sorted, int *perm, ZoneRec *zrec, void *tmp, size_t tmp_bytes,
                       int *zoff, int *zend, unsigned long long *evals_core, hipStream_t st);
// per tile of 64 core_pp points and zone bin, the candidate range of the core-zone kernel
void launch_core_ranges(int core_pp, int nwl, const int *zoff, const int *zend, const unsigned long long *key_sorted,
                        int2 *rng_core, hipStream_t st);
size_t zone_sort_tmp_bytes(int64_t m);
void launch_zero_f32(float *p, int64_t n, hipStream_t st);
void launch_add_clear_f32(float *dst, float *src, int64_t n, hipStream_t st);
// the lines of the narrow width bins (half-widths up to NARROW_HMAX points), tiles of
// NARROW_TILE points; rng [tile][nbins] candidate ranges from launch_tile_ranges
void launch_opacity32_narrow(const DevSetup &S, int nbins, const int2 *rng, const CoefB *cb, const double *wlc_s,
                             const int4 *meta_s, float *lineop, hipStream_t st);
void launch_opacity32_core(const DevSetup &S, int core_pp, const int2 *rng_core, const ZoneRec *zrec, const CoefA *ca,
                           const CoefB *cb, const float *dmaxw, float *lineop, hipStream_t st);
void launch_opacity64(const DevSetup &S, int ntiles, int nbins, int rstride, const int2 *rng, const Coef64 *coef,
                      const double *wlc_s, const int4 *meta_s, double *lineop, hipStream_t st);
void launch_linstr(const DevSetup &S, const DevLines &L, int list_type, const int *flag, float *linstr,
                   hipStream_t st);
void launch_radtran32(const DevSetup &S, const float *lop, double *spl, double *spc, double *sfl, double *sfc,
                      hipStream_t st);
void launch_radtran64(const DevSetup &S, const double *lop, double *spl, double *spc, double *sfl, double *sfc,
                      hipStream_t st);

// polarised path (fsg_magnetic.hip)
void launch_fvoigt_table(float2 *tab, const double *anode, double dv, hipStream_t st);
void launch_magpass1(const DevSetup &S, const DevLines &L, const MagLine &Z, double bsplit, int *flag, int *il,
                     int *ir, int *ic, unsigned *key, int *bin_count, int *bin_hmax, int batched, hipStream_t st);
size_t magcoef_size(int fp32);
void launch_magfill(const DevSetup &S, const DevLines &L, const MagLine &Z, int64_t nret, const int *order,
                    const int *flag, const int *il, const int *ir, const int *ic, void *coef, int fp32, double *wlc_s,
                    int4 *meta_s, int4 *comp_s, hipStream_t st);
void launch_magopacity(const DevSetup &S, const float2 *ftab, int ntiles, int nbins, int rstride, const int2 *rng,
                       const void *coef, const double *wlc_s, const int4 *meta_s, const int4 *comp_s,
                       const MagLine &Z, void *mlop, int fp32, hipStream_t st);
void launch_stokes(const DevSetup &S, const void *lop, const void *mlop, int fp32, const double *geo, double *spz,
                   double *sfl, hipStream_t st);

// reverse mode (fsg_adjoint.hip)
size_t rt_vjp_scratch_per_point(const DevSetup &S);
void launch_rt_vjp(const DevSetup &S, const void *lop, int fp32, const double *gsfl, const double *gsfc, int p_lo,
                   int np, double *scr, double *glop, double *gcont, double *gtemp, hipStream_t st);
// the reverse mode's tiles: 128 points, two per lane (64-point tiles leave fewer points to
// the general path but double the matrix-core kernel's per-tile work: 112 against 108.5 ms
// for the differentiable batch's forward + adjoint)
constexpr int VJP_PPL = 2;
constexpr int VJP_TILE = 64 * VJP_PPL;
// the reverse mode's records of a chunk; win [line] and dmaxv [line][ntp / DPW] (the largest
// far-wing threshold of each wavefront's depths) feed the test that splits the two kernels
void launch_fill64_vjp(const DevSetup &S, const DevLines &L, int list_type, int64_t nret, const int *order,
                       const int *flag, const int *il, const int *ir, const int *ic, Coef64 *coef, CoefV *coefv,
                       double *wlc_s, int4 *meta_s, int *slot_s, VjpWin *win, double *dmaxv, CoefR *coefr,
                       hipStream_t st);
void launch_lineop_vjp(const DevSetup &S, int ntiles, int nbins, int rstride, const int2 *rng, const Coef64 *coef,
                       const CoefV *coefv, const CoefR *coefr, const double *wlc_s, const int4 *meta_s,
                       const VjpWin *win, const double *dmaxv, const double *glop, double *sl, double *gtemp,
                       hipStream_t st);
void launch_lineop_vjp_scatter(const DevSetup &S, int64_t nret, const int *order, const int *atm, const int *slot_s,
                               const double *sl, double *g_loggf, double *g_fpf, hipStream_t st);

// convolution and disk integration (fsg_convol.hip)
struct ConvolSetup {
  int nwl, nmu, wmode;
  double wfirst, wlast, dstep, vsini, vmacro;
  const double *xmu, *wmu;  // host arrays [nmu]
};
struct ConvRow {  // one convolved spectrum: out[i] = sum_j in[clamp(i + j - h)] kern[koff + j]
  const double *in;
  double *out;
  int koff, nk, h;
};
void convol_gauss_kernel(int nw, double fwhm, int wmode, double wfirst, double wlast, double dstep,
                         std::vector<double> &kern, int &h);
void diskinteg_kernels(const ConvolSetup &cs, std::vector<double> &kern, std::vector<int> &rot_off,
                       std::vector<int> &rot_n, std::vector<int> &mac_off, int &mac_n);
// ContOpac [nwl][ntau] (the host's layout, in device memory) -> [ntau][nwl]
void launch_transpose_cont(int64_t nwl, int ntau, const double *src, double *dst, hipStream_t st);
void launch_conv_rows(int nwl, int nrows, int maxk, const ConvRow *d_rows, const double *d_kern, hipStream_t st);

// continuum opacity (fsg_contop.hip): the H- free-free table is CONT_HM_NW wavelengths x
// CONT_HM_NT temperatures, the H- bound-free cross-section CONT_HMBF_N points; tab holds
// ln(91.134 / lambda_k) [NW], ln(K_ff) [NW][NT], theta [NT], lambda_bf [N], sigma_bf [N], then
// ln(lambda) of the H2- free-free table [CONT_H2M_NW] (from CONT_TAB_H2M on); dep holds the
// rows below, each [ntau]: the depth state and populations, then ContopInit's factors of HOP,
// HE1OP and HE2OP (src/contop.f90:377-470; BoltnHe1 for n = 4..27) and of COOLOP's members
// (src/contop.f90:337-420: hc/kT in cm, kT, the level Boltzmann factors of C I, Mg I, Si I and
// Fe I, the pseudo-continuum exponentials of C I, Mg I and Si I, ln K of the H2- table at each
// depth, the NH table's temperature interval and weight and Fraction(NH) / Rho) and the
// populations COOLOP reads.  mask bit s-1 switches source s; bit 20 is Ifop(1) = -1 (HOP
// without levels 1-6).
constexpr int CONT_HM_NW = 22, CONT_HM_NT = 11, CONT_HMBF_N = 85, CONT_H2M_NW = 18;
constexpr int CONT_TAB_H2M = CONT_HM_NW + CONT_HM_NW * CONT_HM_NT + CONT_HM_NT + 2 * CONT_HMBF_N;
// then the Si II log-opacity table of SI2OPINIT [200 wavenumbers][51 temperatures]
constexpr int CONT_SI2_NW = 200, CONT_SI2_NT = 51, CONT_TAB_SI2 = CONT_TAB_H2M + CONT_H2M_NW;
constexpr int COOL_NC1 = 25, COOL_NMG1 = 15, COOL_NSI1 = 33, COOL_NFE1 = 48;
constexpr int LUKE_NC2 = 34, LUKE_NMG2 = 14, HOT_NEDGE = 60, HOT_NFPF = 21;
enum ContRow {
  CR_T, CR_HKT, CR_TKEV, CR_THETA, CR_XNE, CR_RHO, CR_FH1, CR_FH2, CR_FHMIN, CR_FHE1, CR_FHMOL, CR_TLOG,
  CR_BOLTH, CR_FREETH = CR_BOLTH + 8, CR_BOLTEXH, CR_EXLIMH,
  CR_BOLTHE1, CR_BOLTNHE1 = CR_BOLTHE1 + 10, CR_FREETHE1 = CR_BOLTNHE1 + 24, CR_BOLTEXHE1, CR_EXLIMHE1,
  CR_BOLTHE2, CR_FREETHE2 = CR_BOLTHE2 + 9, CR_BOLTEXHE2, CR_EXLIMHE2,
  CR_HCKT, CR_TK,
  CR_BOLTC1, CR_BOLTMG1 = CR_BOLTC1 + COOL_NC1, CR_BOLTSI1 = CR_BOLTMG1 + COOL_NMG1,
  CR_BOLTFE1 = CR_BOLTSI1 + COOL_NSI1, CR_PSEXP0 = CR_BOLTFE1 + COOL_NFE1, CR_PSEXP1 = CR_PSEXP0 + 6,
  CR_H2MLK = CR_PSEXP1 + 6, CR_NHN = CR_H2MLK + CONT_H2M_NW, CR_NHW, CR_NHNRHO,
  CR_FC1, CR_FMG1, CR_FAL1, CR_FSI1, CR_FFE1, CR_FCH, CR_FOH, CR_FRH1, CR_FRHE1,
  // LUKEOP: level Boltzmann factors of C II, Mg II, N I, Ca II, the Si II table's temperature
  // interval, weight and long-wavelength factor, the populations
  CR_BOLTC2, CR_BOLTMG2 = CR_BOLTC2 + LUKE_NC2, CR_BOLTN1 = CR_BOLTMG2 + LUKE_NMG2, CR_BOLTCA2 = CR_BOLTN1 + 2,
  CR_SI2IT = CR_BOLTCA2 + 2, CR_SI2TF, CR_SI2BN, CR_FC2, CR_FMG2, CR_FSI2, CR_FN1, CR_FO1, CR_FCA2,
  // HOTOP: exp(excitation / kT) of the 60 edges, the free-free sums of the ions of charge 1-5,
  // FractionPF of C I-IV, N I-V, O I-VI, Ne I-VI
  CR_HOTEXP, CR_HOTFF = CR_HOTEXP + HOT_NEDGE, CR_XNFP = CR_HOTFF + 5,
  CR_NROW = CR_XNFP + HOT_NFPF
};
constexpr unsigned CONT_HOP_NOLEV = 1u << 20;
void launch_contop(int nwl, int ntau, const double *wave, double wfirst, const double *dep, const double *tab,
                   const double *cstd, int motype, unsigned mask, int add, double *contT, hipStream_t st);
// COOLOP (Ifop 9) added into contT after launch_contop: dep and tab as above; xs holds
// coolop_scratch_per_point() doubles per wavelength (the cross-sections of C I, Mg I, Si I)
void launch_coolop(int nwl, int ntau, const double *wave, double wfirst, const double *dep, const double *tab,
                   const double *cstd, int motype, double *xs, double *contT, hipStream_t st);
constexpr int coolop_scratch_per_point() { return COOL_NC1 + COOL_NMG1 + COOL_NSI1; }
void contop_tables(std::vector<double> &tab);  // host: the table block of launch_contop
// host: COOLOP's rows of dep (T, Hkt and Rho rows filled first); pops: FractionPF of C I,
// Mg I, Al I, Si I, Fe I, CH, OH, then Fraction of NH, H I, He I, each [ntau]
void coolop_rows(int ntau, const float *const pops[10], double *dep);
// host: LUKEOP's rows (after coolop_rows' depth state); pops: FractionPF of C II, Mg II,
// Si II, N I, O I, Ca II, each [ntau]
void lukeop_rows(int ntau, const float *const pops[6], double *dep);
// host: HOTOP's rows; frac [ntau][8][6] Fraction of stages I-VI of C, N, O, Ne, Mg, Si, S,
// Fe; fpf [ntau][21] FractionPF of C I-IV, N I-V, O I-VI, Ne I-VI
void hotop_rows(int ntau, const float *frac, const float *fpf, double *dep);
// LUKEOP (Ifop 10) and HOTOP (Ifop 11) added into contT: mask bit 9 and 10 as launch_contop's;
// xs holds luke_scratch_per_point() doubles per wavelength
void launch_lukehot(int nwl, int ntau, const double *wave, double wfirst, const double *dep, const double *tab,
                    const double *cstd, int motype, unsigned mask, double *xs, double *contT, hipStream_t st);
constexpr int luke_scratch_per_point() { return LUKE_NC2 + LUKE_NMG2; }
void launch_mu_sum(int nwl, int nmu, const double *rl, const double *rc, const double *w, double *sfl, double *sfc,
                   hipStream_t st);

}  // namespace fsg
