/*
 * fsgpu.h -- C-ABI of the MI355X line-opacity and formal-solution library (libfsgpu.so).
 *
 * The library replaces the hot-path module procedures of FastSpec with HIP kernels for
 * gfx950.  Every entry point takes plain pointers and sizes; no C++ or torch types cross
 * the boundary.  The Fortran host binds these symbols with INTERFACE ... BIND(C) blocks
 * (fastspec_amd/fortran/fsgpu_host.f90), the Python side with ctypes
 * (fastspec_amd/_lib.py).
 *
 * Conventions
 *   - Return value 0 = success.  Non-zero = failure; fsg_last_error() then returns the
 *     message (the reference's ErrStop text where the reference has one).
 *   - Array layouts follow the Fortran host (column-major, depth fastest):
 *       ContOpac/LineOpac(Ntau,Nwl)  -> C array [nwl][ntau]
 *       FractionPF(Maxspl,Ntau)      -> C array [ntau][pf_ld]   (REAL(4))
 *       SpL/SpC(Nmu,Nwl)             -> C array [nwl][nmu]
 *       SpZ(Nmu,4,Nwl)               -> C array [nwl][4][nmu]
 *       MagLineOpac(6,Ntau,Nwl)      -> C array [nwl][ntau][6]
 *   - Species indices (LinSpc) and EOS slots (IdxSpc) are 1-based, as in the host.
 *   - Wavelengths of the grid are offsets from Wfirst (the reference's Wave array);
 *     line wavelengths are absolute air wavelengths in Angstrom (Lines%Wl).
 *   - One context per process and device; calls on one context are not re-entrant.
 *   - The host owns every host buffer; the library owns device memory.
 */
#ifndef FSGPU_H
#define FSGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* arithmetic of the line-opacity stage */
#define FSG_FP64 0   /* plain path: FP64 Voigt table interpolation and FP64 sums        */
#define FSG_FP32 1   /* batched path: FP32 Voigt in registers, Kahan-compensated sums  */

/* line-list types (LineListType in the reference, src/shared.f90:403) */
#define FSG_LIST_ATOMIC    0
#define FSG_LIST_MOLECULAR 1
#define FSG_LIST_MIXED     2   /* .z / Zeeman lists: per-line molecule flag */

/* Everything LineopCalc/RadTranCalc read from module globals (src/shared.f90:392-548,
 * src/synth.f90:747-1095, src/synth.f90:1825-1901).  Pointers are host memory and are
 * copied by fsg_init; they need not outlive the call. */
typedef struct fsg_setup {
  int32_t ntau, nwl, nmu;
  /* wavelength grid (src/synth.f90:56-74) */
  int32_t wmode;                 /* 0 linear, 1 logarithmic                          */
  double wfirst, dstep, rstep, wstart;
  const double *wave;            /* [nwl] offsets from wfirst                        */
  /* atmosphere (AtmosData) */
  int32_t motype;                /* 0 column mass, 1 tau scale                       */
  const double *rhox, *temp, *xne, *rho;   /* [ntau]                                 */
  const double *contop;          /* [nwl][ntau] ContOpac                             */
  const double *contop_std;      /* [ntau] ContOpacStd (Motype 1) or NULL            */
  /* EOS outputs (REAL(4) in the host) */
  int32_t pf_ld;                 /* leading dimension of fraction_pf (Maxspl)         */
  int32_t nslot;                 /* number of valid EOS slots (<= pf_ld)              */
  const float *fraction_pf;      /* [ntau][pf_ld] FractionPF                         */
  const float *pot_ion;          /* [nslot] PotIon                                   */
  const float *xamass;           /* [nslot] Xamass                                   */
  const float *frac_h1, *frac_he1, *frac_hmol;  /* [ntau] Fraction of H I, He I, H2  */
  /* global species table: LinSpc -> Ion, IdxSpc (1-based) */
  int32_t nspc;
  const int32_t *spc_ion;        /* [nspc] Ion(ispc)                                 */
  const int32_t *spc_slot;       /* [nspc] IdxSpc(ispc)                              */
  int32_t he1_slot;              /* IdxSpc(Ix%He1), 0 if absent                      */
  /* tabulated He I Stark lines (src/hestark.f90:43-60) */
  int32_t nhe;                   /* 0..255                                           */
  const double *he_wl, *he_elo;  /* [nhe] match table WlcHe, EloHe                   */
  double he_wl_tol, he_elo_tol;  /* HeMatchTol, HeEloTol                             */
  const double *he_gam, *he_shf; /* [nhe][ntau] HeGam(Ntau,NheTab), HeShf            */
  /* Config scalars (src/shared.f90:416-425) */
  double eps_opac, vwfudge, h2brd, vwmol, vmicro;
  int32_t molbrd;
  /* limb angles and flux weights (Xmu, Wmu from MuQuadrature) */
  const double *xmu, *wmu;       /* [nmu]                                            */
  /* library options */
  int32_t precision;             /* FSG_FP64 or FSG_FP32                             */
  int32_t batched;               /* 1: lines binned by window width (batched path);
                                    0: one bin, the reference's index-range scan     */
  int32_t device;                /* HIP device ordinal                               */
  /* spherical models (Radius, Atmos%Height: src/shared.f90:520,527); the ray geometry
   * of SphereInit (src/solver.f90:45-120) is built by fsg_init */
  double radius;                 /* Radius [cm]; <= 0 selects plane-parallel transfer */
  const double *height;          /* [ntau] layer heights [cm], decreasing inward;
                                    NULL when radius <= 0                           */
} fsg_setup;

/* One line list in structure-of-arrays form, after the host's reader conversions
 * (src/ioutils.f90:1055-1076): gf, Grad, Gqst, Gvdw are linear REAL(4) values. */
typedef struct fsg_lines {
  int64_t n;
  const double *wl;              /* [n] Lines%Wl, ascending                          */
  const float *elo, *gf, *grad;  /* [n]                                              */
  const float *gqst, *gvdw;      /* [n]; NULL for molecular lists                    */
  const int32_t *spc;            /* [n] LinSpc (1-based global species index)        */
  const uint8_t *is_mol;         /* [n] IsMolecule(El) for mixed lists, else NULL    */
  /* Zeeman patterns in CSR form for magnetic lists (src/shared.f90:473-481),
   * shifts in units of the reference (before the Bsplit*lambda^2 scaling) */
  const int32_t *pi_off;         /* [n+1] offsets into spl_pi/str_pi                 */
  const float *spl_pi, *str_pi;
  const int32_t *si_off;         /* [n+1] offsets into spl_si/str_si (blue then red) */
  const float *spl_si, *str_si;
} fsg_lines;

/* counters reported by the line-opacity stage */
typedef struct fsg_lineop_stats {
  int64_t nlines_in;             /* Nlines on entry                                  */
  int64_t nlines_use;            /* NlinesUse: retained lines                        */
  int64_t nwide;                 /* Nwide                                            */
  int64_t nblocks;               /* coefficient block fills: the reference's block
                                    refills; on the device the line chunks of at most
                                    nlines_lim coefficient records                   */
  int64_t nsegments;             /* wavelength segments (1 on the device)            */
  int64_t evals;                 /* sum over retained lines of window pixels x Ntau  */
  double ms_pass1, ms_opacity;   /* device time of the two stages                    */
  /* FP32 batched path only (zero otherwise): the evaluations inside the lines' core
   * zones (|lambda - lambda_c| < max_d dthr, Voigt table / continued fractions; the rest
   * are far-wing series on the matrix cores) and the device time of the far-wing and
   * core-zone kernels, from their first launch to their last (they run concurrently) */
  int64_t evals_core;
  double ms_far, ms_core;
  /* FP32 batched path: the launch geometry the autotuner chose (points per core-zone
   * tile, LDS ring slots of the far-wing kernel) */
  int32_t core_tile, far_ring;
} fsg_lineop_stats;

typedef struct fsg_ctx fsg_ctx;

/* Set-up for one synthesis: copies grid, atmosphere, EOS tables and ContOpac to the
 * device, builds the Voigt table and the per-depth kernel constants (the device side of
 * Init, src/synth.f90:22-311: VoigtInit src/voigt.f90:42, XlineInit src/opacity.f90:65).
 * LineOpac starts at zero (src/synth.f90:91). */
int fsg_init(const fsg_setup *setup, fsg_ctx **ctx_out);

/* Releases all device memory of the context. */
int fsg_finalize(fsg_ctx *ctx);

/* Message of the last failed call on this thread ("" if none). */
const char *fsg_last_error(void);

/* Non-magnetic line opacity of one list, accumulated into the device-resident LineOpac:
 * replaces LineopCalc(ifile, NlinesLim) (src/synth.f90:747-1095).
 *   nlines_lim   block size NlinesMax(LimitRAM) computed by the host (src/opacity.f90:713)
 *   linstr_out   [n] central line depths (LinStr, src/synth.f90:1774) or NULL
 *   flags/ileft/iright_out  [n] LinFlag, Ileft, Iright of the input lines (1-based
 *                indices, before PACK) or NULL
 *   stats        or NULL */
int fsg_lineop(fsg_ctx *ctx, int32_t list_type, const fsg_lines *lines, int64_t nlines_lim,
               float *linstr_out, int32_t *flags_out, int32_t *ileft_out,
               int32_t *iright_out, fsg_lineop_stats *stats);

/* Magnetic line opacity of one Zeeman list into the device-resident MagLineOpac:
 * replaces MagLineopCalc(ifile, NlinesLim) (src/synth.f90:1098-1455).
 *   bfield[3], bframe  BFIELD / BFRAME (only |B| enters here, src/synth.f90:1150-1154) */
int fsg_maglineop(fsg_ctx *ctx, const fsg_lines *lines, int64_t nlines_lim,
                  const double *bfield, int32_t bframe, float *linstr_out,
                  int32_t *flags_out, int32_t *ileft_out, int32_t *iright_out,
                  fsg_lineop_stats *stats);

/* Copies LineOpac (Ntau,Nwl) between host and device, e.g. for SAVE flag 2
 * (src/ioutils.f90:1759-1771) and the OPACITY keyword (src/fastspec.f90:215). */
int fsg_get_lineop(fsg_ctx *ctx, double *lineop_out);
int fsg_set_lineop(fsg_ctx *ctx, const double *lineop_in);
int fsg_get_maglineop(fsg_ctx *ctx, double *maglineop_out);

/* Scalar plane-parallel formal solution: replaces RadTranCalc (src/synth.f90:1825-1901,
 * LocPrfPP src/solver.f90:123-179).  spl_out/spc_out [nwl][nmu] SpL/SpC, or NULL to
 * keep the intensities on the device for fsg_flux. */
int fsg_radtran(fsg_ctx *ctx, double *spl_out, double *spc_out);

/* Polarised plane-parallel formal solution: replaces MagRadTranCalc
 * (src/synth.f90:1904-2111, StokesPrfPP src/solver.f90:265-394).
 *   bfield [3] or [1+2*nmu] for bframe 2;  spz_out [nwl][4][nmu], spc_out [nwl][nmu] */
int fsg_magradtran(fsg_ctx *ctx, const double *bfield, int32_t nbfield, int32_t bframe,
                   double *spz_out, double *spc_out);

/* Disk integration of the last formal solution, Sf = sum_mu Wmu I_mu: the flux sums of
 * DiskInteg without convolution (src/convol.f90:67-246, lines 233-234).  Uses Stokes I
 * after fsg_magradtran (src/fastspec.f90:437-440).  sfl_out/sfc_out [nwl]. */
int fsg_flux(fsg_ctx *ctx, double *sfl_out, double *sfc_out);

/* Convolution and disk integration of the last formal solution: replaces IntegConvol
 * (src/convol.f90:27-64) with DiskInteg (67-246) and ConvolGau (249-300).
 *   cimode 1: Gaussian broadening (FWHM vgau, km/s) of every intensity spectrum on the
 *             device: SpL and SpC, or the four Stokes spectra and SpC after fsg_magradtran;
 *   cimode 2: disk integration of SpL (Stokes I after fsg_magradtran) and SpC with the
 *             rotational kernel of each annulus (vsini), radial-tangential macroturbulence
 *             (vmacro) and Gaussian broadening of the flux (vgau); SfL and SfC on the device
 *             are replaced (read them with fsg_flux).
 * The annuli come from the context's mu and flux weights (the host's Xmu and Wmu of
 * MuQuadrature for its DiskQuad); vgau is the host's Vgau after its resolving-power
 * conversion (src/fastspec.f90:103). */
int fsg_integ_convol(fsg_ctx *ctx, int32_t cimode, double vsini, double vmacro, double vgau);

/* fsg_integ_convol without allocation: the same stages, kernels and results, with every work
 * area (annulus spectra, kernel taps, row table, weights) carved from one static device
 * array of 64 MiB per GPU instead of buffers the context grows with hipMalloc.  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid allocation).  A call
 * that needs more than 64 MiB (cimode 2: 32 * nmu * nwl bytes plus the taps, about 299,000
 * points at 7 mu) fails with a message and changes nothing.  All calls on one GPU share the
 * array: calls that overlap in time on one GPU race. */
int fsg_integ_convol_noalloc(fsg_ctx *ctx, int32_t cimode, double vsini, double vmacro, double vgau);

/* The intensity spectra of the last formal solution as they stand on the device (after
 * fsg_integ_convol with cimode 1, the broadened ones), in the host's layouts for WriteOut
 * (src/ioutils.f90:1751-1875): spl_out [nwl][nmu] SpL, or spz_out [nwl][4][nmu] SpZ after
 * fsg_magradtran; spc_out [nwl][nmu] SpC.  NULL pointers are skipped. */
int fsg_get_spectra(fsg_ctx *ctx, double *spl_out, double *spz_out, double *spc_out);

/* Device-resident variants for callers that keep data in HBM (the PyTorch wrapper and
 * the benchmark).  A device line list is uploaded once and reused. */
typedef struct fsg_dlines fsg_dlines;
int fsg_lines_upload(fsg_ctx *ctx, int32_t list_type, const fsg_lines *lines,
                     fsg_dlines **out);
int fsg_lines_free(fsg_dlines *dl);
int fsg_lineop_dev(fsg_ctx *ctx, fsg_dlines *dl, int64_t nlines_lim, fsg_lineop_stats *stats);
/* MagLineopCalc (src/synth.f90:1098-1455) on a device list uploaded with its Zeeman
 * patterns (lines->pi_off/si_off set); the list type is taken as FSG_LIST_MIXED. */
int fsg_maglineop_dev(fsg_ctx *ctx, fsg_dlines *dl, int64_t nlines_lim, const double *bfield,
                      int32_t bframe, fsg_lineop_stats *stats);
/* Replaces the gf column (linear REAL(4) values, 10**log gf) of a device list, from host
 * memory or (on_device != 0) from device memory; windows and flags follow on the next
 * fsg_lineop_dev.  No counterpart in the reference, whose lists are read once: it serves
 * gradient-based fitting, where gf changes between syntheses. */
int fsg_lines_set_gf(fsg_dlines *dl, const float *gf, int32_t on_device);

/* Line-list ingest.  dir names either a FastSpec HDF5 list file (recognised by its
 * signature; its /Lines datasets are read through libhdf5, loaded at run time from
 * FSG_HDF5_LIB or the usual names, the gzip+shuffle chunks inflated by the reader's threads)
 * or a sidecar directory: the /Lines group as raw little-endian column files <COLUMN>.bin
 * (W f8, EL S<k>, ION i2, E, LGGF, GRAD, GQST, GVDW f4; W, E, LGGF, GRAD for molecular
 * lists) and a header.txt of key=value lines (format=fastspec-lines-1, name, wl_vac, wmin,
 * wmax, nl, hflag, columns=NAME:type,...).  Both give the same lines bit for bit.
 * Replaces ReadHDF (src/ioutils.f90:545-1100) for atomic and molecular lists: whole-file
 * cull on WMIN/WMAX, Vac2Air when WL_VAC = 1, the sort check on the full W column, the
 * LowerBoundReal8/UpperBoundReal8 selection of [wstart, wend], the ion filter (ION <= 5)
 * and the REAL(4) conversions gf = 10**LGGF, GRAD 10** for 0 < x < 20, GQST and GVDW 10**
 * when negative.  Species (EL, ION) map to the host's species slots through a table; a
 * molecular list takes mol_slot.  A list with JLO, JUP, GLO, GUP columns gets its Zeeman
 * patterns from ZeemanPattern (src/utils.f90:901-1004; a Lande factor above 90 marks a
 * line without splitting) and is read only by fsg_lines_stream.  A line with an ISOT mass
 * number above 0 has its gf scaled by the REAL(4) isotopic fraction of the host's isotope
 * table (src/ioutils.f90:940-962). */
typedef struct fsg_species_entry {
  char el[8];       /* element or molecule symbol, blank- or NUL-padded        */
  int32_t ion;      /* ionisation stage, 1 = neutral                           */
  int32_t slot;     /* 1-based species slot (LinSpc)                           */
  int32_t is_mol;   /* IsMolecule(El): the list becomes a mixed list           */
} fsg_species_entry;
typedef struct fsg_isotope_entry {
  char el[8];       /* element symbol                                          */
  int32_t a;        /* mass number                                             */
  double frac;      /* isotopic fraction (IsotFrac)                            */
} fsg_isotope_entry;
/* the selection only: 0-based first line and line count inside [wstart, wend];
 * list_type receives FSG_LIST_MOLECULAR or FSG_LIST_ATOMIC.  Host only. */
int fsg_lines_scan(const char *dir, double wstart, double wend, int32_t *list_type, int64_t *first,
                   int64_t *count);
/* lines [first, first + count) converted into caller arrays of count entries (gqst, gvdw,
 * is_mol may be NULL for molecular lists), read by nthreads host threads; kept receives
 * the lines written after the ion filter.  Host only. */
int fsg_lines_read(const char *dir, int64_t first, int64_t count, const fsg_species_entry *species,
                   int32_t nspecies, const fsg_isotope_entry *isotopes, int32_t nisotopes, int32_t mol_slot,
                   int32_t nthreads, double *wl, float *elo, float *gf, float *grad, float *gqst, float *gvdw,
                   int32_t *spc, uint8_t *is_mol, int64_t *kept);
/* scan + read straight into a new device list: chunks of up to 2^22 lines (environment
 * FSG_INGEST_CHUNK overrides the count) are converted by
 * nthreads host threads into one of two pinned buffers while the other one's columns copy
 * to HBM on the library stream.  No counterpart in the reference (whose list stays on the
 * host); the device list is the one fsg_lines_upload would make from the same lines. */
int fsg_lines_stream(fsg_ctx *ctx, const char *dir, double wstart, double wend, const fsg_species_entry *species,
                     int32_t nspecies, const fsg_isotope_entry *isotopes, int32_t nisotopes, int32_t mol_slot,
                     int32_t nthreads, fsg_dlines **out, int64_t *kept);

/* Reverse mode (no counterpart in the reference, which has no derivatives).  For a loss
 * L(SfL, SfC) of the flux sums of the last plane-parallel fsg_radtran:
 *   fsg_flux_vjp    from dL/dSfL, dL/dSfC [nwl] computes and keeps on the device
 *                   dL/dLineOpac; returns, where non-NULL, dL/dLineOpac and dL/dContOpac
 *                   [nwl][ntau] and the source-function part of dL/dT [ntau].
 *   fsg_lineop_vjp  for one list of the synthesis (its windows and flags held at their
 *                   forward values): dL/d(log10 gf) [n] of its lines, and adds the line
 *                   opacity's parts of dL/d(ln FractionPF) [ntau][pf_ld] and dL/dT [ntau].
 * Derivatives with respect to T are taken at fixed ContOpac, FractionPF, Xne and
 * perturber densities (the host's EOS and continuum chain them further). */
int fsg_flux_vjp(fsg_ctx *ctx, const double *gsfl, const double *gsfc, double *glop_out,
                 double *gcont_out, double *gtemp_out);
int fsg_lineop_vjp(fsg_ctx *ctx, fsg_dlines *dl, int64_t nlines_lim, double *gloggf_out,
                   double *gfpf_inout, double *gtemp_inout);

/* Device-pointer variants for callers whose tensors live in HBM (the PyTorch wrapper: with
 * torch loaded first the library shares its HIP runtime and device).  They replace the same
 * seams as the host variants (the flux sums of DiskInteg, src/convol.f90:233-234, and the
 * reverse mode above); the pointer arguments named *_dev are device memory of the
 * context's device, written or read on the library stream, which is synchronised before
 * return (producers of *_dev inputs must have finished).
 *   fsg_flux_dev        SfL, SfC [nwl] to device memory (either may be NULL)
 *   fsg_flux_vjp_dev    dL/dSfL, dL/dSfC [nwl] from device memory; gcont_dev: dL/dContOpac
 *                       [ntau][nwl] (lambda fastest) to device memory, or NULL; gtemp_out host
 *                       [ntau] (batch: [nb][ntau]) or NULL
 *   fsg_lineop_vjp_dev  dL/d(log10 gf) [n] to device memory; gfpf_inout, gtemp_inout host */
int fsg_flux_dev(fsg_ctx *ctx, double *sfl_dev, double *sfc_dev);
int fsg_flux_vjp_dev(fsg_ctx *ctx, const double *gsfl_dev, const double *gsfc_dev, double *gcont_dev,
                     double *gtemp_out);
int fsg_lineop_vjp_dev(fsg_ctx *ctx, fsg_dlines *dl, int64_t nlines_lim, double *gloggf_dev,
                       double *gfpf_inout, double *gtemp_inout);

/* A new atmosphere state on an existing context: the results of the host's EOS and
 * continuum (EosSolve, src/synth.f90:334-406; ContopCalc, src/synth.f90:642-674) for a new
 * temperature or abundance vector, replacing in place what fsg_init uploaded.  No
 * counterpart in the reference, which synthesises one atmosphere per run; it serves fits
 * that move T and abundances (the host maps them through its EOS and continuum and chains
 * their Jacobians onto the reverse mode's dL/dT, dL/d ln FractionPF and dL/dContOpac).
 * Arrays as in fsg_setup; NULL members keep the context's values.  rhox only for
 * plane-parallel models (the rays of a spherical one are built from it); contop_std only for
 * Motype 1 models; a context with He I Stark tables needs he_gam with a new temperature.
 * atm: the atmosphere of a batch context (0 for an ordinary context).  Windows, flags and
 * formal solutions of earlier calls are dropped: line opacity must be computed again. */
typedef struct fsg_atmosphere {
  const double *temp, *xne, *rho;        /* [ntau] */
  const double *rhox;                    /* [ntau] column mass (plane-parallel models) */
  const double *contop;                  /* [nwl][ntau] ContOpac(Ntau, Nwl) */
  const double *contop_std;              /* [ntau] */
  const float *fraction_pf;              /* [ntau][pf_ld] */
  const float *frac_h1, *frac_he1, *frac_hmol;  /* [ntau] */
  const double *he_gam;                  /* [nhe][ntau] */
} fsg_atmosphere;
int fsg_set_atmosphere(fsg_ctx *ctx, int32_t atm, const fsg_atmosphere *a);

/* Continuum opacity on the device: ContopCalc's loop over the grid (src/synth.f90:663-669,
 * Xcont at Air2Vac of each wavelength, divided by ContOpacStd for Motype 1) for the sources
 * of Xcont (src/contop.f90:267-314) that need only hydrogen, helium and electron
 * populations, COOLOP's atoms and molecules, LUKEOP's and HOTOP's ions: all thirteen.
 * ifop_mask bit s-1 switches source s of the reference's Ifop numbering: 1 HOP, 2 H2PLOP,
 * 3 HMINOP, 4 HRAYOP, 5 HE1OP, 6 HE2OP, 7 HEMIOP, 8 HERAOP, 9 COOLOP, 10 LUKEOP, 11 HOTOP,
 * 12 ELECOP, 13 H2RAOP.  Bit FSG_CONTOP_HOP_NOLEV (with bit 0) is the reference's Ifop(1) = -1, set
 * by ContopSetup when a hydrogen line list supplies the bound-free edges of levels 1-6
 * (src/synth.f90:600): HOP then omits those levels.  T, Xne and Rho are the context's
 * (fsg_init / fsg_set_atmosphere); the populations are the EOS rows the sources read, as
 * ContopSetup indexes them (src/synth.f90:608-628); ContopInit's depth factors
 * (src/contop.f90:319-492) are formed from them on the host.  add = 0 replaces the
 * context's ContOpac with the device sources' sum; add = 1 adds the sum to it (the host's
 * part, from Xcont with those Ifop flags off, given at fsg_init or fsg_set_atmosphere).
 * contop_out [nwl][ntau], if non-NULL, receives the resulting ContOpac.  Windows and formal
 * solutions of earlier calls are dropped.  Ordinary contexts only. */
#define FSG_CONTOP_HOP_NOLEV (1 << 20)
typedef struct fsg_contop_eos {
  const float *fpf_h1;     /* [ntau] FractionPF(IxH1, :)   H I   (HOP, H2PLOP, HMINOP, HRAYOP) */
  const float *fpf_h2;     /* [ntau] FractionPF(IxH2, :)   H II  (HOP, H2PLOP) */
  const float *frac_hmin;  /* [ntau] Fraction(IxHmin, :)   H-    (HMINOP) */
  const float *fpf_he1;    /* [ntau] FractionPF(IxHe1, :)  He I  (HE1OP, HEMIOP, HERAOP) */
  const float *frac_hmol;  /* [ntau] Fraction(IxHmol, :)   H2    (H2RAOP) */
  const float *frac_he2;   /* [ntau] Fraction(IxHe2, :)    He II (HE1OP free-free) */
  const float *fpf_he2;    /* [ntau] FractionPF(IxHe2, :)  He II (HE2OP) */
  const float *fpf_he3;    /* [ntau] FractionPF(IxHe3, :)  He III (HE2OP free-free) */
  /* COOLOP (src/contop.f90:1544-1569) */
  const float *fpf_c1;     /* [ntau] FractionPF(IxC1, :)   C I   (C1OP) */
  const float *fpf_mg1;    /* [ntau] FractionPF(IxMg1, :)  Mg I  (MG1OP) */
  const float *fpf_al1;    /* [ntau] FractionPF(IxAl1, :)  Al I  (AL1OP) */
  const float *fpf_si1;    /* [ntau] FractionPF(IxSi1, :)  Si I  (SI1OP) */
  const float *fpf_fe1;    /* [ntau] FractionPF(IxFe1, :)  Fe I  (FE1OP) */
  const float *fpf_ch;     /* [ntau] FractionPF(IxCH, :)   CH    (CHOP) */
  const float *fpf_oh;     /* [ntau] FractionPF(IxOH, :)   OH    (OHOP) */
  const float *frac_nh;    /* [ntau] Fraction(IxNH, :)     NH    (NHOP) */
  const float *frac_h1;    /* [ntau] Fraction(IxH1, :)     H I   (H2COLLOP) */
  const float *frac_he1;   /* [ntau] Fraction(IxHe1, :)    He I  (H2COLLOP; with frac_hmol, H2COLLOP and H2MINOP) */
  /* LUKEOP (src/contop.f90:2422-2440) */
  const float *fpf_c2;     /* [ntau] FractionPF(IxC2, :)   C II  (C2OP) */
  const float *fpf_mg2;    /* [ntau] FractionPF(IxMg2, :)  Mg II (MG2OP) */
  const float *fpf_si2;    /* [ntau] FractionPF(IxSi2, :)  Si II (SI2OP) */
  const float *fpf_n1;     /* [ntau] FractionPF(IxN1, :)   N I   (N1OP) */
  const float *fpf_o1;     /* [ntau] FractionPF(IxO1, :)   O I   (O1OP) */
  const float *fpf_ca2;    /* [ntau] FractionPF(IxCa2, :)  Ca II (CA2OP) */
  /* HOTOP (src/contop.f90:2729-2763): the rows ContopInit takes through IxHot
   * (src/contop.f90:471-490) */
  const float *hot_frac;   /* [ntau][8][6] Fraction of stages I-VI of C, N, O, Ne, Mg, Si, S, Fe */
  const float *hot_fpf;    /* [ntau][21] FractionPF of C I-IV, N I-V, O I-VI, Ne I-VI */
} fsg_contop_eos;
int fsg_contop(fsg_ctx *ctx, const fsg_contop_eos *eos, int32_t ifop_mask, int32_t add, double *contop_out);

/* Batch contexts: many model atmospheres synthesised in one launch sequence (the
 * differentiable batch of many small syntheses).  No counterpart in the reference, which
 * synthesises one atmosphere per run (src/fastspec.f90:241-461); per atmosphere the stages
 * and results are those of an ordinary context.  The atmospheres' grids are concatenated,
 * each starting at a multiple of 256 points (fsg_batch_offsets); on a batch context every
 * array that the entry points above describe as [nwl] or [ntau][nwl] spans the concatenated
 * grid (points between atmospheres lie in no window and are to be ignored).  All set-ups
 * share precision, Ntau, the mu quadrature, the species and EOS slot tables and the device;
 * each is a Motype 0, plane-parallel model without He I Stark tables.  A batch context
 * takes fsg_lineop_dev on batch lists, fsg_radtran, fsg_flux, fsg_get_spectra,
 * fsg_get_lineop, fsg_set_lineop, fsg_reset_lineop, fsg_flux_vjp (gtemp_out [nb][ntau])
 * and fsg_lineop_vjp (gfpf_inout [nb][ntau][pf_ld], gtemp_inout [nb][ntau]); the other
 * entry points refuse it. */
int fsg_init_batch(const fsg_setup *setups, int32_t nb, fsg_ctx **ctx_out);
/* off_out[nb + 1]: the first point (0-based) of each atmosphere in the concatenated grid,
 * then the grid's point count. */
int fsg_batch_offsets(const fsg_ctx *ctx, int64_t *off_out);
/* A batch list for fsg_lineop_dev: line k is line index[k] of the device list src (a list
 * of the same device; not a Zeeman list) for atmosphere atm[k].  Atmosphere indices must
 * not decrease, and wavelengths must not decrease within an atmosphere (each atmosphere's
 * lines are a list as ReadHDF returns it). */
int fsg_lines_batch(fsg_ctx *batch, const fsg_dlines *src, const int64_t *index, const int32_t *atm, int64_t n,
                    fsg_dlines **out);

/* Device pointers to the context's resident arrays (valid until fsg_finalize):
 *   which = 0 LineOpac [ntau][nwl] (lambda fastest; FP32 or FP64 by precision)
 *           1 SpL [nmu][nwl] FP64   2 SpC [nmu][nwl] FP64
 *           3 SfL [nwl] FP64        4 SfC [nwl] FP64 (Stokes I flux after fsg_magradtran)
 *           5 MagLineOpac [6][ntau][nwl] (FP32 or FP64 by precision)
 *           6 SpZ [4][nmu][nwl] FP64 */
int fsg_device_ptr(fsg_ctx *ctx, int32_t which, void **ptr_out);
/* Stream the library launches on (hipStream_t as void*). */
int fsg_stream(fsg_ctx *ctx, void **stream_out);
/* Zeroes the device-resident LineOpac (and MagLineOpac). */
int fsg_reset_lineop(fsg_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* FSGPU_H */
