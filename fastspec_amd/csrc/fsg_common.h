// fsg_common.h -- constants and device helpers shared by the HIP kernels of libfsgpu.
//
// Physical constants follow the reference (src/shared.f90:19-33) and the line-opacity
// kernel constants src/opacity.f90:34-38.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsg {

constexpr double kVc = 299792.458;
constexpr double kHc = 6.62607015e-27;
constexpr double kKbolt = 1.380649e-16;
constexpr double kA0 = 5.29177210903e-11;
constexpr double kM0 = 1.66053906660e-27;
constexpr double kCe = 4.803204712e-10;
constexpr double kMe = 9.1093837139e-28;
constexpr double kMh = 1.673532499e-24;
constexpr double kHie = 13.598434005;
constexpr double kKz = 4.668644778e-13;
constexpr double kPi = 3.14159265358979323846;  // acos(-1d0)
constexpr double kErg2ev = 1e-7 / 1.602176634e-19;
constexpr double kEv2Ang = 8065.544005e-8;
constexpr double kDWstart = 0.2, kDWstep = 2.0, kDwide = 5e3;  // src/shared.f90:430

constexpr double kXlnC = kVc * 1e13;                 // speed of light [A/s]
constexpr double kXlnC4Pi = 4.0 * kXlnC * kPi;
constexpr double kXlnKM0 = kKbolt * 1e-7 / kM0;
constexpr double kXlnDop = 2.0 * kKbolt / (kM0 * 1e3) * 1e16;
constexpr double kSqrtPiInv = 0.56418958354775628695;
constexpr double kPlkC1 = kHc * kVc * 1e13 / kKbolt;

// Voigt table geometry (src/voigt.f90:19-26): H(a,v) on v in [0,8] step 8/1000 and
// sqrt-spaced a nodes a_j = ((j-1) sqrt(8)/400)^2, j = 1..402
constexpr int NVV = 1001, NVA = 401;
constexpr int TAB_LD = NVV + 1;  // 1002 entries per a column

// Depths per wavefront in the opacity kernels; a block holds ceil(ntau/DPW) wavefronts and
// per-line coefficient records are padded to ntp = DPW * ceil(ntau/DPW) depths.
constexpr int DPW = 9;
constexpr int PPL = 2;          // wavelength points per lane
constexpr int TILE = 64 * PPL;  // wavelength points per tile
constexpr int MAX_BINS = 40;
constexpr int WEAK_KEY = 63;    // sort key of rejected lines
constexpr int HE_KEY = 62;      // sort key of retained lines with a tabulated He I Stark shift

// Per-(line, depth) coefficients of the batched FP32 path, two records kept in separate
// arrays (CoefA for every evaluation, CoefB only on the general path).
//   Far-wing expansion (used when every lane of the wavefront has v^2 >= 400 (a^2 + 1/2)):
//     with u = 1/v^2 the reference's one-pole fraction a (s + 1/2)/(s^2 - v^2 + a^2 + 1/4)/sqrt(pi)
//     equals K u (1 + b u)/(1 + al u + b^2 u^2), K = LinCop a/sqrt(pi), al = 2a^2 - 1,
//     b = a^2 + 1/2, expanded to K u (f0 + f1 u + f2 u^2), f0 = 1 (the next term is below
//     1.1e-7 relative there).  With t = 1/dl^2 and w = 1/vv^2 (u = w t) the profile is
//     t (k0 + t (k1 + t k2)), k_n = K f_n w^(n+1); dthr = the |dl| above which it applies.
//   General path: vv = 1/Doppler width, a, a^2, a^2 + 1/4, K, LinCop, ua (a-interpolation
//   weight), ia (a-cell of the table, 0 = outside the table box).
// per-line record of the FP32 matrix-core path: window, core zone and centre (32 bytes)
struct alignas(16) LineRec {
  int il, ir, zl, zr;
  double wc;
  int pad0, pad1;
};

// per-line record of the core-zone kernel, in core-zone order: zone ∩ window [lo, hi], the
// line's position j in the chunk and its centre wavelength
struct alignas(32) ZoneRec {
  int lo, hi, j, pad0;
  double wc, pad1;
};

struct alignas(16) CoefA {
  float k0, k1, k2, dthr;
};
struct alignas(32) CoefB {
  float vv, a, a2, c1, ka, cop, ua;
  int ia;
};
using Coef32 = CoefB;  // the general path's view
// FP64 plain path: the reference's quantities LinVvgt, LinAvgt, LinCop and the a-state
struct alignas(64) Coef64 {
  double vv, a, cop, ua;
  int ia, pad0;
  double pad1, pad2, pad3;
};
// Reverse mode, far wings: where the one-pole fraction's series holds (|dlambda| >= dthr),
// a (line, depth) contributes c = sum_n k_n T^n and dc/dT = sum_n m_n T^n (n = 1..3) with
// T = 1/dlambda^2 shared by all depths of a point (u = 1/v^2 = T / LinVvgt^2 folded into the
// coefficients).
struct alignas(64) CoefV {
  double k1, k2, k3, m1, m2, m3, dthr, pad;
};

// Per-(line, depth) record of the reverse mode's general path, in the precision it is
// evaluated in (FP32), with the far-wing threshold dthr (FP64, the value of CoefV) that decides
// which kernel takes the depth.  48 bytes: one wavefront loads nine of them with three
// 16-byte loads, one record per lane.
struct alignas(16) CoefR {
  double dthr, vv;
  float a, cop, ua, aida, p1, p2, p3;  // p1..p3: d ln LinCop/dT, d ln LinVvgt/dT, d LinAvgt/dT
  int ia;
};
static_assert(sizeof(CoefR) == 48, "three 16-byte loads per record");

// Per-line record of the reverse mode's far-wing test: the wavelengths of the first and last
// points of the line's window, as the grid holds them.
struct alignas(16) VjpWin {
  double wlo, whi;
};

// Everything the kernels read from the set-up (device pointers + scalars).
struct DevSetup {
  int ntau, ntp, nwl, nmu, wmode, motype, pf_ld, nslot, nspc, he1_slot, nhe, molbrd;
  double wfirst, dstep, rstep, wstart;
  double eps_opac, vwfudge, h2brd, vwmol, vmicro, he_wl_tol, he_elo_tol;
  const double *wave;        // [nwl]
  const double *contT;       // [ntau][nwl] ContOpac, wavelength fastest
  const double *cstd;        // [ntau] ContOpacStd (ones if Motype 0)
  const double *temp, *rhox, *xne, *rho;
  const float *fpf;          // [ntau][pf_ld]
  const float *pot_ion, *xamass;
  const int *spc_ion, *spc_slot;
  const double *he_wl, *he_elo, *he_gam;  // [nhe], [nhe][ntau]
  const double *he_shf;      // [nhe][ntp]: HeShf padded with zeros to ntp depths
  const double *xln;         // [6][ntau] XlineInit depth constants
  double vturb2;
  double xlnz;               // sqrt(pi) e^2/(m c^2) 1e-18 (src/opacity.f90:36)
  const double *xmu, *wmu;
  // Voigt table and its node arrays
  const float *vtab;         // [NVA+1][TAB_LD]
  const double *anode, *aida;
  double idv, idt, a1, ia1;
  // spherical models: the rays of SphereInit, per mu up to 2 ntau points
  int sph;
  const int *ntau_mu, *graz_mu;  // [nmu] points on the ray, ray misses the deepest layer
  const int *map_mu;             // [nmu][2 ntau] model layer (0-based) of each ray point
  const double *rhox_mu;         // [nmu][2 ntau] column mass along the ray
  // batch contexts (several model atmospheres in one launch sequence): the set-up of the
  // concatenated grid carries the atmospheres' own set-ups (batch[nbatch]) and the
  // atmosphere of each point of the concatenated grid (pmap[nwl]); an atmosphere's own
  // set-up has goff = its first point in the concatenated grid (0 outside batches)
  int goff, nbatch;
  const DevSetup *batch;
  const int *pmap;
};

// Each atmosphere's points start at a multiple of SEG_ALIGN in the concatenated grid of a
// batch, so that no tile of the opacity kernels (TILE, 64 or 128 points) and no wavefront
// of the per-point kernels spans two atmospheres.
constexpr int SEG_ALIGN = 256;

struct DevLines {
  int64_t n;
  const double *wl;
  const float *elo, *gf, *grad, *gqst, *gvdw;
  const int *spc;
  const uint8_t *is_mol;
  const int *atm;  // batch lists: the atmosphere (0-based) of each line, else nullptr
};

// Zeeman patterns of a device-resident list (src/ioutils.f90:1012-1031): per line i the
// pi components [pi_off[i], pi_off[i+1]) and the sigma components [si_off[i], si_off[i+1]),
// blue half first.  spl_* are shifts per unit Bsplit lambda^2; sp_* receive the scaled
// shifts of the current field (REAL(4), src/synth.f90:1193-1197).
struct MagLine {
  const int *pi_off, *si_off;
  const float *spl_pi, *str_pi, *spl_si, *str_si;
  float *sp_pi, *sp_si;
};

// the polarised opacity kernel: one wavelength point per lane, DPW depths per wavefront
constexpr int MTILE = 64;
// narrow windows: width bins whose largest half-width is at most NARROW_HMAX points go to the
// narrow-line kernel (tiles of NARROW_TILE points x all depths, up to 128 depths)
constexpr int NARROW_HMAX = 16;
constexpr int NARROW_TILE = 8;
constexpr int NZB = 32;  // core-zone width bins of the FP32 path (the last one holds lines without a core zone)

}  // namespace fsg
