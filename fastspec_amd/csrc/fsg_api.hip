// fsg_api.hip -- the C-ABI of libfsgpu (include/fsgpu.h): device context, uploads, the
// orchestration of the line-opacity stage (pass 1, width binning, coefficient fill,
// tiled accumulation) and of the formal solution.
//
// The stage order follows LineopCalc (src/synth.f90:747-1095) and RadTranCalc
// (src/synth.f90:1825-1901); IntegConvol's flux sums (src/convol.f90:233-234) are fused
// into the formal-solution kernel.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <atomic>
#include <mutex>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fsgpu.h"
#include "fsg_ingest.h"
#include "fsg_contop_tables.h"
#include "fsg_kernels.h"

using namespace fsg;

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
  g_err = msg;
  return 1;
}

#define CK(call)                                                                              \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_));       \
  } while (0)

// CK for set-up code that owns a partly built context: releases it before returning
#define CKF(ctx, call)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      fail(std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_));              \
      fsg_finalize(ctx);                                                                      \
      return 1;                                                                               \
    }                                                                                         \
  } while (0)

// grow-only device buffer
struct DBuf {
  void *p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= n) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) n = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  template <typename T>
  T *as() const {
    return (T *)p;
  }
};

}  // namespace

// Work space of fsg_integ_convol_noalloc, the entry point that allocates nothing: one
// array per GPU, held from the library's load until the process ends.  In host code the
// array's name is only a stand-in; its device address on each GPU comes from
// hipGetSymbolAddress (fsg_init, once per device, with that device current) and is the only
// address passed to kernels and copies.  Calls on one GPU share it and race if they overlap.
constexpr size_t kStaticScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// records the device address of `scratch` on the current device dev (once per device)
int static_scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return fail("fsg_init: device ordinal beyond the static work-space table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(std::string("hipGetSymbolAddress of the static work space: ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = p;
  return 0;
}

int static_scratch(int dev, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_ptr[dev]) return fail("static work space not set up on this device");
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

// pass 1 of a line list: bin populations and maximal half-widths, evaluation count
struct LinePrep {
  int nbins = 0;
  int64_t nmain = 0, nhe = 0, nret = 0;
  int hcnt[64], hhmax[64];
  unsigned long long hu[4];
};

// identities of device line lists (never reused within a process)
static std::atomic<uint64_t> g_list_id{1};

struct fsg_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  int precision = FSG_FP64, batched = 0;
  DevSetup S{};
  std::vector<void *> owned;
  void *lineop = nullptr;   // [ntau][nwl] float or double
  double *spl = nullptr, *spc = nullptr, *sfl = nullptr, *sfc = nullptr;
  int have_rt = 0;
  DBuf flag, il, ir, ic, key, key2, idx, idx2, coef, wlc, meta, rng, rng2, tmp, pos, wlcl, rflag;
  DBuf zkey, zkey2, zval, zperm, zrec;  // FP32 path: lines ordered by core-zone width
  DBuf rngn;                             // FP32 path: candidate ranges of the narrow-line tiles
  DBuf fcap, foff, ftmp, flist;          // FP32 path: per-tile line lists of the matrix-core kernel
  // convolution and disk integration: host copies of mu, weights, the grid's last point;
  // device scratch for the convolved rows, kernels and row tables
  std::vector<double> hxmu, hwmu;
  double wlast = 0.0;
  DBuf cv_a, cv_b, cv_kern, cv_rows, cv_w;
  int *bincnt = nullptr, *binhmax = nullptr, *binoff = nullptr;
  unsigned long long *u64 = nullptr;
  hipEvent_t ev[4] = {};
  // FP32 path: the core-zone kernel runs on a second stream beside the far-wing kernel and
  // accumulates into lop2 (zero between calls), added to lineop after the last chunk
  hipStream_t st2 = nullptr;
  hipEvent_t evf = nullptr, evj = nullptr;
  hipEvent_t evt[4] = {};  // first and last far-wing launch (on st), core-zone launch (on st2)
  float *lop2 = nullptr;
  bool lop2_dirty = false;  // a core launch has not been followed by its add (an error path)
  // launch geometry of the FP32 opacity kernels: core-zone tile of 64 core_pp points (fixed:
  // it sets how lines group into the compensated sums, so it never depends on timing;
  // FSG_CORE_TILE=128 selects 128-point tiles), far_ring LDS slots of the far-wing kernel
  // (chosen by the autotuner on the first chunk of the first FP32 line-opacity call; it does
  // not change any result)
  int core_pp = 1, far_ring = 4;
  bool tuned = false;
  // polarised path, allocated on first use
  void *maglineop = nullptr;  // [6][ntau][nwl] MagLineOpac, wavelength fastest
  float2 *ftab = nullptr;     // Faraday-Voigt table [NVA+1][TAB_LD]
  double *spz = nullptr;      // [4][nmu][nwl] Stokes I, Q, U, V
  double *geo = nullptr;      // [nmu][5] field geometry
  int magnetic = 0;           // the last formal solution was the polarised one
  DBuf comp;
  // spherical rays on the host (field geometry of the polarised solution)
  std::vector<int> ntau_mu;
  std::vector<double> muloc;  // [nmu][2 ntau]
  // reverse mode: dL/dLineOpac [ntau][nwl] of the last fsg_flux_vjp, and work buffers
  double *glop = nullptr;
  int have_glop = 0;
  DBuf vscr, vgcont, vsl, vslot, vout;
  // batch contexts: one context per atmosphere (its own set-up on the device) and the first
  // point of each atmosphere in the concatenated grid (nb + 1 entries, the last the total)
  std::vector<fsg_ctx *> kids;
  std::vector<int64_t> boff;
  // the list (identity, gf generation) whose pass-1 results (flags, windows, bin order) the
  // pass-1 buffers hold, for the reverse mode of the same synthesis; id 0: none
  uint64_t p1_id = 0, p1_gen = 0;
  LinePrep p1;
  // per-depth H I, He I and H2 number fractions behind XlineInit's constants (kept for
  // fsg_set_atmosphere)
  const float *fh1 = nullptr, *fhe1 = nullptr, *fhmol = nullptr;
  DBuf ctab;  // fsg_contop: per-depth state and the H- tables
  DBuf cxs;   // fsg_contop: COOLOP's per-wavelength cross-sections of C I, Mg I and Si I
};

struct fsg_dlines {
  uint64_t id = g_list_id++;
  uint64_t gen = 0;  // bumped when the gf column changes
  fsg_ctx *ctx = nullptr;
  int list_type = 0;
  DevLines L{};
  MagLine Z{};      // Zeeman patterns (magnetic lists only)
  int zeeman = 0;
  std::vector<void *> owned;
};

namespace {

// ------------------------------------------------------------------------------------
// Autotuner of the FP32 opacity launch geometry: the matrix-core kernel's LDS ring depth (2 or
// 4 group slots), which changes no result.  The core-zone tile is not tuned: it sets the
// grouping of the compensated sums, and a choice made from wall-clock times would make the
// FP32 bits depend on the run.  Each ring depth runs twice on the chunk at hand as the
// accumulation stage runs it -- both kernels at once on the two streams, into scratch
// accumulators -- and the shortest wall time wins.  Choices are kept per process for a
// workload signature (device, depth count, grid size and chunk size in powers of two, core
// tile) and, with FSG_AUTOTUNE_CACHE set to a file name, in that file ("signature core_pp
// far_ring" lines).  FSG_AUTOTUNE=0 keeps the default (4 ring slots) without timing.
// ------------------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::string, std::pair<int, int>> g_tune;
bool g_tune_file_read = false;

int ilog2_64(int64_t v) {
  int b = 0;
  while (v > 1) {
    v >>= 1;
    b++;
  }
  return b;
}

std::string tune_signature(const fsg_ctx *c, int64_t m) {
  hipDeviceProp_t prop{};
  (void)hipGetDeviceProperties(&prop, c->dev);
  std::ostringstream k;
  k << prop.gcnArchName << "/cu" << prop.multiProcessorCount << "/ntau" << c->S.ntau << "/nwl2^" << ilog2_64(c->S.nwl)
    << "/lines2^" << ilog2_64(m) << "/cp" << c->core_pp;
  std::string s = k.str();
  for (char &ch : s)
    if (ch == ' ' || ch == ':') ch = '_';
  return s;
}

void tune_read_file() {  // under g_tune_mu
  if (g_tune_file_read) return;
  g_tune_file_read = true;
  const char *fn = std::getenv("FSG_AUTOTUNE_CACHE");
  if (!fn || !*fn) return;
  std::ifstream in(fn);
  std::string key;
  int pp = 0, ring = 0;
  while (in >> key >> pp >> ring)
    if ((pp == 1 || pp == 2) && (ring == 2 || ring == 4)) g_tune[key] = {pp, ring};
}

void tune_write_file(const std::string &key, int pp, int ring) {  // under g_tune_mu
  const char *fn = std::getenv("FSG_AUTOTUNE_CACHE");
  if (!fn || !*fn) return;
  std::ofstream out(fn, std::ios::app);
  out << key << " " << pp << " " << ring << "\n";
}

}  // namespace

// the chunk's coefficient records, line lists and zone order are complete on c->st
static int autotune(fsg_ctx *c, int ntiles, int64_t m, const CoefA *ca, const CoefB *cb, const float *dmaxw,
                    const LineRec *lrec, const int *zoff, const int *zend) {
  const std::string key = tune_signature(c, m);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_read_file();
    auto it = g_tune.find(key);
    if (it != g_tune.end() && it->second.first == c->core_pp) {
      c->far_ring = it->second.second;
      return 0;
    }
  }
  const char *en = std::getenv("FSG_AUTOTUNE");
  if (en && en[0] == '0') return 0;
  const DevSetup &S = c->S;
  hipStream_t s = c->st;
  // every candidate pair runs as in the accumulation stage: the far-wing kernel on the
  // library stream and the core-zone kernel on the second stream at once, each into its own
  // scratch accumulator (lop2, and tune_scr released afterwards); the pair's wall time counts
  DBuf tune_scr;
  const size_t scr_bytes = sizeof(float) * (size_t)S.ntau * S.nwl;
  if (tune_scr.ensure(scr_bytes) != hipSuccess) {  // no room to time the pair: keep the defaults
    (void)hipGetLastError();
    return 0;
  }
  CK(hipMemsetAsync(tune_scr.p, 0, scr_bytes, s));
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  CK(hipEventCreateWithFlags(&e2, hipEventDisableTiming));
  const int rings[2] = {2, 4}, pps[1] = {c->core_pp};
  float best = 1e30f;
  int bring = c->far_ring, bpp = c->core_pp;
  int rc = 0;
  for (int ip = 0; ip < 1 && !rc; ip++) {
    launch_core_ranges(pps[ip], S.nwl, zoff, zend, c->zkey2.as<unsigned long long>(), c->rng2.as<int2>(), s);
    for (int ir = 0; ir < 2 && !rc; ir++) {
      for (int rep = 0; rep < 2 && !rc; rep++) {
        float ms = 0.f;
        if (hipEventRecord(e0, s) != hipSuccess || hipStreamWaitEvent(c->st2, e0, 0) != hipSuccess) {
          rc = fail("autotune: event record failed");
          break;
        }
        launch_opacity32(S, ntiles, 0, 0, nullptr, ca, cb, dmaxw, lrec, nullptr, nullptr, c->foff.as<long long>(),
                         c->flist.as<int>(), m, c->lop2, 0, rings[ir], s);
        launch_opacity32_core(S, pps[ip], c->rng2.as<int2>(), c->zrec.as<ZoneRec>(), ca, cb, dmaxw,
                              tune_scr.as<float>(), c->st2);
        if (hipEventRecord(e2, c->st2) != hipSuccess || hipStreamWaitEvent(s, e2, 0) != hipSuccess ||
            hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = fail("autotune: timing failed");
          break;
        }
        if (ms < best) {
          best = ms;
          bring = rings[ir];
          bpp = pps[ip];
        }
      }
    }
  }
  launch_zero_f32(c->lop2, (int64_t)S.ntau * S.nwl, s);
  (void)hipStreamSynchronize(s);
  tune_scr.release();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipEventDestroy(e2);
  if (rc) return rc;
  CK(hipGetLastError());
  c->far_ring = bring;
  c->core_pp = bpp;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = {c->core_pp, c->far_ring};
  tune_write_file(key, c->core_pp, c->far_ring);
  return 0;
}

namespace {

template <typename T>
int upload(fsg_ctx *c, const T *h, size_t n, const T **dst, std::vector<void *> &owned) {
  if (!h || n == 0) {
    *dst = nullptr;
    return 0;
  }
  void *p = nullptr;
  CK(hipMalloc(&p, n * sizeof(T)));
  owned.push_back(p);
  CK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, c->st));
  *dst = (const T *)p;
  return 0;
}

int alloc_zero(fsg_ctx *c, size_t bytes, void **out) {
  void *p = nullptr;
  CK(hipMalloc(&p, bytes));
  c->owned.push_back(p);
  CK(hipMemsetAsync(p, 0, bytes, c->st));
  *out = p;
  return 0;
}

size_t lop_elem(const fsg_ctx *c) { return c->precision == FSG_FP32 ? sizeof(float) : sizeof(double); }

// The rays of SphereInit (src/solver.f90:45-120).  A ray with impact parameter
// p = R_top sqrt(1 - mu^2) crosses the layers r_k = Radius + Height_k; its column-mass steps
// are the radial ones scaled by (r_k + r_k-1) / (z_k + z_k-1), z = sqrt((r - p)(r + p)).
// A ray with p above the deepest layer (grazing) is tangent below layer n: the tangent
// shell adds chord * dm/dr, and the far half revisits layers n..1 with mirrored steps.
// Outputs per mu (row stride 2 ntau): point count, grazing flag, 0-based layer of each
// point, column mass and the local ray-to-radial cosine (sign flipped beyond the tangent).
struct SphereRays {
  std::vector<int> ntau_mu, graz, map;
  std::vector<double> rhox, muloc;
};

int sphere_rays(int N, int M, double radius, const double *height, const double *rhox, const double *xmu,
                SphereRays &R) {
  const int L = 2 * N;
  std::vector<double> Rk(N);
  for (int k = 0; k < N; k++) Rk[k] = radius + height[k];
  R.ntau_mu.assign(M, 0);
  R.graz.assign(M, 0);
  R.map.assign((size_t)M * L, 0);
  R.rhox.assign((size_t)M * L, 0.0);
  R.muloc.assign((size_t)M * L, 0.0);
  for (int im = 0; im < M; im++) {
    const double mu = xmu[im];
    if (mu <= 0.0) return fail("SphereInit: mu = 0 rays are unsupported for spherical models");
    const double P = Rk[0] * std::sqrt((1.0 - mu) * (1.0 + mu));
    const bool graz = P > Rk[N - 1];
    int n = N;
    if (graz) {
      n = 1;
      while (n < N && P < Rk[n]) n++;  // n = number of layers above the tangent point
      if (n < 2)
        return fail("SphereInit: grazing ray tangent in the top model layer; remove extreme mu values close to 0");
    }
    int *mp = R.map.data() + (size_t)im * L;
    double *rm = R.rhox.data() + (size_t)im * L, *ml = R.muloc.data() + (size_t)im * L;
    rm[0] = rhox[0] / mu;
    double z1 = std::sqrt((Rk[0] - P) * (Rk[0] + P));
    ml[0] = z1 / Rk[0];
    mp[0] = 0;
    for (int k = 1; k < n; k++) {
      const double z = std::sqrt((Rk[k] - P) * (Rk[k] + P));
      rm[k] = rm[k - 1] + (rhox[k] - rhox[k - 1]) * (Rk[k] + Rk[k - 1]) / (z + z1);
      ml[k] = z / Rk[k];
      mp[k] = k;
      z1 = z;
    }
    R.graz[im] = graz ? 1 : 0;
    if (!graz) {
      R.ntau_mu[im] = N;
      continue;
    }
    const double chord = 2.0 * std::sqrt((Rk[n - 1] - P) * (Rk[n - 1] + P));
    rm[n] = rm[n - 1] + chord * (rhox[n] - rhox[n - 1]) / (Rk[n - 1] - Rk[n]);
    mp[n] = n - 1;
    ml[n] = -ml[n - 1];
    for (int k = n + 1; k < 2 * n; k++) {  // point k mirrors point 2n - 1 - k
      rm[k] = rm[k - 1] + (rm[2 * n - k] - rm[2 * n - k - 1]);
      mp[k] = 2 * n - 1 - k;
      ml[k] = -ml[2 * n - 1 - k];
    }
    R.ntau_mu[im] = 2 * n;
  }
  return 0;
}

// Lines per accumulation chunk: NlinesLim, lowered so that the chunk's per-line device
// buffers (per_line bytes each) fit in 85 % of the free device memory plus what the
// context's reusable buffers already hold.
int64_t chunk_lines(fsg_ctx *c, int64_t nret, int64_t nlim, size_t per_line, int magnetic) {
  int64_t m = std::min(nret, nlim);
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
    size_t avail = fr + c->coef.n + c->wlc.n + c->meta.n + (magnetic ? c->comp.n : 0);
    int64_t fit = (int64_t)((double)avail * 0.85 / (double)per_line);
    m = std::min(m, fit);
  }
  return std::max<int64_t>(m, 1);
}

// group tables of one chunk [s0, s1) of the bin-sorted retained lines: per group the offset
// of its overlap with the chunk (chunk-relative), the overlap's size and the group's
// maximal window half-width
void chunk_tables(const std::vector<int64_t> &goff, const std::vector<int64_t> &gcnt, const std::vector<int> &ghmax,
                  int64_t s0, int64_t s1, std::vector<int> &out) {
  const size_t ng = goff.size();
  out.assign(3 * ng, 0);
  for (size_t b = 0; b < ng; b++) {
    const int64_t lo = std::max(goff[b], s0), hi = std::min(goff[b] + gcnt[b], s1);
    out[b] = hi > lo ? (int)(lo - s0) : 0;
    out[ng + b] = hi > lo ? (int)(hi - lo) : 0;
    out[2 * ng + b] = ghmax[b];
  }
}

}  // namespace

extern "C" {

const char *fsg_last_error(void) { return g_err.c_str(); }

int fsg_init(const fsg_setup *s, fsg_ctx **ctx_out) {
  g_err.clear();
  if (!s || !ctx_out) return fail("fsg_init: NULL argument");
  *ctx_out = nullptr;
  if (s->ntau < 3 || s->ntau > DPW * 16) return fail("fsg_init: Ntau must be in 3..144");
  if (s->nwl < 2) return fail("Init: the wavelength grid has fewer than two points");
  if (s->nmu < 1 || s->nmu > 16) return fail("fsg_init: Nmu must be in 1..16");
  if (s->precision != FSG_FP64 && s->precision != FSG_FP32) return fail("fsg_init: unknown precision");
  if (!s->wave || !s->rhox || !s->temp || !s->xne || !s->rho || !s->contop || !s->fraction_pf ||
      !s->pot_ion || !s->xamass || !s->frac_h1 || !s->frac_he1 || !s->frac_hmol || !s->spc_ion ||
      !s->spc_slot || !s->xmu || !s->wmu)
    return fail("fsg_init: required array is NULL");
  if (s->motype == 1 && !s->contop_std) return fail("fsg_init: Motype 1 needs ContOpacStd");
  if (s->nslot < 1 || s->nslot > s->pf_ld) return fail("fsg_init: EOS slot count exceeds FractionPF leading dimension");
  for (int i = 0; i < s->nspc; i++)
    if (s->spc_slot[i] < 1 || s->spc_slot[i] > s->nslot || s->spc_ion[i] < 1)
      return fail("fsg_init: species table refers to an EOS slot out of range");
  if (s->wmode != 0 && s->wmode != 1) return fail("fsg_init: unknown wavelength grid mode");
  if (s->nhe < 0 || s->nhe > 255 || (s->nhe > 0 && (!s->he_wl || !s->he_elo || !s->he_gam || !s->he_shf)))
    return fail("fsg_init: bad He I Stark tables");
  int ndev = 0;
  CK(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return fail("fsg_init: no HIP device");
  if (s->device < 0 || s->device >= ndev) return fail("fsg_init: device ordinal out of range");

  fsg_ctx *c = new fsg_ctx();
  c->dev = s->device;
  c->precision = s->precision;
  c->batched = s->batched ? 1 : 0;
  {
    const char *ct = std::getenv("FSG_CORE_TILE");
    if (ct && std::atoi(ct) == 128) c->core_pp = 2;
  }
  CKF(c, hipSetDevice(c->dev));
  if (static_scratch_init(c->dev)) {
    fsg_finalize(c);
    return 1;
  }
  CKF(c, hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
  for (auto &e : c->ev) CKF(c, hipEventCreate(&e));
  if (c->precision == FSG_FP32) {
    CKF(c, hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
    CKF(c, hipEventCreateWithFlags(&c->evf, hipEventDisableTiming));
    CKF(c, hipEventCreateWithFlags(&c->evj, hipEventDisableTiming));
    for (auto &e : c->evt) CKF(c, hipEventCreate(&e));
  }

  DevSetup &S = c->S;
  const int N = s->ntau, Nw = s->nwl;
  S.ntau = N;
  S.nwl = Nw;
  S.nmu = s->nmu;
  S.wmode = s->wmode;
  S.motype = s->motype;
  S.pf_ld = s->pf_ld;
  S.nslot = s->nslot;
  S.nspc = s->nspc;
  S.he1_slot = s->he1_slot;
  S.nhe = s->nhe;
  S.molbrd = s->molbrd;
  S.wfirst = s->wfirst;
  S.dstep = s->dstep;
  S.rstep = s->rstep;
  S.wstart = s->wstart;
  S.eps_opac = s->eps_opac;
  S.vwfudge = s->vwfudge;
  S.h2brd = s->h2brd;
  S.vwmol = s->vwmol;
  S.vmicro = s->vmicro;
  S.he_wl_tol = s->he_wl_tol;
  S.he_elo_tol = s->he_elo_tol;
  int rc = 0;
  rc |= upload(c, s->wave, Nw, &S.wave, c->owned);
  rc |= upload(c, s->temp, N, &S.temp, c->owned);
  rc |= upload(c, s->rhox, N, &S.rhox, c->owned);
  rc |= upload(c, s->xne, N, &S.xne, c->owned);
  rc |= upload(c, s->rho, N, &S.rho, c->owned);
  rc |= upload(c, s->fraction_pf, (size_t)N * s->pf_ld, &S.fpf, c->owned);
  rc |= upload(c, s->pot_ion, s->nslot, &S.pot_ion, c->owned);
  rc |= upload(c, s->xamass, s->nslot, &S.xamass, c->owned);
  rc |= upload(c, s->spc_ion, s->nspc, &S.spc_ion, c->owned);
  rc |= upload(c, s->spc_slot, s->nspc, &S.spc_slot, c->owned);
  rc |= upload(c, s->xmu, s->nmu, &S.xmu, c->owned);
  rc |= upload(c, s->wmu, s->nmu, &S.wmu, c->owned);
  c->hxmu.assign(s->xmu, s->xmu + s->nmu);
  c->hwmu.assign(s->wmu, s->wmu + s->nmu);
  c->wlast = s->wfirst + s->wave[Nw - 1];
  S.ntp = DPW * ((N + DPW - 1) / DPW);
  std::vector<double> shf((size_t)std::max(s->nhe, 1) * S.ntp, 0.0);  // zero-padded to ntp depths
  if (s->nhe > 0) {
    rc |= upload(c, s->he_wl, s->nhe, &S.he_wl, c->owned);
    rc |= upload(c, s->he_elo, s->nhe, &S.he_elo, c->owned);
    rc |= upload(c, s->he_gam, (size_t)s->nhe * N, &S.he_gam, c->owned);
    for (int e = 0; e < s->nhe; e++)
      for (int k = 0; k < N; k++) shf[(size_t)e * S.ntp + k] = s->he_shf[(size_t)e * N + k];
  }
  rc |= upload(c, shf.data(), shf.size(), &S.he_shf, c->owned);
  std::vector<double> cstd(N, 1.0);
  if (s->motype == 1) std::memcpy(cstd.data(), s->contop_std, sizeof(double) * N);
  rc |= upload(c, cstd.data(), N, &S.cstd, c->owned);
  // ContOpac(Ntau, Nwl) -> [ntau][nwl]: wavelength fastest for coalesced per-lambda lanes
  // (copied as the host holds it, then transposed on the device)
  if (!rc) {
    void *ct = nullptr;
    DBuf raw;
    if (alloc_zero(c, sizeof(double) * (size_t)N * Nw, &ct) || raw.ensure(sizeof(double) * (size_t)N * Nw) != hipSuccess) {
      fsg_finalize(c);
      return fail("fsg_init: no device memory for ContOpac");
    }
    CKF(c, hipMemcpyAsync(raw.p, s->contop, sizeof(double) * (size_t)N * Nw, hipMemcpyHostToDevice, c->st));
    launch_transpose_cont(Nw, N, raw.as<double>(), (double *)ct, c->st);
    S.contT = (const double *)ct;
    CKF(c, hipStreamSynchronize(c->st));
    raw.release();
  }
  // spherical models: the rays of SphereInit
  if (s->radius > 0.0 && s->nmu > 0) {
    if (!s->height) {
      fsg_finalize(c);
      return fail("fsg_init: a spherical model (radius > 0) needs the layer heights");
    }
    SphereRays R;
    if (sphere_rays(N, s->nmu, s->radius, s->height, s->rhox, s->xmu, R)) {
      std::string e = g_err;
      fsg_finalize(c);
      return fail(e);
    }
    rc |= upload(c, R.ntau_mu.data(), R.ntau_mu.size(), &S.ntau_mu, c->owned);
    rc |= upload(c, R.graz.data(), R.graz.size(), &S.graz_mu, c->owned);
    rc |= upload(c, R.map.data(), R.map.size(), &S.map_mu, c->owned);
    rc |= upload(c, R.rhox.data(), R.rhox.size(), &S.rhox_mu, c->owned);
    S.sph = 1;
    c->ntau_mu = R.ntau_mu;
    c->muloc = R.muloc;
  }
  // per-depth perturber densities and XlineInit
  rc |= upload(c, s->frac_h1, N, &c->fh1, c->owned);
  rc |= upload(c, s->frac_he1, N, &c->fhe1, c->owned);
  rc |= upload(c, s->frac_hmol, N, &c->fhmol, c->owned);
  if (rc) {
    fsg_finalize(c);
    return 1;
  }
  void *xln = nullptr;
  if (alloc_zero(c, sizeof(double) * 6 * N, &xln)) {
    fsg_finalize(c);
    return 1;
  }
  launch_xline_init(N, S.temp, S.xne, c->fh1, c->fhe1, c->fhmol, s->h2brd, (double *)xln, c->st);
  S.xln = (const double *)xln;
  S.vturb2 = 1e26 / kXlnC / kXlnC * (s->vmicro * s->vmicro);
  S.xlnz = std::sqrt(kPi) * (kCe * kCe) / (kMe * (kVc * kVc)) * 1e-18;
  // Voigt table nodes (src/voigt.f90:48-62) and the table itself on the device
  {
    double dv = 8.0 / (double)(NVV - 1), dt = std::sqrt(8.0) / (double)(NVA - 1);
    std::vector<double> anode(NVA + 1), aida(NVA + 1);
    for (int ia = 0; ia <= NVA; ia++) anode[ia] = ((double)ia * dt) * ((double)ia * dt);
    for (int ia = 0; ia < NVA; ia++) aida[ia] = 1.0 / (anode[ia + 1] - anode[ia]);
    aida[NVA] = 0.0;
    S.idv = 1.0 / dv;
    S.idt = 1.0 / dt;
    S.a1 = anode[1];
    S.ia1 = aida[0];
    if (upload(c, anode.data(), anode.size(), &S.anode, c->owned) ||
        upload(c, aida.data(), aida.size(), &S.aida, c->owned)) {
      fsg_finalize(c);
      return 1;
    }
    void *tab = nullptr;
    if (alloc_zero(c, sizeof(float) * (NVA + 1) * TAB_LD, &tab)) {
      fsg_finalize(c);
      return 1;
    }
    launch_voigt_table((float *)tab, S.anode, dv, c->st);
    S.vtab = (const float *)tab;
    CKF(c, hipStreamSynchronize(c->st));
  }
  if (c->precision == FSG_FP32 && alloc_zero(c, sizeof(float) * (size_t)N * Nw, (void **)&c->lop2)) {
    fsg_finalize(c);
    return 1;
  }
  if (alloc_zero(c, lop_elem(c) * (size_t)N * Nw, &c->lineop) ||
      alloc_zero(c, sizeof(int) * 64, (void **)&c->bincnt) ||
      alloc_zero(c, sizeof(int) * 64, (void **)&c->binhmax) ||
      alloc_zero(c, sizeof(int) * (3 * (MAX_BINS + 1) + 2 * NZB), (void **)&c->binoff) ||
      alloc_zero(c, sizeof(unsigned long long) * 4, (void **)&c->u64)) {
    fsg_finalize(c);
    return 1;
  }
  CKF(c, hipGetLastError());
  CKF(c, hipStreamSynchronize(c->st));
  *ctx_out = c;
  return 0;
}

int fsg_finalize(fsg_ctx *c) {
  if (!c) return 0;
  for (fsg_ctx *k : c->kids) fsg_finalize(k);
  (void)hipSetDevice(c->dev);
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (void *p : c->owned) (void)hipFree(p);
  for (DBuf *b : {&c->flag, &c->il, &c->ir, &c->ic, &c->key, &c->key2, &c->idx, &c->idx2, &c->coef, &c->wlc,
                  &c->meta, &c->rng, &c->rng2, &c->tmp, &c->pos, &c->wlcl, &c->rflag, &c->comp, &c->vscr,
                  &c->vgcont, &c->vsl, &c->vslot, &c->vout, &c->zkey, &c->zkey2, &c->zval, &c->zperm, &c->zrec, &c->rngn, &c->fcap, &c->foff, &c->ftmp, &c->flist, &c->ctab, &c->cxs, &c->cv_a,
                  &c->cv_b, &c->cv_kern, &c->cv_rows, &c->cv_w})
    b->release();
  if (c->st2) (void)hipStreamSynchronize(c->st2);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->evt)
    if (e) (void)hipEventDestroy(e);
  if (c->evf) (void)hipEventDestroy(c->evf);
  if (c->evj) (void)hipEventDestroy(c->evj);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return 0;
}

// Batch contexts.  Each atmosphere gets an ordinary context (its set-up, XlineInit constants,
// ContOpac and grid on the device); the batch context itself is an ordinary context over the
// concatenated grid (atmosphere b at points [boff[b], boff[b] + nwl_b), each start a multiple
// of SEG_ALIGN; padding points continue the grid's last step, repeat its last ContOpac
// column and never lie inside a window), whose set-up carries the atmospheres' set-ups and
// the atmosphere of every point.
int fsg_init_batch(const fsg_setup *setups, int32_t nb, fsg_ctx **ctx_out) {
  g_err.clear();
  if (!setups || !ctx_out) return fail("fsg_init_batch: NULL argument");
  *ctx_out = nullptr;
  if (nb < 1) return fail("fsg_init_batch: no atmospheres");
  const fsg_setup &s0 = setups[0];
  if (s0.ntau < 3 || s0.nmu < 1 || s0.nmu > 16 || !s0.xmu || !s0.wmu || s0.nspc < 0 || s0.nslot < 1)
    return fail("fsg_init_batch: bad first set-up");
  const int N = s0.ntau;
  auto same = [](const void *a, const void *b, size_t bytes) {
    return bytes == 0 || (a && b && std::memcmp(a, b, bytes) == 0);
  };
  std::vector<int64_t> off(nb + 1, 0);
  for (int b = 0; b < nb; b++) {
    const fsg_setup &s = setups[b];
    if (s.precision != s0.precision || s.ntau != N || s.nmu != s0.nmu || s.device != s0.device ||
        s.batched != s0.batched || s.nspc != s0.nspc || s.nslot != s0.nslot || s.pf_ld != s0.pf_ld)
      return fail("fsg_init_batch: the set-ups differ in precision, depth count, mu count, device or species tables");
    if (!same(s.xmu, s0.xmu, sizeof(double) * s.nmu) || !same(s.wmu, s0.wmu, sizeof(double) * s.nmu))
      return fail("fsg_init_batch: the set-ups differ in their mu quadrature");
    if (!same(s.spc_ion, s0.spc_ion, sizeof(int32_t) * s.nspc) ||
        !same(s.spc_slot, s0.spc_slot, sizeof(int32_t) * s.nspc) ||
        !same(s.pot_ion, s0.pot_ion, sizeof(float) * s.nslot) || !same(s.xamass, s0.xamass, sizeof(float) * s.nslot))
      return fail("fsg_init_batch: the set-ups differ in their species tables");
    if (s.motype != 0) return fail("fsg_init_batch: batches need Motype 0 models");
    if (s.radius > 0.0) return fail("fsg_init_batch: batches need plane-parallel models");
    if (s.nhe != 0) return fail("fsg_init_batch: batches take no He I Stark tables");
    if (s.nwl < 2 || !s.wave || !s.contop) return fail("fsg_init_batch: a set-up has no grid or ContOpac");
    off[b + 1] = off[b] + ((int64_t)s.nwl + SEG_ALIGN - 1) / SEG_ALIGN * SEG_ALIGN;
  }
  const int64_t total = off[nb];
  if (total > (int64_t)1 << 30) return fail("fsg_init_batch: the concatenated grid is too large");
  std::vector<fsg_ctx *> kids;
  auto drop = [&kids]() {
    for (fsg_ctx *k : kids) fsg_finalize(k);
  };
  for (int b = 0; b < nb; b++) {
    fsg_ctx *k = nullptr;
    if (fsg_init(&setups[b], &k)) {
      std::string e = g_err;
      drop();
      return fail("fsg_init_batch: atmosphere " + std::to_string(b) + ": " + e);
    }
    kids.push_back(k);
  }
  // the concatenated grid (each atmosphere's points relative to its own Wfirst) and ContOpac
  std::vector<double> wave((size_t)total), cont((size_t)total * N);
  std::vector<int> pmap((size_t)total);
  for (int b = 0; b < nb; b++) {
    const fsg_setup &s = setups[b];
    const int64_t o = off[b], nw = s.nwl;
    std::memcpy(wave.data() + o, s.wave, sizeof(double) * nw);
    std::memcpy(cont.data() + o * N, s.contop, sizeof(double) * nw * N);
    const double step = s.wave[nw - 1] - s.wave[nw - 2];
    for (int64_t p = nw; p < off[b + 1] - o; p++) {
      wave[o + p] = s.wave[nw - 1] + (double)(p - nw + 1) * step;
      std::memcpy(cont.data() + (o + p) * N, s.contop + (nw - 1) * N, sizeof(double) * N);
    }
    std::fill(pmap.begin() + o, pmap.begin() + off[b + 1], b);
  }
  fsg_setup sup = s0;
  sup.nwl = (int32_t)total;
  sup.wave = wave.data();
  sup.contop = cont.data();
  fsg_ctx *c = nullptr;
  if (fsg_init(&sup, &c)) {
    std::string e = g_err;
    drop();
    return fail("fsg_init_batch: " + e);
  }
  c->kids = kids;
  c->boff = off;
  std::vector<DevSetup> tab(nb);
  for (int b = 0; b < nb; b++) {
    tab[b] = kids[b]->S;
    tab[b].goff = (int)off[b];
  }
  int rc = upload(c, tab.data(), tab.size(), &c->S.batch, c->owned);
  rc |= upload(c, pmap.data(), pmap.size(), &c->S.pmap, c->owned);
  if (rc || hipStreamSynchronize(c->st) != hipSuccess) {
    std::string e = g_err;
    fsg_finalize(c);
    return fail("fsg_init_batch: " + e);
  }
  c->S.nbatch = nb;
  *ctx_out = c;
  return 0;
}

}  // extern "C"

namespace {

// The depth-dependent inputs of an ordinary context replaced on its stream from host arrays
// in the layouts of fsg_setup (NULL members keep their values).
int set_atmosphere_one(fsg_ctx *c, const fsg_atmosphere *a) {
  const DevSetup &S = c->S;
  const size_t N = (size_t)S.ntau, Nw = (size_t)S.nwl;
  hipStream_t s = c->st;
  auto put = [s](const void *src, const void *dst, size_t bytes) {
    return src ? hipMemcpyAsync(const_cast<void *>(dst), src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  CK(put(a->temp, S.temp, sizeof(double) * N));
  CK(put(a->xne, S.xne, sizeof(double) * N));
  CK(put(a->rho, S.rho, sizeof(double) * N));
  CK(put(a->rhox, S.rhox, sizeof(double) * N));
  CK(put(a->fraction_pf, S.fpf, sizeof(float) * N * (size_t)S.pf_ld));
  CK(put(a->frac_h1, c->fh1, sizeof(float) * N));
  CK(put(a->frac_he1, c->fhe1, sizeof(float) * N));
  CK(put(a->frac_hmol, c->fhmol, sizeof(float) * N));
  CK(put(a->contop_std, S.cstd, sizeof(double) * N));
  CK(put(a->he_gam, S.he_gam, sizeof(double) * (size_t)S.nhe * N));
  DBuf raw;
  if (a->contop) {  // ContOpac(Ntau, Nwl) as the host holds it, transposed on the device
    CK(raw.ensure(sizeof(double) * N * Nw));
    CK(put(a->contop, raw.p, sizeof(double) * N * Nw));
    launch_transpose_cont((int64_t)Nw, (int)N, raw.as<double>(), const_cast<double *>(S.contT), s);
  }
  // XlineInit's per-depth constants follow T, Xne and the perturber fractions
  if (a->temp || a->xne || a->frac_h1 || a->frac_he1 || a->frac_hmol)
    launch_xline_init(S.ntau, S.temp, S.xne, c->fh1, c->fhe1, c->fhmol, S.h2brd, const_cast<double *>(S.xln), s);
  const hipError_t e1 = hipGetLastError(), e2 = hipStreamSynchronize(s);
  raw.release();
  CK(e1);
  CK(e2);
  return 0;
}

}  // namespace

extern "C" {

int fsg_set_atmosphere(fsg_ctx *c, int32_t atm, const fsg_atmosphere *a) {
  g_err.clear();
  if (!c || !a) return fail("fsg_set_atmosphere: NULL argument");
  if (atm < 0 || atm >= std::max(c->S.nbatch, 1)) return fail("fsg_set_atmosphere: atmosphere index out of range");
  fsg_ctx *k = c->S.batch ? c->kids[atm] : c;
  if (a->contop_std && k->S.motype != 1) return fail("fsg_set_atmosphere: ContOpacStd belongs to Motype 1 models");
  if (a->rhox && k->S.sph) return fail("fsg_set_atmosphere: the rays of a spherical model are fixed at fsg_init");
  if (a->he_gam && k->S.nhe == 0) return fail("fsg_set_atmosphere: the context has no He I Stark tables");
  if (a->temp && k->S.nhe > 0 && !a->he_gam)
    return fail("fsg_set_atmosphere: a new temperature needs the He I Stark widths at it (he_gam)");
  CK(hipSetDevice(c->dev));
  if (set_atmosphere_one(k, a)) return 1;
  if (c->S.batch) {
    // the batch's own ContOpac over the concatenated grid: this atmosphere's segment, its
    // padding points repeating the last column
    const size_t N = (size_t)c->S.ntau, tot = (size_t)c->S.nwl;
    const size_t o = (size_t)c->boff[atm], seg = (size_t)(c->boff[atm + 1] - c->boff[atm]);
    const size_t nw = (size_t)k->S.nwl;
    if (a->contop) {
      std::vector<double> h(N * seg);
      for (size_t kk = 0; kk < N; kk++)
        for (size_t p = 0; p < seg; p++) h[kk * seg + p] = a->contop[std::min(p, nw - 1) * N + kk];
      CK(hipMemcpy2DAsync(const_cast<double *>(c->S.contT) + o, sizeof(double) * tot, h.data(), sizeof(double) * seg,
                          sizeof(double) * seg, N, hipMemcpyHostToDevice, c->st));
      CK(hipStreamSynchronize(c->st));
    }
    // the batch context's own set-up is the first atmosphere's
    if (atm == 0) {
      fsg_atmosphere b = *a;
      b.contop = nullptr;
      if (set_atmosphere_one(c, &b)) return 1;
    }
  }
  // windows, flags, the formal solution and the reverse-mode state of earlier calls belong
  // to the old atmosphere
  for (fsg_ctx *x : {c, k}) {
    x->p1_id = 0;
    x->have_rt = 0;
    x->have_glop = 0;
  }
  return 0;
}

int fsg_contop(fsg_ctx *c, const fsg_contop_eos *e, int32_t ifop_mask, int32_t add, double *contop_out) {
  g_err.clear();
  if (!c || !e) return fail("fsg_contop: NULL argument");
  if (c->S.batch) return fail("fsg_contop: a batch context takes each atmosphere's ContOpac through fsg_set_atmosphere");
  const unsigned mask = (unsigned)ifop_mask & ~CONT_HOP_NOLEV;
  if ((ifop_mask & FSG_CONTOP_HOP_NOLEV) && !(mask & 1u))
    return fail("fsg_contop: FSG_CONTOP_HOP_NOLEV qualifies HOP (bit 0), which is not switched on");
  // all thirteen sources of Xcont (Ifop 1-13)
  const unsigned on_device = 0x1FFFu;
  if (mask & ~on_device) {
    std::string m = "fsg_contop: continuum source(s)";
    for (int b = 0; b < 32; b++)
      if ((mask & ~on_device) >> b & 1u) m += " " + std::to_string(b + 1);
    return fail(m + " are not evaluated on the device (the host's Xcont keeps them)");
  }
  auto on = [mask](int ifop) { return (mask >> (ifop - 1)) & 1u; };
  if (((on(1) || on(2) || on(3) || on(4)) && !e->fpf_h1) || ((on(1) || on(2)) && !e->fpf_h2) ||
      (on(3) && !e->frac_hmin) || ((on(5) || on(7) || on(8)) && !e->fpf_he1) || (on(13) && !e->frac_hmol) ||
      (on(5) && !e->frac_he2) || (on(6) && (!e->fpf_he2 || !e->fpf_he3)))
    return fail("fsg_contop: a population the selected sources need is NULL");
  const float *const cool_pops[10] = {e->fpf_c1, e->fpf_mg1, e->fpf_al1, e->fpf_si1, e->fpf_fe1,
                                      e->fpf_ch,  e->fpf_oh,  e->frac_nh, e->frac_h1, e->frac_he1};
  if (on(9)) {
    if (!e->frac_hmol) return fail("fsg_contop: COOLOP needs frac_hmol");
    for (const float *q : cool_pops)
      if (!q) return fail("fsg_contop: COOLOP needs the C I, Mg I, Al I, Si I, Fe I, CH, OH, NH, H I and He I rows");
  }
  const float *const luke_pops[6] = {e->fpf_c2, e->fpf_mg2, e->fpf_si2, e->fpf_n1, e->fpf_o1, e->fpf_ca2};
  if (on(10))
    for (const float *q : luke_pops)
      if (!q) return fail("fsg_contop: LUKEOP needs the C II, Mg II, Si II, N I, O I and Ca II rows");
  if (on(11) && (!e->hot_frac || !e->hot_fpf))
    return fail("fsg_contop: HOTOP needs hot_frac and hot_fpf");
  CK(hipSetDevice(c->dev));
  const DevSetup &S = c->S;
  const size_t N = (size_t)S.ntau, Nw = (size_t)S.nwl;
  hipStream_t s = c->st;
  // per-depth state as ContopSetup forms it (src/synth.f90:589-596): T, Xne and Rho from the
  // context's atmosphere
  std::vector<double> T(N), xne(N), rho(N);
  CK(hipMemcpyAsync(T.data(), S.temp, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(xne.data(), S.xne, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(rho.data(), S.rho, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  std::vector<double> tab;
  contop_tables(tab);
  std::vector<double> blk(CR_NROW * N + tab.size(), 0.0);
  auto col = [&](int r) { return blk.data() + (size_t)r * N; };
  const float *pops[5] = {e->fpf_h1, e->fpf_h2, e->frac_hmin, e->fpf_he1, e->frac_hmol};
  for (size_t k = 0; k < N; k++) {
    const double tk = kKbolt * T[k];
    col(CR_T)[k] = T[k];
    col(CR_HKT)[k] = kHc / tk;
    col(CR_TKEV)[k] = tk * kErg2ev;
    col(CR_THETA)[k] = 5040.0 / T[k];
    col(CR_XNE)[k] = xne[k];
    col(CR_RHO)[k] = rho[k];
    for (int q = 0; q < 5; q++) col(CR_FH1 + q)[k] = pops[q] ? (double)pops[q][k] : 0.0;
    col(CR_TLOG)[k] = std::log(T[k]);
  }
  // ContopInit's factors of HOP, HE1OP and HE2OP (src/contop.f90:377-470), in its order of
  // operations with the host's libm; the EOS rows are REAL(4), so the logarithms of He I's
  // FractionPF are single-precision ones, as the reference takes them
  const double *Tkev = col(CR_TKEV);
  if (on(1)) {
    for (size_t j = 0; j < N; j++) {
      const double nh1 = (double)e->fpf_h1[j];
      for (int n = 1; n <= 8; n++)
        col(CR_BOLTH + n - 1)[j] =
            std::exp(-13.595 * (1.0 - 1.0 / (double)(n * n)) / Tkev[j]) * 2.0 * (double)(n * n) * nh1 / rho[j];
      col(CR_FREETH)[j] = xne[j] * (double)e->fpf_h2[j] / (std::sqrt(T[j]) * rho[j]);
      const double xr = nh1 / 13.595 * Tkev[j] / rho[j];
      col(CR_BOLTEXH)[j] = std::exp(-13.427 / Tkev[j]) * xr;
      col(CR_EXLIMH)[j] = std::exp(-13.595 / Tkev[j]) * xr;
    }
  }
  if (on(5)) {
    for (size_t j = 0; j < N; j++) {
      const float fhe1 = e->fpf_he1[j];
      const double lf = (double)std::log(fhe1), lr = std::log(rho[j]);
      for (int n = 1; n <= 10; n++)
        col(CR_BOLTHE1 + n - 1)[j] = std::exp(-kHe1Chi[n - 1] / Tkev[j] + lf - lr) * kHe1G[n - 1];
      for (int n = 4; n <= 27; n++)
        col(CR_BOLTNHE1 + n - 4)[j] =
            std::exp(-24.587 * (1.0 - 1.0 / (double)(n * n)) / Tkev[j] + lf - lr) * 4.0 * (double)(n * n);
      col(CR_FREETHE1)[j] = xne[j] * 1.0e-10 * (double)e->frac_he2[j] * 1.0e-10 / rho[j] / std::sqrt(T[j]) * 1.0e-10;
      const double xrlog = std::log((double)fhe1 * (2.0 / 13.595) * Tkev[j] / rho[j]);
      col(CR_BOLTEXHE1)[j] = std::exp(-23.730 / Tkev[j] + xrlog);
      col(CR_EXLIMHE1)[j] = std::exp(-24.587 / Tkev[j] + xrlog);
    }
  }
  if (on(6)) {
    for (size_t j = 0; j < N; j++) {
      const double fhe2 = (double)e->fpf_he2[j];
      for (int n = 1; n <= 9; n++) {
        const double bltarg = (54.403 - 54.403 / (double)(n * n)) / Tkev[j] + std::log(rho[j]);
        col(CR_BOLTHE2 + n - 1)[j] =
            (fhe2 == 0.0 || bltarg > 80.0) ? 0.0 : std::exp(-bltarg) * 2.0 * (double)(n * n) * fhe2;
      }
      col(CR_FREETHE2)[j] = xne[j] * (double)e->fpf_he3[j] / std::sqrt(T[j]) / rho[j];
      const double xrlog = std::log(Tkev[j] / 13.595 / rho[j]);
      const double bltlog = 53.859 / Tkev[j] - xrlog, exllog = 54.403 / Tkev[j] - xrlog;
      col(CR_BOLTEXHE2)[j] = (fhe2 == 0.0 || bltlog > 80.0) ? 0.0 : fhe2 * std::exp(-bltlog);
      col(CR_EXLIMHE2)[j] = (fhe2 == 0.0 || exllog > 80.0) ? 0.0 : fhe2 * std::exp(-exllog);
    }
  }
  // ContopInit's factors of COOLOP's members (src/contop.f90:337-420)
  if (on(9)) coolop_rows((int)N, cool_pops, blk.data());
  // and of LUKEOP's members and HOTOP (src/contop.f90:337-366, 462-490)
  if (on(10)) lukeop_rows((int)N, luke_pops, blk.data());
  if (on(11)) hotop_rows((int)N, e->hot_frac, e->hot_fpf, blk.data());
  std::copy(tab.begin(), tab.end(), blk.begin() + CR_NROW * N);
  CK(c->ctab.ensure(sizeof(double) * blk.size()));
  double *dblk = c->ctab.as<double>();
  CK(hipMemcpyAsync(dblk, blk.data(), sizeof(double) * blk.size(), hipMemcpyHostToDevice, s));
  const unsigned kmask = mask | ((ifop_mask & FSG_CONTOP_HOP_NOLEV) ? CONT_HOP_NOLEV : 0u);
  launch_contop(S.nwl, S.ntau, S.wave, S.wfirst, dblk, dblk + CR_NROW * N, S.cstd, S.motype, kmask, add ? 1 : 0,
                const_cast<double *>(S.contT), s);
  if (on(9) || on(10) || on(11))
    CK(c->cxs.ensure(sizeof(double) * std::max(coolop_scratch_per_point(), luke_scratch_per_point()) * Nw));
  if (on(9))
    launch_coolop(S.nwl, S.ntau, S.wave, S.wfirst, dblk, dblk + CR_NROW * N, S.cstd, S.motype, c->cxs.as<double>(),
                  const_cast<double *>(S.contT), s);
  if (on(10) || on(11))
    launch_lukehot(S.nwl, S.ntau, S.wave, S.wfirst, dblk, dblk + CR_NROW * N, S.cstd, S.motype, mask,
                   c->cxs.as<double>(), const_cast<double *>(S.contT), s);
  CK(hipGetLastError());
  if (contop_out) {  // [ntau][nwl] -> ContOpac(Ntau, Nwl)
    std::vector<double> h(N * Nw);
    CK(hipMemcpyAsync(h.data(), S.contT, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (size_t k = 0; k < N; k++)
      for (size_t iw = 0; iw < Nw; iw++) contop_out[iw * N + k] = h[k * Nw + iw];
  }
  CK(hipStreamSynchronize(s));
  // windows and formal solutions of earlier calls belong to the old continuum
  c->p1_id = 0;
  c->have_rt = 0;
  c->have_glop = 0;
  return 0;
}

int fsg_batch_offsets(const fsg_ctx *c, int64_t *off_out) {
  g_err.clear();
  if (!c || !off_out) return fail("fsg_batch_offsets: NULL argument");
  if (!c->S.batch) return fail("fsg_batch_offsets: not a batch context");
  std::copy(c->boff.begin(), c->boff.end(), off_out);
  return 0;
}

int fsg_lines_batch(fsg_ctx *c, const fsg_dlines *src, const int64_t *index, const int32_t *atm, int64_t n,
                    fsg_dlines **out) {
  g_err.clear();
  if (!c || !src || !out || (n > 0 && (!index || !atm))) return fail("fsg_lines_batch: NULL argument");
  *out = nullptr;
  if (!c->S.batch) return fail("fsg_lines_batch: not a batch context");
  if (src->zeeman) return fail("fsg_lines_batch: batches take no Zeeman lists");
  if (src->L.atm) return fail("fsg_lines_batch: the source is itself a batch list");
  if (n < 0) return fail("fsg_lines_batch: negative line count");
  const int64_t ns = src->L.n;
  std::vector<double> w((size_t)std::max<int64_t>(ns, 1));
  CK(hipSetDevice(c->dev));
  if (ns > 0) CK(hipMemcpy(w.data(), src->L.wl, sizeof(double) * ns, hipMemcpyDeviceToHost));
  for (int64_t k = 0; k < n; k++) {
    if (index[k] < 0 || index[k] >= ns) return fail("fsg_lines_batch: line index out of range");
    if (atm[k] < 0 || atm[k] >= c->S.nbatch) return fail("fsg_lines_batch: atmosphere index out of range");
    if (k > 0 && atm[k] < atm[k - 1]) return fail("fsg_lines_batch: atmosphere indices must not decrease");
    if (k > 0 && atm[k] == atm[k - 1] && w[index[k]] < w[index[k - 1]])
      return fail("ReadHDF: line list is not sorted in wavelength");
  }
  fsg_dlines *d = new fsg_dlines();
  d->ctx = c;
  d->list_type = src->list_type;
  d->L.n = n;
  const size_t m = (size_t)std::max<int64_t>(n, 1);
  auto dev = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    d->owned.push_back(p);
    return p;
  };
  double *wl = (double *)dev(m * 8);
  float *elo = (float *)dev(m * 4), *gf = (float *)dev(m * 4), *grad = (float *)dev(m * 4);
  int *spc = (int *)dev(m * 4);
  float *gq = src->L.gqst ? (float *)dev(m * 4) : nullptr, *gv = src->L.gvdw ? (float *)dev(m * 4) : nullptr;
  uint8_t *im = src->L.is_mol ? (uint8_t *)dev(m) : nullptr;
  int64_t *idx = (int64_t *)dev(m * 8);
  int *da = (int *)dev(m * 4);
  if (!wl || !elo || !gf || !grad || !spc || (src->L.gqst && !gq) || (src->L.gvdw && !gv) || (src->L.is_mol && !im) ||
      !idx || !da) {
    fsg_lines_free(d);
    return fail("fsg_lines_batch: out of device memory");
  }
  d->L.wl = wl;
  d->L.elo = elo;
  d->L.gf = gf;
  d->L.grad = grad;
  d->L.spc = spc;
  d->L.gqst = gq;
  d->L.gvdw = gv;
  d->L.is_mol = im;
  d->L.atm = da;
  std::vector<int> ha(atm, atm + n);
  int rc = 0;
  if (n > 0) {
    if (hipMemcpyAsync(idx, index, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipMemcpyAsync(da, ha.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->st) != hipSuccess)
      rc = fail("fsg_lines_batch: copy to the device failed");
    if (!rc) {
      launch_gather_lines(src->L, idx, n, wl, elo, gf, grad, gq, gv, spc, im, c->st);
      if (hipGetLastError() != hipSuccess) rc = fail("fsg_lines_batch: gather kernel failed");
    }
  }
  if (hipStreamSynchronize(c->st) != hipSuccess && !rc) rc = fail("fsg_lines_batch: gather failed");
  if (rc) {
    fsg_lines_free(d);
    return 1;
  }
  *out = d;
  return 0;
}

int fsg_lines_upload(fsg_ctx *c, int32_t list_type, const fsg_lines *h, fsg_dlines **out) {
  g_err.clear();
  if (!c || !h || !out) return fail("fsg_lines_upload: NULL argument");
  *out = nullptr;
  if (list_type < 0 || list_type > 2) return fail("fsg_lines_upload: unknown line-list type");
  if (h->n < 0 || h->n > 2147483647LL) return fail("fsg_lines_upload: line count out of range");
  if (h->n > 0 && (!h->wl || !h->elo || !h->gf || !h->grad || !h->spc))
    return fail("fsg_lines_upload: required line column is NULL");
  if (h->n > 0 && list_type != FSG_LIST_MOLECULAR && (!h->gqst || !h->gvdw))
    return fail("fsg_lines_upload: atomic and mixed lists need Gqst and Gvdw");
  for (int64_t i = 1; i < h->n; i++)
    if (h->wl[i] < h->wl[i - 1]) return fail("ReadHDF: line list is not sorted by wavelength");
  for (int64_t i = 0; i < h->n; i++)
    if (h->spc[i] < 1 || h->spc[i] > c->S.nspc) return fail("LineopCalc: species index out of range");
  CK(hipSetDevice(c->dev));
  fsg_dlines *d = new fsg_dlines();
  d->ctx = c;
  d->list_type = list_type;
  d->L.n = h->n;
  size_t n = (size_t)h->n;
  int rc = 0;
  rc |= upload(c, h->wl, n, &d->L.wl, d->owned);
  rc |= upload(c, h->elo, n, &d->L.elo, d->owned);
  rc |= upload(c, h->gf, n, &d->L.gf, d->owned);
  rc |= upload(c, h->grad, n, &d->L.grad, d->owned);
  rc |= upload(c, h->spc, n, (const int32_t **)&d->L.spc, d->owned);
  if (list_type != FSG_LIST_MOLECULAR) {
    rc |= upload(c, h->gqst, n, &d->L.gqst, d->owned);
    rc |= upload(c, h->gvdw, n, &d->L.gvdw, d->owned);
  }
  if (list_type == FSG_LIST_MIXED && h->is_mol) rc |= upload(c, h->is_mol, n, &d->L.is_mol, d->owned);
  if (h->pi_off && h->si_off && n > 0 && !rc) {
    // Zeeman patterns (CSR): offsets start at 0 and grow; sigma groups hold a blue and a
    // red half of equal size
    const char *bad = nullptr;
    if (h->pi_off[0] != 0 || h->si_off[0] != 0) bad = "offsets must start at 0";
    for (size_t i = 0; i < n && !bad; i++) {
      int32_t np = h->pi_off[i + 1] - h->pi_off[i], ns = h->si_off[i + 1] - h->si_off[i];
      if (np < 0 || ns < 0) bad = "offsets must not decrease";
      else if (ns % 2) bad = "sigma components must come in blue/red pairs";
    }
    const int64_t npi = h->pi_off[n], nsi = h->si_off[n];
    if (!bad && ((npi > 0 && (!h->spl_pi || !h->str_pi)) || (nsi > 0 && (!h->spl_si || !h->str_si))))
      bad = "component arrays are NULL";
    if (bad) {
      fsg_lines_free(d);
      return fail(std::string("fsg_lines_upload: Zeeman pattern: ") + bad);
    }
    float *sp = nullptr;
    rc |= upload(c, h->pi_off, n + 1, &d->Z.pi_off, d->owned);
    rc |= upload(c, h->si_off, n + 1, &d->Z.si_off, d->owned);
    rc |= upload(c, h->spl_pi, (size_t)npi, &d->Z.spl_pi, d->owned);
    rc |= upload(c, h->str_pi, (size_t)npi, &d->Z.str_pi, d->owned);
    rc |= upload(c, h->spl_si, (size_t)nsi, &d->Z.spl_si, d->owned);
    rc |= upload(c, h->str_si, (size_t)nsi, &d->Z.str_si, d->owned);
    if (!rc && hipMalloc((void **)&sp, sizeof(float) * (size_t)std::max<int64_t>(npi + nsi, 1)) != hipSuccess)
      rc = fail("fsg_lines_upload: out of device memory for the Zeeman shifts");
    if (sp) {
      d->owned.push_back(sp);
      d->Z.sp_pi = sp;
      d->Z.sp_si = sp + npi;
    }
    d->zeeman = 1;
  }
  if (rc || hipStreamSynchronize(c->st) != hipSuccess) {
    fsg_lines_free(d);
    return 1;
  }
  *out = d;
  return 0;
}

static int species_table(const fsg_species_entry *species, int32_t nspecies, int nspc, SpeciesMap &m, bool &any_mol) {
  any_mol = false;
  if (nspecies < 0 || (nspecies > 0 && !species)) return fail("fsg_lines: species table is NULL");
  for (int32_t i = 0; i < nspecies; i++) {
    const fsg_species_entry &e = species[i];
    std::string el(e.el, strnlen(e.el, sizeof(e.el)));
    while (!el.empty() && el.back() == ' ') el.pop_back();
    while (!el.empty() && el.front() == ' ') el.erase(el.begin());
    if (e.slot < 1 || (nspc > 0 && e.slot > nspc)) return fail("LineopCalc: species index out of range");
    m[SpeciesKey(el, e.ion)] = {e.slot, (uint8_t)(e.is_mol != 0)};
    any_mol |= e.is_mol != 0;
  }
  return 0;
}

static int isotope_table(const fsg_isotope_entry *isotopes, int32_t n, IsotopeMap &m) {
  if (n < 0 || (n > 0 && !isotopes)) return fail("fsg_lines: isotope table is NULL");
  for (int32_t i = 0; i < n; i++) {
    std::string el(isotopes[i].el, strnlen(isotopes[i].el, sizeof(isotopes[i].el)));
    while (!el.empty() && el.back() == ' ') el.pop_back();
    while (!el.empty() && el.front() == ' ') el.erase(el.begin());
    m[SpeciesKey(el, isotopes[i].a)] = isotopes[i].frac;
  }
  return 0;
}

int fsg_lines_scan(const char *dir, double wstart, double wend, int32_t *list_type, int64_t *first, int64_t *count) {
  g_err.clear();
  if (!dir || !first || !count) return fail("fsg_lines_scan: NULL argument");
  Sidecar sc;
  std::string err;
  if (sidecar_open(dir, sc, err) || sidecar_select(sc, wstart, wend, *first, *count, err)) return fail(err);
  if (list_type) *list_type = sc.molecular ? FSG_LIST_MOLECULAR : FSG_LIST_ATOMIC;
  return 0;
}

int fsg_lines_read(const char *dir, int64_t first, int64_t count, const fsg_species_entry *species, int32_t nspecies,
                   const fsg_isotope_entry *isotopes, int32_t nisotopes, int32_t mol_slot, int32_t nthreads,
                   double *wl, float *elo, float *gf, float *grad, float *gqst, float *gvdw, int32_t *spc,
                   uint8_t *is_mol, int64_t *kept) {
  g_err.clear();
  if (!dir || !kept || (count > 0 && (!wl || !elo || !gf || !grad || !spc))) return fail("fsg_lines_read: NULL argument");
  Sidecar sc;
  std::string err;
  if (sidecar_open(dir, sc, err)) return fail(err);
  if (first < 0 || count < 0 || first + count > sc.nl) return fail("fsg_lines_read: lines outside the file");
  if (!sc.molecular && count > 0 && (!gqst || !gvdw)) return fail("fsg_lines_read: atomic lists need Gqst and Gvdw");
  SpeciesMap m;
  bool any_mol = false;
  if (species_table(species, nspecies, 0, m, any_mol)) return 1;
  IsotopeMap im;
  if (isotope_table(isotopes, nisotopes, im)) return 1;
  IngestOut o{wl, elo, gf, grad, gqst, gvdw, spc, is_mol};
  if (sidecar_convert(sc, first, count, m, im, mol_slot, nthreads, o, *kept, err)) return fail(err);
  return 0;
}

int fsg_lines_stream(fsg_ctx *c, const char *dir, double wstart, double wend, const fsg_species_entry *species,
                     int32_t nspecies, const fsg_isotope_entry *isotopes, int32_t nisotopes, int32_t mol_slot,
                     int32_t nthreads, fsg_dlines **out, int64_t *kept) {
  g_err.clear();
  if (!c || !dir || !out || !kept) return fail("fsg_lines_stream: NULL argument");
  *out = nullptr;
  *kept = 0;
  Sidecar sc;
  std::string err;
  int64_t first = 0, count = 0;
  if (sidecar_open(dir, sc, err) || sidecar_select(sc, wstart, wend, first, count, err)) return fail(err);
  SpeciesMap m;
  bool any_mol = false;
  if (species_table(species, nspecies, c->S.nspc, m, any_mol)) return 1;
  IsotopeMap im;
  if (isotope_table(isotopes, nisotopes, im)) return 1;
  if (sc.molecular && (mol_slot < 1 || mol_slot > c->S.nspc)) return fail("LineopCalc: species index out of range");
  if (sc.zeeman) {
    // magnetic lists (a few thousand lines) are converted on the host with their Zeeman
    // patterns and uploaded as fsg_lines_upload does
    const size_t n = (size_t)std::max<int64_t>(count, 1);
    std::vector<double> wl(n);
    std::vector<float> elo(n), gf(n), grad(n), gq(n), gv(n);
    std::vector<int32_t> spc(n);
    std::vector<uint8_t> mol(n);
    ZeemanOut z;
    IngestOut o{wl.data(), elo.data(), gf.data(), grad.data(), gq.data(), gv.data(), spc.data(), mol.data()};
    int64_t kn = 0;
    if (sidecar_convert(sc, first, count, m, im, mol_slot, nthreads, o, kn, err, &z)) return fail(err);
    fsg_lines h{};
    h.n = kn;
    h.wl = wl.data();
    h.elo = elo.data();
    h.gf = gf.data();
    h.grad = grad.data();
    h.gqst = gq.data();
    h.gvdw = gv.data();
    h.spc = spc.data();
    h.is_mol = mol.data();
    h.pi_off = z.pi_off.data();
    h.spl_pi = z.spl_pi.data();
    h.str_pi = z.str_pi.data();
    h.si_off = z.si_off.data();
    h.spl_si = z.spl_si.data();
    h.str_si = z.str_si.data();
    if (fsg_lines_upload(c, FSG_LIST_MIXED, &h, out)) return 1;
    *kept = kn;
    return 0;
  }
  CK(hipSetDevice(c->dev));
  fsg_dlines *d = new fsg_dlines();
  d->ctx = c;
  d->list_type = sc.molecular ? FSG_LIST_MOLECULAR : (any_mol ? FSG_LIST_MIXED : FSG_LIST_ATOMIC);
  const size_t nmax = (size_t)std::max<int64_t>(count, 1);
  double *dwl = nullptr;
  float *delo = nullptr, *dgf = nullptr, *dgrad = nullptr, *dgq = nullptr, *dgv = nullptr;
  int32_t *dspc = nullptr;
  uint8_t *dmol = nullptr;
  auto dalloc = [&](void **p, size_t bytes) -> bool {
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return false;
    d->owned.push_back(*p);
    return true;
  };
  bool ok = dalloc((void **)&dwl, nmax * 8) && dalloc((void **)&delo, nmax * 4) && dalloc((void **)&dgf, nmax * 4) &&
            dalloc((void **)&dgrad, nmax * 4) && dalloc((void **)&dspc, nmax * 4);
  if (ok && !sc.molecular) ok = dalloc((void **)&dgq, nmax * 4) && dalloc((void **)&dgv, nmax * 4);
  if (ok && d->list_type == FSG_LIST_MIXED) ok = dalloc((void **)&dmol, nmax);
  if (!ok) {
    fsg_lines_free(d);
    return fail("fsg_lines_stream: out of device memory for the line list");
  }
  // two pinned buffer sets; chunk k converts into set k % 2 while set (k - 1) % 2 copies
  int64_t cl = 1LL << 22;  // lines per pinned chunk; FSG_INGEST_CHUNK overrides it
  if (const char *e = std::getenv("FSG_INGEST_CHUNK")) cl = std::max<int64_t>(1, std::atoll(e));
  const int64_t chunk = std::min<int64_t>(std::max<int64_t>(count, 1), cl);
  const size_t per = (size_t)chunk * (8 + 4 * 6 + 1);
  char *pin[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  int rc = 0;
  for (int b = 0; b < 2 && !rc; b++) {
    if (hipHostMalloc((void **)&pin[b], per, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&done[b], hipEventDisableTiming) != hipSuccess)
      rc = fail("fsg_lines_stream: pinned staging buffers unavailable");
  }
  int64_t total = 0;
  for (int64_t k = 0, off = first; !rc && off < first + count; k++, off += chunk) {
    const int b = (int)(k & 1);
    if (k >= 2 && hipEventSynchronize(done[b]) != hipSuccess) {
      rc = fail("fsg_lines_stream: copy to the device failed");
      break;
    }
    char *p = pin[b];
    IngestOut o;
    o.wl = reinterpret_cast<double *>(p);
    o.elo = reinterpret_cast<float *>(p + chunk * 8);
    o.gf = o.elo + chunk;
    o.grad = o.gf + chunk;
    o.gqst = o.grad + chunk;
    o.gvdw = o.gqst + chunk;
    o.spc = reinterpret_cast<int32_t *>(o.gvdw + chunk);
    o.is_mol = reinterpret_cast<uint8_t *>(o.spc + chunk);
    const int64_t n = std::min<int64_t>(chunk, first + count - off);
    int64_t kn = 0;
    if (sidecar_convert(sc, off, n, m, im, mol_slot, nthreads, o, kn, err)) {
      rc = fail(err);
      break;
    }
    hipStream_t s = c->st;
    bool cp = hipMemcpyAsync(dwl + total, o.wl, kn * 8, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(delo + total, o.elo, kn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(dgf + total, o.gf, kn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(dgrad + total, o.grad, kn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(dspc + total, o.spc, kn * 4, hipMemcpyHostToDevice, s) == hipSuccess;
    if (cp && dgq) cp = hipMemcpyAsync(dgq + total, o.gqst, kn * 4, hipMemcpyHostToDevice, s) == hipSuccess &&
                        hipMemcpyAsync(dgv + total, o.gvdw, kn * 4, hipMemcpyHostToDevice, s) == hipSuccess;
    if (cp && dmol) cp = hipMemcpyAsync(dmol + total, o.is_mol, kn, hipMemcpyHostToDevice, s) == hipSuccess;
    if (!cp || hipEventRecord(done[b], s) != hipSuccess) {
      rc = fail("fsg_lines_stream: copy to the device failed");
      break;
    }
    total += kn;
  }
  if (hipStreamSynchronize(c->st) != hipSuccess && !rc) rc = fail("fsg_lines_stream: copy to the device failed");
  for (int b = 0; b < 2; b++) {
    if (pin[b]) (void)hipHostFree(pin[b]);
    if (done[b]) (void)hipEventDestroy(done[b]);
  }
  if (rc) {
    fsg_lines_free(d);
    return 1;
  }
  d->L.n = total;
  d->L.wl = dwl;
  d->L.elo = delo;
  d->L.gf = dgf;
  d->L.grad = dgrad;
  d->L.spc = dspc;
  d->L.gqst = dgq;
  d->L.gvdw = dgv;
  d->L.is_mol = dmol;
  *out = d;
  *kept = total;
  return 0;
}

int fsg_lines_free(fsg_dlines *d) {
  if (!d) return 0;
  (void)hipSetDevice(d->ctx->dev);
  (void)hipStreamSynchronize(d->ctx->st);
  for (void *p : d->owned) (void)hipFree(p);
  delete d;
  return 0;
}

// The line-opacity stage on a device-resident list.  All retained lines are processed in
// one pass: the coefficient arrays live in HBM (Ntau x 32 B per retained line in FP32).
// When nlines_lim is smaller than the list, MarkWideLines and PACK follow the reference's
// multi-block rules; the accumulated opacity does not depend on the block partition.
}  // extern "C"

namespace {

// pass 1 of LineopCalc on a device list, the evaluation count and the stable bin-major
// order of the lines (retained lines keep their index order per bin); returns the bin
// populations.  Retained lines: width bins 0..nbins-1, then the He I Stark lines (HE_KEY).

int scalar_pass1(fsg_ctx *c, fsg_dlines *dl, LinePrep &P) {
  c->p1_id = 0;  // the buffers are overwritten from here on
  const int64_t n = dl->L.n;
  const DevSetup &S = c->S;
  const int N = S.ntau;
  hipStream_t s = c->st;
  CK(c->flag.ensure(n * sizeof(int)));
  CK(c->il.ensure(n * sizeof(int)));
  CK(c->ir.ensure(n * sizeof(int)));
  CK(c->ic.ensure(n * sizeof(int)));
  CK(c->key.ensure(n * sizeof(unsigned)));
  CK(c->key2.ensure(n * sizeof(unsigned)));
  CK(c->idx.ensure(n * sizeof(int)));
  CK(c->idx2.ensure(n * sizeof(int)));
  CK(hipMemsetAsync(c->bincnt, 0, sizeof(int) * 64, s));
  CK(hipMemsetAsync(c->binhmax, 0, sizeof(int) * 64, s));
  CK(hipMemsetAsync(c->u64, 0, sizeof(unsigned long long) * 4, s));
  CK(hipEventRecord(c->ev[0], s));
  launch_pass1(S, dl->L, dl->list_type, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(), c->ic.as<int>(),
               c->key.as<unsigned>(), c->bincnt, c->binhmax, c->batched, s);
  launch_count_evals(n, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(), N, c->u64, s);
  launch_iota(c->idx.as<int>(), n, s);
  size_t tbytes = 0;
  CK(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, c->key.as<unsigned>(), c->key2.as<unsigned>(),
                                        c->idx.as<int>(), c->idx2.as<int>(), (int)n, 0, 6, s));
  CK(c->tmp.ensure(tbytes));
  CK(hipcub::DeviceRadixSort::SortPairs(c->tmp.p, tbytes, c->key.as<unsigned>(), c->key2.as<unsigned>(),
                                        c->idx.as<int>(), c->idx2.as<int>(), (int)n, 0, 6, s));
  CK(hipMemcpyAsync(P.hcnt, c->bincnt, sizeof(P.hcnt), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(P.hhmax, c->binhmax, sizeof(P.hhmax), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(P.hu, c->u64, sizeof(P.hu), hipMemcpyDeviceToHost, s));
  CK(hipEventRecord(c->ev[1], s));
  CK(hipStreamSynchronize(s));
  P.nmain = 0;
  P.nbins = 0;
  for (int b = 0; b < HE_KEY; b++) {
    P.nmain += P.hcnt[b];
    if (P.hcnt[b] > 0) P.nbins = b + 1;
  }
  P.nhe = P.hcnt[HE_KEY];
  P.nret = P.nmain + P.nhe;
  if (P.nbins > MAX_BINS) return fail("LineopCalc: too many window-width bins");
  c->p1_id = dl->id;
  c->p1_gen = dl->gen;
  c->p1 = P;
  return 0;
}

// pass 1 of the list, or the results still held from the last pass 1 of the same list with
// the same gf (the forward synthesis before its reverse mode)
int scalar_pass1_cached(fsg_ctx *c, fsg_dlines *dl, LinePrep &P) {
  if (c->p1_id == dl->id && c->p1_gen == dl->gen) {
    P = c->p1;
    return 0;
  }
  return scalar_pass1(c, dl, P);
}

// groups of the bin-sorted retained lines: offsets, sizes and maximal half-widths of the
// nbins width bins and, last, the He I group
void group_tables(const LinePrep &P, std::vector<int64_t> &goff, std::vector<int64_t> &gcnt, std::vector<int> &ghmax) {
  const int ng = P.nbins + 1;
  goff.resize(ng);
  gcnt.resize(ng);
  ghmax.resize(ng);
  int64_t o = 0;
  for (int b = 0; b < ng; b++) {
    int key = b < P.nbins ? b : HE_KEY;
    goff[b] = o;
    gcnt[b] = P.hcnt[key];
    ghmax[b] = P.hhmax[key];
    o += P.hcnt[key];
  }
}

}  // namespace

extern "C" {

int fsg_lineop_dev(fsg_ctx *c, fsg_dlines *dl, int64_t nlines_lim, fsg_lineop_stats *stats) {
  g_err.clear();
  if (!c || !dl) return fail("fsg_lineop_dev: NULL argument");
  if (nlines_lim < 1) return fail("fsg_lineop_dev: NlinesLim must be positive");
  if ((c->S.batch != nullptr) != (dl->L.atm != nullptr))
    return fail("fsg_lineop_dev: batch contexts take batch lists (fsg_lines_batch) and only they do");
  CK(hipSetDevice(c->dev));
  fsg_lineop_stats st{};
  const int64_t n = dl->L.n;
  st.nlines_in = n;
  st.nsegments = 1;
  if (n == 0) {
    if (stats) *stats = st;
    return 0;
  }
  const DevSetup &S = c->S;
  hipStream_t s = c->st;
  LinePrep P;
  if (scalar_pass1(c, dl, P)) return 1;
  const int nbins = P.nbins;
  const int64_t nret = P.nret;
  unsigned long long *hu = P.hu;  // [4]: evaluations, wide lines
  st.nlines_use = nret;
  st.evals = (int64_t)hu[0];
  // MarkWideLines over the packed (single block, atomic/mixed) or full list; a statistic
  // here (every line is accumulated over its whole window), not counted for batch lists
  const int64_t nstep = std::min<int64_t>(n, nlines_lim);
  const int pack = (nstep == n) && (dl->list_type == FSG_LIST_ATOMIC || dl->list_type == FSG_LIST_MIXED);
  if (!S.batch) {
    CK(c->wlcl.ensure(n * sizeof(double)));
    CK(c->rflag.ensure(n * sizeof(int)));
    CK(c->pos.ensure(n * sizeof(int)));
    if (pack) {
      launch_retained_flag(n, c->flag.as<int>(), c->rflag.as<int>(), s);
      size_t sb = 0;
      CK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, c->rflag.as<int>(), c->pos.as<int>(), (int)n, s));
      CK(c->tmp.ensure(sb));
      CK(hipcub::DeviceScan::ExclusiveSum(c->tmp.p, sb, c->rflag.as<int>(), c->pos.as<int>(), (int)n, s));
    }
    launch_wlc_list(n, dl->L.wl, S.wfirst, c->flag.as<int>(), c->pos.as<int>(), c->wlcl.as<double>(), pack, s);
    launch_mark_wide(S, n, dl->L.wl, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(), c->wlcl.as<double>(),
                     pack ? nret : n, nstep, c->u64 + 1, s);
  }
  // Accumulation over chunks of the bin-sorted retained lines: the coefficient records of
  // at most `chunk` lines are resident at once (the device analogue of the reference's
  // NlinesStep blocks, src/synth.f90:1011-1088).  Groups 0..nbins-1 are the width bins,
  // group nbins the He I lines.
  bool timed = false;  // ms_opacity runs from the first accumulation launch to the last
  int nfar = 0, ncore = 0;  // FP32 far-wing and core-zone launches, for their timings
  if (nret > 0) {
    const int ng = nbins + 1;
    std::vector<int64_t> goff, gcnt;
    std::vector<int> ghmax;
    group_tables(P, goff, gcnt, ghmax);
    const size_t csz = c->precision == FSG_FP32 ? sizeof(CoefA) + sizeof(CoefB) : sizeof(Coef64);
    // FP32 adds one float per (line, wavefront depth group), the group's largest far-wing
    // threshold, and the line's core zone (int2)
    const size_t wsz =
        c->precision == FSG_FP32 ? (size_t)(S.ntp / DPW) * sizeof(float) + sizeof(int2) + sizeof(LineRec) : 0;
    const int64_t chunk = chunk_lines(c, nret, nlines_lim, S.ntp * csz + wsz + sizeof(double) + sizeof(int4), 0);
    CK(c->coef.ensure((size_t)chunk * (S.ntp * csz + wsz) + 6 * 32 + sizeof(LineRec)));  // + alignment, empty record
    CK(c->wlc.ensure((size_t)chunk * sizeof(double)));
    CK(c->meta.ensure((size_t)chunk * sizeof(int4)));
    const int ntiles = (S.nwl + TILE - 1) / TILE;
    CK(c->rng.ensure((size_t)ntiles * ng * sizeof(int2)));
    // FP32 batched path: the leading width bins whose half-widths stay below NARROW_HMAX
    // points go to the narrow-line kernel (up to 128 depths)
    int nbn = 0;
    if (c->precision == FSG_FP32 && c->batched && S.ntau <= 128)
      while (nbn < nbins && ghmax[nbn] <= NARROW_HMAX) nbn++;
    const int ntiles_n = (S.nwl + NARROW_TILE - 1) / NARROW_TILE;
    if (nbn > 0) CK(c->rngn.ensure((size_t)ntiles_n * nbn * sizeof(int2)));
    size_t ztmp = 0;
    if (c->precision == FSG_FP32) {
      CK(c->rng2.ensure((size_t)((S.nwl + 63) / 64) * NZB * sizeof(int2)));  // core tiles of 64 points at least
      CK(c->zkey.ensure((size_t)chunk * sizeof(unsigned long long)));
      CK(c->zkey2.ensure((size_t)chunk * sizeof(unsigned long long)));
      CK(c->zval.ensure((size_t)chunk * sizeof(int)));
      CK(c->zperm.ensure((size_t)chunk * sizeof(int)));
      CK(c->zrec.ensure((size_t)(chunk + 1) * sizeof(ZoneRec)));  // + the spare record the kernel reads ahead
      ztmp = zone_sort_tmp_bytes(chunk);
      CK(c->tmp.ensure(ztmp));
    }
    int *zoff = c->binoff + 3 * (MAX_BINS + 1), *zend = zoff + NZB;
    const int64_t nlop = (int64_t)S.ntau * S.nwl;
    bool core_pending = false;  // a core-zone launch on st2 may still read the chunk buffers
    if (c->precision == FSG_FP32 && c->lop2_dirty) {
      CK(hipStreamSynchronize(c->st2));
      launch_zero_f32(c->lop2, nlop, s);
      c->lop2_dirty = false;
    }
    std::vector<std::vector<int>> tables;  // staged for upload; alive until the final sync
    for (int64_t s0 = 0; s0 < nret; s0 += chunk) {
      if (core_pending) CK(hipStreamWaitEvent(s, c->evj, 0));  // before the chunk buffers are reused
      const int64_t s1 = std::min(nret, s0 + chunk), m = s1 - s0;
      tables.emplace_back();
      std::vector<int> &packed = tables.back();
      chunk_tables(goff, gcnt, ghmax, s0, s1, packed);
      CK(hipMemcpyAsync(c->binoff, packed.data(), sizeof(int) * packed.size(), hipMemcpyHostToDevice, s));
      // carve the chunk's records out of the coefficient buffer, each array 32-byte aligned
      char *cur = static_cast<char *>(c->coef.p);
      auto carve = [&cur](size_t bytes) {
        char *p = cur;
        cur += (bytes + 31) & ~(size_t)31;
        return p;
      };
      CoefA *ca = reinterpret_cast<CoefA *>(carve((size_t)m * S.ntp * sizeof(CoefA)));
      CoefB *cb = reinterpret_cast<CoefB *>(carve((size_t)m * S.ntp * sizeof(CoefB)));
      float *dmaxw = reinterpret_cast<float *>(carve((size_t)m * (S.ntp / DPW) * sizeof(float)));
      int2 *zone = reinterpret_cast<int2 *>(carve((size_t)m * sizeof(int2)));
      LineRec *lrec = reinterpret_cast<LineRec *>(carve((size_t)(m + 1) * sizeof(LineRec)));  // + an empty one
      const int *order = c->idx2.as<int>() + s0;
      // chunk positions below mnarrow hold the lines of the narrow bins
      const int64_t mnarrow = nbn > 0 ? std::min<int64_t>(std::max<int64_t>(goff[nbn] - s0, 0), m) : 0;
      if (c->precision == FSG_FP32)
        launch_fill32(S, dl->L, dl->list_type, m, order, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(),
                      c->ic.as<int>(), ca, cb, dmaxw, zone, lrec, c->wlc.as<double>(), c->meta.as<int4>(), s);
      else
        launch_fill64(S, dl->L, dl->list_type, m, order, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(),
                      c->ic.as<int>(), c->coef.as<Coef64>(), c->wlc.as<double>(), c->meta.as<int4>(), s);
      launch_tile_ranges(ntiles, TILE, ng, S.nwl, c->binoff, c->binoff + ng, c->binoff + 2 * ng,
                         c->meta.as<int4>(), c->rng.as<int2>(), s);
      if (mnarrow > 0)
        launch_tile_ranges(ntiles_n, NARROW_TILE, nbn, S.nwl, c->binoff, c->binoff + ng, c->binoff + 2 * ng,
                           c->meta.as<int4>(), c->rngn.as<int2>(), s);
      if (c->precision == FSG_FP32)  // the chunk's lines by core-zone width, and per-tile candidate ranges
        launch_zone_order(m, mnarrow, std::min<int64_t>(std::max<int64_t>(goff[nbins] - s0, 0), m), ntiles, S.nwl, S.ntau,
                          c->meta.as<int4>(), zone, c->wlc.as<double>(), c->zkey.as<unsigned long long>(),
                          c->zval.as<int>(), c->zkey2.as<unsigned long long>(), c->zperm.as<int>(),
                          c->zrec.as<ZoneRec>(), c->tmp.p, ztmp, zoff, zend, c->u64 + 2, s);
      // the range table has ng groups per tile: the first nbins are read by the main launch
      // (stride ng), the last one by the He launch (offset nbins, stride ng)
      const bool has_main = s0 < goff[nbins], has_he = s1 > goff[nbins];
      // the far-wing and core-zone kernels take the bins from nbn on (rng offset by nbn
      // groups, stride ng); the narrow bins have their own kernel
      const bool has_wide = s1 > goff[nbn] && s0 < goff[nbins];
      const int nbw = nbins - nbn;
      const int2 *rngw = c->rng.as<int2>() + nbn;
      if (c->precision == FSG_FP32 && has_wide && nbw > 0) {
        if ((S.ntau + 15) / 16 <= MFMA_MAX_NQ) {
          // the matrix-core kernel's line lists; their total size comes back to the host
          const size_t ftmp = far_scan_tmp_bytes(ntiles);
          CK(c->fcap.ensure((size_t)(ntiles + 1) * sizeof(long long)));
          CK(c->foff.ensure((size_t)(ntiles + 1) * sizeof(long long)));
          CK(c->ftmp.ensure(ftmp));
          launch_far_cap(ntiles, nbw, ng, rngw, c->fcap.as<long long>(), c->foff.as<long long>(), c->ftmp.p, ftmp,
                         lrec, m, s);
          long long ftotal = 0;
          CK(hipMemcpyAsync(&ftotal, c->foff.as<long long>() + ntiles, sizeof(long long), hipMemcpyDeviceToHost, s));
          CK(hipStreamSynchronize(s));
          CK(c->flist.ensure((size_t)ftotal * sizeof(int)));
          launch_far_lists(ntiles, S.nwl, nbw, ng, rngw, lrec, c->foff.as<long long>(), c->flist.as<int>(), s);
          if (!c->tuned) {
            if (autotune(c, ntiles, m, ca, cb, dmaxw, lrec, zoff, zend)) return 1;
            c->tuned = true;
          }
          launch_core_ranges(c->core_pp, S.nwl, zoff, zend, c->zkey2.as<unsigned long long>(), c->rng2.as<int2>(), s);
        }
      }
      if (!timed) CK(hipEventRecord(c->ev[2], s));
      timed = true;
      if (c->precision == FSG_FP32) {
        if (mnarrow > 0)
          launch_opacity32_narrow(S, nbn, c->rngn.as<int2>(), cb, c->wlc.as<double>(), c->meta.as<int4>(),
                                  (float *)c->lineop, s);
        if (has_wide && nbw > 0) {
          // the core zones on st2 (into lop2) while the far wings run on st
          CK(hipEventRecord(c->evf, s));
          CK(hipStreamWaitEvent(c->st2, c->evf, 0));
          if (!nfar) CK(hipEventRecord(c->evt[0], s));
          const int mf = launch_opacity32(S, ntiles, nbw, ng, rngw, ca, cb, dmaxw, lrec,
                                          c->wlc.as<double>(), c->meta.as<int4>(), c->foff.as<long long>(),
                                          c->flist.as<int>(), m, (float *)c->lineop, 0, c->far_ring, s);
          CK(hipEventRecord(c->evt[1], s));
          nfar++;
          if (mf) {
            if (!ncore) CK(hipEventRecord(c->evt[2], c->st2));
            launch_opacity32_core(S, c->core_pp, c->rng2.as<int2>(), c->zrec.as<ZoneRec>(), ca, cb, dmaxw, c->lop2,
                                  c->st2);
            CK(hipEventRecord(c->evt[3], c->st2));
            CK(hipEventRecord(c->evj, c->st2));
            ncore++;
            core_pending = true;
            c->lop2_dirty = true;
          }
        }
        if (has_he)
          launch_opacity32(S, ntiles, 1, ng, c->rng.as<int2>() + nbins, ca, cb, dmaxw, lrec, c->wlc.as<double>(),
                           c->meta.as<int4>(), nullptr, nullptr, 0, (float *)c->lineop, 1, c->far_ring, s);
      } else {
        if (has_main && nbins > 0)
          launch_opacity64(S, ntiles, nbins, ng, c->rng.as<int2>(), c->coef.as<Coef64>(), c->wlc.as<double>(),
                           c->meta.as<int4>(), (double *)c->lineop, s);
        if (has_he)
          launch_opacity64(S, ntiles, 1, ng, c->rng.as<int2>() + nbins, c->coef.as<Coef64>(),
                           c->wlc.as<double>(), c->meta.as<int4>(), (double *)c->lineop, s);
      }
      st.nblocks++;
    }
    if (core_pending) {  // lop2 is zero again afterwards (allocated zeroed)
      CK(hipStreamWaitEvent(s, c->evj, 0));
      launch_add_clear_f32((float *)c->lineop, c->lop2, nlop, s);
      c->lop2_dirty = false;
    }
  }
  if (!timed) CK(hipEventRecord(c->ev[2], s));
  CK(hipEventRecord(c->ev[3], s));
  CK(hipGetLastError());
  CK(hipMemcpyAsync(hu, c->u64, sizeof(P.hu), hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  st.nwide = (int64_t)hu[1];
  float ms1 = 0.f, ms2 = 0.f;
  CK(hipEventElapsedTime(&ms1, c->ev[0], c->ev[1]));
  CK(hipEventElapsedTime(&ms2, c->ev[2], c->ev[3]));
  st.ms_pass1 = ms1;
  st.ms_opacity = ms2;
  st.evals_core = (int64_t)hu[2];
  if (nfar) {
    CK(hipEventElapsedTime(&ms1, c->evt[0], c->evt[1]));
    st.ms_far = ms1;
  }
  if (ncore) {
    CK(hipEventElapsedTime(&ms1, c->evt[2], c->evt[3]));
    st.ms_core = ms1;
  }
  if (c->precision == FSG_FP32) {
    st.core_tile = 64 * c->core_pp;
    st.far_ring = c->far_ring;
  }
  c->have_rt = 0;
  c->have_glop = 0;
  if (stats) *stats = st;
  return 0;
}

}  // extern "C"

namespace {

// host copies of pass 1's results for the input lines: LinFlag, Ileft, Iright and, on
// request, the central line depths LinStr (src/synth.f90:1774-1822)
int lineop_outputs(fsg_ctx *c, fsg_dlines *dl, int list_type, float *linstr_out, int32_t *flags_out,
                   int32_t *ileft_out, int32_t *iright_out) {
  const int64_t n = dl->L.n;
  if (n == 0) return 0;
  size_t b = (size_t)n * sizeof(int32_t);
  if (linstr_out) {
    DBuf ls;
    if (ls.ensure(n * sizeof(float)) != hipSuccess)
      return fail("LineopCalc: out of device memory allocating the line strength array");
    launch_linstr(c->S, dl->L, list_type, c->flag.as<int>(), ls.as<float>(), c->st);
    int rc = 0;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->st) != hipSuccess ||
        hipMemcpy(linstr_out, ls.p, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail("line strength kernel failed");
    ls.release();
    if (rc) return rc;
  }
  if (flags_out && hipMemcpy(flags_out, c->flag.p, b, hipMemcpyDeviceToHost) != hipSuccess) return fail("D2H flags");
  if (ileft_out && hipMemcpy(ileft_out, c->il.p, b, hipMemcpyDeviceToHost) != hipSuccess) return fail("D2H ileft");
  if (iright_out && hipMemcpy(iright_out, c->ir.p, b, hipMemcpyDeviceToHost) != hipSuccess) return fail("D2H iright");
  return 0;
}

// MagLineOpac, the Faraday-Voigt table and the Stokes buffers, on first use
int ensure_mag_buffers(fsg_ctx *c) {
  if (c->maglineop) return 0;
  const size_t N = (size_t)c->S.ntau, Nw = (size_t)c->S.nwl, M = (size_t)c->S.nmu;
  void *tab = nullptr;
  if (alloc_zero(c, lop_elem(c) * 6 * N * Nw, &c->maglineop) ||
      alloc_zero(c, sizeof(float2) * (NVA + 1) * TAB_LD, &tab) ||
      alloc_zero(c, sizeof(double) * 4 * M * Nw, (void **)&c->spz) ||
      alloc_zero(c, sizeof(double) * 5 * M * 2 * N, (void **)&c->geo))
    return 1;
  c->ftab = (float2 *)tab;
  launch_fvoigt_table(c->ftab, c->S.anode, 8.0 / (double)(NVV - 1), c->st);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(c->st));
  return 0;
}

}  // namespace

extern "C" {

int fsg_lineop(fsg_ctx *c, int32_t list_type, const fsg_lines *lines, int64_t nlines_lim, float *linstr_out,
               int32_t *flags_out, int32_t *ileft_out, int32_t *iright_out, fsg_lineop_stats *stats) {
  g_err.clear();
  if (!c || !lines) return fail("fsg_lineop: NULL argument");
  if (c->S.batch) return fail("fsg_lineop: batch contexts take batch lists (fsg_lines_batch, fsg_lineop_dev)");
  fsg_dlines *dl = nullptr;
  if (fsg_lines_upload(c, list_type, lines, &dl)) return 1;
  int rc = fsg_lineop_dev(c, dl, nlines_lim, stats);
  if (rc == 0) rc = lineop_outputs(c, dl, list_type, linstr_out, flags_out, ileft_out, iright_out);
  fsg_lines_free(dl);
  return rc;
}

// The polarised line-opacity stage on a device-resident Zeeman list: pass 1 with the
// window search, width binning, coefficient fill and the six-profile accumulation into
// MagLineOpac.  MarkWideLines and PACK follow the reference's rules for the counters; the
// accumulated opacity does not depend on the block partition.
int fsg_maglineop_dev(fsg_ctx *c, fsg_dlines *dl, int64_t nlines_lim, const double *bfield, int32_t bframe,
                      fsg_lineop_stats *stats) {
  g_err.clear();
  if (!c || !dl || !bfield) return fail("fsg_maglineop_dev: NULL argument");
  if (c->S.batch) return fail("fsg_maglineop_dev: batch contexts have no polarised path");
  if (nlines_lim < 1) return fail("fsg_maglineop_dev: NlinesLim must be positive");
  if (bframe < 0 || bframe > 2) return fail("fsg_maglineop_dev: BFRAME must be 0, 1 or 2");
  const int64_t n = dl->L.n;
  if (n > 0 && !dl->zeeman) return fail("MagLineopCalc: line list has no Zeeman patterns");
  CK(hipSetDevice(c->dev));
  if (ensure_mag_buffers(c)) return 1;
  fsg_lineop_stats st{};
  st.nlines_in = n;
  st.nsegments = 1;
  if (n == 0) {
    if (stats) *stats = st;
    return 0;
  }
  const DevSetup &S = c->S;
  const int N = S.ntau;
  hipStream_t s = c->st;
  // Zeeman splitting per unit shift (src/synth.f90:1150-1154): |B| for BFRAME 0, the
  // first BFIELD value (the field modulus) otherwise
  const double bsplit = (bframe == 0)
                            ? kKz * std::sqrt(bfield[0] * bfield[0] + bfield[1] * bfield[1] + bfield[2] * bfield[2])
                            : kKz * std::fabs(bfield[0]);
  c->p1_id = 0;  // the pass-1 buffers now hold the Zeeman list's results
  CK(c->flag.ensure(n * sizeof(int)));
  CK(c->il.ensure(n * sizeof(int)));
  CK(c->ir.ensure(n * sizeof(int)));
  CK(c->ic.ensure(n * sizeof(int)));
  CK(c->key.ensure(n * sizeof(unsigned)));
  CK(c->key2.ensure(n * sizeof(unsigned)));
  CK(c->idx.ensure(n * sizeof(int)));
  CK(c->idx2.ensure(n * sizeof(int)));
  CK(hipMemsetAsync(c->bincnt, 0, sizeof(int) * 64, s));
  CK(hipMemsetAsync(c->binhmax, 0, sizeof(int) * 64, s));
  CK(hipMemsetAsync(c->u64, 0, sizeof(unsigned long long) * 4, s));
  CK(hipEventRecord(c->ev[0], s));
  launch_magpass1(S, dl->L, dl->Z, bsplit, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(), c->ic.as<int>(),
                  c->key.as<unsigned>(), c->bincnt, c->binhmax, c->batched, s);
  launch_count_evals(n, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(), N, c->u64, s);
  launch_iota(c->idx.as<int>(), n, s);
  size_t tbytes = 0;
  CK(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, c->key.as<unsigned>(), c->key2.as<unsigned>(),
                                        c->idx.as<int>(), c->idx2.as<int>(), (int)n, 0, 6, s));
  CK(c->tmp.ensure(tbytes));
  CK(hipcub::DeviceRadixSort::SortPairs(c->tmp.p, tbytes, c->key.as<unsigned>(), c->key2.as<unsigned>(),
                                        c->idx.as<int>(), c->idx2.as<int>(), (int)n, 0, 6, s));
  int hcnt[64], hhmax[64];
  unsigned long long hu[4];
  CK(hipMemcpyAsync(hcnt, c->bincnt, sizeof(hcnt), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(hhmax, c->binhmax, sizeof(hhmax), hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(hu, c->u64, sizeof(hu), hipMemcpyDeviceToHost, s));
  CK(hipEventRecord(c->ev[1], s));
  CK(hipStreamSynchronize(s));
  int64_t nret = 0;
  int nbins = 0;
  for (int b = 0; b < WEAK_KEY; b++) {
    nret += hcnt[b];
    if (hcnt[b] > 0) nbins = b + 1;
  }
  if (nbins > MAX_BINS) return fail("fsg_maglineop_dev: too many window-width bins");
  st.nlines_use = nret;
  st.evals = (int64_t)hu[0];
  // MarkWideLines (PACK happens when the whole list fits one block)
  const int64_t nstep = std::min<int64_t>(n, nlines_lim);
  const int pack = (nstep == n);
  CK(c->wlcl.ensure(n * sizeof(double)));
  CK(c->rflag.ensure(n * sizeof(int)));
  CK(c->pos.ensure(n * sizeof(int)));
  if (pack) {
    launch_retained_flag(n, c->flag.as<int>(), c->rflag.as<int>(), s);
    size_t sb = 0;
    CK(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, c->rflag.as<int>(), c->pos.as<int>(), (int)n, s));
    CK(c->tmp.ensure(sb));
    CK(hipcub::DeviceScan::ExclusiveSum(c->tmp.p, sb, c->rflag.as<int>(), c->pos.as<int>(), (int)n, s));
  }
  launch_wlc_list(n, dl->L.wl, S.wfirst, c->flag.as<int>(), c->pos.as<int>(), c->wlcl.as<double>(), pack, s);
  launch_mark_wide(S, n, dl->L.wl, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(), c->wlcl.as<double>(),
                   pack ? nret : n, nstep, c->u64 + 1, s);
  const int fp32 = c->precision == FSG_FP32;
  bool timed = false;
  if (nret > 0) {
    std::vector<int64_t> goff(nbins), gcnt(nbins);
    std::vector<int> ghmax(nbins);
    int64_t o = 0;
    for (int b = 0; b < nbins; b++) {
      goff[b] = o;
      gcnt[b] = hcnt[b];
      ghmax[b] = hhmax[b];
      o += hcnt[b];
    }
    const size_t rec = magcoef_size(fp32);
    const int64_t chunk = chunk_lines(c, nret, nlines_lim, S.ntp * rec + sizeof(double) + 2 * sizeof(int4), 1);
    CK(c->coef.ensure((size_t)chunk * S.ntp * rec));
    CK(c->wlc.ensure((size_t)chunk * sizeof(double)));
    CK(c->meta.ensure((size_t)chunk * sizeof(int4)));
    CK(c->comp.ensure((size_t)chunk * sizeof(int4)));
    const int ntiles = (S.nwl + MTILE - 1) / MTILE;
    CK(c->rng.ensure((size_t)ntiles * nbins * sizeof(int2)));
    std::vector<std::vector<int>> tables;  // staged for upload; alive until the final sync
    for (int64_t s0 = 0; s0 < nret; s0 += chunk) {
      const int64_t s1 = std::min(nret, s0 + chunk), m = s1 - s0;
      tables.emplace_back();
      chunk_tables(goff, gcnt, ghmax, s0, s1, tables.back());
      CK(hipMemcpyAsync(c->binoff, tables.back().data(), sizeof(int) * tables.back().size(), hipMemcpyHostToDevice,
                        s));
      launch_magfill(S, dl->L, dl->Z, m, c->idx2.as<int>() + s0, c->flag.as<int>(), c->il.as<int>(),
                     c->ir.as<int>(), c->ic.as<int>(), c->coef.p, fp32, c->wlc.as<double>(), c->meta.as<int4>(),
                     c->comp.as<int4>(), s);
      launch_tile_ranges(ntiles, MTILE, nbins, S.nwl, c->binoff, c->binoff + nbins, c->binoff + 2 * nbins,
                         c->meta.as<int4>(), c->rng.as<int2>(), s);
      if (!timed) CK(hipEventRecord(c->ev[2], s));
      timed = true;
      launch_magopacity(S, c->ftab, ntiles, nbins, nbins, c->rng.as<int2>(), c->coef.p, c->wlc.as<double>(),
                        c->meta.as<int4>(), c->comp.as<int4>(), dl->Z, c->maglineop, fp32, s);
      st.nblocks++;
    }
  }
  if (!timed) CK(hipEventRecord(c->ev[2], s));
  CK(hipEventRecord(c->ev[3], s));
  CK(hipGetLastError());
  CK(hipMemcpyAsync(hu, c->u64, sizeof(hu), hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  st.nwide = (int64_t)hu[1];
  float ms1 = 0.f, ms2 = 0.f;
  CK(hipEventElapsedTime(&ms1, c->ev[0], c->ev[1]));
  CK(hipEventElapsedTime(&ms2, c->ev[2], c->ev[3]));
  st.ms_pass1 = ms1;
  st.ms_opacity = ms2;
  c->have_rt = 0;
  c->have_glop = 0;
  if (stats) *stats = st;
  return 0;
}

int fsg_maglineop(fsg_ctx *c, const fsg_lines *lines, int64_t nlines_lim, const double *bfield, int32_t bframe,
                  float *linstr_out, int32_t *flags_out, int32_t *ileft_out, int32_t *iright_out,
                  fsg_lineop_stats *stats) {
  g_err.clear();
  if (!c || !lines || !bfield) return fail("fsg_maglineop: NULL argument");
  if (c->S.batch) return fail("fsg_maglineop: batch contexts have no polarised path");
  if (lines->n > 0 && (!lines->pi_off || !lines->si_off))
    return fail("MagLineopCalc: line list has no Zeeman patterns");
  fsg_dlines *dl = nullptr;
  if (fsg_lines_upload(c, FSG_LIST_MIXED, lines, &dl)) return 1;
  int rc = fsg_maglineop_dev(c, dl, nlines_lim, bfield, bframe, stats);
  if (rc == 0) rc = lineop_outputs(c, dl, FSG_LIST_MIXED, linstr_out, flags_out, ileft_out, iright_out);
  fsg_lines_free(dl);
  return rc;
}

int fsg_lines_set_gf(fsg_dlines *dl, const float *gf, int32_t on_device) {
  g_err.clear();
  if (!dl || (!gf && dl->L.n > 0)) return fail("fsg_lines_set_gf: NULL argument");
  if (dl->L.n == 0) return 0;
  fsg_ctx *c = dl->ctx;
  CK(hipSetDevice(c->dev));
  dl->gen++;  // windows and flags of the list change with gf
  CK(hipMemcpyAsync((void *)dl->L.gf, gf, sizeof(float) * (size_t)dl->L.n,
                    on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->st));
  CK(hipStreamSynchronize(c->st));
  return 0;
}

// Reverse mode of RadTranCalc + the flux sums, in wavelength chunks whose checkpoint fits
// the scratch budget.
}  // extern "C"

namespace {

// fsg_flux_vjp with dL/dSfL, dL/dSfC in host memory or (dev_in) in device memory
int flux_vjp_run(fsg_ctx *c, const double *gsfl, const double *gsfc, double *glop_out, double *gcont_out,
                 double *gtemp_out, bool dev_in, double *gcont_dev = nullptr) {
  g_err.clear();
  if (!c || !gsfl || !gsfc) return fail("fsg_flux_vjp: NULL argument");
  if (!c->have_rt || c->magnetic) return fail("fsg_flux_vjp: needs the scalar formal solution of fsg_radtran");
  if (c->S.sph) return fail("fsg_flux_vjp: the reverse mode covers plane-parallel models");
  CK(hipSetDevice(c->dev));
  const DevSetup &S = c->S;
  const size_t N = (size_t)S.ntau, Nw = (size_t)S.nwl;
  hipStream_t s = c->st;
  if (!c->glop && alloc_zero(c, sizeof(double) * N * Nw, (void **)&c->glop)) return 1;
  CK(c->vgcont.ensure(sizeof(double) * N * Nw));
  const size_t NB = (size_t)std::max(S.nbatch, 1);  // dL/dT rows: one per atmosphere of a batch
  CK(c->vout.ensure(sizeof(double) * (2 * Nw + NB * N)));
  double *dg = c->vout.as<double>(), *dgc = dg + Nw, *dgt = dgc + Nw;
  const hipMemcpyKind kin = dev_in ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  CK(hipMemcpyAsync(dg, gsfl, sizeof(double) * Nw, kin, s));
  CK(hipMemcpyAsync(dgc, gsfc, sizeof(double) * Nw, kin, s));
  CK(hipMemsetAsync(dgt, 0, sizeof(double) * NB * N, s));
  const size_t per = rt_vjp_scratch_per_point(S);
  // chunks of whole wavefronts (a batch's atmospheres start at multiples of 64 points) within
  // 8 GiB of checkpoint scratch: 1.2 million points of 72 depths and 7 angles in one launch,
  // so that the wavefronts of the last round do not leave most of the chip idle
  const int64_t chunk = std::min<int64_t>((int64_t)Nw, std::max<int64_t>(64, (int64_t)((size_t)8 << 30) / (int64_t)per) & ~(int64_t)63);
  CK(c->vscr.ensure(per * (size_t)chunk));
  for (int64_t p0 = 0; p0 < (int64_t)Nw; p0 += chunk) {
    const int np = (int)std::min<int64_t>(chunk, (int64_t)Nw - p0);
    launch_rt_vjp(S, c->lineop, c->precision == FSG_FP32, dg, dgc, (int)p0, np, c->vscr.as<double>(), c->glop,
                  c->vgcont.as<double>(), dgt, s);
  }
  CK(hipGetLastError());
  for (int which = 0; which < 2; which++) {  // [ntau][nwl] -> (Ntau, Nwl) layout
    double *out = which ? gcont_out : glop_out;
    if (!out) continue;
    std::vector<double> h(N * Nw);
    CK(hipMemcpyAsync(h.data(), which ? c->vgcont.p : (void *)c->glop, sizeof(double) * h.size(),
                      hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    for (size_t k = 0; k < N; k++)
      for (size_t iw = 0; iw < Nw; iw++) out[iw * N + k] = h[k * Nw + iw];
  }
  if (gcont_dev) CK(hipMemcpyAsync(gcont_dev, c->vgcont.p, sizeof(double) * N * Nw, hipMemcpyDeviceToDevice, s));
  if (gtemp_out) CK(hipMemcpyAsync(gtemp_out, dgt, sizeof(double) * NB * N, hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  c->have_glop = 1;
  return 0;
}

}  // namespace

extern "C" {

int fsg_flux_vjp(fsg_ctx *c, const double *gsfl, const double *gsfc, double *glop_out, double *gcont_out,
                 double *gtemp_out) {
  return flux_vjp_run(c, gsfl, gsfc, glop_out, gcont_out, gtemp_out, false);
}

int fsg_flux_vjp_dev(fsg_ctx *c, const double *gsfl_dev, const double *gsfc_dev, double *gcont_dev,
                     double *gtemp_out) {
  return flux_vjp_run(c, gsfl_dev, gsfc_dev, nullptr, nullptr, gtemp_out, true, gcont_dev);
}

// Reverse mode of LineopCalc for one list from the resident dL/dLineOpac: pass 1 again
// (same windows and flags), then per chunk of coefficient records the per-(line, depth)
// sums of dL/dLineOpac x contribution and the temperature derivative of the contributions.
}  // extern "C"

namespace {

// fsg_lineop_vjp with dL/d log gf written to host or (dev_out) device memory
int lineop_vjp_run(fsg_ctx *c, fsg_dlines *dl, int64_t nlines_lim, double *gloggf_out, double *gfpf_inout,
                   double *gtemp_inout, bool dev_out) {
  g_err.clear();
  if (!c || !dl) return fail("fsg_lineop_vjp: NULL argument");
  if (nlines_lim < 1) return fail("fsg_lineop_vjp: NlinesLim must be positive");
  if (!c->have_glop) return fail("fsg_lineop_vjp: call fsg_flux_vjp first");
  CK(hipSetDevice(c->dev));
  const DevSetup &S = c->S;
  const int64_t n = dl->L.n;
  const size_t N = (size_t)S.ntau;
  hipStream_t s = c->st;
  if ((c->S.batch != nullptr) != (dl->L.atm != nullptr))
    return fail("fsg_lineop_vjp: batch contexts take batch lists (fsg_lines_batch) and only they do");
  // a batch has one dL/d ln FractionPF block and one dL/dT row per atmosphere
  const size_t NB = (size_t)std::max(S.nbatch, 1);
  const size_t nfpf = NB * N * (size_t)S.pf_ld;
  CK(c->vout.ensure(sizeof(double) * ((size_t)std::max<int64_t>(n, 1) + nfpf + NB * N)));
  double *dgf = c->vout.as<double>(), *dfpf = dgf + std::max<int64_t>(n, 1), *dgt = dfpf + nfpf;
  CK(hipMemsetAsync(c->vout.p, 0, sizeof(double) * ((size_t)std::max<int64_t>(n, 1) + nfpf + NB * N), s));
  if (n > 0) {
    LinePrep P;
    if (scalar_pass1_cached(c, dl, P)) return 1;
    if (P.nret > 0) {
      const int ng = P.nbins + 1;
      std::vector<int64_t> goff, gcnt;
      std::vector<int> ghmax;
      group_tables(P, goff, gcnt, ghmax);
      const int nwv = S.ntp / DPW;
      const size_t per = S.ntp * (sizeof(Coef64) + sizeof(CoefV) + sizeof(CoefR) + sizeof(double)) + sizeof(double) +
                         sizeof(int4) + sizeof(int) + sizeof(VjpWin) + nwv * sizeof(double);
      const int64_t chunk = chunk_lines(c, P.nret, nlines_lim, per, 0);
      // the chunk's Coef64 records, its far-wing series records, its general records, the
      // window records and the wavefronts' largest thresholds
      CK(c->coef.ensure((size_t)chunk * (S.ntp * (sizeof(Coef64) + sizeof(CoefV) + sizeof(CoefR)) + sizeof(VjpWin) +
                                         nwv * sizeof(double))));
      CoefV *coefv = reinterpret_cast<CoefV *>(c->coef.as<char>() + (size_t)chunk * S.ntp * sizeof(Coef64));
      CoefR *coefr = reinterpret_cast<CoefR *>(coefv + (size_t)chunk * S.ntp);
      VjpWin *vwin = reinterpret_cast<VjpWin *>(coefr + (size_t)chunk * S.ntp);
      double *dmaxv = reinterpret_cast<double *>(vwin + chunk);
      CK(c->wlc.ensure((size_t)chunk * sizeof(double)));
      CK(c->meta.ensure((size_t)chunk * sizeof(int4)));
      CK(c->vsl.ensure((size_t)chunk * S.ntp * sizeof(double)));
      CK(c->vslot.ensure((size_t)chunk * sizeof(int)));
      const int ntiles = (S.nwl + VJP_TILE - 1) / VJP_TILE;
      CK(c->rng.ensure((size_t)ntiles * ng * sizeof(int2)));
      std::vector<std::vector<int>> tables;
      for (int64_t s0 = 0; s0 < P.nret; s0 += chunk) {
        const int64_t s1 = std::min(P.nret, s0 + chunk), m = s1 - s0;
        tables.emplace_back();
        chunk_tables(goff, gcnt, ghmax, s0, s1, tables.back());
        CK(hipMemcpyAsync(c->binoff, tables.back().data(), sizeof(int) * tables.back().size(),
                          hipMemcpyHostToDevice, s));
        const int *order = c->idx2.as<int>() + s0;
        launch_fill64_vjp(S, dl->L, dl->list_type, m, order, c->flag.as<int>(), c->il.as<int>(), c->ir.as<int>(),
                          c->ic.as<int>(), c->coef.as<Coef64>(), coefv, c->wlc.as<double>(), c->meta.as<int4>(),
                          c->vslot.as<int>(), vwin, dmaxv, coefr, s);
        launch_tile_ranges(ntiles, VJP_TILE, ng, S.nwl, c->binoff, c->binoff + ng, c->binoff + 2 * ng,
                           c->meta.as<int4>(), c->rng.as<int2>(), s);
        CK(hipMemsetAsync(c->vsl.p, 0, (size_t)m * S.ntp * sizeof(double), s));
        launch_lineop_vjp(S, ntiles, ng, ng, c->rng.as<int2>(), c->coef.as<Coef64>(), coefv, coefr,
                          c->wlc.as<double>(), c->meta.as<int4>(), vwin, dmaxv, c->glop, c->vsl.as<double>(), dgt, s);
        launch_lineop_vjp_scatter(S, m, order, dl->L.atm, c->vslot.as<int>(), c->vsl.as<double>(), dgf, dfpf, s);
      }
      CK(hipGetLastError());
    }
  }
  // the read-backs are ordered on the library stream after the clear and the kernels (the
  // stream is non-blocking, so a null-stream copy would not wait for them)
  std::vector<double> h(nfpf + NB * N);
  if (gloggf_out && n > 0)
    CK(hipMemcpyAsync(gloggf_out, dgf, sizeof(double) * (size_t)n,
                      dev_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  CK(hipMemcpyAsync(h.data(), dfpf, sizeof(double) * (nfpf + NB * N), hipMemcpyDeviceToHost, s));
  CK(hipStreamSynchronize(s));
  if (gfpf_inout)
    for (size_t i = 0; i < nfpf; i++) gfpf_inout[i] += h[i];
  if (gtemp_inout)
    for (size_t k = 0; k < NB * N; k++) gtemp_inout[k] += h[nfpf + k];
  return 0;
}

}  // namespace

extern "C" {

int fsg_lineop_vjp(fsg_ctx *c, fsg_dlines *dl, int64_t nlines_lim, double *gloggf_out, double *gfpf_inout,
                   double *gtemp_inout) {
  return lineop_vjp_run(c, dl, nlines_lim, gloggf_out, gfpf_inout, gtemp_inout, false);
}

int fsg_lineop_vjp_dev(fsg_ctx *c, fsg_dlines *dl, int64_t nlines_lim, double *gloggf_dev, double *gfpf_inout,
                       double *gtemp_inout) {
  return lineop_vjp_run(c, dl, nlines_lim, gloggf_dev, gfpf_inout, gtemp_inout, true);
}

int fsg_reset_lineop(fsg_ctx *c) {
  if (!c) return fail("fsg_reset_lineop: NULL context");
  CK(hipSetDevice(c->dev));
  CK(hipMemsetAsync(c->lineop, 0, lop_elem(c) * (size_t)c->S.ntau * c->S.nwl, c->st));
  if (c->maglineop) CK(hipMemsetAsync(c->maglineop, 0, lop_elem(c) * 6 * (size_t)c->S.ntau * c->S.nwl, c->st));
  CK(hipStreamSynchronize(c->st));
  c->have_rt = 0;
  c->have_glop = 0;
  return 0;
}

int fsg_get_lineop(fsg_ctx *c, double *out) {
  g_err.clear();
  if (!c || !out) return fail("fsg_get_lineop: NULL argument");
  CK(hipSetDevice(c->dev));
  const int N = c->S.ntau, Nw = c->S.nwl;
  size_t tot = (size_t)N * Nw;
  if (c->precision == FSG_FP32) {
    std::vector<float> h(tot);
    CK(hipMemcpyAsync(h.data(), c->lineop, tot * sizeof(float), hipMemcpyDeviceToHost, c->st));
    CK(hipStreamSynchronize(c->st));
    for (int k = 0; k < N; k++)
      for (size_t iw = 0; iw < (size_t)Nw; iw++) out[iw * N + k] = (double)h[(size_t)k * Nw + iw];
  } else {
    std::vector<double> h(tot);
    CK(hipMemcpyAsync(h.data(), c->lineop, tot * sizeof(double), hipMemcpyDeviceToHost, c->st));
    CK(hipStreamSynchronize(c->st));
    for (int k = 0; k < N; k++)
      for (size_t iw = 0; iw < (size_t)Nw; iw++) out[iw * N + k] = h[(size_t)k * Nw + iw];
  }
  return 0;
}

int fsg_set_lineop(fsg_ctx *c, const double *in) {
  g_err.clear();
  if (!c || !in) return fail("fsg_set_lineop: NULL argument");
  CK(hipSetDevice(c->dev));
  const int N = c->S.ntau, Nw = c->S.nwl;
  size_t tot = (size_t)N * Nw;
  if (c->precision == FSG_FP32) {
    std::vector<float> h(tot);
    for (int k = 0; k < N; k++)
      for (size_t iw = 0; iw < (size_t)Nw; iw++) h[(size_t)k * Nw + iw] = (float)in[iw * N + k];
    CK(hipMemcpyAsync(c->lineop, h.data(), tot * sizeof(float), hipMemcpyHostToDevice, c->st));
    CK(hipStreamSynchronize(c->st));
  } else {
    std::vector<double> h(tot);
    for (int k = 0; k < N; k++)
      for (size_t iw = 0; iw < (size_t)Nw; iw++) h[(size_t)k * Nw + iw] = in[iw * N + k];
    CK(hipMemcpyAsync(c->lineop, h.data(), tot * sizeof(double), hipMemcpyHostToDevice, c->st));
    CK(hipStreamSynchronize(c->st));
  }
  c->have_rt = 0;
  c->have_glop = 0;
  return 0;
}

int fsg_get_maglineop(fsg_ctx *c, double *out) {
  g_err.clear();
  if (!c || !out) return fail("fsg_get_maglineop: NULL argument");
  CK(hipSetDevice(c->dev));
  const size_t N = (size_t)c->S.ntau, Nw = (size_t)c->S.nwl, tot = 6 * N * Nw;
  if (!c->maglineop) {
    std::memset(out, 0, sizeof(double) * tot);
    return 0;
  }
  // device [6][ntau][nwl] -> MagLineOpac(6, Ntau, Nwl)
  std::vector<double> h(tot);
  if (c->precision == FSG_FP32) {
    std::vector<float> f(tot);
    CK(hipMemcpyAsync(f.data(), c->maglineop, tot * sizeof(float), hipMemcpyDeviceToHost, c->st));
    CK(hipStreamSynchronize(c->st));
    for (size_t i = 0; i < tot; i++) h[i] = (double)f[i];
  } else {
    CK(hipMemcpyAsync(h.data(), c->maglineop, tot * sizeof(double), hipMemcpyDeviceToHost, c->st));
    CK(hipStreamSynchronize(c->st));
  }
  for (size_t q = 0; q < 6; q++)
    for (size_t k = 0; k < N; k++)
      for (size_t iw = 0; iw < Nw; iw++) out[(iw * N + k) * 6 + q] = h[(q * N + k) * Nw + iw];
  return 0;
}

static int ensure_rt_buffers(fsg_ctx *c) {
  if (c->spl) return 0;
  const size_t Nw = (size_t)c->S.nwl, M = (size_t)c->S.nmu;
  if (alloc_zero(c, sizeof(double) * M * Nw, (void **)&c->spl) ||
      alloc_zero(c, sizeof(double) * M * Nw, (void **)&c->spc) ||
      alloc_zero(c, sizeof(double) * Nw, (void **)&c->sfl) || alloc_zero(c, sizeof(double) * Nw, (void **)&c->sfc))
    return 1;
  return 0;
}

static int copy_spec_out(fsg_ctx *c, const double *dev, double *host) {
  const size_t Nw = (size_t)c->S.nwl, M = (size_t)c->S.nmu;
  std::vector<double> h(M * Nw);
  CK(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
  CK(hipStreamSynchronize(c->st));
  for (size_t m = 0; m < M; m++)
    for (size_t iw = 0; iw < Nw; iw++) host[iw * M + m] = h[m * Nw + iw];
  return 0;
}

int fsg_radtran(fsg_ctx *c, double *spl_out, double *spc_out) {
  g_err.clear();
  if (!c) return fail("fsg_radtran: NULL context");
  CK(hipSetDevice(c->dev));
  if (ensure_rt_buffers(c)) return 1;
  if (c->precision == FSG_FP32)
    launch_radtran32(c->S, (const float *)c->lineop, c->spl, c->spc, c->sfl, c->sfc, c->st);
  else
    launch_radtran64(c->S, (const double *)c->lineop, c->spl, c->spc, c->sfl, c->sfc, c->st);
  CK(hipGetLastError());
  c->have_rt = 1;
  c->magnetic = 0;
  if (spl_out && copy_spec_out(c, c->spl, spl_out)) return 1;
  if (spc_out && copy_spec_out(c, c->spc, spc_out)) return 1;
  CK(hipStreamSynchronize(c->st));
  return 0;
}

int fsg_magradtran(fsg_ctx *c, const double *bfield, int32_t nbfield, int32_t bframe, double *spz_out,
                   double *spc_out) {
  g_err.clear();
  if (!c || !bfield) return fail("fsg_magradtran: NULL argument");
  if (c->S.batch) return fail("fsg_magradtran: batch contexts have no polarised path");
  const int M = c->S.nmu;
  if (bframe < 0 || bframe > 2) return fail("fsg_magradtran: BFRAME must be 0, 1 or 2");
  if (bframe == 2 && nbfield < 1 + 2 * M) return fail("BFRAME 2 needs 1+2*Nmu BFIELD values");
  if (bframe != 2 && nbfield < 3) return fail("fsg_magradtran: BFIELD needs 3 values");
  CK(hipSetDevice(c->dev));
  if (ensure_rt_buffers(c) || ensure_mag_buffers(c)) return 1;
  // field geometry (src/synth.f90:1935-1994, 2017-2090): sin^2 and cos^2 of the
  // inclination gamma, its cosine, and sin/cos of twice the azimuth chi; per mu in
  // plane-parallel geometry, per ray point on spherical rays
  std::vector<double> xmu(M);
  CK(hipMemcpy(xmu.data(), c->S.xmu, sizeof(double) * M, hipMemcpyDeviceToHost));
  const double d2r = kPi / 180.0;
  auto angles = [&](int im, double coslon, double *gm) {
    double Sphi, Cphi, Spsi, Cpsi;
    if (bframe == 1) {
      Sphi = std::sin(bfield[1] * d2r);
      Cphi = std::cos(bfield[1] * d2r);
      Spsi = std::sin(bfield[2] * d2r);
      Cpsi = std::cos(bfield[2] * d2r);
    } else if (bframe == 2) {
      Sphi = std::sin(bfield[2 * im + 1] * d2r);
      Cphi = std::cos(bfield[2 * im + 1] * d2r);
      Spsi = std::sin(bfield[2 * im + 2] * d2r);
      Cpsi = std::cos(bfield[2 * im + 2] * d2r);
    } else {
      // stellar frame (Br, Bm, Bp) projected with the cosine between the line of sight
      // and the local radial direction
      const double Fldmod = std::sqrt(bfield[0] * bfield[0] + bfield[1] * bfield[1] + bfield[2] * bfield[2]);
      if (Fldmod > 0.0) {
        const double Coslon = coslon, Sinlon = std::sqrt(1.0 - Coslon * Coslon);
        const double Bx = bfield[0] * Sinlon + bfield[2] * Coslon, By = bfield[1];
        const double Bz = bfield[0] * Coslon - bfield[2] * Sinlon;
        const double Bt = std::sqrt(Bx * Bx + By * By);
        Sphi = Bt / Fldmod;
        Cphi = Bz / Fldmod;
        if (Bt > 0.0) {
          Spsi = By / Bt;
          Cpsi = Bx / Bt;
        } else {
          Spsi = 0.0;
          Cpsi = 1.0;
        }
      } else {
        Sphi = 0.0;
        Cphi = 1.0;
        Spsi = 0.0;
        Cpsi = 1.0;
      }
    }
    gm[0] = Sphi * Sphi;
    gm[1] = Cphi * Cphi;
    gm[2] = Cphi;
    gm[3] = 2.0 * Spsi * Cpsi;
    gm[4] = Cpsi * Cpsi - Spsi * Spsi;
  };
  std::vector<double> g;
  if (c->S.sph) {
    const int L = 2 * c->S.ntau;
    g.assign((size_t)5 * M * L, 0.0);
    for (int im = 0; im < M; im++)
      for (int k = 0; k < c->ntau_mu[im]; k++)
        angles(im, c->muloc[(size_t)im * L + k], g.data() + ((size_t)im * L + k) * 5);
  } else {
    g.assign((size_t)5 * M, 0.0);
    for (int im = 0; im < M; im++) angles(im, xmu[im], g.data() + (size_t)5 * im);
  }
  CK(hipMemcpyAsync(c->geo, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, c->st));
  // continuum intensities and their flux (the scalar solution of LineOpac + ContOpac is
  // computed alongside and replaced by the Stokes I flux below)
  const int fp32 = c->precision == FSG_FP32;
  if (fp32)
    launch_radtran32(c->S, (const float *)c->lineop, c->spl, c->spc, c->sfl, c->sfc, c->st);
  else
    launch_radtran64(c->S, (const double *)c->lineop, c->spl, c->spc, c->sfl, c->sfc, c->st);
  launch_stokes(c->S, c->lineop, c->maglineop, fp32, c->geo, c->spz, c->sfl, c->st);
  CK(hipGetLastError());
  CK(hipStreamSynchronize(c->st));
  c->have_rt = 1;
  c->magnetic = 1;
  if (spz_out) {
    const size_t Nw = (size_t)c->S.nwl;
    std::vector<double> h((size_t)4 * M * Nw);
    CK(hipMemcpy(h.data(), c->spz, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < 4; q++)
      for (size_t m = 0; m < (size_t)M; m++)
        for (size_t iw = 0; iw < Nw; iw++) spz_out[(iw * 4 + q) * M + m] = h[(q * M + m) * Nw + iw];
  }
  if (spc_out && copy_spec_out(c, c->spc, spc_out)) return 1;
  return 0;
}

int fsg_get_spectra(fsg_ctx *c, double *spl_out, double *spz_out, double *spc_out) {
  g_err.clear();
  if (!c) return fail("fsg_get_spectra: NULL context");
  if (!c->have_rt) return fail("fsg_get_spectra: no formal solution computed since the last opacity change");
  if (spl_out && c->magnetic) return fail("fsg_get_spectra: the last formal solution was polarised (read SpZ)");
  if (spz_out && !c->magnetic) return fail("fsg_get_spectra: the last formal solution was not polarised");
  CK(hipSetDevice(c->dev));
  if (spl_out && copy_spec_out(c, c->spl, spl_out)) return 1;
  if (spz_out) {
    const size_t Nw = (size_t)c->S.nwl, M = (size_t)c->S.nmu;
    std::vector<double> h(4 * M * Nw);
    CK(hipMemcpyAsync(h.data(), c->spz, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    CK(hipStreamSynchronize(c->st));
    for (size_t q = 0; q < 4; q++)
      for (size_t m = 0; m < M; m++)
        for (size_t iw = 0; iw < Nw; iw++) spz_out[(iw * 4 + q) * M + m] = h[(q * M + m) * Nw + iw];
  }
  if (spc_out && copy_spec_out(c, c->spc, spc_out)) return 1;
  return 0;
}

int fsg_flux_dev(fsg_ctx *c, double *sfl_dev, double *sfc_dev) {
  g_err.clear();
  if (!c) return fail("fsg_flux_dev: NULL context");
  if (!c->have_rt) return fail("fsg_flux: no formal solution computed since the last opacity change");
  CK(hipSetDevice(c->dev));
  const size_t Nw = (size_t)c->S.nwl;
  if (sfl_dev) CK(hipMemcpyAsync(sfl_dev, c->sfl, Nw * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  if (sfc_dev) CK(hipMemcpyAsync(sfc_dev, c->sfc, Nw * sizeof(double), hipMemcpyDeviceToDevice, c->st));
  CK(hipStreamSynchronize(c->st));
  return 0;
}

int fsg_flux(fsg_ctx *c, double *sfl_out, double *sfc_out) {
  g_err.clear();
  if (!c) return fail("fsg_flux: NULL context");
  if (!c->have_rt) return fail("fsg_flux: no formal solution computed since the last opacity change");
  CK(hipSetDevice(c->dev));
  const size_t Nw = (size_t)c->S.nwl;
  if (sfl_out) CK(hipMemcpyAsync(sfl_out, c->sfl, Nw * sizeof(double), hipMemcpyDeviceToHost, c->st));
  if (sfc_out) CK(hipMemcpyAsync(sfc_out, c->sfc, Nw * sizeof(double), hipMemcpyDeviceToHost, c->st));
  CK(hipStreamSynchronize(c->st));
  return 0;
}

}  // extern "C"

namespace {

// IntegConvol on the device.  The work areas -- the rotated / macroturbulent annulus spectra
// (a, b: [2][nmu][nwl]; IntegConvol(1) broadens into a), the kernel taps, the row table and
// the flux weights -- come either from the context's grow-only buffers (fsg_integ_convol)
// or, for fsg_integ_convol_noalloc, from the static array `scratch` of this device.  Every
// part of a work area that a kernel reads was written earlier in the same call, so nothing
// needs clearing between calls.
int integ_convol_run(fsg_ctx *c, int32_t cimode, double vsini, double vmacro, double vgau, bool noalloc,
                     const char *name) {
  g_err.clear();
  const std::string nm(name);
  if (!c) return fail(nm + ": NULL context");
  if (c->S.batch) return fail(nm + ": convolve the atmospheres of a batch one by one");
  if (!c->have_rt) return fail(nm + ": no formal solution computed since the last opacity change");
  if (vsini < 0.0 || vmacro < 0.0 || vgau < 0.0) return fail("Fastspec: invalid value(s) of convolution parameters");
  if (cimode != 1 && cimode != 2) return fail(nm + ": CImode must be 1 or 2");
  CK(hipSetDevice(c->dev));
  const DevSetup &S = c->S;
  const int M = S.nmu, Nw = S.nwl;
  hipStream_t st = c->st;
  // kernel taps: tap 0 is the identity (a copy), then the Gaussian, then DiskInteg's
  std::vector<double> kern{1.0};
  std::vector<double> gk;
  int gh = 0, goff = 0;
  if (vgau > 0.0) {
    convol_gauss_kernel(Nw, vgau, S.wmode, S.wfirst, c->wlast, S.dstep, gk, gh);
    goff = (int)kern.size();
    kern.insert(kern.end(), gk.begin(), gk.end());
  }
  const int ng = (int)gk.size();
  std::vector<double> dk;
  std::vector<int> roff, rn, moff;
  int mn = 0, base = 0;
  if (cimode == 2) {
    ConvolSetup cs{Nw, M, S.wmode, S.wfirst, c->wlast, S.dstep, vsini, vmacro, c->hxmu.data(), c->hwmu.data()};
    diskinteg_kernels(cs, dk, roff, rn, moff, mn);
    base = (int)kern.size();
    kern.insert(kern.end(), dk.begin(), dk.end());
  }
  // rows of the spectra: Stokes I..V blocks after a polarised solution, else SpL; then SpC
  const int nlrow = c->magnetic ? 4 * M : M;
  double *lbase = c->magnetic ? c->spz : c->spl;
  const size_t na = cimode == 1 ? (vgau > 0.0 ? (size_t)(nlrow + M) * Nw : 0) : (size_t)2 * M * Nw;
  const size_t nbw = cimode == 2 ? (size_t)2 * M * Nw : 0;
  const size_t nrow = cimode == 1 ? (vgau > 0.0 ? (size_t)(nlrow + M) : 0)
                                  : (size_t)2 * M + (vmacro > 0.0 ? 2 * M : 0) + (vgau > 0.0 ? 2 : 0);
  double *a = nullptr, *b = nullptr, *dkern = nullptr, *dw = nullptr;
  ConvRow *drows = nullptr;
  if (!noalloc) {
    if (na) CK(c->cv_a.ensure(sizeof(double) * na));
    if (nbw) CK(c->cv_b.ensure(sizeof(double) * nbw));
    CK(c->cv_kern.ensure(sizeof(double) * kern.size()));
    CK(c->cv_rows.ensure(sizeof(ConvRow) * std::max<size_t>(nrow, 1)));
    CK(c->cv_w.ensure(sizeof(double) * M));
    a = c->cv_a.as<double>();
    b = c->cv_b.as<double>();
    dkern = c->cv_kern.as<double>();
    drows = c->cv_rows.as<ConvRow>();
    dw = c->cv_w.as<double>();
  } else {
    char *sp = nullptr;
    if (static_scratch(c->dev, (void **)&sp)) return 1;
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t oa = 0, ob = oa + up(sizeof(double) * na), ok = ob + up(sizeof(double) * nbw),
                 orw = ok + up(sizeof(double) * kern.size()), ow = orw + up(sizeof(ConvRow) * nrow),
                 total = ow + up(sizeof(double) * M);
    if (total > kStaticScratchBytes) {
      char msg[256];
      std::snprintf(msg, sizeof msg,
                    "%s: this convolution needs %zu bytes of work space (%d points, %d rows of spectra), more "
                    "than the %zu of the static buffer; use fsg_integ_convol",
                    name, total, Nw, nlrow + M, (size_t)kStaticScratchBytes);
      return fail(msg);
    }
    a = (double *)(sp + oa);
    b = (double *)(sp + ob);
    dkern = (double *)(sp + ok);
    drows = (ConvRow *)(sp + orw);
    dw = (double *)(sp + ow);
  }
  std::vector<ConvRow> rows;
  struct Launch {
    size_t first;
    int n, maxk;
  };
  std::vector<Launch> launches;
  auto add_launch = [&](size_t first) {
    int maxk = 1;
    for (size_t i = first; i < rows.size(); i++) maxk = std::max(maxk, rows[i].nk);
    launches.push_back(Launch{first, (int)(rows.size() - first), maxk});
  };
  if (cimode == 1) {
    // IntegConvol(1): every intensity spectrum broadened into a, then copied back
    if (vgau > 0.0) {
      for (int r = 0; r < nlrow; r++)
        rows.push_back(ConvRow{lbase + (size_t)r * Nw, a + (size_t)r * Nw, goff, ng, gh});
      for (int m = 0; m < M; m++)
        rows.push_back(ConvRow{c->spc + (size_t)m * Nw, a + (size_t)(nlrow + m) * Nw, goff, ng, gh});
      add_launch(0);
    }
  } else {
    // DiskInteg: rotation, macroturbulence, the weighted sum over mu, Gaussian broadening;
    // annulus spectra in a: [2][M][Nw] (line, continuum), rotated or copied
    for (int t = 0; t < 2; t++)
      for (int m = 0; m < M; m++) {
        const double *in = t == 0 ? lbase + (size_t)m * Nw : c->spc + (size_t)m * Nw;
        double *out = a + ((size_t)t * M + m) * Nw;
        if (roff[m] >= 0) rows.push_back(ConvRow{in, out, base + roff[m], rn[m], (int)((rn[m] - 1.0) / 2.0)});
        else rows.push_back(ConvRow{in, out, 0, 1, 0});
      }
    add_launch(0);
    double *fin = a;
    if (vmacro > 0.0) {
      const size_t first = rows.size();
      for (int t = 0; t < 2; t++)
        for (int m = 0; m < M; m++)
          rows.push_back(ConvRow{a + ((size_t)t * M + m) * Nw, b + ((size_t)t * M + m) * Nw, base + moff[m], mn,
                                 (int)((mn - 1.0) / 2.0)});
      add_launch(first);
      fin = b;
    }
    if (vgau > 0.0) {  // the flux sums land in the free buffer, broadened into SfL, SfC
      double *fr = fin == a ? b : a;
      const size_t first = rows.size();
      rows.push_back(ConvRow{fr, c->sfl, goff, ng, gh});
      rows.push_back(ConvRow{fr + Nw, c->sfc, goff, ng, gh});
      add_launch(first);
    }
  }
  if (rows.size() != nrow) return fail(nm + ": internal error (row count)");
  CK(hipMemcpyAsync(dkern, kern.data(), sizeof(double) * kern.size(), hipMemcpyHostToDevice, st));
  if (!rows.empty()) CK(hipMemcpyAsync(drows, rows.data(), sizeof(ConvRow) * rows.size(), hipMemcpyHostToDevice, st));
  CK(hipMemcpyAsync(dw, c->hwmu.data(), sizeof(double) * M, hipMemcpyHostToDevice, st));
  if (cimode == 1) {
    if (!launches.empty()) {
      launch_conv_rows(Nw, launches[0].n, launches[0].maxk, drows, dkern, st);
      CK(hipMemcpyAsync(lbase, a, sizeof(double) * (size_t)nlrow * Nw, hipMemcpyDeviceToDevice, st));
      CK(hipMemcpyAsync(c->spc, a + (size_t)nlrow * Nw, sizeof(double) * (size_t)M * Nw, hipMemcpyDeviceToDevice,
                        st));
    }
  } else {
    launch_conv_rows(Nw, launches[0].n, launches[0].maxk, drows, dkern, st);
    double *fin = a;
    size_t li = 1;
    if (vmacro > 0.0) {
      launch_conv_rows(Nw, launches[li].n, launches[li].maxk, drows + launches[li].first, dkern, st);
      li++;
      fin = b;
    }
    if (vgau > 0.0) {
      double *fr = fin == a ? b : a;
      launch_mu_sum(Nw, M, fin, fin + (size_t)M * Nw, dw, fr, fr + Nw, st);
      launch_conv_rows(Nw, launches[li].n, launches[li].maxk, drows + launches[li].first, dkern, st);
    } else {
      launch_mu_sum(Nw, M, fin, fin + (size_t)M * Nw, dw, c->sfl, c->sfc, st);
    }
  }
  CK(hipGetLastError());
  CK(hipStreamSynchronize(st));  // the host tables above stay alive until here
  return 0;
}

}  // namespace

extern "C" {

int fsg_integ_convol(fsg_ctx *c, int32_t cimode, double vsini, double vmacro, double vgau) {
  return integ_convol_run(c, cimode, vsini, vmacro, vgau, false, "fsg_integ_convol");
}

int fsg_integ_convol_noalloc(fsg_ctx *c, int32_t cimode, double vsini, double vmacro, double vgau) {
  return integ_convol_run(c, cimode, vsini, vmacro, vgau, true, "fsg_integ_convol_noalloc");
}

int fsg_device_ptr(fsg_ctx *c, int32_t which, void **ptr_out) {
  if (!c || !ptr_out) return fail("fsg_device_ptr: NULL argument");
  if (which != 0 && ensure_rt_buffers(c)) return 1;
  if (which >= 5 && ensure_mag_buffers(c)) return 1;
  switch (which) {
    case 0: *ptr_out = c->lineop; break;
    case 1: *ptr_out = c->spl; break;
    case 2: *ptr_out = c->spc; break;
    case 3: *ptr_out = c->sfl; break;
    case 4: *ptr_out = c->sfc; break;
    case 5: *ptr_out = c->maglineop; break;
    case 6: *ptr_out = c->spz; break;
    default: return fail("fsg_device_ptr: unknown array");
  }
  return 0;
}

int fsg_stream(fsg_ctx *c, void **stream_out) {
  if (!c || !stream_out) return fail("fsg_stream: NULL argument");
  *stream_out = (void *)c->st;
  return 0;
}

}  // extern "C"
