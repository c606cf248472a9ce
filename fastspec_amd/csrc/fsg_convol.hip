// fsg_convol.hip -- convolution and disk integration of the intensity spectra on the device.
//
// Reference routines restated here (file:line in the reference tree):
//   IntegConvol                          src/convol.f90:27-64
//   DiskInteg                            src/convol.f90:67-246
//   ConvolGau                            src/convol.f90:249-300
//   SemiArea                             src/utils.f90:1435-1445
//
// Every convolution of DiskInteg and ConvolGau has the same form: the spectrum is padded at
// both ends with its end values and out[i] = sum_j in[clamp(i + j - h)] kern[j], the taps
// summed in ascending order.  One launch convolves many rows (limb angles, line and
// continuum, Stokes parameters), each row with its own kernel: a block owns 256 outputs of a
// row and stages the padded input window and the kernel in LDS.  The kernels (at most a
// few thousand taps) are built on the host in FP64 with the reference's formulas; all device
// arithmetic is FP64 without contraction, in the reference's order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fsg_device.h"
#include "fsg_kernels.h"

namespace fsg {

constexpr int CONV_TILE = 256;
constexpr int CONV_LDS_TAPS = 3072;  // kernels up to this length are staged in LDS

#pragma clang fp contract(off)
__global__ void __launch_bounds__(CONV_TILE) k_conv_rows(int nwl, const ConvRow *__restrict__ rows,
                                                         const double *__restrict__ kern, int staged) {
  extern __shared__ double lds[];
  const ConvRow r = rows[blockIdx.y];
  const int i0 = blockIdx.x * CONV_TILE, i = i0 + threadIdx.x;
  const double *kp = kern + r.koff;
  if (staged) {  // every row of the launch has at most CONV_LDS_TAPS taps
    double *win = lds, *kw = lds + CONV_TILE + r.nk - 1;
    for (int t = threadIdx.x; t < CONV_TILE + r.nk - 1; t += CONV_TILE) {
      const int k = min(max(i0 + t - r.h, 0), nwl - 1);
      win[t] = r.in[k];
    }
    for (int t = threadIdx.x; t < r.nk; t += CONV_TILE) kw[t] = kp[t];
    __syncthreads();
    if (i < nwl) {
      double s = 0.0;
      for (int j = 0; j < r.nk; j++) s = s + win[threadIdx.x + j] * kw[j];
      r.out[i] = s;
    }
  } else if (i < nwl) {  // long kernels: taps and inputs through the caches
    double s = 0.0;
    for (int j = 0; j < r.nk; j++) {
      const int k = min(max(i + j - r.h, 0), nwl - 1);
      s = s + r.in[k] * kp[j];
    }
    r.out[i] = s;
  }
}

// SfL, SfC = sum over mu of the annulus spectra times the weights, mu ascending from zero
__global__ void k_mu_sum(int nwl, int nmu, const double *__restrict__ rl, const double *__restrict__ rc,
                         const double *__restrict__ w, double *__restrict__ sfl, double *__restrict__ sfc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwl) return;
  double a = 0.0, b = 0.0;
  for (int m = 0; m < nmu; m++) {
    a = a + rl[(size_t)m * nwl + i] * w[m];
    b = b + rc[(size_t)m * nwl + i] * w[m];
  }
  sfl[i] = a;
  sfc[i] = b;
}
#pragma clang fp contract(on)

static inline unsigned cblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ------------------------------------------------------------------------------------
// host: kernels of DiskInteg / ConvolGau
// ------------------------------------------------------------------------------------
static double semi_area(double R, double V) {
  if (R <= 0.0) return 0.0;
  const double X = std::fmax(-R, std::fmin(R, V));
  return 0.5 * (X * std::sqrt(std::fmax(0.0, (R - X) * (R + X))) + R * R * std::asin(X / R));
}

// normalised Gaussian taps exp(max(-20, -(k/sigma)^2/2)), k = -half .. half
static void gauss_taps(int half, double sigma, std::vector<double> &out) {
  const int n = 2 * half + 1;
  out.assign(n, 0.0);
  double norm = 0.0;
  for (int i = 1; i <= n; i++) {
    const double q = (double)(i - half - 1) / sigma;
    out[i - 1] = std::exp(std::fmax(-20.0, -0.5 * (q * q)));
    norm = norm + out[i - 1];
  }
  for (double &v : out) v = v / norm;
}

void convol_gauss_kernel(int nw, double fwhm, int wmode, double wfirst, double wlast, double dstep,
                         std::vector<double> &kern, int &h) {
  double sigma = wmode == 0 ? fwhm * (wfirst + wlast) / kVc / 2.0 / dstep : fwhm / dstep;
  sigma = sigma / (2.0 * std::sqrt(2.0 * std::log(2.0)));
  int j = std::min((int)(10.0 * sigma), (int)((nw - 3.0) / 2.0));
  j = std::max(3, j);
  gauss_taps(j, sigma, kern);
  h = (int)((2 * j + 1 - 1.0) / 2.0);
}

void diskinteg_kernels(const ConvolSetup &cs, std::vector<double> &kern, std::vector<int> &rot_off,
                       std::vector<int> &rot_n, std::vector<int> &mac_off, int &mac_n) {
  const int M = cs.nmu;
  // annulus boundaries from the flux weights: each annulus encloses the projected area of
  // its weight (MuQuadrature, src/utils.f90:1273-1333)
  std::vector<double> R(M + 1);
  R[0] = 0.0;
  for (int i = 1; i < M; i++) R[i] = std::sqrt(R[i - 1] * R[i - 1] + cs.wmu[i - 1]);
  R[M] = 1.0;
  const double Dv = cs.wmode == 0 ? 2.0 * cs.dstep / (cs.wfirst + cs.wlast) * kVc : cs.dstep;
  const double Sigma = cs.vmacro / Dv / std::sqrt(2.0);
  int Nmk = std::min((int)(10.0 * Sigma), (int)((cs.nwl - 3) / 2));
  Nmk = std::max(3, Nmk);
  const int Nm = 2 * Nmk + 1;
  kern.clear();
  rot_off.assign(M, -1);
  rot_n.assign(M, 0);
  mac_off.assign(M, -1);
  mac_n = Nm;
  std::vector<double> t, mr, mt;
  for (int imu = 1; imu <= M; imu++) {
    if (cs.vsini > 0.0) {
      const double R1 = cs.vsini * R[imu - 1], R2 = cs.vsini * R[imu];
      const int Nrk = 2 * (int)(R2 / Dv) + 3;
      t.assign(Nrk, 0.0);
      double norm = 0.0;
      for (int i = 1; i <= Nrk; i++) {
        const double V = std::fabs(Dv * (double)(i - 1 - (Nrk - 1) / 2));
        t[i - 1] = semi_area(R2, V + 0.5 * Dv) - semi_area(R2, V - 0.5 * Dv) - semi_area(R1, V + 0.5 * Dv) +
                   semi_area(R1, V - 0.5 * Dv);
        norm = norm + t[i - 1];
      }
      if (R2 > 0.5 * Dv) {  // all velocities inside the central pixel: no convolution
        rot_off[imu - 1] = (int)kern.size();
        rot_n[imu - 1] = Nrk;
        for (double v : t) kern.push_back(v / norm);
      }
    }
    if (cs.vmacro > 0.0) {
      const double mu = cs.xmu[imu - 1];
      const double Sigr = Sigma * mu, Sigt = Sigma * std::sqrt(1.0 - mu * mu);
      if (Sigr > 0.0) gauss_taps(Nmk, Sigr, mr);
      else {
        mr.assign(Nm, 0.0);
        mr[Nmk] = 1.0;
      }
      if (Sigt > 0.0) gauss_taps(Nmk, Sigt, mt);
      else {
        mt.assign(Nm, 0.0);
        mt[Nmk] = 1.0;
      }
      mac_off[imu - 1] = (int)kern.size();
      for (int i = 0; i < Nm; i++) kern.push_back(0.5 * mr[i] + 0.5 * mt[i]);
    }
  }
}

// one launch over nrows rows whose table (in, out, kernel offset, taps, h) is already on
// the device; maxk is the longest kernel among them
void launch_conv_rows(int nwl, int nrows, int maxk, const ConvRow *d_rows, const double *d_kern, hipStream_t st) {
  if (nrows <= 0) return;
  const int lk = maxk <= CONV_LDS_TAPS ? std::max(maxk, 1) : 0;
  const size_t lds = lk ? sizeof(double) * (CONV_TILE + 2 * lk - 1) : 0;
  k_conv_rows<<<dim3(cblk(nwl, CONV_TILE), (unsigned)nrows), CONV_TILE, lds, st>>>(nwl, d_rows, d_kern, lk ? 1 : 0);
}

void launch_mu_sum(int nwl, int nmu, const double *rl, const double *rc, const double *w, double *sfl, double *sfc,
                   hipStream_t st) {
  k_mu_sum<<<cblk(nwl, 256), 256, 0, st>>>(nwl, nmu, rl, rc, w, sfl, sfc);
}

}  // namespace fsg
