// Monte Carlo propagation and the likelihoods beside the Gaussian one:
//   - x = eps sqrt(vx) + mx for standard-normal draws and its adjoint
//     (aep_models.py:160-180, base_models.py:309-332, 373-388);
//   - the Gaussian tilted normaliser over S Monte Carlo samples (lik_layers.py:139-152);
//   - the probit tilted normaliser, 2-D and over samples (lik_layers.py:285-411);
//   - the GPSSM transition normaliser over samples (aep_models.py:1344-1372).
// A sample-averaged normaliser is the log-mean-exp of the per-sample ones; each (n, d)
// entry is one thread that walks its S samples twice (maximum, then weights) in sample
// order.  Scalar results (log Z, sum of dv) are reduced per block and finished by one
// thread in block order, so every result is deterministic.
#include "gpx_common.h"

namespace {

constexpr double kTwoPi = 6.283185307179586;
constexpr int kMaxBlocks = 64;

// numpy.polynomial.hermite.hermgauss(10), the rule of lik_layers.py:298-301 with
// GH_DEGREE = 10 (geepee/config.py:12)
__constant__ double kGhX[10] = {-3.4361591188377374, -2.5327316742327897, -1.7566836492998816,
                                -1.0366108297895136, -0.3429013272237046, 0.3429013272237046,
                                1.0366108297895136,  1.7566836492998816,  2.5327316742327897,
                                3.4361591188377374};
__constant__ double kGhW[10] = {7.640432855232641e-06, 0.0013436457467812324, 0.033874394455481106,
                                0.2401386110823147,    0.6108626337353258,    0.6108626337353258,
                                0.2401386110823147,    0.033874394455481106,  0.0013436457467812324,
                                7.640432855232641e-06};

__global__ __launch_bounds__(256) void mc_inputs_kernel(long total, long nq, const double *__restrict__ eps,
                                                        const double *__restrict__ mx,
                                                        const double *__restrict__ vx, double *__restrict__ x) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= total) return;
  const long i = e % nq;
  x[e] = eps[e] * sqrt(vx[i]) + mx[i];
}

__global__ __launch_bounds__(256) void mc_reparam_kernel(int S, long nq, const double *__restrict__ dx,
                                                         const double *__restrict__ eps,
                                                         const double *__restrict__ vx, double *__restrict__ dmx,
                                                         double *__restrict__ dvx) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i >= nq) return;
  double dm = 0.0, de = 0.0;
  for (int s = 0; s < S; ++s) {
    const double g = dx[s * nq + i];
    dm += g;
    de += g * eps[s * nq + i];
  }
  dmx[i] = dm;
  dvx[i] = de / (2.0 * sqrt(vx[i]));
}

// Block partials [logZ, sum dv] -> scal[0] = lz_scale * sum logZ + lz_add,
// scal[1] = (sum dv * dv_mul + dv_add) * dv_scale.
// One wave, lane b holding block b's partials (nblk <= 64), then fixed-order wave sums.
__global__ __launch_bounds__(64) void lik_finish(int nblk, const double *__restrict__ part, double lz_scale,
                                                 double lz_add, const double *__restrict__ sn, double alpha,
                                                 double dv_add, double dv_scale, double *__restrict__ scal) {
  const int b = threadIdx.x;
  double lz = b < nblk ? part[2 * b] : 0.0, sdv = b < nblk ? part[2 * b + 1] : 0.0;
  lz = gpx_wave_sum(lz);
  sdv = gpx_wave_sum(sdv);
  if (b != 0) return;
  scal[0] = lz_scale * lz + lz_add;
  if (sn) scal[1] = (sdv * 2.0 * exp(2.0 * sn[0]) / alpha + dv_add) * dv_scale;
}

__device__ __forceinline__ void block_partials(double lz, double sdv, double *part) {
  __shared__ double red[16];
  lz = gpx_block_sum(lz, red);
  sdv = gpx_block_sum(sdv, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lz;
    part[2 * blockIdx.x + 1] = sdv;
  }
}

// Gaussian likelihood over S samples: m, v are (S, B, D); y is (B, D).
__global__ __launch_bounds__(256) void gauss_mc_kernel(int S, long n, const double *__restrict__ m,
                                                       const double *__restrict__ v, const double *__restrict__ y,
                                                       const double *__restrict__ sn, double alpha, double scale,
                                                       double *__restrict__ dm, double *__restrict__ dv,
                                                       double *__restrict__ part) {
  const double sn2 = exp(2.0 * sn[0]);
  const double cst = 0.5 * log(kTwoPi * sn2 / alpha) - 0.5 * alpha * log(kTwoPi * sn2);
  double lz = 0.0, sdv = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double ye = y[e];
    double mxl = -INFINITY;
    for (int s = 0; s < S; ++s) {
      const double vt = v[s * n + e] + sn2 / alpha, r = ye - m[s * n + e];
      mxl = fmax(mxl, -0.5 * (log(kTwoPi * vt) + r * r / vt) + cst);
    }
    double se = 0.0;
    for (int s = 0; s < S; ++s) {
      const double vt = v[s * n + e] + sn2 / alpha, r = ye - m[s * n + e];
      se += exp(-0.5 * (log(kTwoPi * vt) + r * r / vt) + cst - mxl);
    }
    lz += mxl + log(se) - log((double)S);
    for (int s = 0; s < S; ++s) {
      const double vt = v[s * n + e] + sn2 / alpha, r = ye - m[s * n + e];
      const double w = exp(-0.5 * (log(kTwoPi * vt) + r * r / vt) + cst - mxl) / se;
      const double gv = w * (-0.5 / vt + 0.5 * r * r / (vt * vt));
      dm[s * n + e] = scale * (w * r / vt);
      dv[s * n + e] = scale * gv;
      sdv += gv;
    }
  }
  block_partials(lz, sdv, part);
}

// Probit likelihood (y in {-1, +1}).  three = 0: m, v, dm, dv are (B, D); three = 1: they are
// (S, B, D) and the per-sample normalisers are log-mean-exp averaged.  dmall (optional,
// same shape) = dm - 2 dv m for a moment-matched layer's adjoint.
__device__ __forceinline__ void probit_point(double m, double v, double y, double alpha, bool three,
                                             double &lz, double &gm, double &gv) {
  if (alpha == 1.0) {
    const double t = y * m / sqrt(1.0 + v);
    const double Z = 0.5 * (1.0 + erf(t / sqrt(2.0)));
    lz = log(Z + 1e-16);
    const double g = 1.0 / (Z + 1e-16) / sqrt(kTwoPi) * exp(-t * t / 2.0);
    gm = g * (y / sqrt(1.0 + v));
    gv = g * (-0.5 * y * m / ((1.0 + v) * sqrt(1.0 + v)));
    return;
  }
  const double eps = three ? 1e-16 : 1e-8;
  double Zt = 0.0, am = 0.0, av = 0.0;
  const double sv = sqrt(2.0 * v);
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const double ts = kGhX[k] * sv + m;
    const double pdf = 0.5 * (1.0 + erf(y * ts / sqrt(2.0))) + eps;
    Zt += pow(pdf, alpha) * kGhW[k];
    const double a = pow(pdf, alpha - 1.0) * exp(-ts * ts / 2.0);
    am += kGhW[k] * a;
    av += kGhW[k] * (a * kGhX[k]);
  }
  const double pi = 3.141592653589793;
  Zt = Zt / sqrt(pi);
  lz = log(Zt);
  gm = am * y * alpha / pi / sqrt(2.0) / Zt;
  gv = av * y * alpha / pi / sqrt(2.0) / sv / Zt;
}

__global__ __launch_bounds__(256) void probit_kernel(int S, int three, long n, const double *__restrict__ m,
                                                     const double *__restrict__ v, const double *__restrict__ y,
                                                     double alpha, double scale, double *__restrict__ dm,
                                                     double *__restrict__ dv, double *__restrict__ dmall,
                                                     double *__restrict__ part) {
  double lz = 0.0;
  const double eps = three ? 1e-16 : (alpha == 1.0 ? 0.0 : 1e-8);
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double ye = y[e];
    if (!three) {
      double l, gm, gv;
      probit_point(m[e], v[e], ye, alpha, false, l, gm, gv);
      lz += l;
      gm += eps;
      gv += eps;
      dm[e] = scale * gm;
      dv[e] = scale * gv;
      if (dmall) dmall[e] = scale * gm - 2.0 * scale * gv * m[e];
      continue;
    }
    double mxl = -INFINITY;
    for (int s = 0; s < S; ++s) {
      double l, gm, gv;
      probit_point(m[s * n + e], v[s * n + e], ye, alpha, true, l, gm, gv);
      mxl = fmax(mxl, l);
    }
    double se = 0.0;
    for (int s = 0; s < S; ++s) {
      double l, gm, gv;
      probit_point(m[s * n + e], v[s * n + e], ye, alpha, true, l, gm, gv);
      se += exp(l - mxl);
    }
    lz += mxl + log(se) - log((double)S);
    for (int s = 0; s < S; ++s) {
      double l, gm, gv;
      probit_point(m[s * n + e], v[s * n + e], ye, alpha, true, l, gm, gv);
      const double w = exp(l - mxl) / se;
      const double a = w * gm + (alpha == 1.0 ? 0.0 : eps), b = w * gv + (alpha == 1.0 ? 0.0 : eps);
      dm[s * n + e] = scale * a;
      dv[s * n + e] = scale * b;
      if (dmall) dmall[s * n + e] = scale * a - 2.0 * scale * b * m[s * n + e];
    }
  }
  block_partials(lz, 0.0, part);
}

// GPSSM transition over S samples: mp, vp are (S, B, D) propagated means / variances,
// mt, vt the (B, D) cavities of the next state.  Outputs scaled by `scale`.
__global__ __launch_bounds__(256) void transition_mc_kernel(int S, long n, const double *__restrict__ mp,
                                                            const double *__restrict__ vp,
                                                            const double *__restrict__ mt,
                                                            const double *__restrict__ vt,
                                                            const double *__restrict__ sn, double alpha,
                                                            double scale, double *__restrict__ dmp,
                                                            double *__restrict__ dvp, double *__restrict__ dmt,
                                                            double *__restrict__ dvt, double *__restrict__ part) {
  const double sn2 = exp(2.0 * sn[0]);
  const double cst = -0.5 * alpha * log(kTwoPi * sn2);
  double lz = 0.0, sdv = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double mte = mt[e], vte = vt[e];
    double mxl = -INFINITY;
    for (int s = 0; s < S; ++s) {
      const double vs = vte + vp[s * n + e] + sn2 / alpha, md = mte - mp[s * n + e];
      mxl = fmax(mxl, -0.5 * md * md / vs - 0.5 * log(1.0 + alpha * (vte + vp[s * n + e]) / sn2) + cst);
    }
    double se = 0.0;
    for (int s = 0; s < S; ++s) {
      const double vs = vte + vp[s * n + e] + sn2 / alpha, md = mte - mp[s * n + e];
      se += exp(-0.5 * md * md / vs - 0.5 * log(1.0 + alpha * (vte + vp[s * n + e]) / sn2) + cst - mxl);
    }
    lz += mxl + log(se) - log((double)S);
    double gmt = 0.0, gvt = 0.0;
    for (int s = 0; s < S; ++s) {
      const double vs = vte + vp[s * n + e] + sn2 / alpha, md = mte - mp[s * n + e];
      const double w =
          scale *
          (exp(-0.5 * md * md / vs - 0.5 * log(1.0 + alpha * (vte + vp[s * n + e]) / sn2) + cst - mxl) / se);
      const double a = w * md / vs, b = w * (-0.5 / vs + 0.5 * md * md / (vs * vs));
      dmp[s * n + e] = a;
      dvp[s * n + e] = b;
      gmt -= a;
      gvt += b;
      sdv += b;
    }
    dmt[e] = gmt;
    dvt[e] = gvt;
  }
  block_partials(lz, sdv, part);
}

inline int lik_blocks(long n) {
  int nblk = (int)gpx_cdiv(n, 1024);
  return nblk > kMaxBlocks ? kMaxBlocks : (nblk < 1 ? 1 : nblk);
}

}  // namespace

extern "C" {

int gpx_mc_inputs(int S, int N, int Q, const double *eps, const double *mx, const double *vx, double *x,
                  void *stream) {
  if (S < 1 || N < 1 || Q < 1 || !eps || !mx || !vx || !x) return GPX_EINVAL;
  const long nq = (long)N * Q, total = (long)S * nq;
  hipLaunchKernelGGL(mc_inputs_kernel, dim3(gpx_cdiv(total, 256)), dim3(256), 0, (hipStream_t)stream, total, nq,
                     eps, mx, vx, x);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_mc_reparam_bwd(int S, int N, int Q, const double *dx, const double *eps, const double *vx, double *dmx,
                       double *dvx, void *stream) {
  if (S < 1 || N < 1 || Q < 1 || !dx || !eps || !vx || !dmx || !dvx) return GPX_EINVAL;
  const long nq = (long)N * Q;
  hipLaunchKernelGGL(mc_reparam_kernel, dim3(gpx_cdiv(nq, 256)), dim3(256), 0, (hipStream_t)stream, S, nq, dx,
                     eps, vx, dmx, dvx);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_gauss_tilted_mc(int S, int B, int D, const double *m, const double *v, const double *y, const double *sn,
                        double alpha, double scale, double *dm, double *dv, double *scal, void *stream) {
  if (S < 1 || B < 1 || D < 1 || !(alpha > 0) || !m || !v || !y || !sn || !dm || !dv || !scal) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)B * D;
  const int nblk = lik_blocks(n);
  hipLaunchKernelGGL(gauss_mc_kernel, dim3(nblk), dim3(256), 0, st, S, n, m, v, y, sn, alpha, scale, dm, dv,
                     scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(lik_finish, dim3(1), dim3(64), 0, st, nblk, (const double *)(scal + 2), 1.0, 0.0, sn, alpha,
                     (double)B * D * (1.0 - alpha), scale, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_probit_tilted(int S, int three, int B, int D, const double *m, const double *v, const double *y,
                      double alpha, double scale, double *dm, double *dv, double *dmall, double *scal,
                      void *stream) {
  if (S < 1 || B < 1 || D < 1 || !(alpha > 0) || !m || !v || !y || !dm || !dv || !scal) return GPX_EINVAL;
  if (!three && S != 1) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)B * D;
  const int nblk = lik_blocks(n);
  hipLaunchKernelGGL(probit_kernel, dim3(nblk), dim3(256), 0, st, S, three, n, m, v, y, alpha, scale, dm, dv,
                     dmall, scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(lik_finish, dim3(1), dim3(64), 0, st, nblk, (const double *)(scal + 2), 1.0, 0.0,
                     (const double *)nullptr, alpha, 0.0, 0.0, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_transition_tilted_mc(int S, int B, int D, const double *mp, const double *vp, const double *mt,
                             const double *vt, const double *sn, double alpha, double scale, double *dmp,
                             double *dvp, double *dmt, double *dvt, double *scal, void *stream) {
  if (S < 1 || B < 1 || D < 1 || !(alpha > 0) || !mp || !vp || !mt || !vt || !sn || !dmp || !dvp || !dmt ||
      !dvt || !scal)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)B * D;
  const int nblk = lik_blocks(n);
  hipLaunchKernelGGL(transition_mc_kernel, dim3(nblk), dim3(256), 0, st, S, n, mp, vp, mt, vt, sn, alpha, scale,
                     dmp, dvp, dmt, dvt, scal + 2);
  GPX_CHECK_LAUNCH();
  // dsn = sum(dvp) 2 sn2 / alpha + scale B D (1 - alpha), where sum(dvp) is already scaled
  hipLaunchKernelGGL(lik_finish, dim3(1), dim3(64), 0, st, nblk, (const double *)(scal + 2), scale, 0.0, sn, alpha,
                     scale * (double)B * D * (1.0 - alpha), 1.0, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

}  // extern "C"
