// Latent-variable (x-factor) terms of SGPLVM and SGPSSM: cavity inputs, their gradients
// and the x log-partition contributions.  Element-wise over N x Din plus one scalar
// reduction, so each is a single grid-stride launch.
#include "gpx_common.h"

// Workgroups of the x-gradient kernels; scal holds 2 + 2 * GPX_XF_BLOCKS doubles.
#define GPX_XF_BLOCKS 64

namespace {

// SGPLVM get_cavity_x (aep_models.py:840-861) for rows idx (or all rows).
__global__ void lvm_cav_kernel(int B, int Q, const long *__restrict__ idx, const double *__restrict__ x1,
                               const double *__restrict__ x2, double alpha, double px1, double px2,
                               int nat, double *__restrict__ mcav, double *__restrict__ vcav) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)B * Q;
       e += (long)gridDim.x * blockDim.x) {
    const long r = idx ? idx[e / Q] : e / Q;
    const int q = (int)(e % Q);
    const double f1 = x1[r * Q + q], f2 = exp(2.0 * x2[r * Q + q]);
    double t1, t2;
    if (nat) {
      t1 = px1 + (1.0 - alpha) * f1;
      t2 = px2 + (1.0 - alpha) * f2;
    } else {
      t1 = px1 + (f1 / f2 - px1) * (1.0 - alpha);
      t2 = px2 + (1.0 / f2 - px2) * (1.0 - alpha);
    }
    vcav[e] = 1.0 / t2;
    mcav[e] = t1 / t2;
  }
}

// SGPLVM x gradients and log-partition terms (aep_models.py:784-805, 817-838, 863-867;
// base_models.py:913-929).  g1/g2 are N x Q (rows outside idx left untouched).
// scal[0] = phi_prior_x + sc_cav phi(cav) + sc_post phi(post).
__global__ __launch_bounds__(256) void lvm_grad_kernel(int N, int B, int Q, const long *__restrict__ idx,
                                                        const double *__restrict__ x1,
                                                        const double *__restrict__ x2,
                                                        const double *__restrict__ gmx,
                                                        const double *__restrict__ gvx, double alpha,
                                                        double pm, double pv, int nat, double sc_cav,
                                                        double sc_post, double *__restrict__ g1,
                                                        double *__restrict__ g2, double *__restrict__ scal) {
  __shared__ double red[16];
  const double px1 = pm / pv, px2 = 1.0 / pv;
  double pc = 0.0, pp = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)B * Q; e += (long)gridDim.x * blockDim.x) {
    const long r = idx ? idx[e / Q] : e / Q;
    const int q = (int)(e % Q);
    const double f1 = x1[r * Q + q], f2 = exp(2.0 * x2[r * Q + q]);
    double t1, t2, post1, post2;
    if (nat) {
      t1 = px1 + (1.0 - alpha) * f1;
      t2 = px2 + (1.0 - alpha) * f2;
      post1 = px1 + f1;
      post2 = px2 + f2;
    } else {
      t1 = px1 + (f1 / f2 - px1) * (1.0 - alpha);
      t2 = px2 + (1.0 / f2 - px2) * (1.0 - alpha);
      post1 = f1 / f2;
      post2 = 1.0 / f2;
    }
    const double vc = 1.0 / t2, mc = t1 / t2;
    const double vp = 1.0 / post2, mp = post1 / post2;
    pc += 0.5 * (mc * mc / vc + log(vc));
    pp += 0.5 * (mp * mp / vp + log(vp));
    const double dmc = sc_cav * (mc / vc) + gmx[e];
    const double dvc = sc_cav * 0.5 * (-mc * mc / (vc * vc) + 1.0 / vc) + gvx[e];
    const double dmp = sc_post * (mp / vp);
    const double dvp = sc_post * 0.5 * (-mp * mp / (vp * vp) + 1.0 / vp);
    // compute_cav_grad_x
    const double c1 = mc / vc, c2 = 1.0 / vc;
    double dt1 = (1.0 - alpha) * dmc / c2;
    double dt2 = (1.0 - alpha) * (-dmc * c1 / (c2 * c2) - dvc / (c2 * c2));
    double h1, h2;
    if (nat) {
      dt2 *= 2.0 * f2;
      h1 = dmp / post2;
      h2 = (-dmp * post1 / (post2 * post2) - dvp / (post2 * post2)) * (2.0 * f2);
    } else {
      const double m0 = f1, v0 = f2;
      const double a = dt1 / v0, b = -dt1 * m0 / (v0 * v0) - dt2 / (v0 * v0);
      dt1 = a;
      dt2 = b * 2.0 * f2;
      h1 = dmp;
      h2 = dvp * (2.0 * f2);
    }
    g1[r * Q + q] = dt1 + h1;
    g2[r * Q + q] = dt2 + h2;
  }
  pc = gpx_block_sum(pc, red);
  pp = gpx_block_sum(pp, red);
  if (threadIdx.x == 0) {
    scal[2 + 2 * blockIdx.x] = pc;
    scal[3 + 2 * blockIdx.x] = pp;
  }
}

// scal[0] from the per-block partials, summed in block order.
// One wave, lane b holding block b's partials (nblk <= 64), then fixed-order wave sums.
__global__ __launch_bounds__(64) void lvm_grad_finish(int nblk, int N, int Q, double pm, double pv, double sc_cav,
                                                      double sc_post, double *__restrict__ scal) {
  const int b = threadIdx.x;
  double pc = b < nblk ? scal[2 + 2 * b] : 0.0, pp = b < nblk ? scal[3 + 2 * b] : 0.0;
  pc = gpx_wave_sum(pc);
  pp = gpx_wave_sum(pp);
  if (b != 0) return;
  const double prior = 0.5 * (pm * pm / pv + log(pv)) * (double)N * Q;
  scal[0] = prior + sc_cav * pc + sc_post * pp;
}

// SGPSSM cavity of every x_t (aep_models.py:1376-1387 with base_models.py:1719-1725).
__global__ void ssm_cav_kernel(int N, int Q, const double *__restrict__ xf1, const double *__restrict__ x2,
                               double alpha, double px1, double px2, double *__restrict__ cm,
                               double *__restrict__ cv) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)N * Q;
       e += (long)gridDim.x * blockDim.x) {
    const int t = (int)(e / Q);
    const double f1 = xf1[e], f2 = exp(2.0 * x2[e]);
    const double w = (t == 0 || t == N - 1) ? 2.0 : 3.0;
    const double p1 = w * f1 + (t == 0 ? px1 : 0.0), p2 = w * f2 + (t == 0 ? px2 : 0.0);
    const double c1 = p1 - alpha * f1, c2 = p2 - alpha * f2;
    cm[e] = c1 / (c2 + 1e-16);
    cv[e] = 1.0 / (c2 + 1e-16);
  }
}

// SGPSSM x-factor gradients and log-partitions (aep_models.py:1170-1196, 1208-1315,
// 1389-1437).  Window [s, e) of emission rows; transitions s..e-2 -> s+1..e-1.
// dup/vup: emission input grads (rows s..e-1, leading dim ldu), dmt/dvt: transition
// target grads (rows s+1..e-1, ld Q), dmn/dvn: dynamics input grads (rows s..e-2, ld ldn).
__global__ __launch_bounds__(256) void ssm_grad_kernel(
    int N, int Q, int s, int e, const double *__restrict__ xf1, const double *__restrict__ x2,
    double alpha, double px1, double px2, const double *__restrict__ dup, const double *__restrict__ vup,
    int ldu, const double *__restrict__ dmt, const double *__restrict__ dvt,
    const double *__restrict__ dmn, const double *__restrict__ dvn, int ldn,
    double *__restrict__ g1, double *__restrict__ g2, double *__restrict__ scal) {
  __shared__ double red[16];
  double pp = 0.0, pc = 0.0;
  const double ia = 1.0 / alpha;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (long)N * Q; i += (long)gridDim.x * blockDim.x) {
    const int t = (int)(i / Q), q = (int)(i % Q);
    const double f1 = xf1[i], f2 = exp(2.0 * x2[i]);
    const bool end = (t == 0 || t == N - 1);
    const double w = end ? 2.0 : 3.0;
    const double p1 = w * f1 + (t == 0 ? px1 : 0.0), p2 = w * f2 + (t == 0 ? px2 : 0.0);
    const double c1 = p1 - alpha * f1, c2 = p2 - alpha * f2;
    const double sxp = -(1.0 - ia) + (t < N - 1 ? ia : 0.0) + (t >= 1 ? ia : 0.0);
    const double sc = -ia + (t < N - 1 ? -ia : 0.0) + (t >= 1 ? -ia : 0.0);
    const double sxc = end ? 2.0 - alpha : 3.0 - alpha;
    pp += sxp * 0.5 * (p1 * p1 / p2 - log(p2));
    pc += sc * 0.5 * (c1 * c1 / c2 - log(c2));
    const double dp1 = p1 / p2, dp2 = -0.5 * p1 * p1 / (p2 * p2) - 0.5 / p2;
    double a1 = w * sxp * dp1;
    double a2 = 2.0 * w * sxp * dp2 * f2;
    const double dc1 = c1 / c2, dc2 = -0.5 * c1 * c1 / (c2 * c2) - 0.5 / c2;
    a1 += sc * dc1 * sxc;
    a2 += sc * dc2 * sxc * 2.0 * f2;
    double z1 = 0.0, z2 = 0.0;
    const double ic2 = 1.0 / c2, ic22 = ic2 * ic2;
    if (t >= s && t < e) {
      const double dm = dup[(long)(t - s) * ldu + q], dv = vup[(long)(t - s) * ldu + q];
      z1 += dm * ic2;
      z2 += -dm * c1 * ic22 - dv * ic22;
    }
    if (t >= s + 1 && t < e) {
      const double dm = dmt[(long)(t - s - 1) * Q + q], dv = dvt[(long)(t - s - 1) * Q + q];
      z1 += dm * ic2;
      z2 += -dm * c1 * ic22 - dv * ic22;
    }
    if (t >= s && t < e - 1) {
      const double dm = dmn[(long)(t - s) * ldn + q], dv = dvn[(long)(t - s) * ldn + q];
      z1 += dm * ic2;
      z2 += -dm * c1 * ic22 - dv * ic22;
    }
    a1 += z1 * sxc;
    a2 += z2 * sxc * 2.0 * f2;
    g1[i] = a1;
    g2[i] = a2;
  }
  pp = gpx_block_sum(pp, red);
  pc = gpx_block_sum(pc, red);
  if (threadIdx.x == 0) {
    scal[2 + 2 * blockIdx.x] = pp;
    scal[3 + 2 * blockIdx.x] = pc;
  }
}

// One wave, lane b holding block b's partials (nblk <= 64), then fixed-order wave sums.
__global__ __launch_bounds__(64) void ssm_grad_finish(int nblk, int Q, double px1, double px2,
                                                      double *__restrict__ scal) {
  const int b = threadIdx.x;
  double pp = b < nblk ? scal[2 + 2 * b] : 0.0, pc = b < nblk ? scal[3 + 2 * b] : 0.0;
  pp = gpx_wave_sum(pp);
  pc = gpx_wave_sum(pc);
  if (b != 0) return;
  const double m0 = px1 / px2, v0 = 1.0 / px2;
  scal[0] = 0.5 * Q * (m0 * m0 / v0 + log(v0)) + pp + pc;
}

// SGPSSM posterior of every x_t (base_models.py:1562-1576 over the tied factors of
// :1705-1728): three factors inside the chain, two at its ends, the prior at t = 0.
__global__ void ssm_post_kernel(int N, int Q, const double *__restrict__ xf1, const double *__restrict__ x2,
                                double px1, double px2, double *__restrict__ pm, double *__restrict__ pv) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)N * Q;
       e += (long)gridDim.x * blockDim.x) {
    const int t = (int)(e / Q);
    const double w = (t == 0 || t == N - 1) ? 2.0 : 3.0;
    const double p1 = w * xf1[e] + (t == 0 ? px1 : 0.0), p2 = w * exp(2.0 * x2[e]) + (t == 0 ? px2 : 0.0);
    pm[e] = p1 / p2;
    pv[e] = 1.0 / p2;
  }
}

// SDGPR_H hidden-variable factors (aep_models.py:1620-1752): each hidden value has two
// tied factors, so its posterior natural parameters are 2 (h1, exp(2 h2)) and its cavity
// (2 - alpha) (h1, exp(2 h2)).  Cavity mean / variance of every row.
__global__ void hidden_cav_kernel(long n, const double *__restrict__ h1, const double *__restrict__ h2,
                                  double alpha, double *__restrict__ cm, double *__restrict__ cv) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const double c1 = h1[e] * (2.0 - alpha), c2 = exp(2.0 * h2[e]) * (2.0 - alpha);
    cm[e] = c1 / c2;
    cv[e] = 1.0 / c2;
  }
}

// Gradients w.r.t. the hidden factors through the posterior, the cavity and the cavity
// inputs (compute_grads_hidden, aep_models.py:1620-1660) and the posterior / cavity
// log-partitions (compute_phi_posterior_h / cavity_h, :1662-1700), block partials in
// scal[2 + b].
__global__ __launch_bounds__(256) void hidden_grad_kernel(long n, const double *__restrict__ h1,
                                                          const double *__restrict__ h2, double alpha,
                                                          const double *__restrict__ dmc,
                                                          const double *__restrict__ dvc, double *__restrict__ g1,
                                                          double *__restrict__ g2, double *__restrict__ scal) {
  __shared__ double red[16];
  const double sp = -(1.0 - 1.0 / alpha), sc = -1.0 / alpha;
  double phi = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double f2 = exp(2.0 * h2[e]);
    const double p1 = 2.0 * h1[e], p2 = 2.0 * f2;
    const double c1 = (2.0 - alpha) * h1[e], c2 = (2.0 - alpha) * f2;
    const double d1p = 2.0 * (p1 / p2), d2p = -p1 * p1 / (p2 * p2) - 1.0 / p2;
    const double d1c = (2.0 - alpha) * (c1 / c2), d2c = (2.0 - alpha) * (-0.5 * c1 * c1 / (c2 * c2) - 0.5 / c2);
    const double d1z = (2.0 - alpha) * (dmc[e] / c2);
    const double d2z = (2.0 - alpha) * (-dmc[e] * c1 / (c2 * c2) - dvc[e] / (c2 * c2));
    g1[e] = d1z + sc * d1c + sp * d1p;
    g2[e] = 2.0 * (d2z + sc * d2c + sp * d2p) * f2;
    phi += sp * (0.5 * (p1 * p1 / p2 - log(p2))) + sc * (0.5 * (c1 * c1 / c2 - log(c2)));
  }
  phi = gpx_block_sum(phi, red);
  if (threadIdx.x == 0) scal[2 + blockIdx.x] = phi;
}

// One wave, lane b holding block b's partial (nblk <= 64), then a fixed-order wave sum.
__global__ __launch_bounds__(64) void hidden_grad_finish(int nblk, double *__restrict__ scal) {
  const int b = threadIdx.x;
  const double t = gpx_wave_sum(b < nblk ? scal[2 + b] : 0.0);
  if (b == 0) scal[0] = t;
}

}  // namespace

extern "C" {

int gpx_ssm_x_posterior(int N, int Q, const double *xf1, const double *x2, double prior_mean, double prior_var,
                        double *pm, double *pv, void *stream) {
  if (N < 2 || Q < 1 || !xf1 || !x2 || !pm || !pv || !(prior_var > 0)) return GPX_EINVAL;
  unsigned g = gpx_cdiv((long)N * Q, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(ssm_post_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, N, Q, xf1, x2,
                     prior_mean / prior_var, 1.0 / prior_var, pm, pv);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_hidden_cavity(int N, int D, const double *h1, const double *h2, double alpha, double *cm, double *cv,
                      void *stream) {
  if (N < 1 || D < 1 || !h1 || !h2 || !cm || !cv || !(alpha > 0)) return GPX_EINVAL;
  const long n = (long)N * D;
  unsigned g = gpx_cdiv(n, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(hidden_cav_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, n, h1, h2, alpha, cm, cv);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_hidden_grads(int N, int D, const double *h1, const double *h2, double alpha, const double *dmc,
                     const double *dvc, double *g1, double *g2, double *scal, void *stream) {
  if (N < 1 || D < 1 || !h1 || !h2 || !dmc || !dvc || !g1 || !g2 || !scal || !(alpha > 0)) return GPX_EINVAL;
  const long n = (long)N * D;
  unsigned nblk = gpx_cdiv(n, 1024);
  if (nblk > GPX_XF_BLOCKS) nblk = GPX_XF_BLOCKS;
  if (nblk < 1) nblk = 1;
  hipLaunchKernelGGL(hidden_grad_kernel, dim3(nblk), dim3(256), 0, (hipStream_t)stream, n, h1, h2, alpha, dmc,
                     dvc, g1, g2, scal);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(hidden_grad_finish, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)nblk, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}


int gpx_lvm_x_cavity(int B, int Q, const long *idx, const double *x1, const double *x2, double alpha,
                     double prior_mean, double prior_var, int nat_param, double *mcav, double *vcav,
                     void *stream) {
  if (B < 1 || Q < 1 || !x1 || !x2 || !mcav || !vcav || !(prior_var > 0)) return GPX_EINVAL;
  unsigned g = gpx_cdiv((long)B * Q, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(lvm_cav_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, B, Q, idx, x1, x2, alpha,
                     prior_mean / prior_var, 1.0 / prior_var, nat_param, mcav, vcav);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_lvm_x_grads(int N, int B, int Q, const long *idx, const double *x1, const double *x2,
                    const double *gmx, const double *gvx, double alpha, double prior_mean,
                    double prior_var, int nat_param, double *g1, double *g2, double *scal,
                    void *stream) {
  if (N < 1 || B < 1 || Q < 1 || !x1 || !x2 || !gmx || !gvx || !g1 || !g2 || !scal) return GPX_EINVAL;
  const double sc_cav = -(double)N / B / alpha;
  const double sc_post = -(double)N / B * (1.0 - 1.0 / alpha);
  unsigned nblk = gpx_cdiv((long)B * Q, 256);
  if (nblk > GPX_XF_BLOCKS) nblk = GPX_XF_BLOCKS;
  hipLaunchKernelGGL(lvm_grad_kernel, dim3(nblk), dim3(256), 0, (hipStream_t)stream, N, B, Q, idx, x1, x2, gmx,
                     gvx, alpha, prior_mean, prior_var, nat_param, sc_cav, sc_post, g1, g2, scal);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(lvm_grad_finish, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)nblk, N, Q, prior_mean,
                     prior_var, sc_cav, sc_post, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_ssm_x_cavity(int N, int Q, const double *xf1, const double *x2, double alpha, double prior_mean,
                     double prior_var, double *cm, double *cv, void *stream) {
  if (N < 2 || Q < 1 || !xf1 || !x2 || !cm || !cv || !(prior_var > 0)) return GPX_EINVAL;
  unsigned g = gpx_cdiv((long)N * Q, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(ssm_cav_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, N, Q, xf1, x2, alpha,
                     prior_mean / prior_var, 1.0 / prior_var, cm, cv);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_ssm_x_grads(int N, int Q, int s, int e, const double *xf1, const double *x2, double alpha,
                    double prior_mean, double prior_var, const double *dup, const double *vup, int ldu,
                    const double *dmt, const double *dvt, const double *dmn, const double *dvn, int ldn,
                    double *g1, double *g2, double *scal, void *stream) {
  if (N < 2 || Q < 1 || s < 0 || e > N || e - s < 2 || !xf1 || !x2 || !dup || !vup || !dmt || !dvt ||
      !dmn || !dvn || !g1 || !g2 || !scal || ldu < Q || ldn < Q)
    return GPX_EINVAL;
  unsigned nblk = gpx_cdiv((long)N * Q, 256);
  if (nblk > GPX_XF_BLOCKS) nblk = GPX_XF_BLOCKS;
  hipLaunchKernelGGL(ssm_grad_kernel, dim3(nblk), dim3(256), 0, (hipStream_t)stream, N, Q, s, e, xf1, x2, alpha,
                     prior_mean / prior_var, 1.0 / prior_var, dup, vup, ldu, dmt, dvt, dmn, dvn, ldn, g1, g2,
                     scal);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(ssm_grad_finish, dim3(1), dim3(64), 0, (hipStream_t)stream, (int)nblk, Q,
                     prior_mean / prior_var, 1.0 / prior_var, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

}  // extern "C"
