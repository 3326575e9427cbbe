// Variational free energy (VFE) family: the terms that differ from the AEP path
// (geepee/vfe_models.py, geepee/lik_layers.py):
//   - KL(q(u) || p(u)) of a layer (vfe_models.py:303-324);
//   - expected log-likelihoods, Gaussian and probit, 2-D and over Monte Carlo samples
//     (lik_layers.py:183-236, 412-455), the linear-Gaussian emission (:629-676) and the
//     GPSSM transition (vfe_models.py:1085-1120);
//   - the latent-variable terms: SGPLVM posterior inputs, KL(q(x)) and gradients
//     (vfe_models.py:772-855, base_models.py:895-929), SGPSSM entropy and gradients
//     (vfe_models.py:1030-1060, base_models.py:1730-1752).
// Element-wise work is a grid-stride loop; every scalar is reduced per block in a fixed
// order and finished by one thread in block order, so results are deterministic.
#include "gpx_common.h"

namespace {

constexpr double kTwoPi = 6.283185307179586;
constexpr double kPi = 3.141592653589793;
constexpr int kBlocks = 64;

__constant__ double kGhX[10] = {-3.4361591188377374, -2.5327316742327897, -1.7566836492998816,
                                -1.0366108297895136, -0.3429013272237046, 0.3429013272237046,
                                1.0366108297895136,  1.7566836492998816,  2.5327316742327897,
                                3.4361591188377374};
__constant__ double kGhW[10] = {7.640432855232641e-06, 0.0013436457467812324, 0.033874394455481106,
                                0.2401386110823147,    0.6108626337353258,    0.6108626337353258,
                                0.2401386110823147,    0.033874394455481106,  0.0013436457467812324,
                                7.640432855232641e-06};

inline int nblocks(long n) {
  int b = (int)gpx_cdiv(n, 1024);
  return b > kBlocks ? kBlocks : (b < 1 ? 1 : b);
}

__device__ __forceinline__ void partials2(double a, double b, double *part) {
  __shared__ double red[16];
  a = gpx_block_sum(a, red);
  b = gpx_block_sum(b, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

// scal[0] = s0 * sum(part even) + a0, scal[1] = s1 * sum(part odd) + a1 (when want1)
// One wave, lane b holding block b's partials (nblk <= 64), then fixed-order wave sums.
__global__ __launch_bounds__(64) void finish2(int nblk, const double *__restrict__ part, double s0, double a0,
                                              double s1, double a1, int want1, double *__restrict__ scal) {
  const int b = threadIdx.x;
  double x = b < nblk ? part[2 * b] : 0.0, y = b < nblk ? part[2 * b + 1] : 0.0;
  x = gpx_wave_sum(x);
  y = gpx_wave_sum(y);
  if (b != 0) return;
  scal[0] = s0 * x + a0;
  if (want1) scal[1] = s1 * y + a1;
}

// KL(q(u)||p(u)) = 0.5 (D logdet Kuu - sum_d logdet Su_d - D M + sum_d <Kuu^-1, Splusmm_d>)
__global__ __launch_bounds__(1024) void vfe_kl_kernel(int D, int M, const double *__restrict__ ldK,
                                                      const double *__restrict__ ldS, double sgnS,
                                                      const double *__restrict__ Ki,
                                                      const double *__restrict__ Spm, double *__restrict__ out) {
  __shared__ double red[16];
  const long MM = (long)M * M;
  double tr = 0.0;
  for (long e = threadIdx.x; e < D * MM; e += blockDim.x) tr += Ki[e % MM] * Spm[e];
  tr = gpx_block_sum(tr, red);
  if (threadIdx.x == 0) {
    double lp = 0.0;
    for (int d = 0; d < D; ++d) lp += ldS[d];
    out[0] = 0.5 * ((double)D * ldK[0] - sgnS * lp - (double)D * M + tr);
  }
}

// Gaussian expected log-likelihood; S = samples (1 with three = 0).
__global__ __launch_bounds__(256) void gauss_exp_kernel(int S, int three, long n, const double *__restrict__ m,
                                                        const double *__restrict__ v, const double *__restrict__ y,
                                                        const double *__restrict__ sn, double scale,
                                                        double *__restrict__ dm, double *__restrict__ dv,
                                                        double *__restrict__ part) {
  const double sn2 = exp(2.0 * sn[0]);
  const double t1 = -0.5 * log(kTwoPi * sn2);
  double ll = 0.0, ds = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double ye = y[e];
    if (!three) {
      const double q = ye * ye - 2.0 * ye * m[e] + m[e] * m[e] + v[e];
      ll += t1 + (-0.5 / sn2 * q);
      ds += -1.0 + 1.0 / sn2 * q;
      dm[e] = scale * (1.0 / sn2 * (ye - m[e]));
      dv[e] = scale * (-0.5 / sn2);
      continue;
    }
    double mean = 0.0, dsum = 0.0;
    for (int s = 0; s < S; ++s) {
      const double r = ye - m[s * n + e];
      const double q = r * r + v[s * n + e];
      mean += t1 - 0.5 * q / sn2;
      dsum += -1.0 + q / sn2;
      dm[s * n + e] = scale * (r / sn2 / S);
      dv[s * n + e] = scale * (-0.5 / sn2 / S);
    }
    ll += mean / S;
    ds += dsum / S;
  }
  partials2(ll, ds, part);
}

// log Phi(x) and phi(x) / Phi(x).  Below zero through the scaled complementary error
// function, Phi(x) = erfcx(z) exp(-z^2) / 2 with z = -x / sqrt(2), which neither underflows
// nor loses digits in the lower tail; above zero log Phi(x) = log1p(-Phi(-x)).
__device__ __forceinline__ void log_ndtr_ratio(double x, double &lp, double &r) {
  const double z = -x * 0.7071067811865476;
  if (z > 0.0) {
    const double ex = erfcx(z);
    lp = log(0.5 * ex) - z * z;
    r = 0.7978845608028654 / ex;  // sqrt(2 / pi) / erfcx(z)
  } else {
    const double tail = 0.5 * erfc(-z);  // Phi(-x)
    lp = log1p(-tail);
    r = exp(-0.5 * x * x) * 0.3989422804014327 / (1.0 - tail);
  }
}

__global__ __launch_bounds__(256) void probit_exp_kernel(int S, int three, long n, const double *__restrict__ m,
                                                         const double *__restrict__ v,
                                                         const double *__restrict__ y, double scale,
                                                         double *__restrict__ dm, double *__restrict__ dv,
                                                         double *__restrict__ part) {
  double ll = 0.0;
  const int SS = three ? S : 1;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double ye = y[e];
    double lsum = 0.0;
    for (int s = 0; s < SS; ++s) {
      const double me = m[s * n + e], ve = v[s * n + e];
      const double sv = sqrt(2.0 * ve);
      double l = 0.0, gm = 0.0, gv = 0.0;
#pragma unroll
      for (int k = 0; k < 10; ++k) {
        const double w = kGhW[k] / sqrt(kPi);
        const double t = (kGhX[k] * sv + me) * ye;
        double lp, r;
        log_ndtr_ratio(t, lp, r);
        l += w * lp;
        const double g = ye * w * r;
        gm += g;
        gv += g * (0.5 * kGhX[k] * sqrt(2.0 / ve));
      }
      lsum += l;
      dm[s * n + e] = scale * (gm / SS);
      dv[s * n + e] = scale * (gv / SS);
    }
    ll += lsum / SS;
  }
  partials2(ll, 0.0, part);
}

// GPSSM transition expected log-likelihood (outputs already scaled)
__global__ __launch_bounds__(256) void transition_exp_kernel(int S, int three, long n,
                                                             const double *__restrict__ mp,
                                                             const double *__restrict__ vp,
                                                             const double *__restrict__ mt,
                                                             const double *__restrict__ vt,
                                                             const double *__restrict__ sn, double scale,
                                                             double *__restrict__ dmp, double *__restrict__ dvp,
                                                             double *__restrict__ dmt, double *__restrict__ dvt,
                                                             double *__restrict__ part) {
  const double sn2 = exp(2.0 * sn[0]);
  const double t1 = -0.5 * log(kTwoPi * sn2);
  const int K = three ? S : 1;
  double ll = 0.0, ds = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double a = mt[e], b = vt[e];
    double gmt = 0.0, gvt = 0.0;
    for (int s = 0; s < K; ++s) {
      const double p = mp[s * n + e];
      const double t2 = -0.5 / sn2 * (a * a + b - 2.0 * a * p + p * p + vp[s * n + e]);
      ll += t1 + t2;
      ds += -1.0 - 2.0 * t2;
      if (three) {
        const double g = scale / sn2 * (a - p) / K, h = -scale * 0.5 / sn2 / K;
        dmp[s * n + e] = g;
        dvp[s * n + e] = h;
        gmt -= g;
        gvt += h;
      } else {
        const double g = -scale / sn2 * (a - p);
        dmp[e] = -g;
        dvp[e] = -scale * 0.5 / sn2;
        gmt = g;
        gvt = -scale * 0.5 / sn2;
      }
    }
    dmt[e] = gmt;
    dvt[e] = gvt;
  }
  partials2(ll, ds, part);
}

// Linear-Gaussian emission expected log-likelihood: per row r = y - C mx, dmx = C^T R^-1 r,
// dvx = -diag(C^T R^-1 C) / 2 (scaled); block partials of sum r_a^2 (Dout), sum r_a mx_b
// (Dout x Din) and sum vx_q (Din).
constexpr int EX_DMAX = 16;
__global__ __launch_bounds__(256) void emission_exp_kernel(int Nb, int Dout, int Din, const double *__restrict__ C,
                                                           const double *__restrict__ Rp,
                                                           const double *__restrict__ y, int ldy,
                                                           const double *__restrict__ mx, int ldm,
                                                           const double *__restrict__ vx, int ldv, double scale,
                                                           double *__restrict__ dmx, double *__restrict__ dvx,
                                                           double *__restrict__ work) {
  __shared__ double red[16];
  const int W = Dout + Dout * Din + Din;
  double acc[EX_DMAX + EX_DMAX * EX_DMAX + EX_DMAX];
  for (int i = 0; i < W; ++i) acc[i] = 0.0;
  for (int n = blockIdx.x * 256 + threadIdx.x; n < Nb; n += gridDim.x * 256) {
    double r[EX_DMAX];
    for (int a = 0; a < Dout; ++a) {
      double c = 0.0;
      for (int b = 0; b < Din; ++b) c += C[a * Din + b] * mx[(long)n * ldm + b];
      r[a] = y[(long)n * ldy + a] - c;
      acc[a] += r[a] * r[a];
      for (int b = 0; b < Din; ++b) acc[Dout + a * Din + b] += r[a] * mx[(long)n * ldm + b];
    }
    for (int q = 0; q < Din; ++q) {
      acc[Dout + Dout * Din + q] += vx[(long)n * ldv + q];
      double g = 0.0, h = 0.0;
      for (int a = 0; a < Dout; ++a) {
        const double iR = 1.0 / exp(2.0 * Rp[a]);
        g += C[a * Din + q] * iR * r[a];
        h += C[a * Din + q] * iR * C[a * Din + q];
      }
      dmx[(long)n * Din + q] = scale * g;
      dvx[(long)n * Din + q] = scale * (-0.5 * h);
    }
  }
  for (int i = 0; i < W; ++i) {
    const double s = gpx_block_sum(acc[i], red);
    if (threadIdx.x == 0) work[(long)blockIdx.x * W + i] = s;
  }
}

__global__ void emission_exp_finish(int nblk, int Nb, int Dout, int Din, const double *__restrict__ C,
                                    const double *__restrict__ Rp, double scale, const double *__restrict__ work,
                                    double *__restrict__ dC, double *__restrict__ dR, double *__restrict__ scal) {
  if (threadIdx.x != 0) return;
  const int W = Dout + Dout * Din + Din;
  double S[EX_DMAX + EX_DMAX * EX_DMAX + EX_DMAX];
  for (int i = 0; i < W; ++i) {
    double t = 0.0;
    for (int b = 0; b < nblk; ++b) t += work[(long)b * W + i];
    S[i] = t;
  }
  const double *sr2 = S, *srm = S + Dout, *sv = S + Dout + Dout * Din;
  double lz = -0.5 * Nb * Dout * log(kTwoPi), lr = 0.0, quad = 0.0, vterm = 0.0;
  for (int a = 0; a < Dout; ++a) {
    const double R = exp(2.0 * Rp[a]);
    lr += log(R);
    quad += sr2[a] / R;
  }
  for (int q = 0; q < Din; ++q) {
    double crc = 0.0;
    for (int a = 0; a < Dout; ++a) crc += C[a * Din + q] * C[a * Din + q] / exp(2.0 * Rp[a]);
    vterm += sv[q] * crc;
  }
  lz += -0.5 * Nb * lr - 0.5 * quad - 0.5 * vterm;
  scal[0] = lz * scale;
  for (int a = 0; a < Dout; ++a) {
    const double R = exp(2.0 * Rp[a]);
    double cvc = 0.0;
    for (int q = 0; q < Din; ++q) cvc += C[a * Din + q] * sv[q] * C[a * Din + q];
    const double g = -0.5 * Nb / R + 0.5 * sr2[a] / (R * R) + 0.5 * cvc / (R * R);
    dR[a] = g * 2.0 * R * scale;
    for (int b = 0; b < Din; ++b)
      dC[a * Din + b] = (1.0 / R * srm[a * Din + b] - 1.0 / R * (C[a * Din + b] * sv[b])) * scale;
  }
}

// SGPLVM posterior of rows idx (NULL = all rows): natural parameters prior + factor, or the
// factors as mean / variance (base_models.py:895-911, 765-775).
__device__ __forceinline__ void lvm_post(double f1, double x2, double px1, double px2, int nat, double &p1,
                                         double &p2, double &f2) {
  f2 = exp(2.0 * x2);
  if (nat) {
    p1 = px1 + f1;
    p2 = px2 + f2;
  } else {
    p1 = f1 / f2;
    p2 = 1.0 / f2;
  }
}

__global__ void lvm_post_kernel(int B, int Q, const long *__restrict__ idx, const double *__restrict__ x1,
                                const double *__restrict__ x2, double px1, double px2, int nat,
                                double *__restrict__ mx, double *__restrict__ vx) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)B * Q; e += (long)gridDim.x * blockDim.x) {
    const long j = e / Q, q = e % Q, r = idx ? idx[j] : j;
    double p1, p2, f2;
    lvm_post(x1[r * Q + q], x2[r * Q + q], px1, px2, nat, p1, p2, f2);
    mx[e] = p1 / p2;
    vx[e] = 1.0 / p2;
  }
}

// KL(q(x) || N(m0, v0)) over the batch rows and the gradients w.r.t. x1, x2
// (compute_KL_x vfe_models.py:847-855, compute_posterior_grad_x base_models.py:913-929).
__global__ __launch_bounds__(256) void lvm_vfe_grad_kernel(int B, int Q, const long *__restrict__ idx,
                                                           const double *__restrict__ x1,
                                                           const double *__restrict__ x2,
                                                           const double *__restrict__ dmx,
                                                           const double *__restrict__ dvx, double m0, double v0,
                                                           int nat, double scale_x, double *__restrict__ g1,
                                                           double *__restrict__ g2, double *__restrict__ part) {
  const double px1 = m0 / v0, px2 = 1.0 / v0;
  double kl = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < (long)B * Q; e += (long)gridDim.x * 256) {
    const long j = e / Q, q = e % Q, r = idx ? idx[j] : j;
    double p1, p2, f2;
    lvm_post(x1[r * Q + q], x2[r * Q + q], px1, px2, nat, p1, p2, f2);
    const double m = p1 / p2, v = 1.0 / p2;
    kl += 0.5 * (log(v0) - log(v) + (v + (m - m0) * (m - m0)) / v0 - 1.0);
    const double dm = dmx[e] + scale_x * ((m - m0) / v0);
    const double dv = dvx[e] + scale_x * (-0.5 / v + 0.5 / v0);
    if (nat) {
      g1[r * Q + q] = dm / p2;
      g2[r * Q + q] = (-dm * p1 / (p2 * p2) - dv / (p2 * p2)) * (2.0 * f2);
    } else {
      g1[r * Q + q] = dm;
      g2[r * Q + q] = dv * (2.0 * f2);
    }
  }
  partials2(kl, 0.0, part);
}

// SGPSSM VFE: gradients w.r.t. the x factors of the window rows [s, e) and the entropy.
// dup / vup: emission input grads (rows s..e-1, leading dim ldu); dmt / dvt: transition
// target grads (rows s+1..e-1, ld Q); dmn / dvn: dynamics input grads (rows s..e-2, ld ldn).
__global__ __launch_bounds__(256) void ssm_vfe_grad_kernel(int N, int Q, int s, int e,
                                                           const double *__restrict__ xf1,
                                                           const double *__restrict__ x2, double px1, double px2,
                                                           const double *__restrict__ dup,
                                                           const double *__restrict__ vup, int ldu,
                                                           const double *__restrict__ dmt,
                                                           const double *__restrict__ dvt,
                                                           const double *__restrict__ dmn,
                                                           const double *__restrict__ dvn, int ldn,
                                                           double scale_ent, double *__restrict__ g1,
                                                           double *__restrict__ g2, double *__restrict__ part) {
  const int Be = e - s;
  double ent = 0.0;
  for (long k = blockIdx.x * 256L + threadIdx.x; k < (long)Be * Q; k += (long)gridDim.x * 256) {
    const int j = (int)(k / Q), q = (int)(k % Q), t = s + j;
    const double w = (t == 0 || t == N - 1) ? 2.0 : 3.0;
    const double f2 = exp(2.0 * x2[(long)t * Q + q]);
    const double p1 = w * xf1[(long)t * Q + q] + (t == 0 ? px1 : 0.0);
    const double p2 = w * f2 + (t == 0 ? px2 : 0.0);
    const double pv = 1.0 / p2;
    ent += 0.5 * log(pv);
    double dm = dup[(long)j * ldu + q], dv = vup[(long)j * ldu + q] + scale_ent * 0.5 / pv;
    if (j >= 1) {
      dm += dmt[(long)(j - 1) * Q + q];
      dv += dvt[(long)(j - 1) * Q + q];
    }
    if (j <= Be - 2) {
      dm += dmn[(long)j * ldn + q];
      dv += dvn[(long)j * ldn + q];
    }
    g1[(long)t * Q + q] = dm / p2 * w;
    g2[(long)t * Q + q] = (-dm * p1 / (p2 * p2) - dv / (p2 * p2)) * w * 2.0 * f2;
  }
  partials2(ent, 0.0, part);
}

}  // namespace

extern "C" {

int gpx_vfe_kl(int D, int M, const double *ldK, const double *ldS, double sgnS, const double *Kuuinv,
               const double *Splusmm, double *out, void *stream) {
  if (D < 1 || M < 1 || !ldK || !ldS || !Kuuinv || !Splusmm || !out) return GPX_EINVAL;
  hipLaunchKernelGGL(vfe_kl_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, D, M, ldK, ldS, sgnS, Kuuinv,
                     Splusmm, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_gauss_lik_exp(int S, int three, int B, int D, const double *m, const double *v, const double *y,
                      const double *sn, double scale, double *dm, double *dv, double *scal, void *stream) {
  if (S < 1 || B < 1 || D < 1 || (!three && S != 1) || !m || !v || !y || !sn || !dm || !dv || !scal)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)B * D;
  const int nb = nblocks(n);
  hipLaunchKernelGGL(gauss_exp_kernel, dim3(nb), dim3(256), 0, st, S, three, n, m, v, y, sn, scale, dm, dv,
                     scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(finish2, dim3(1), dim3(64), 0, st, nb, (const double *)(scal + 2), scale, 0.0, scale, 0.0, 1,
                     scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_probit_lik_exp(int S, int three, int B, int D, const double *m, const double *v, const double *y,
                       double scale, double *dm, double *dv, double *scal, void *stream) {
  if (S < 1 || B < 1 || D < 1 || (!three && S != 1) || !m || !v || !y || !dm || !dv || !scal) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)B * D;
  const int nb = nblocks(n);
  hipLaunchKernelGGL(probit_exp_kernel, dim3(nb), dim3(256), 0, st, S, three, n, m, v, y, scale, dm, dv,
                     scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(finish2, dim3(1), dim3(64), 0, st, nb, (const double *)(scal + 2), scale, 0.0, 0.0, 0.0, 0,
                     scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_transition_lik_exp(int S, int three, int B, int D, const double *mp, const double *vp, const double *mt,
                           const double *vt, const double *sn, double scale, double *dmp, double *dvp,
                           double *dmt, double *dvt, double *scal, void *stream) {
  if (S < 1 || B < 1 || D < 1 || (!three && S != 1) || !mp || !vp || !mt || !vt || !sn || !dmp || !dvp ||
      !dmt || !dvt || !scal)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)B * D;
  const int nb = nblocks(n);
  const double K = three ? S : 1;
  hipLaunchKernelGGL(transition_exp_kernel, dim3(nb), dim3(256), 0, st, S, three, n, mp, vp, mt, vt, sn, scale,
                     dmp, dvp, dmt, dvt, scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(finish2, dim3(1), dim3(64), 0, st, nb, (const double *)(scal + 2), scale / K, 0.0,
                     scale / K, 0.0, 1, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

long gpx_emission_lik_exp_workspace(int Dout, int Din) { return (long)kBlocks * (Dout + Dout * Din + Din); }

int gpx_emission_lik_exp(int Nb, int Dout, int Din, const double *C, const double *Rp, const double *y, int ldy,
                         const double *mx, int ldm, const double *vx, int ldv, double scale, double *dmx,
                         double *dvx, double *dC, double *dR, double *scal, double *work, void *stream) {
  if (Nb < 1 || Dout < 1 || Din < 1 || Dout > EX_DMAX || Din > EX_DMAX || !C || !Rp || !y || !mx || !vx ||
      ldy < Dout || ldm < Din || ldv < Din || !dmx || !dvx || !dC || !dR || !scal || !work)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  int nb = (int)gpx_cdiv(Nb, 256);
  if (nb > kBlocks) nb = kBlocks;
  hipLaunchKernelGGL(emission_exp_kernel, dim3(nb), dim3(256), 0, st, Nb, Dout, Din, C, Rp, y, ldy, mx, ldm, vx,
                     ldv, scale, dmx, dvx, work);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(emission_exp_finish, dim3(1), dim3(64), 0, st, nb, Nb, Dout, Din, C, Rp, scale,
                     (const double *)work, dC, dR, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_lvm_x_post(int B, int Q, const long *idx, const double *x1, const double *x2, double prior_mean,
                   double prior_var, int nat_param, double *mx, double *vx, void *stream) {
  if (B < 1 || Q < 1 || !x1 || !x2 || !mx || !vx || !(prior_var > 0)) return GPX_EINVAL;
  unsigned g = gpx_cdiv((long)B * Q, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(lvm_post_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, B, Q, idx, x1, x2,
                     prior_mean / prior_var, 1.0 / prior_var, nat_param, mx, vx);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_lvm_vfe_x_grads(int B, int Q, const long *idx, const double *x1, const double *x2, const double *dmx,
                        const double *dvx, double prior_mean, double prior_var, int nat_param, double scale_x,
                        double *g1, double *g2, double *scal, void *stream) {
  if (B < 1 || Q < 1 || !x1 || !x2 || !dmx || !dvx || !g1 || !g2 || !scal || !(prior_var > 0)) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const int nb = nblocks((long)B * Q);
  hipLaunchKernelGGL(lvm_vfe_grad_kernel, dim3(nb), dim3(256), 0, st, B, Q, idx, x1, x2, dmx, dvx, prior_mean,
                     prior_var, nat_param, scale_x, g1, g2, scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(finish2, dim3(1), dim3(64), 0, st, nb, (const double *)(scal + 2), scale_x, 0.0, 0.0, 0.0, 0,
                     scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_ssm_vfe_x_grads(int N, int Q, int s, int e, const double *xf1, const double *x2, double prior_mean,
                        double prior_var, const double *dup, const double *vup, int ldu, const double *dmt,
                        const double *dvt, const double *dmn, const double *dvn, int ldn, double scale_ent,
                        double *g1, double *g2, double *scal, void *stream) {
  if (N < 2 || Q < 1 || s < 0 || e > N || e - s < 2 || !xf1 || !x2 || !dup || !vup || !dmt || !dvt || !dmn ||
      !dvn || !g1 || !g2 || !scal || ldu < Q || ldn < Q || !(prior_var > 0))
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const int Be = e - s;
  const int nb = nblocks((long)Be * Q);
  hipLaunchKernelGGL(ssm_vfe_grad_kernel, dim3(nb), dim3(256), 0, st, N, Q, s, e, xf1, x2, prior_mean / prior_var,
                     1.0 / prior_var, dup, vup, ldu, dmt, dvt, dmn, dvn, ldn, scale_ent, g1, g2, scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(finish2, dim3(1), dim3(64), 0, st, nb, (const double *)(scal + 2), scale_ent,
                     scale_ent * Be * Q * (0.5 + 0.5 * log(kTwoPi)), 0.0, 0.0, 0, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

}  // extern "C"
