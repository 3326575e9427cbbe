// SE kernel matrices, their adjoints, and the small per-layer / likelihood kernels of
// the AEP objective.  All arrays are row-major FP64 device memory; log
// hyper-parameters (ls, sf, sn, R) are read from the device so no host sync is needed.
#include "gpx_common.h"

namespace {

// ---------------------------------------------------------------------------
// SE kernel: compute_kernel (kernels.py:10-22), K = sf2 exp(-sum_q (x-z)^2 / (2 l_q))
// ---------------------------------------------------------------------------
__global__ void se_kernel(int N, int M, int Q, const double *__restrict__ x, int ldx,
                          const double *__restrict__ z, const double *__restrict__ ls,
                          const double *__restrict__ sf, double jitter, double *__restrict__ K) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)N * M) return;
  const int n = (int)(e / M), m = (int)(e % M);
  double r2 = 0.0;
  for (int q = 0; q < Q; ++q) {
    const double d = x[(long)n * ldx + q] - z[m * Q + q];
    r2 += d * d / exp(2.0 * ls[q]);
  }
  double k = exp(2.0 * sf[0]) * exp(-0.5 * r2);
  if (n == m) k += jitter;
  K[e] = k;
}

// kfucompDer (kernels.py:381-399) with L = dK * K, pass 1 over slabs of KFU_NB rows, one
// slab per workgroup staged in LDS (rows past N as zeros): part[blk][m] = [colsum_m,
// sum_n L x_nq (Q), sum_n L (z_mq - x_nq)^2 (Q)], one thread per (m, q') entry with the
// slab's rows unrolled so their LDS reads issue together;
// dx_nq = sum_m L_nm (z_mq - x_nq) / l_q.
constexpr int KFU_NB = 16;
__global__ __launch_bounds__(256) void kfu_bwd_partial(int N, int M, int Q, const double *__restrict__ dK,
                                                       const double *__restrict__ K,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ z,
                                                       const double *__restrict__ ls,
                                                       double *__restrict__ part,
                                                       double *__restrict__ dx) {
  __shared__ double xs[KFU_NB][17];
  __shared__ double Ls[KFU_NB][257];
  const int n0 = blockIdx.x * KFU_NB, nb = min(KFU_NB, N - n0), tid = threadIdx.x;
  for (int e = tid; e < KFU_NB * Q; e += 256) {
    const int i = e / Q, q = e % Q;
    xs[i][q] = i < nb ? x[(long)(n0 + i) * Q + q] : 0.0;
  }
  for (int e = tid; e < KFU_NB * M; e += 256) {
    const int i = e / M, m = e % M;
    const long o = (long)(n0 + i) * M + m;
    Ls[i][m] = i < nb ? dK[o] * K[o] : 0.0;
  }
  __syncthreads();
  const int W = 2 * Q + 1;
  for (int e = tid; e < M * (Q + 1); e += 256) {
    const int m = e / (Q + 1), q1 = e % (Q + 1);
    double *P = part + ((long)blockIdx.x * M + m) * W;
    if (q1 == 0) {
      double cs = 0.0;
#pragma unroll
      for (int i = 0; i < KFU_NB; ++i) cs += Ls[i][m];
      P[0] = cs;
    } else {
      const int q = q1 - 1;
      const double zq = z[m * Q + q];
      double lx = 0.0, ld2 = 0.0;
#pragma unroll
      for (int i = 0; i < KFU_NB; ++i) {
        const double L = Ls[i][m], xv = xs[i][q], d = zq - xv;
        lx += L * xv;
        ld2 += L * d * d;
      }
      P[1 + q] = lx;
      P[1 + Q + q] = ld2;
    }
  }
  if (!dx) return;
  for (int e = tid; e < nb * Q; e += 256) {
    const int i = e / Q, q = e % Q;
    double s = 0.0;
    for (int m = 0; m < M; ++m) s += Ls[i][m] * (z[m * Q + q] - xs[i][q]);
    dx[(long)(n0 + i) * Q + q] = s / exp(2.0 * ls[q]);
  }
}

// Pass 2: workgroup g sums the slabs' partials (in slab order) for the rows
// m in [g mb, (g + 1) mb), mb = 256 / (2Q + 1), one thread per partial column with its
// loads issued KFU_LB at a time; then dZ_mq = -(z_mq colsum_m - Lx_mq) / l_q, and
// colsum_m, Ld2_mq go to `red` (M x (Q + 1)) for pass 3.
constexpr int KFU_LB = 16;
__global__ __launch_bounds__(256) void kfu_bwd_final(int nblk, int M, int Q, const double *__restrict__ part,
                                                     const double *__restrict__ z,
                                                     const double *__restrict__ ls, double *__restrict__ dZ,
                                                     double *__restrict__ red) {
  __shared__ double sm[256];
  const int W = 2 * Q + 1, mb = 256 / W, m0 = blockIdx.x * mb, tid = threadIdx.x;
  const int ncol = min(mb, M - m0) * W;
  const long bs = (long)M * W;
  if (tid < ncol) {
    const double *P = part + (long)m0 * W + tid;
    double s = 0.0;
    for (int b0 = 0; b0 < nblk; b0 += KFU_LB) {
      double v[KFU_LB];
#pragma unroll
      for (int u = 0; u < KFU_LB; ++u) v[u] = b0 + u < nblk ? P[(b0 + u) * bs] : 0.0;
#pragma unroll
      for (int u = 0; u < KFU_LB; ++u)
        if (b0 + u < nblk) s += v[u];
    }
    sm[tid] = s;
  }
  __syncthreads();
  for (int e = tid; e < min(mb, M - m0) * Q; e += 256) {
    const int ml = e / Q, q = e % Q, m = m0 + ml;
    dZ[m * Q + q] = -(z[m * Q + q] * sm[ml * W] - sm[ml * W + 1 + q]) / exp(2.0 * ls[q]);
    red[m * (Q + 1) + 1 + q] = sm[ml * W + 1 + Q + q];
  }
  if (tid < min(mb, M - m0)) red[(m0 + tid) * (Q + 1)] = sm[tid * W];
}

// pass 3, one wave per output j <= Q, lanes over m and a fixed-order wave reduction:
// dsf2 = sum_m colsum_m / sf2, dell_q = sum_m Ld2_mq / ell^3.
__global__ __launch_bounds__(64) void kfu_bwd_scalars(int M, int Q, const double *__restrict__ red,
                                                      const double *__restrict__ ls,
                                                      const double *__restrict__ sf, double *__restrict__ dsf2,
                                                      double *__restrict__ dell) {
  const int j = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int m = lane; m < M; m += 64) s += red[m * (Q + 1) + j];
  s = gpx_wave_sum(s);
  if (lane == 0) {
    if (j == 0)
      dsf2[0] = s / exp(2.0 * sf[0]);
    else
      dell[j - 1] = s / exp(3.0 * ls[j - 1]);
  }
}

// d_trace_MKzz_dhypers (kernels.py:447-475) with Kzz = SE(z, z) (no jitter):
//   gsf = sum_ab M_ab K_ab, gls_q = sum_ab M_ab K_ab (z_aq - z_bq)^2 / (2 l_q),
//   gz_aq = -sum_b (M_ab + M_ba) K_ab (z_aq - z_bq) / l_q.
// One wave per row a (four per workgroup), its lanes over b; part[a] = [gsf_a, gls_a (Q)].
// Z (q-major, so lanes b read consecutive words) and 1/l_q are staged in LDS once per
// workgroup; QT = Q fixes the per-lane arrays.
template <int QT>
__global__ __launch_bounds__(256) void dtrace_kernel(int M, const double *__restrict__ Mm,
                                                     const double *__restrict__ z,
                                                     const double *__restrict__ ls,
                                                     const double *__restrict__ sf,
                                                     double *__restrict__ gz, double *__restrict__ part) {
  __shared__ double zs[QT][256];
  __shared__ double ils[QT];
  const int tid = threadIdx.x, lane = tid & 63, a = blockIdx.x * 4 + (tid >> 6);
  for (int e = tid; e < M * QT; e += 256) zs[e % QT][e / QT] = z[e];
  if (tid < QT) ils[tid] = 1.0 / exp(2.0 * ls[tid]);
  __syncthreads();
  if (a >= M) return;  // whole waves, after the only barrier
  const double sf2 = exp(2.0 * sf[0]);
  double il[QT], za[QT], gl[QT], gzq[QT];
#pragma unroll
  for (int q = 0; q < QT; ++q) {
    il[q] = ils[q];
    za[q] = zs[q][a];
    gl[q] = gzq[q] = 0.0;
  }
  double g0 = 0.0;
  for (int b = lane; b < M; b += 64) {
    double r2 = 0.0, d[QT];
#pragma unroll
    for (int q = 0; q < QT; ++q) {
      d[q] = za[q] - zs[q][b];
      r2 += d[q] * d[q] * il[q];
    }
    const double K = sf2 * exp(-0.5 * r2);
    const double mab = Mm[(long)a * M + b], mba = Mm[(long)b * M + a];
    g0 += mab * K;
#pragma unroll
    for (int q = 0; q < QT; ++q) {
      gl[q] += 0.5 * mab * K * d[q] * d[q] * il[q];
      gzq[q] += (mab + mba) * K * d[q];
    }
  }
  g0 = gpx_wave_sum(g0);
#pragma unroll
  for (int q = 0; q < QT; ++q) {
    gl[q] = gpx_wave_sum(gl[q]);
    gzq[q] = gpx_wave_sum(gzq[q]);
  }
  if (lane == 0) part[(long)a * (QT + 1)] = g0;
#pragma unroll
  for (int q = 0; q < QT; ++q)
    if (lane == q) {
      part[(long)a * (QT + 1) + 1 + q] = gl[q];
      gz[a * QT + q] = -gzq[q] * il[q];
    }
}

// One wave per output column j <= Q: lanes stride over the M rows, then a fixed-order
// wave reduction (a single thread per column summing M dependent loads was 16 us).
__global__ __launch_bounds__(64) void dtrace_final(int M, int Q, const double *__restrict__ part,
                                                   double *__restrict__ gsf, double *__restrict__ gls) {
  const int j = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int a = lane; a < M; a += 64) s += part[(long)a * (Q + 1) + j];
  s = gpx_wave_sum(s);
  if (lane == 0) {
    if (j == 0) gsf[0] = s;
    else gls[j - 1] = s;
  }
}

// One block per column, threads stride over the rows (fixed-order block reduction).
__global__ __launch_bounds__(256) void colsum_tall_kernel(int rows, int cols, const double *__restrict__ P,
                                                          double alpha, double beta, double *__restrict__ out) {
  __shared__ double red[16];
  const int j = blockIdx.x;
  double s = 0.0;
#pragma unroll 8
  for (long i = threadIdx.x; i < rows; i += 256) s += P[i * cols + j];
  s = gpx_block_sum(s, red);
  if (threadIdx.x == 0) out[j] = (beta == 0.0) ? alpha * s : alpha * s + beta * out[j];
}

// Stage 1 of the two-stage column sum: block (j, c) sums rows [c*chunk, (c+1)*chunk) of
// column j into part[c * cols + j] (strided loads, fixed-order block reduction).
__global__ __launch_bounds__(256) void colsum_chunk_kernel(int rows, int cols, int chunk,
                                                           const double *__restrict__ P,
                                                           double *__restrict__ part) {
  __shared__ double red[16];
  const int j = blockIdx.x, c = blockIdx.y;
  const long r0 = (long)c * chunk, r1 = min((long)rows, r0 + chunk);
  double s0 = 0.0, s1 = 0.0;
  long i = r0 + threadIdx.x;
  for (; i + 256 < r1; i += 512) {
    s0 += P[i * cols + j];
    s1 += P[(i + 256) * cols + j];
  }
  if (i < r1) s0 += P[i * cols + j];
  const double s = gpx_block_sum(s0 + s1, red);
  if (threadIdx.x == 0) part[(long)c * cols + j] = s;
}

__global__ void colsum_kernel(int rows, int cols, const double *__restrict__ P, double alpha,
                              double beta, double *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  // unrolled: eight loads in flight ahead of the (same-order) adds
#pragma unroll 8
  for (int i = 0; i < rows; ++i) s += P[(long)i * cols + j];
  out[j] = (beta == 0.0) ? alpha * s : alpha * s + beta * out[j];
}

// ---------------------------------------------------------------------------
// q(u) parameter plumbing (base_models.py:645-653; aep_models.py:267-280)
// ---------------------------------------------------------------------------
__device__ __forceinline__ long triu_index(int M, int i, int j) {
  return (long)i * M - (long)i * (i - 1) / 2 + (j - i);
}

__global__ void eta_to_R_kernel(int D, int M, const double *__restrict__ eta, double *__restrict__ R) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long MM = (long)M * M;
  if (e >= D * MM) return;
  const int d = (int)(e / MM), i = (int)((e % MM) / M), j = (int)(e % M);
  double v = 0.0;
  if (j >= i) {
    const double t = eta[(long)d * (M * (M + 1) / 2) + triu_index(M, i, j)];
    v = (i == j) ? exp(t) : t;
  }
  R[e] = v;
}

// out[d, triu(i,j)] = (sum_{k>=i} R_ik (G_kj + G_jk)) * (i == j ? R_ii : 1)
// One wave per upper 16 x 16 output tile (I, J >= I) of one d, four per workgroup: the
// product R (G + G^T) over the k-blocks K >= I (R is upper triangular) as FP64 MFMAs fed
// straight from global memory (lane l: A = R[I*16 + l%16][k], B = G[k][J*16 + l%16] +
// G[J*16 + l%16][k] for k = K*16 + 4s + l/16).  No LDS, no barriers.
__global__ __launch_bounds__(256) void pack_dR_kernel(int D, int M, const double *__restrict__ R,
                                                      const double *__restrict__ G, double *__restrict__ out) {
  const int lane = threadIdx.x & 63, lr = lane >> 4, lc = lane & 15;
  const int MT = (M + 15) / 16, T = MT * (MT + 1) / 2;
  int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= T) return;
  int I = 0;
  while (t >= MT - I) {
    t -= MT - I;
    ++I;
  }
  const int J = I + t, d = blockIdx.y;
  const double *Rd = R + (long)d * M * M, *Gd = G + (long)d * M * M;
  const int i = I * 16 + lc, j = J * 16 + lc;
  gpx_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 2
  for (int K = I; K < MT; ++K) {
    double av[4], bv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      // loads at clamped indices, unconditional (so they issue together), then zeroed
      // where the padding needs it by a 0/1 weight (a select lets the compiler sink the
      // load back under a branch); the clamped values are finite
      const int k = K * 16 + 4 * s + lr;
      const int ic = min(i, M - 1), jc = min(j, M - 1), kc = min(k, M - 1);
      const double a = Rd[(long)ic * M + kc], b = Gd[(long)kc * M + jc] + Gd[(long)jc * M + kc];
      av[s] = a * ((i < M && k < M) ? 1.0 : 0.0);
      bv[s] = b * ((j < M && k < M) ? 1.0 : 0.0);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) acc = gpx_mfma16(av[s], bv[s], acc);
  }
  const long Kp = (long)M * (M + 1) / 2;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int ii = I * 16 + lr + 4 * r, jj = J * 16 + lc;
    if (ii >= M || jj >= M || jj < ii) continue;
    const long rowoff = (long)ii * M - (long)ii * (ii - 1) / 2;
    out[(long)d * Kp + rowoff + (jj - ii)] = (ii == jj) ? acc[r] * Rd[(long)ii * M + ii] : acc[r];
  }
}

__global__ void lincomb_kernel(long n, double a, const double *__restrict__ x, long nx, double b,
                               const double *__restrict__ y, long ny, double c,
                               const double *__restrict__ w, long nw, double *__restrict__ out) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double v = a * x[i % nx];
    if (y) v += b * y[i % ny];
    if (w) v += c * w[i % nw];
    out[i] = v;
  }
}

__global__ void sym_kernel(int batch, int M, double a, const double *__restrict__ X, double *__restrict__ out) {
  const long MM = (long)M * M;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= batch * MM) return;
  const long base = (e / MM) * MM;
  const int i = (int)((e % MM) / M), j = (int)(e % M);
  out[e] = a * (X[base + (long)i * M + j] + X[base + (long)j * M + i]);
}

// v[n,d] = sf2 + sum_a T[d][n][a] kfu[n][a]   (aep_models.py:155-157)
// One wave per (n, d), its lanes over a (coalesced rows of T and kfu).
__global__ __launch_bounds__(256) void det_var_kernel(int N, int D, int M, const double *__restrict__ T,
                                                      const double *__restrict__ kfu, const double *__restrict__ sf,
                                                      double *__restrict__ v) {
  const long e = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (e >= (long)N * D) return;
  const int n = (int)(e / D), d = (int)(e % D);
  const double *t = T + ((long)d * N + n) * M, *k = kfu + (long)n * M;
  double s = 0.0;
  for (int a = lane; a < M; a += 64) s += t[a] * k[a];
  s = gpx_wave_sum(s);
  if (lane == 0) v[e] = exp(2.0 * sf[0]) + s;
}

// v = sf2 + Bpsi2 - m^2   (aep_models.py:196-198)
__global__ void mm_var_kernel(long n, const double *__restrict__ m, const double *__restrict__ Bp,
                              const double *__restrict__ sf, double *__restrict__ v) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n) return;
  v[e] = exp(2.0 * sf[0]) + Bp[e] - m[e] * m[e];
}

// out[b] = sum_i x[b][i] y[b][i]
__global__ void batched_dot_kernel(int batch, int n, const double *__restrict__ x,
                                   const double *__restrict__ y, double *__restrict__ out) {
  __shared__ double red[16];
  const int b = blockIdx.x;
  double s = 0.0;
#pragma unroll 4
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += x[(long)b * n + i] * y[(long)b * n + i];
  s = gpx_block_sum(s, red);
  if (threadIdx.x == 0) out[b] = s;
}

// ---------------------------------------------------------------------------
// Gauss_Layer.compute_log_Z + backprop_grads (lik_layers.py:104-181), 2-D branch.
// scal[0] = logZ (with the alpha constant), scal[1] = dsn (scaled).
// dm/dv outputs are scaled by `scale`; dmall = dm - 2 dv m when requested.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gauss_tilted_kernel(int B, int D, const double *__restrict__ m,
                                                           const double *__restrict__ v,
                                                           const double *__restrict__ y,
                                                           const double *__restrict__ sn, double alpha,
                                                           double scale, double *__restrict__ dm,
                                                           double *__restrict__ dv,
                                                           double *__restrict__ dmall,
                                                           double *__restrict__ part) {
  __shared__ double red[16];
  const double sn2 = exp(2.0 * sn[0]);
  const double two_pi = 6.283185307179586;
  double lz = 0.0, sdv = 0.0;
  const long n = (long)B * D;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double vt = v[e] + sn2 / alpha;
    const double r = y[e] - m[e];
    lz += -0.5 * (log(two_pi * vt) + r * r / vt);
    const double gm = r / vt, gv = -0.5 / vt + 0.5 * r * r / (vt * vt);
    sdv += gv;
    dm[e] = scale * gm;
    dv[e] = scale * gv;
    if (dmall) dmall[e] = scale * gm - 2.0 * scale * gv * m[e];
  }
  lz = gpx_block_sum(lz, red);
  sdv = gpx_block_sum(sdv, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lz;
    part[2 * blockIdx.x + 1] = sdv;
  }
}

// One wave, lane b holding block b's partials (nblk <= 64), then fixed-order wave sums.
__global__ __launch_bounds__(64) void gauss_tilted_finish(int nblk, int B, int D, const double *__restrict__ part,
                                                          const double *__restrict__ sn, double alpha,
                                                          double scale, double *__restrict__ scal) {
  const int b = threadIdx.x;
  double lz = b < nblk ? part[2 * b] : 0.0, sdv = b < nblk ? part[2 * b + 1] : 0.0;
  lz = gpx_wave_sum(lz);
  sdv = gpx_wave_sum(sdv);
  if (b != 0) return;
  const double sn2 = exp(2.0 * sn[0]);
  const double two_pi = 6.283185307179586;
  lz += (double)B * D * (0.5 * log(two_pi * sn2 / alpha) - 0.5 * alpha * log(two_pi * sn2));
  scal[0] = lz;
  scal[1] = (sdv * 2.0 * sn2 / alpha + (double)B * D * (1.0 - alpha)) * scale;
}

// SGPSSM.compute_transition_tilted, 2-D branch (aep_models.py:1317-1343).
// Outputs are already scaled; dmall = dmprop - 2 dvprop mprop for the MM adjoint.
// Up to 64 blocks; part[2b], part[2b+1] = block b's (logZ, sum dv) partials.
__global__ __launch_bounds__(256) void transition_kernel(int B, int D, const double *__restrict__ mp,
                                                          const double *__restrict__ vp,
                                                          const double *__restrict__ mt,
                                                          const double *__restrict__ vt,
                                                          const double *__restrict__ sn, double alpha,
                                                          double scale, double *__restrict__ dmall,
                                                          double *__restrict__ dvp,
                                                          double *__restrict__ dmt,
                                                          double *__restrict__ dvt,
                                                          double *__restrict__ part) {
  __shared__ double red[16];
  const double sn2 = exp(2.0 * sn[0]);
  const double two_pi = 6.283185307179586;
  double lz = 0.0, sdv = 0.0;
  const long n = (long)B * D;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const double vs = vt[e] + vp[e] + sn2 / alpha;
    const double md = mt[e] - mp[e];
    lz += -0.5 * md * md / vs - 0.5 * log(1.0 + alpha * (vt[e] + vp[e]) / sn2) -
          0.5 * alpha * log(two_pi * sn2);
    const double gv = scale * (-0.5 / vs + 0.5 * md * md / (vs * vs));
    const double gmt = scale * (-md / vs);
    sdv += gv;
    dvt[e] = gv;
    dvp[e] = gv;
    dmt[e] = gmt;
    dmall[e] = -gmt - 2.0 * gv * mp[e];
  }
  lz = gpx_block_sum(lz, red);
  sdv = gpx_block_sum(sdv, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lz;
    part[2 * blockIdx.x + 1] = sdv;
  }
}

// One wave, lane b holding block b's partials (nblk <= 64), then fixed-order wave sums.
__global__ __launch_bounds__(64) void transition_finish(int nblk, int B, int D, const double *__restrict__ part,
                                                        const double *__restrict__ sn, double alpha,
                                                        double scale, double *__restrict__ scal) {
  const int b = threadIdx.x;
  double lz = b < nblk ? part[2 * b] : 0.0, sdv = b < nblk ? part[2 * b + 1] : 0.0;
  lz = gpx_wave_sum(lz);
  sdv = gpx_wave_sum(sdv);
  if (b != 0) return;
  const double sn2 = exp(2.0 * sn[0]);
  scal[0] = scale * lz;
  scal[1] = sdv * 2.0 * sn2 / alpha + scale * (double)B * D * (1.0 - alpha);
}

// Gauss_Emis.compute_emission_tilted (lik_layers.py:573-627) for Dout <= 8, one thread
// per time step; DT = Dout fixed at compile time, so Vy, its Cholesky factor, the factor's
// inverse and Vy^{-1} live in registers (fully unrolled, lower triangles packed).
// part[blk] collects [logZ, dC (Dout x Din), dR (Dout)] sums over the block's 64 steps
// (unscaled), summed with wave reductions.
constexpr int EM_T = 64;
constexpr int EM_DMAX = 8;
template <int DT>
__global__ __launch_bounds__(EM_T) void emission_kernel(int Nb, int Din, const double *__restrict__ C,
                                                        const double *__restrict__ Rp,
                                                        const double *__restrict__ y, int ldy,
                                                        const double *__restrict__ mx, int ldm,
                                                        const double *__restrict__ vx, int ldv,
                                                        double alpha, double scale,
                                                        double *__restrict__ dmx, double *__restrict__ dvx,
                                                        double *__restrict__ part) {
  constexpr int NL = DT * (DT + 1) / 2;
  __shared__ double Cs[EM_DMAX * EM_DMAX];
  __shared__ double Rs[EM_DMAX];
  const int tid = threadIdx.x, t = blockIdx.x * EM_T + tid;
  for (int e = tid; e < DT * Din; e += EM_T) Cs[e] = C[e];
  if (tid < DT) Rs[tid] = exp(2.0 * Rp[tid]);
  __syncthreads();
  const bool ok = t < Nb;
  const int tt = ok ? t : Nb - 1;  // padding steps compute a real step's values, weighted 0
  const double wgt = ok ? 1.0 : 0.0;
  const double *vt = vx + (long)tt * ldv, *mt = mx + (long)tt * ldm;
#define LI(a, b) ((a) * ((a) + 1) / 2 + (b))
  // Vy = diag(R/alpha) + C diag(vx) C^T (lower triangle)
  double L[NL];
#pragma unroll
  for (int a = 0; a < DT; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      double sum = 0.0;
      for (int c = 0; c < Din; ++c) sum += Cs[a * Din + c] * vt[c] * Cs[b * Din + c];
      if (a == b) sum += Rs[a] / alpha;
      L[LI(a, b)] = sum;
    }
  // Cholesky (lower) in place
  double logdetV = 0.0;
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    double sum = L[LI(k, k)];
#pragma unroll
    for (int j = 0; j < k; ++j) sum -= L[LI(k, j)] * L[LI(k, j)];
    const double lkk = sqrt(sum);
    L[LI(k, k)] = lkk;
    logdetV += 2.0 * log(lkk);
#pragma unroll
    for (int i = k + 1; i < DT; ++i) {
      double v = L[LI(i, k)];
#pragma unroll
      for (int j = 0; j < k; ++j) v -= L[LI(i, j)] * L[LI(k, j)];
      L[LI(i, k)] = v / lkk;
    }
  }
  // r = y - C mx;  u = Vy^{-1} r
  double r[DT], u[DT];
#pragma unroll
  for (int a = 0; a < DT; ++a) {
    double v = y[(long)tt * ldy + a];
    for (int c = 0; c < Din; ++c) v -= Cs[a * Din + c] * mt[c];
    r[a] = v;
  }
#pragma unroll
  for (int a = 0; a < DT; ++a) {  // forward
    double v = r[a];
#pragma unroll
    for (int j = 0; j < a; ++j) v -= L[LI(a, j)] * u[j];
    u[a] = v / L[LI(a, a)];
  }
#pragma unroll
  for (int a = DT - 1; a >= 0; --a) {  // backward
    double v = u[a];
#pragma unroll
    for (int j = a + 1; j < DT; ++j) v -= L[LI(j, a)] * u[j];
    u[a] = v / L[LI(a, a)];
  }
  // Li = L^{-1} (lower), then Vinv = Li^T Li (lower triangle; symmetric)
  double Li[NL];
#pragma unroll
  for (int j = 0; j < DT; ++j)
#pragma unroll
    for (int i = j; i < DT; ++i) {
      double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int k = j; k < i; ++k) v -= L[LI(i, k)] * Li[LI(k, j)];
      Li[LI(i, j)] = v / L[LI(i, i)];
    }
  double Vi[NL];
#pragma unroll
  for (int a = 0; a < DT; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      double v = 0.0;
#pragma unroll
      for (int k = a; k < DT; ++k) v += Li[LI(k, a)] * Li[LI(k, b)];
      Vi[LI(a, b)] = v;
    }
#define VI(a, b) ((a) >= (b) ? Vi[LI(a, b)] : Vi[LI(b, a)])
  double quad = 0.0, sumlogR = 0.0;
#pragma unroll
  for (int a = 0; a < DT; ++a) {
    quad += r[a] * u[a];
    sumlogR += log(Rs[a]);
  }
  // logdet(I + alpha C V C^T / R) = logdet(Vy) + Dout log(alpha) - sum log R
  const double vlog = logdetV + DT * log(alpha) - sumlogR;
  double part0 = -0.5 * quad - 0.5 * vlog - 0.5 * DT * alpha * log(6.283185307179586) - 0.5 * alpha * sumlogR;
  part0 = gpx_wave_sum(wgt * part0);
  if (tid == 0) part[(long)blockIdx.x * (DT * Din + DT + 1)] = part0;
  // dSigma = -Vinv/2 + u u^T / 2
  const long pb = (long)blockIdx.x * (DT * Din + DT + 1);
#pragma unroll
  for (int a = 0; a < DT; ++a) {
    const double dr = gpx_wave_sum(wgt * (-0.5 * VI(a, a) + 0.5 * u[a] * u[a]));
    if (tid == 0) part[pb + 1 + DT * Din + a] = dr;
    for (int c = 0; c < Din; ++c) {
      double v = 0.0;
#pragma unroll
      for (int b = 0; b < DT; ++b) v += (-0.5 * VI(a, b) + 0.5 * u[a] * u[b]) * Cs[b * Din + c];
      const double dc = gpx_wave_sum(wgt * (u[a] * mt[c] + 2.0 * v * vt[c]));
      if (tid == 0) part[pb + 1 + a * Din + c] = dc;
    }
  }
  if (ok) {
    for (int c = 0; c < Din; ++c) {
      double gm = 0.0, gv = 0.0;
#pragma unroll
      for (int a = 0; a < DT; ++a) {
        gm += u[a] * Cs[a * Din + c];
#pragma unroll
        for (int b = 0; b < DT; ++b)
          gv += (-0.5 * VI(a, b) + 0.5 * u[a] * u[b]) * Cs[a * Din + c] * Cs[b * Din + c];
      }
      dmx[(long)t * Din + c] = gm * scale;
      dvx[(long)t * Din + c] = gv * scale;
    }
  }
#undef VI
#undef LI
}

// logZ, dC, dR from the block partials (the dR chain of lik_layers.py:610-613).
__global__ void emission_final(int nblk, int Nb, int Dout, int Din, const double *__restrict__ part,
                               const double *__restrict__ Rp, double alpha, double scale,
                               double *__restrict__ dC, double *__restrict__ dR, double *__restrict__ scal) {
  const int W = Dout * Din + Dout + 1;
  for (int i = threadIdx.x; i < W; i += blockDim.x) {
    double s = 0.0;
#pragma unroll 8
    for (int b = 0; b < nblk; ++b) s += part[(long)b * W + i];
    if (i == 0) {
      scal[0] = s * scale;
    } else if (i <= Dout * Din) {
      dC[i - 1] = s * scale;
    } else {
      const int a = i - 1 - Dout * Din;
      const double R = exp(2.0 * Rp[a]);
      const double g = s / alpha + 0.5 * Nb * (1.0 - alpha) / R;
      dR[a] = g * 2.0 * R * scale;
    }
  }
}


// out[d][n][m] = w[n, d] * X[n][m]   (the dv-weighted rows of aep_models.py:464, 490)
__global__ void rowscale_kernel(int N, int D, int M, const double *__restrict__ w,
                                const double *__restrict__ X, double *__restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long NM = (long)N * M;
  if (e >= D * NM) return;
  const int d = (int)(e / NM);
  const long r = e % NM;
  const int n = (int)(r / M);
  out[e] = w[(long)n * D + d] * X[r];
}

// out[n][m] = alpha sum_d w[n, d] T[d][n][m] + beta out[n][m]   (aep_models.py:455)
__global__ void wsum_kernel(int N, int D, int M, double alpha, const double *__restrict__ w,
                            const double *__restrict__ T, double beta, double *__restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long NM = (long)N * M;
  if (e >= NM) return;
  const int n = (int)(e / M);
  double s = 0.0;
#pragma unroll 8
  for (int d = 0; d < D; ++d) s += w[(long)n * D + d] * T[(long)d * NM + e];
  out[e] = (beta == 0.0) ? alpha * s : alpha * s + beta * out[e];
}

// out = dm - 2 dv m   (aep_models.py:235)
__global__ void dmall_kernel(long n, const double *__restrict__ dm, const double *__restrict__ dv,
                             const double *__restrict__ m, double *__restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < n) out[e] = dm[e] - 2.0 * dv[e] * m[e];
}

// dls = dell * exp(ls) + 2 gls;  dsf = 2 exp(2 sf) (dsf2 + sumdv) + 2 gsf
// (aep_models.py:244-247, 296-298, 456-460, 502-504)
__global__ void hyper_kernel(int Q, const double *__restrict__ ls, const double *__restrict__ sf,
                             const double *__restrict__ dell, const double *__restrict__ gls,
                             const double *__restrict__ dsf2, const double *__restrict__ sumdv,
                             const double *__restrict__ gsf, double *__restrict__ dls,
                             double *__restrict__ dsf) {
  const int q = threadIdx.x;
  if (q < Q) dls[q] = dell[q] * exp(ls[q]) + 2.0 * gls[q];
  if (q == 0) dsf[0] = 2.0 * exp(2.0 * sf[0]) * (dsf2[0] + sumdv[0]) + 2.0 * gsf[0];
}

// phi = (D/2) ldK + sp/2 sum_d (sp_sign ldP_d + dotP_d) + sc/2 sum_d (sc_sign ldC_d + dotC_d)
// (aep_models.py:62-114)
__global__ void phi_kernel(int D, const double *__restrict__ ldK, const double *__restrict__ ldP,
                           double sgnP, const double *__restrict__ dotP, const double *__restrict__ ldC,
                           double sgnC, const double *__restrict__ dotC, double sp, double sc,
                           double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double p = 0.0, c = 0.0;
  for (int d = 0; d < D; ++d) {
    p += sgnP * ldP[d] + dotP[d];
    c += sgnC * ldC[d] + dotC[d];
  }
  out[0] = D * 0.5 * ldK[0] + sp * 0.5 * p + sc * 0.5 * c;
}

// Natural-parameter phi with the quadratic terms taken from theta2: Sigma^-1 mu = theta2
// and Sigma-hat^-1 mu-hat = beta theta2, so mu^T Sigma^-1 mu = mu . theta2 and
// mu-hat^T Sigma-hat^-1 mu-hat = beta mu-hat . theta2.  One workgroup; the partial sums
// and their reduction run in a fixed order.
__global__ __launch_bounds__(256) void phi_nat_kernel(int D, int M, const double *__restrict__ ldK,
                                                      const double *__restrict__ ldP,
                                                      const double *__restrict__ ldC,
                                                      const double *__restrict__ mu,
                                                      const double *__restrict__ muhat,
                                                      const double *__restrict__ th2, double beta,
                                                      double sp, double sc, double *__restrict__ out) {
  __shared__ double red[16];
  double p = 0.0, c = 0.0;
  const long n = (long)D * M;
  for (long e = threadIdx.x; e < n; e += 256) {
    p = fma(mu[e], th2[e], p);
    c = fma(muhat[e], th2[e], c);
  }
  p = gpx_block_sum(p, red);
  c = gpx_block_sum(c, red);
  if (threadIdx.x == 0) {
    double lp = 0.0, lc = 0.0;
    for (int d = 0; d < D; ++d) {
      lp -= ldP[d];
      lc -= ldC[d];
    }
    out[0] = D * 0.5 * ldK[0] + sp * 0.5 * (lp + p) + sc * 0.5 * (lc + beta * c);
  }
}

// my = C mf, vy_d = sum_c C_dc^2 vf_c + R_d (diagonal of lik_layers.py:558-571)
__global__ void emission_predict_kernel(int Nb, int Dout, int Din, const double *__restrict__ C,
                                        const double *__restrict__ Rp, const double *__restrict__ mf,
                                        const double *__restrict__ vf, double *__restrict__ my,
                                        double *__restrict__ vy) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)Nb * Dout) return;
  const int n = (int)(e / Dout), d = (int)(e % Dout);
  double m = 0.0, v = exp(2.0 * Rp[d]);
  for (int c = 0; c < Din; ++c) {
    const double cc = C[d * Din + c];
    m += cc * mf[(long)n * Din + c];
    v += cc * cc * vf[(long)n * Din + c];
  }
  my[e] = m;
  vy[e] = v;
}

// Kuu of several layers of one M in one launch: K[l] = sf2_l exp(-r2 / 2) + jitter I over
// the layer's inducing inputs z_l (M x Q_l), the arithmetic of se_kernel with x = z.
constexpr int kKuuLayers = 16;
struct KuuLayers {
  const double *z[kKuuLayers], *ls[kKuuLayers], *sf[kKuuLayers];
  int Q[kKuuLayers];
};
__global__ __launch_bounds__(256) void kuu_batched_kernel(int nl, KuuLayers L, int M, double jitter,
                                                          double *__restrict__ K) {
  const long MM = (long)M * M;
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= nl * MM) return;
  const int l = (int)(e / MM), n = (int)((e % MM) / M), m = (int)(e % M);
  const double *z = L.z[l], *ls = L.ls[l];
  const int Q = L.Q[l];
  double r2 = 0.0;
  for (int q = 0; q < Q; ++q) {
    const double d = z[n * Q + q] - z[m * Q + q];
    r2 += d * d / exp(2.0 * ls[q]);
  }
  double k = exp(2.0 * L.sf[l][0]) * exp(-0.5 * r2);
  if (n == m) k += jitter;
  K[e] = k;
}

// Posterior and cavity precisions of several layers of one M from X = R^T R (the
// layers' output dims stacked) and their Kuu^-1 (stacked): layer l with D_l outputs at
// rows x_l of X writes prec[2 x_l + d] = X[x_l + d] + Ki[l] and prec[2 x_l + D_l + d] =
// beta X[x_l + d] + Ki[l].  One thread per element of X.
constexpr int kPrecLayers = 16;
struct PrecLayers {
  int D[kPrecLayers];
};
__global__ __launch_bounds__(256) void prec_form_kernel(int nl, PrecLayers L, int M, const double *__restrict__ X,
                                                        const double *__restrict__ Ki, double beta,
                                                        double *__restrict__ prec) {
  const long MM = (long)M * M;
  const long e = blockIdx.x * 256L + threadIdx.x;
  int row = (int)(e / MM), l = 0, x0 = 0;
  while (l < nl && row >= x0 + L.D[l]) x0 += L.D[l++];
  if (l >= nl) return;  // past the last layer's rows
  const long m = e - (long)row * MM;
  const double x = X[e], k = Ki[l * MM + m];
  const int d = row - x0;
  prec[(2L * x0 + d) * MM + m] = x + k;
  prec[(2L * x0 + L.D[l] + d) * MM + m] = beta * x + k;
}

}  // namespace

extern "C" {

// K[n,m] = SE(x_n, z_m) (+ jitter on n == m); x is N rows with leading dim ldx.
int gpx_se_kernel(int N, int M, int Q, const double *x, int ldx, const double *z, const double *ls,
                  const double *sf, double jitter, double *K, void *stream) {
  if (N < 0 || M < 0 || Q < 1 || ldx < Q) return GPX_EINVAL;
  if (N == 0 || M == 0) return GPX_OK;
  if (!x || !z || !ls || !sf || !K) return GPX_EINVAL;
  hipLaunchKernelGGL(se_kernel, dim3(gpx_cdiv((long)N * M, 256)), dim3(256), 0, (hipStream_t)stream, N,
                     M, Q, x, ldx, z, ls, sf, jitter, K);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

long gpx_kfu_bwd_workspace(int N, int M, int Q) {
  return (long)gpx_cdiv(N, KFU_NB) * M * (2 * Q + 1) + (long)M * (Q + 1);
}

// Adjoint of K(x, Z): dsf2 (w.r.t. sf2), dell (w.r.t. ell), dZ (M x Q), dx (N x Q, optional).
int gpx_kfu_bwd(int N, int M, int Q, const double *dK, const double *K, const double *x,
                const double *z, const double *ls, const double *sf, double *dsf2, double *dell,
                double *dZ, double *dx, double *work, void *stream) {
  if (N < 1 || M < 1 || M > 256 || Q < 1 || Q > 16) return GPX_EINVAL;
  if (!dK || !K || !x || !z || !ls || !sf || !dsf2 || !dell || !dZ || !work) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nblk = gpx_cdiv(N, KFU_NB);
  double *red = work + (long)nblk * M * (2 * Q + 1);
  hipLaunchKernelGGL(kfu_bwd_partial, dim3(nblk), dim3(256), 0, st, N, M, Q, dK, K, x, z, ls, work, dx);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(kfu_bwd_final, dim3(gpx_cdiv(M, 256 / (2 * Q + 1))), dim3(256), 0, st, (int)nblk, M, Q,
                     (const double *)work, z, ls, dZ, red);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(kfu_bwd_scalars, dim3(Q + 1), dim3(64), 0, st, M, Q, (const double *)red, ls, sf, dsf2, dell);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

long gpx_dtrace_workspace(int M, int Q) { return (long)M * (Q + 1); }

// d tr(Mm Kzz) / d(lsf, lls, z) for Kzz = SE(z, z): gsf (1), gls (Q), gz (M x Q).
int gpx_dtrace_kzz(int M, int Q, const double *Mm, const double *z, const double *ls,
                   const double *sf, double *gsf, double *gls, double *gz, double *work,
                   void *stream) {
  if (M < 1 || M > 256 || Q < 1 || Q > 16 || !Mm || !z || !ls || !sf || !gsf || !gls || !gz || !work)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  switch (Q) {
#define GPX_DT(QV)                                                                                        \
  case QV:                                                                                                \
    hipLaunchKernelGGL((dtrace_kernel<QV>), dim3(gpx_cdiv(M, 4)), dim3(256), 0, st, M, Mm, z, ls, sf, gz, work); \
    break;
    GPX_DT(1) GPX_DT(2) GPX_DT(3) GPX_DT(4) GPX_DT(5) GPX_DT(6) GPX_DT(7) GPX_DT(8)
    GPX_DT(9) GPX_DT(10) GPX_DT(11) GPX_DT(12) GPX_DT(13) GPX_DT(14) GPX_DT(15) GPX_DT(16)
#undef GPX_DT
  }
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(dtrace_final, dim3(Q + 1), dim3(64), 0, st, M, Q, (const double *)work, gsf, gls);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

// out[j] = alpha * sum_i P[i, j] + beta * out[j]
int gpx_colsum(int rows, int cols, const double *P, double alpha, double beta, double *out, void *stream) {
  if (rows < 0 || cols < 0 || !P || !out) return GPX_EINVAL;
  if (cols == 0) return GPX_OK;
  if (rows >= 512 && cols <= 1024) {
    hipLaunchKernelGGL(colsum_tall_kernel, dim3(cols), dim3(256), 0, (hipStream_t)stream, rows, cols, P, alpha,
                       beta, out);
    GPX_CHECK_LAUNCH();
    return GPX_OK;
  }
  hipLaunchKernelGGL(colsum_kernel, dim3(gpx_cdiv(cols, 256)), dim3(256), 0, (hipStream_t)stream, rows, cols,
                     P, alpha, beta, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

// Chunks of the two-stage column sum and its workspace (doubles).
static int colsum_chunks(int rows) {
  int c = (rows + 2047) / 2048;
  return c > 256 ? 256 : c;
}

long gpx_colsum_workspace(int rows, int cols) {
  if (rows < 0 || cols < 0) return -1;
  return (long)colsum_chunks(rows) * cols;
}

// out[j] = alpha sum_i P[i, j] + beta out[j] for tall P: chunked partial sums into
// `work`, then their fixed-order sum; the same result for the same input and shape.
int gpx_colsum_tall(int rows, int cols, const double *P, double alpha, double beta, double *out, double *work,
                    void *stream) {
  if (rows < 0 || cols < 0 || cols > 65535 || !P || !out || !work) return GPX_EINVAL;
  if (cols == 0) return GPX_OK;
  const int nc = colsum_chunks(rows > 0 ? rows : 1);
  const int chunk = (rows + nc - 1) / nc;
  hipLaunchKernelGGL(colsum_chunk_kernel, dim3(cols, nc), dim3(256), 0, (hipStream_t)stream, rows, cols, chunk, P,
                     work);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(colsum_kernel, dim3(gpx_cdiv(cols, 256)), dim3(256), 0, (hipStream_t)stream, nc, cols,
                     (const double *)work, alpha, beta, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_eta_to_R(int D, int M, const double *eta, double *R, void *stream) {
  if (D < 0 || M < 1 || !eta || !R) return GPX_EINVAL;
  long n = (long)D * M * M;
  if (n == 0) return GPX_OK;
  hipLaunchKernelGGL(eta_to_R_kernel, dim3(gpx_cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, D, M, eta, R);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_kuu_batched(int nl, int M, const double *const *z, const int *Q, const double *const *ls,
                    const double *const *sf, double jitter, double *K, void *stream) {
  if (nl < 1 || nl > kKuuLayers || M < 1 || !z || !Q || !ls || !sf || !K) return GPX_EINVAL;
  KuuLayers L = {};
  for (int l = 0; l < nl; ++l) {
    if (!z[l] || !ls[l] || !sf[l] || Q[l] < 1) return GPX_EINVAL;
    L.z[l] = z[l];
    L.ls[l] = ls[l];
    L.sf[l] = sf[l];
    L.Q[l] = Q[l];
  }
  hipLaunchKernelGGL(kuu_batched_kernel, dim3(gpx_cdiv((long)nl * M * M, 256)), dim3(256), 0, (hipStream_t)stream, nl,
                     L, M, jitter, K);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_prec_form(int nl, const int *D, int M, const double *X, const double *Ki, double beta, double *prec,
                  void *stream) {
  if (nl < 1 || nl > kPrecLayers || !D || M < 1 || !X || !Ki || !prec) return GPX_EINVAL;
  PrecLayers L = {};
  long rows = 0;
  for (int l = 0; l < nl; ++l) {
    if (D[l] < 1) return GPX_EINVAL;
    L.D[l] = D[l];
    rows += D[l];
  }
  hipLaunchKernelGGL(prec_form_kernel, dim3(gpx_cdiv(rows * M * M, 256)), dim3(256), 0, (hipStream_t)stream, nl, L,
                     M, X, Ki, beta, prec);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_pack_dR(int D, int M, const double *R, const double *G, double *out, void *stream) {
  if (D < 0 || M < 1 || !R || !G || !out) return GPX_EINVAL;
  long n = (long)D * M * (M + 1) / 2;
  if (n == 0) return GPX_OK;
  if (D > 65535) return GPX_EINVAL;
  const int MT = (M + 15) / 16;
  hipLaunchKernelGGL(pack_dR_kernel, dim3(gpx_cdiv(MT * (MT + 1) / 2, 4), D), dim3(256), 0, (hipStream_t)stream,
                     D, M, R, G, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

// out[i] = a x[i % nx] + b y[i % ny] + c w[i % nw]  (y, w optional); out may alias inputs.
int gpx_lincomb(long n, double a, const double *x, long nx, double b, const double *y, long ny,
                double c, const double *w, long nw, double *out, void *stream) {
  if (n < 0 || !x || nx < 1 || !out || (y && ny < 1) || (w && nw < 1)) return GPX_EINVAL;
  if (n == 0) return GPX_OK;
  unsigned g = gpx_cdiv(n, 256);
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(lincomb_kernel, dim3(g), dim3(256), 0, (hipStream_t)stream, n, a, x, nx, b, y, ny, c, w,
                     nw, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

// out[b] = a (X[b] + X[b]^T); out must not alias X.
int gpx_sym(int batch, int M, double a, const double *X, double *out, void *stream) {
  if (batch < 0 || M < 0 || !X || !out || X == out) return GPX_EINVAL;
  long n = (long)batch * M * M;
  if (n == 0) return GPX_OK;
  hipLaunchKernelGGL(sym_kernel, dim3(gpx_cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, batch, M, a, X, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_det_var(int N, int D, int M, const double *T, const double *kfu, const double *sf, double *v,
                void *stream) {
  if (N < 0 || D < 0 || M < 1 || !T || !kfu || !sf || !v) return GPX_EINVAL;
  long n = (long)N * D;
  if (n == 0) return GPX_OK;
  hipLaunchKernelGGL(det_var_kernel, dim3(gpx_cdiv(n, 4)), dim3(256), 0, (hipStream_t)stream, N, D, M, T, kfu,
                     sf, v);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_mm_var(long n, const double *m, const double *Bp, const double *sf, double *v, void *stream) {
  if (n < 0 || !m || !Bp || !sf || !v) return GPX_EINVAL;
  if (n == 0) return GPX_OK;
  hipLaunchKernelGGL(mm_var_kernel, dim3(gpx_cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, m, Bp, sf, v);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_batched_dot(int batch, int n, const double *x, const double *y, double *out, void *stream) {
  if (batch < 0 || n < 0 || !x || !y || !out) return GPX_EINVAL;
  if (batch == 0) return GPX_OK;
  hipLaunchKernelGGL(batched_dot_kernel, dim3(batch), dim3(256), 0, (hipStream_t)stream, batch, n, x, y, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

// scal needs 2 + 2 * 64 doubles: [logZ, dsn] followed by per-block partials.
int gpx_gauss_tilted(int B, int D, const double *m, const double *v, const double *y, const double *sn,
                     double alpha, double scale, double *dm, double *dv, double *dmall, double *scal,
                     void *stream) {
  if (B < 1 || D < 1 || !(alpha > 0) || !m || !v || !y || !sn || !dm || !dv || !scal) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  long n = (long)B * D;
  int nblk = (int)gpx_cdiv(n, 1024);
  if (nblk > 64) nblk = 64;
  hipLaunchKernelGGL(gauss_tilted_kernel, dim3(nblk), dim3(256), 0, st, B, D, m, v, y, sn, alpha, scale, dm, dv,
                     dmall, scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(gauss_tilted_finish, dim3(1), dim3(64), 0, st, nblk, B, D, (const double *)(scal + 2), sn,
                     alpha, scale, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_transition_tilted(int B, int D, const double *mp, const double *vp, const double *mt,
                          const double *vt, const double *sn, double alpha, double scale,
                          double *dmall, double *dvp, double *dmt, double *dvt, double *scal,
                          void *stream) {
  if (B < 1 || D < 1 || !(alpha > 0) || !mp || !vp || !mt || !vt || !sn || !dmall || !dvp || !dmt ||
      !dvt || !scal)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  long n = (long)B * D;
  int nblk = (int)gpx_cdiv(n, 1024);
  if (nblk > 64) nblk = 64;
  hipLaunchKernelGGL(transition_kernel, dim3(nblk), dim3(256), 0, st, B, D, mp, vp, mt, vt, sn, alpha, scale, dmall,
                     dvp, dmt, dvt, scal + 2);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(transition_finish, dim3(1), dim3(64), 0, st, nblk, B, D, (const double *)(scal + 2), sn,
                     alpha, scale, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

long gpx_emission_workspace(int Nb, int Dout, int Din) {
  return (long)gpx_cdiv(Nb, EM_T) * (Dout * Din + Dout + 1);
}

int gpx_emission_tilted(int Nb, int Dout, int Din, const double *C, const double *Rp, const double *y,
                        int ldy, const double *mx, int ldm, const double *vx, int ldv, double alpha,
                        double scale, double *dmx, double *dvx, double *dC, double *dR, double *scal,
                        double *work, void *stream) {
  if (Nb < 1 || Dout < 1 || Dout > EM_DMAX || Din < 1 || !(alpha > 0)) return GPX_EINVAL;
  if (!C || !Rp || !y || !mx || !vx || !dmx || !dvx || !dC || !dR || !scal || !work) return GPX_EINVAL;
  if (Dout * Din > EM_DMAX * EM_DMAX || Dout * Din + Dout + 2 > EM_DMAX * EM_DMAX + EM_DMAX + 2)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nblk = gpx_cdiv(Nb, EM_T);
#define GPX_EMIS(DTV)                                                                                         \
  hipLaunchKernelGGL((emission_kernel<DTV>), dim3(nblk), dim3(EM_T), 0, st, Nb, Din, C, Rp, y, ldy, mx, ldm, vx, \
                     ldv, alpha, scale, dmx, dvx, work)
  switch (Dout) {
    case 1: GPX_EMIS(1); break;
    case 2: GPX_EMIS(2); break;
    case 3: GPX_EMIS(3); break;
    case 4: GPX_EMIS(4); break;
    case 5: GPX_EMIS(5); break;
    case 6: GPX_EMIS(6); break;
    case 7: GPX_EMIS(7); break;
    default: GPX_EMIS(8); break;
  }
#undef GPX_EMIS
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(emission_final, dim3(1), dim3(256), 0, st, (int)nblk, Nb, Dout, Din, (const double *)work,
                     Rp, alpha, scale, dC, dR, scal);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_rowscale(int N, int D, int M, const double *w, const double *X, double *out, void *stream) {
  if (N < 0 || D < 0 || M < 0 || !w || !X || !out) return GPX_EINVAL;
  long n = (long)N * D * M;
  if (n == 0) return GPX_OK;
  hipLaunchKernelGGL(rowscale_kernel, dim3(gpx_cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, N, D, M, w, X, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_wsum(int N, int D, int M, double alpha, const double *w, const double *T, double beta,
             double *out, void *stream) {
  if (N < 0 || D < 0 || M < 0 || !w || !T || !out) return GPX_EINVAL;
  long n = (long)N * M;
  if (n == 0) return GPX_OK;
  hipLaunchKernelGGL(wsum_kernel, dim3(gpx_cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, N, D, M, alpha, w, T,
                     beta, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_dmall(long n, const double *dm, const double *dv, const double *m, double *out, void *stream) {
  if (n < 0 || !dm || !dv || !m || !out) return GPX_EINVAL;
  if (n == 0) return GPX_OK;
  hipLaunchKernelGGL(dmall_kernel, dim3(gpx_cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, dm, dv, m, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_hyper_grads(int Q, const double *ls, const double *sf, const double *dell, const double *gls,
                    const double *dsf2, const double *sumdv, const double *gsf, double *dls, double *dsf,
                    void *stream) {
  if (Q < 1 || Q > 1024 || !ls || !sf || !dell || !gls || !dsf2 || !sumdv || !gsf || !dls || !dsf)
    return GPX_EINVAL;
  hipLaunchKernelGGL(hyper_kernel, dim3(1), dim3(Q < 64 ? 64 : ((Q + 63) / 64) * 64), 0, (hipStream_t)stream, Q,
                     ls, sf, dell, gls, dsf2, sumdv, gsf, dls, dsf);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_phi(int D, const double *ldK, const double *ldP, double sgnP, const double *dotP,
            const double *ldC, double sgnC, const double *dotC, double sp, double sc, double *out,
            void *stream) {
  if (D < 1 || !ldK || !ldP || !dotP || !ldC || !dotC || !out) return GPX_EINVAL;
  hipLaunchKernelGGL(phi_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, D, ldK, ldP, sgnP, dotP, ldC, sgnC,
                     dotC, sp, sc, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_phi_nat(int D, int M, const double *ldK, const double *ldP, const double *ldC, const double *mu,
                const double *muhat, const double *theta2, double beta, double sp, double sc, double *out,
                void *stream) {
  if (D < 1 || M < 1 || !ldK || !ldP || !ldC || !mu || !muhat || !theta2 || !out) return GPX_EINVAL;
  hipLaunchKernelGGL(phi_nat_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, D, M, ldK, ldP, ldC, mu,
                     muhat, theta2, beta, sp, sc, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_emission_predict(int Nb, int Dout, int Din, const double *C, const double *Rp, const double *mf,
                         const double *vf, double *my, double *vy, void *stream) {
  if (Nb < 0 || Dout < 1 || Din < 1 || !C || !Rp || !mf || !vf || !my || !vy) return GPX_EINVAL;
  if (Nb == 0) return GPX_OK;
  hipLaunchKernelGGL(emission_predict_kernel, dim3(gpx_cdiv((long)Nb * Dout, 256)), dim3(256), 0,
                     (hipStream_t)stream, Nb, Dout, Din, C, Rp, mf, vf, my, vy);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

}  // extern "C"
