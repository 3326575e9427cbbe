// psi statistics of the ARD SE kernel under Gaussian inputs and their adjoint.
//
// Forward: compute_psi_weave (kernels.py:181-240, loop :201-234)
//   psi1[n,m]   = sf2 prod_q (l/(l+v))^1/2 exp(-(mu-z_m)^2 / (2(l+v)))
//   psi2[n,a,b] = sf2^2 prod_q (l/(l+2v))^1/2
//                 exp(-(z_a-z_b)^2/(4l) - (mu-(z_a+z_b)/2)^2/(2v+l))
// with l = ell^2 = exp(2 ls) and sf2 = exp(2 sf).  psi2 is written exactly
// symmetric: every value is computed once and stored at (a,b) and (b,a).
//
// Adjoint: compute_psi_derivatives -> psi1compDer / psi2compDer (kernels.py:302-309,
// 355-444), restated with per-point sufficient statistics.  With
// L[n,a,b] = dpsi2[n,a,b] psi2[n,a,b] (dpsi2 symmetric), T_na = sum_b L_nab,
// V_nq = sum_ab L_nab z_aq z_bq, S0 = sum_a T_a, S1_q = sum_a T_a z_aq,
// S2_q = (sum_a T_a z_aq^2 + V_q)/2 and d = 1/(2v+l):
//   dmu = -2 d (mu S0 - S1),  dv = 2 d^2 (mu^2 S0 - 2 mu S1 + S2) - d S0,
//   dell (per n) = 2 ell [(v/l) d S0 + d^2 (mu^2 S0 - 2 mu S1 + S2)
//                         + (sum_a T_a z_aq^2 - V_q) / (2 l^2)],
//   dZ_aq = -(z_aq R_a - (Lsum z)_aq)/l + 2 P_aq - z_aq R'_aq - sum_b W_abq z_bq,
// where Lsum = sum_n L_n, R = rowsum(Lsum), W_abq = sum_n L_nab d_nq,
// P_aq = sum_n T_na d_nq mu_nq, R'_aq = sum_n T_na d_nq (the n-reductions run as
// MFMA GEMMs on the host side of this file's entry points).
#include "gpx_common.h"

#include <cstdint>

#define GPX_QMAX 16

namespace {

// ---------------------------------------------------------------------------
// forward
// ---------------------------------------------------------------------------

// PSI1_PTS points per block: their constants (w = 1 / (2 (v + l)) per q, and
// lc = sum_q log(l / (l + v)) / 2) are formed once in LDS, then thread (point, m) sums
// its Q terms; a row of M entries is written contiguously.
constexpr int PSI1_PTS = 4;

__global__ __launch_bounds__(256) void psi1_kernel(int N, int M, int Q, const double *__restrict__ mx,
                                                    const double *__restrict__ vx,
                                                    const double *__restrict__ z,
                                                    const double *__restrict__ ls,
                                                    const double *__restrict__ sf,
                                                    double *__restrict__ psi1) {
  __shared__ double lam[GPX_QMAX], pw[PSI1_PTS][GPX_QMAX], pm[PSI1_PTS][GPX_QMAX], plc[PSI1_PTS];
  const int tid = threadIdx.x, n0 = blockIdx.x * PSI1_PTS;
  if (tid < Q) lam[tid] = exp(2.0 * ls[tid]);
  __syncthreads();
  if (tid < PSI1_PTS * Q) {
    const int pt = tid / Q, q = tid % Q, n = min(n0 + pt, N - 1);
    const double v = vx[(long)n * Q + q];
    pw[pt][q] = 1.0 / (2.0 * (v + lam[q]));
    pm[pt][q] = mx[(long)n * Q + q];
  }
  if (tid < PSI1_PTS) {
    const int n = min(n0 + tid, N - 1);
    double lc = 0.0;
    for (int q = 0; q < Q; ++q) lc += 0.5 * log(lam[q] / (lam[q] + vx[(long)n * Q + q]));
    plc[tid] = lc;
  }
  __syncthreads();
  const double sf2 = exp(2.0 * sf[0]);
  for (int e = tid; e < PSI1_PTS * M; e += 256) {
    const int pt = e / M, m = e % M, n = n0 + pt;
    if (n >= N) break;
    double acc = 0.0;
    for (int q = 0; q < Q; ++q) {
      const double d = pm[pt][q] - z[m * Q + q];
      acc -= d * d * pw[pt][q];
    }
    psi1[(long)n * M + m] = sf2 * exp(acc + plc[pt]);
  }
}

// One block per (32 x 32 tile pair ta >= tb, chunk of PSI2_CH points).  Wave w computes
// the 16 x 16 block (w >> 1, w & 1) of the tile for every point of the chunk, lane l the
// entries (row (l >> 4) + 4r, column l & 15), r < 4.  The per-pair terms zh_q = (z_aq +
// z_bq)/2 and P_ab = sum_q (z_aq - z_bq)^2 / (4 l_q) stay in registers across the points,
// the points' constants are formed once per block in LDS, and an off-diagonal tile's
// mirrored block goes out through a wave-private LDS transpose: after the set-up no wave
// waits for another, and every store writes 128-byte row segments.
// QC: compile-time bound on Q (register arrays).
constexpr int PSI2_CH = 16;

template <int QC>
__global__ __launch_bounds__(256) void psi2_kernel(int N, int M, int Q, const double *__restrict__ mx,
                                                    const double *__restrict__ vx,
                                                    const double *__restrict__ z,
                                                    const double *__restrict__ ls,
                                                    const double *__restrict__ sf,
                                                    double *__restrict__ psi1,
                                                    double *__restrict__ psi2) {
  __shared__ double pmu[PSI2_CH][QC], pdd[PSI2_CH][QC], plc[PSI2_CH];
  __shared__ double pw1[PSI2_CH][QC], plc1[PSI2_CH];  // psi1 constants (tile pair 0 only)
  __shared__ double lam[QC], i4l[QC];
  __shared__ double tr[4][16][17];
  int p = blockIdx.x, ta = 0;
  while (p > ta) { p -= ta + 1; ++ta; }
  const int tb = p;
  const int a0 = ta * 32, b0 = tb * 32, tid = threadIdx.x;
  const int n0 = blockIdx.y * PSI2_CH, nend = min(N, n0 + PSI2_CH);
  if (tid < Q) {
    const double l = exp(2.0 * ls[tid]);
    lam[tid] = l;
    i4l[tid] = 1.0 / (4.0 * l);
  }
  __syncthreads();
  // point constants: mu, d = 1/(2v + l) (zero beyond Q, so the loops below run to QC
  // without branches) and lc = sum_q log(l / (l + 2v)) / 2
  for (int e = tid; e < PSI2_CH * QC; e += 256) {
    const int pt = e / QC, q = e % QC, n = min(n0 + pt, N - 1);
    const bool in = q < Q;
    pmu[pt][q] = in ? mx[(long)n * Q + q] : 0.0;
    pdd[pt][q] = in ? 1.0 / (2.0 * vx[(long)n * Q + q] + lam[q]) : 0.0;
  }
  if (tid < PSI2_CH) {
    const int n = min(n0 + tid, N - 1);
    double lc = 0.0;
    for (int q = 0; q < Q; ++q) lc += 0.5 * log(lam[q] / (lam[q] + 2.0 * vx[(long)n * Q + q]));
    plc[tid] = lc;
  }
  // the blocks of tile pair 0 also write psi1 for their points (one launch for both)
  const bool do_psi1 = blockIdx.x == 0;
  if (do_psi1) {
    for (int e = tid; e < PSI2_CH * Q; e += 256) {
      const int pt = e / Q, q = e % Q, n = min(n0 + pt, N - 1);
      pw1[pt][q] = 1.0 / (2.0 * (vx[(long)n * Q + q] + lam[q]));
    }
    if (tid >= 64 && tid < 64 + PSI2_CH) {
      const int pt = tid - 64, n = min(n0 + pt, N - 1);
      double lc = 0.0;
      for (int q = 0; q < Q; ++q) lc += 0.5 * log(lam[q] / (lam[q] + vx[(long)n * Q + q]));
      plc1[pt] = lc;
    }
  }
  __syncthreads();
  const double sf2 = exp(2.0 * sf[0]), sf4 = sf2 * sf2;
  if (do_psi1) {
    for (int e = tid; e < PSI2_CH * M; e += 256) {
      const int pt = e / M, m = e % M, n = n0 + pt;
      if (n >= nend) break;
      double acc = 0.0;
      for (int q = 0; q < Q; ++q) {
        const double d = pmu[pt][q] - z[m * Q + q];
        acc -= d * d * pw1[pt][q];
      }
      psi1[(long)n * M + m] = sf2 * exp(acc + plc1[pt]);
    }
  }
  const int lane = tid & 63, w = tid >> 6;
  const int ra = (w >> 1) * 16 + (lane >> 4), cb = (w & 1) * 16 + (lane & 15);
  const int b = b0 + cb, bl = min(b, M - 1);
  double zh[4][QC], P[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int al = min(a0 + ra + 4 * r, M - 1);  // padding rows read a real z row; never stored
    double pp = 0.0;
#pragma unroll
    for (int q = 0; q < QC; ++q) {
      if (q < Q) {
        const double x = z[al * Q + q], y = z[bl * Q + q];
        const double dz = x - y;
        zh[r][q] = 0.5 * (x + y);
        pp += dz * dz * i4l[q];
      } else {
        zh[r][q] = 0.0;
      }
    }
    P[r] = pp;
  }
  const bool mirror = ta != tb;
  // mirrored block: lane l writes row b0 + (w & 1) * 16 + (l >> 4) + 4r, columns
  // a0 + (w >> 1) * 16 + (l & 15)
  const int mb = b0 + (w & 1) * 16 + (lane >> 4), ma = a0 + (w >> 1) * 16 + (lane & 15);
  for (int n = n0; n < nend; ++n) {
    const int pt = n - n0;
    const double lc = plc[pt];
    double *out = psi2 + (long)n * M * M;
    double val[4], m[QC], d[QC];
#pragma unroll
    for (int q = 0; q < QC; ++q) {
      m[q] = pmu[pt][q];
      d[q] = pdd[pt][q];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double e = lc - P[r];
#pragma unroll
      for (int q = 0; q < QC; ++q) {
        const double h = m[q] - zh[r][q];
        e -= h * h * d[q];
      }
      val[r] = sf4 * exp(e);
      const int a = a0 + ra + 4 * r;
      if (a < M && b < M) __builtin_nontemporal_store(val[r], out + (long)a * M + b);
    }
    if (mirror) {
#pragma unroll
      for (int r = 0; r < 4; ++r) tr[w][(lane >> 4) + 4 * r][lane & 15] = val[r];
      // the wave reads what its other lanes wrote: order the LDS accesses at wave scope
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int bb = mb + 4 * r;
        const double v = tr[w][lane & 15][(lane >> 4) + 4 * r];
        if (bb < M && ma < M) __builtin_nontemporal_store(v, out + (long)bb * M + ma);
      }
      // and before the next point overwrites the buffer
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

// ---------------------------------------------------------------------------
// adjoint, per-point pass (one block per n, thread b owns column b of L_n)
// ---------------------------------------------------------------------------

constexpr int NRED = 5 * GPX_QMAX + 2;

__global__ __launch_bounds__(256) void psi_bwd_rows_kernel(
    int N, int M, int Q, const double *__restrict__ mx, const double *__restrict__ vx,
    const double *__restrict__ z, const double *__restrict__ ls,
    const double *__restrict__ dpsi1, const double *__restrict__ psi1,
    const double *__restrict__ dpsi2, const double *__restrict__ psi2,
    double *__restrict__ dmx, double *__restrict__ dvx,
    double *__restrict__ Lout,    // N x M x M
    double *__restrict__ Tout,    // N x M
    double *__restrict__ Lpout,   // N x M
    double *__restrict__ Hout,    // N x Q   per-point d/d ell contribution
    double *__restrict__ s0out,   // N x 2   (S0, sum Lp)
    double *__restrict__ Dn,      // N x (Q+1): [1 | d]
    double *__restrict__ Pn,      // N x 2Q : [d*mu | d]
    double *__restrict__ En) {    // N x 2Q : [e | e*mu]
  __shared__ double zs[256][GPX_QMAX];
  __shared__ double red[4][NRED];
  __shared__ double tot[NRED];
  const int n = blockIdx.x, b = threadIdx.x, lane = b & 63, w = b >> 6;
  for (int e = b; e < M * Q; e += 256) zs[e / Q][e % Q] = z[e];
  __syncthreads();
  double mu[GPX_QMAX];
#pragma unroll
  for (int q = 0; q < GPX_QMAX; ++q) mu[q] = (q < Q) ? mx[(long)n * Q + q] : 0.0;
  double T = 0.0, U[GPX_QMAX];
#pragma unroll
  for (int q = 0; q < GPX_QMAX; ++q) U[q] = 0.0;
  double Lp = 0.0;
  if (b < M) {
    const double *P2 = psi2 + (long)n * M * M + b;
    const double *G2 = dpsi2 + (long)n * M * M + b;
    double *Lo = Lout ? Lout + (long)n * M * M + b : nullptr;
    for (int a = 0; a < M; ++a) {
      const double l = G2[(long)a * M] * P2[(long)a * M];
      if (Lo) Lo[(long)a * M] = l;
      T += l;
#pragma unroll
      for (int q = 0; q < GPX_QMAX; ++q)
        if (q < Q) U[q] = fma(l, zs[a][q], U[q]);
    }
    Lp = dpsi1[(long)n * M + b] * psi1[(long)n * M + b];
    Tout[(long)n * M + b] = T;
    Lpout[(long)n * M + b] = Lp;
  }
  // per-thread contributions, reduced over b:
  // [0] S0, [1] sum Lp, [2+q] S1, [2+Q+q] sum T z^2, [2+2Q+q] V,
  // [2+3Q+q] B1 = sum Lp (z-mu), [2+4Q+q] C1 = sum Lp (z-mu)^2
  double c[NRED];
#pragma unroll
  for (int i = 0; i < NRED; ++i) c[i] = 0.0;
  c[0] = T;
  c[1] = Lp;
  if (b < M) {
#pragma unroll
    for (int q = 0; q < GPX_QMAX; ++q)
      if (q < Q) {
        const double zq = zs[b][q], zm = zq - mu[q];
        c[2 + q] = T * zq;
        c[2 + GPX_QMAX + q] = T * zq * zq;
        c[2 + 2 * GPX_QMAX + q] = U[q] * zq;
        c[2 + 3 * GPX_QMAX + q] = Lp * zm;
        c[2 + 4 * GPX_QMAX + q] = Lp * zm * zm;
      }
  }
#pragma unroll
  for (int i = 0; i < NRED; ++i) {
    if (i >= 2 && (i - 2) % GPX_QMAX >= Q) continue;  // unused q slots
    const double s = gpx_wave_sum(c[i]);
    if (lane == 0) red[w][i] = s;
  }
  __syncthreads();
  for (int i = b; i < NRED; i += 256) tot[i] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
  __syncthreads();
  if (b < Q) {
    const int q = b;
    const double S0 = tot[0], A1 = tot[1];
    const double S1 = tot[2 + q], TZ2 = tot[2 + GPX_QMAX + q], V = tot[2 + 2 * GPX_QMAX + q];
    const double B1 = tot[2 + 3 * GPX_QMAX + q], C1 = tot[2 + 4 * GPX_QMAX + q];
    const double l = exp(2.0 * ls[q]);
    const double m = mx[(long)n * Q + q], vv = vx[(long)n * Q + q];
    const double ell = exp(ls[q]);
    const double d = 1.0 / (2.0 * vv + l), e = 1.0 / (vv + l);
    const double S2 = 0.5 * (TZ2 + V);
    const double quad = m * m * S0 - 2.0 * m * S1 + S2;
    const double dmu = -2.0 * d * (m * S0 - S1) + e * B1;
    const double dvv = 2.0 * d * d * quad - d * S0 + 0.5 * e * (e * C1 - A1);
    const double h2 = 2.0 * ell * ((vv / l) * d * S0 + d * d * quad + (TZ2 - V) / (2.0 * l * l));
    const double h1 = e * ell * (e * C1 + (vv / l) * A1);
    dmx[(long)n * Q + q] = dmu;
    dvx[(long)n * Q + q] = dvv;
    Hout[(long)n * Q + q] = h1 + h2;
    Dn[(long)n * (Q + 1) + 1 + q] = d;
    Pn[(long)n * 2 * Q + q] = d * m;
    Pn[(long)n * 2 * Q + Q + q] = d;
    En[(long)n * 2 * Q + q] = e;
    En[(long)n * 2 * Q + Q + q] = e * m;
  }
  if (b == 0) {
    Dn[(long)n * (Q + 1)] = 1.0;
    s0out[2 * n] = tot[0];
    s0out[2 * n + 1] = tot[1];
  }
}

// Final assembly for the hyper-parameter adjoints.  One block per inducing point a
// (threads over b), plus block M which finishes the scalar / per-q sums.
//   LW:  M^2 x (Q+1)  [Lsum | W]          (sum_n L_nab [1 | d_nq])
//   TP:  M x 2Q       [P | R']            (sum_n T_na [d mu | d])
//   LE:  M x 2Q       [sum Lp e | sum Lp e mu]
//   colH: Q (sum_n H), cols0: 2 (sum_n S0, sum_n Lp)
__global__ __launch_bounds__(256) void psi_bwd_final_kernel(
    int M, int Q, const double *__restrict__ z, const double *__restrict__ ls,
    const double *__restrict__ sf, const double *__restrict__ LW,
    const double *__restrict__ TP, const double *__restrict__ LE,
    const double *__restrict__ colH, const double *__restrict__ cols0,
    double *__restrict__ dsf2, double *__restrict__ dell, double *__restrict__ dZ) {
  const int a = blockIdx.x, b = threadIdx.x, lane = b & 63, w = b >> 6;
  if (a == M) {
    if (b == 0) {
      const double sf2 = exp(2.0 * sf[0]);
      dsf2[0] = (cols0[1] + 2.0 * cols0[0]) / sf2;
    }
    return;
  }
  // per-a sums over b: R_a, (Lsum z)_aq, sum_b W_abq z_bq
  double c[2 * GPX_QMAX + 1];
#pragma unroll
  for (int i = 0; i < 2 * GPX_QMAX + 1; ++i) c[i] = 0.0;
  for (int bb = b; bb < M; bb += 256) {
    const double *row = LW + ((long)a * M + bb) * (Q + 1);
    const double ls_ab = row[0];
    c[0] += ls_ab;
#pragma unroll
    for (int q = 0; q < GPX_QMAX; ++q)
      if (q < Q) {
        const double zb = z[bb * Q + q];
        c[1 + q] += ls_ab * zb;
        c[1 + GPX_QMAX + q] += row[1 + q] * zb;
      }
  }
  __shared__ double tot[2 * GPX_QMAX + 1];
  __shared__ double red2[4][2 * GPX_QMAX + 1];
#pragma unroll
  for (int i = 0; i < 2 * GPX_QMAX + 1; ++i) {
    const double s = gpx_wave_sum(c[i]);
    if (lane == 0) red2[w][i] = s;
  }
  __syncthreads();
  for (int i = b; i < 2 * GPX_QMAX + 1; i += 256)
    tot[i] = red2[0][i] + red2[1][i] + red2[2][i] + red2[3][i];
  __syncthreads();
  if (b < Q) {
    const int q = b;
    const double l = exp(2.0 * ls[q]);
    const double zaq = z[a * Q + q];
    const double Ra = tot[0], LZ = tot[1 + q], WZ = tot[1 + GPX_QMAX + q];
    const double P = TP[a * 2 * Q + q], Rp = TP[a * 2 * Q + Q + q];
    const double g2 = -(zaq * Ra - LZ) / l + 2.0 * P - zaq * Rp - WZ;
    const double g1 = -(zaq * LE[a * 2 * Q + q] - LE[a * 2 * Q + Q + q]);
    dZ[a * Q + q] = g1 + g2;
  }
}

// out[j] = sum_i P[i*cols + j] over `rows`: one block per column, fixed-order reduction.
__global__ __launch_bounds__(256) void colsum_kernel(int rows, int cols, const double *__restrict__ P,
                                                     double *__restrict__ out) {
  __shared__ double red[16];
  const int j = blockIdx.x;
  double s = 0.0;
#pragma unroll 8
  for (long i = threadIdx.x; i < rows; i += 256) s += P[i * cols + j];
  s = gpx_block_sum(s, red);
  if (threadIdx.x == 0) out[j] = s;
}

}  // namespace

extern "C" int gpx_gemm_batched(int transA, int transB, int m, int n, int k, double alpha,
                                const double *A, int lda, long strideA, const double *B, int ldb,
                                long strideB, double beta, double *C, int ldc, long strideC,
                                int batch, int splits, double *work, int tile, void *stream);

extern "C" {

// Shared tail of the psi adjoint (materialising and fused paths): from the per-point
// outputs T, Lp (N x M), H (N x Q), s0 (N x 2), Dn, Pn, En and the pair sums LW
// ((M*M) x (Q+1)) finish dsf2, dell (w.r.t. ell) and dZ.  TP, LE (M x 2Q) and cols0 (2)
// are workspace.
// Split-K factor of the two N-long reductions TP = T^T Pn, LE = Lp^T En (M x 2Q outputs,
// few tiles) and the doubles of partials they need.
static int tail_splits(int N) {
  int s = N / 128;
  return s < 1 ? 1 : (s > 64 ? 64 : s);
}

long gpx_psi_tail_workspace(int N, int M, int Q) { return 2L * tail_splits(N) * M * 2 * Q; }

int gpx_psi_bwd_tail(int N, int M, int Q, const double *z, const double *ls, const double *sf,
                     const double *T, const double *Lp, const double *H, const double *s0,
                     const double *Dn, const double *Pn, const double *En, const double *LW,
                     double *TP, double *LE, double *cols0, double *dsf2, double *dell, double *dZ,
                     double *kwork, void *stream) {
  if (N < 1 || M < 1 || M > 256 || Q < 1 || Q > GPX_QMAX || !kwork) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const int ks = tail_splits(N);
  // TP = T^T Pn and LE = Lp^T En as one batch of two split-K products: member 1's operands
  // are addressed from member 0's at the element distances between the buffers
  const long sA = ((intptr_t)Lp - (intptr_t)T) / (long)sizeof(double);
  const long sB = ((intptr_t)En - (intptr_t)Pn) / (long)sizeof(double);
  const long sC = ((intptr_t)LE - (intptr_t)TP) / (long)sizeof(double);
  int rc = gpx_gemm_batched(1, 0, M, 2 * Q, N, 1.0, T, M, sA, Pn, 2 * Q, sB, 0.0, TP, 2 * Q, sC, 2, ks, kwork, 0,
                            stream);
  if (rc) return rc;
  hipLaunchKernelGGL(colsum_kernel, dim3(Q), dim3(256), 0, st, N, Q, H, dell);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(colsum_kernel, dim3(2), dim3(256), 0, st, N, 2, s0, cols0);
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(psi_bwd_final_kernel, dim3(M + 1), dim3(256), 0, st, M, Q, z, ls, sf, LW, TP, LE,
                     (const double *)nullptr, (const double *)cols0, dsf2, dell, dZ);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_psi_qmax(void) { return GPX_QMAX; }

// psi1 (N x M) and, if psi2 != NULL, psi2 (N x M x M) for inputs x_n ~ N(mx_n, diag vx_n).
// ls (Q) and sf (1) are the raw log hyper-parameters on the device.
int gpx_psi_fwd(int N, int M, int Q, const double *mx, const double *vx, const double *z,
                const double *ls, const double *sf, double *psi1, double *psi2, void *stream) {
  if (N < 0 || M < 0 || Q < 1 || Q > GPX_QMAX) return GPX_EINVAL;
  if (N == 0 || M == 0) return GPX_OK;
  if (!mx || !vx || !z || !ls || !sf || !psi1) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (!psi2) {
    hipLaunchKernelGGL(psi1_kernel, dim3(gpx_cdiv(N, PSI1_PTS)), dim3(256), 0, st, N, M, Q, mx, vx, z, ls, sf,
                       psi1);
    GPX_CHECK_LAUNCH();
  } else {
    const int T = (M + 31) / 32;
    const unsigned chunks = gpx_cdiv(N, PSI2_CH);
    if (chunks > 65535) return GPX_EINVAL;
    const dim3 grid(T * (T + 1) / 2, chunks);
    if (Q <= 4)
      hipLaunchKernelGGL(psi2_kernel<4>, grid, dim3(256), 0, st, N, M, Q, mx, vx, z, ls, sf, psi1, psi2);
    else if (Q <= 8)
      hipLaunchKernelGGL(psi2_kernel<8>, grid, dim3(256), 0, st, N, M, Q, mx, vx, z, ls, sf, psi1, psi2);
    else
      hipLaunchKernelGGL(psi2_kernel<GPX_QMAX>, grid, dim3(256), 0, st, N, M, Q, mx, vx, z, ls, sf, psi1, psi2);
    GPX_CHECK_LAUNCH();
  }
  return GPX_OK;
}

// Workspace (doubles) for gpx_psi_bwd.
long gpx_psi_bwd_workspace(int N, int M, int Q) {
  long s = (long)N * M * M;                 // L
  s += 2L * N * M;                          // T, Lp
  s += (long)N * Q + 2L * N;                // H, s0
  s += (long)N * (Q + 1) + 4L * N * Q;      // Dn, Pn, En
  s += (long)M * M * (Q + 1) + 4L * M * Q;  // LW, TP, LE
  s += Q + 2;                               // colH, cols0
  s += gpx_psi_tail_workspace(N, M, Q);     // split-K partials of the tail
  return s;
}

// Adjoint of gpx_psi_fwd given dL/dpsi1 (N x M) and a symmetric dL/dpsi2 (N x M x M):
// dsf2 (1, w.r.t. sf2), dell (Q, w.r.t. ell = exp(ls)), dZ (M x Q), dmx, dvx (N x Q).
int gpx_psi_bwd(int N, int M, int Q, const double *mx, const double *vx, const double *z,
                const double *ls, const double *sf, const double *dpsi1, const double *psi1,
                const double *dpsi2, const double *psi2, double *dsf2, double *dell, double *dZ,
                double *dmx, double *dvx, double *work, void *stream) {
  if (N < 1 || M < 1 || M > 256 || Q < 1 || Q > GPX_QMAX) return GPX_EINVAL;
  if (!mx || !vx || !z || !ls || !sf || !dpsi1 || !psi1 || !dpsi2 || !psi2 || !work) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  double *L = work;
  double *T = L + (long)N * M * M;
  double *Lp = T + (long)N * M;
  double *H = Lp + (long)N * M;
  double *s0 = H + (long)N * Q;
  double *Dn = s0 + 2L * N;
  double *Pn = Dn + (long)N * (Q + 1);
  double *En = Pn + 2L * N * Q;
  double *LW = En + 2L * N * Q;
  double *TP = LW + (long)M * M * (Q + 1);
  double *LE = TP + 2L * M * Q;
  double *colH = LE + 2L * M * Q;
  double *cols0 = colH + Q;
  hipLaunchKernelGGL(psi_bwd_rows_kernel, dim3(N), dim3(256), 0, st, N, M, Q, mx, vx, z, ls, dpsi1,
                     psi1, dpsi2, psi2, dmx, dvx, L, T, Lp, H, s0, Dn, Pn, En);
  GPX_CHECK_LAUNCH();
  // LW = L^T [1 | d]  (M^2 x (Q+1)), reduction over n
  int rc = gpx_gemm_batched(1, 0, M * M, Q + 1, N, 1.0, L, M * M, 0, Dn, Q + 1, 0, 0.0, LW, Q + 1, 0, 1,
                            1, nullptr, 0, stream);
  if (rc) return rc;
  return gpx_psi_bwd_tail(N, M, Q, z, ls, sf, T, Lp, H, s0, Dn, Pn, En, LW, TP, LE, cols0, dsf2, dell, dZ, cols0 + 2,
                          stream);
}

}  // extern "C"
