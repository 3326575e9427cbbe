// Fused moment-matched (uncertain-input) layer: psi2 is never written to HBM.
//
// The psi2 exponent is a K = 2Q+2 inner product of per-point and per-pair features,
//   log psi2[n,a,b] - 4 sf = c_n + P_ab + sum_q [ (2 d_nq mu_nq) zh_abq - d_nq zh_abq^2 ]
// with d = 1/(2v+l), zh = (z_a+z_b)/2, c_n = sum_q [log(l/(l+2v))/2 - d mu^2] and
// P_ab = -sum_q (z_a-z_b)^2/(4l) (the expansion of geepee/kernels.py:219-223), so each
// 16 (pairs) x 16 (points) block of exponents is a short FP64 MFMA GEMM, followed by one
// exp per entry on the vector ALUs, and the result feeds the next MFMA contraction in the
// register layout the matrix core produced it in.
//
// Pairs are visited as 16x16 tiles (A <= B) of the M x M square; a tile is processed in
// 16 "row blocks" (fixed a, 16 consecutive b).  Off-diagonal tiles count each unordered
// pair once with weight 2; diagonal tiles hold the full square with weight 1.
//
// Forward (aep_models.py:182-199): v[n,d] = sf2 + sum_ab Bhat[d,a,b] psi2[n,a,b] - m^2.
// Backward (aep_models.py:201-304 with kernels.py:302-444): per point the statistics
// T_na = sum_b L_nab and V_nq = sum_ab L_nab z_aq z_bq of L = dpsi2 * psi2, and over the
// points dBhat[d] = sum_n dv_nd psi2_n, Lsum = sum_n L_n, W_q = sum_n L_n d_nq, where
// dpsi2_n = sum_d dv_nd Bsym[d].
#include "gpx_common.h"

#include <mutex>

#define FQ_MAX 16
#define FM_MAX 256
#define FKS_MAX 9  // ceil((2*FQ_MAX+2)/4)

namespace {

__device__ __forceinline__ void decode_tile(int t, int T, int &A, int &B) {
  A = 0;
  int rem = t;
  while (rem >= T - A) {
    rem -= T - A;
    ++A;
  }
  B = A + rem;
}

// Work items (point block, split) of the forward and per-point passes on a 1-D grid.
// Workgroups are dealt to the 8 XCDs round-robin by linear index; when the split count S
// divides 8, every block on XCD x serves split x mod S, so that XCD's L2 holds only that
// split's share of the pair operands (B-hat, P_ab) rather than all of them.  Returns false
// for the blocks past the last point block (a whole workgroup, before any barrier).
__device__ __forceinline__ bool split_block(int nblk, int S, int &pb, int &split) {
  const int id = blockIdx.x;
  if (8 % S == 0) {
    const int x = id & 7, per = 8 / S;
    split = x % S;
    pb = (id >> 3) * per + x / S;
  } else {
    split = id / nblk;
    pb = id - split * nblk;
  }
  return pb < nblk;
}
inline unsigned split_grid(long nblk, int S) {
  return (8 % S == 0) ? 8u * gpx_cdiv(nblk, 8 / S) : (unsigned)(nblk * S);
}

// For D > 16, and for N >= 1000, the adjoint's two passes run as one kernel
// (fused_bwd_merged_kernel), whose per-point partials have T slots (T) and one slot per
// tile (V) instead of one per column-tile split.  Below that the two-kernel form is kept.
inline bool fused_merged(int N, int D) { return D > 16 || N >= 1000; }
inline int fused_t_slots(int N, int M, int D, int splits) { return fused_merged(N, D) ? (M + 15) / 16 : splits; }
inline int fused_v_slots(int N, int M, int D, int splits) {
  const int T = (M + 15) / 16;
  return fused_merged(N, D) ? T * (T + 1) / 2 : splits;
}

// Per-point exponent feature k (B operand of the exponent GEMM).
__device__ __forceinline__ double point_feature(int k, int Q, const double *mu, const double *d, double c) {
  if (k < Q) return 2.0 * d[k] * mu[k];
  if (k < 2 * Q) return d[k - Q];
  if (k == 2 * Q) return 1.0;
  if (k == 2 * Q + 1) return c;
  return 0.0;
}

// Per-pair exponent feature k (A operand of the exponent GEMM); zs rows in LDS.
__device__ __forceinline__ double pair_feature(int k, int Q, const double *za, const double *zb,
                                               const double *il4) {
  if (k < Q) return 0.5 * (za[k] + zb[k]);
  if (k < 2 * Q) {
    const double h = 0.5 * (za[k - Q] + zb[k - Q]);
    return -h * h;
  }
  if (k == 2 * Q) {
    double p = 0.0;
    for (int q = 0; q < Q; ++q) {
      const double dz = za[q] - zb[q];
      p -= dz * dz * il4[q];
    }
    return p;
  }
  if (k == 2 * Q + 1) return 1.0;
  return 0.0;
}

// Branch-free pair feature for the hot loops.  Feature k = (lane>>4) + 4s is fixed per lane
// and slice, so its kind is set up once (PairSel); each evaluation reads both z rows at a
// clamped index and selects, so no lane's load sits behind a branch (a load under a
// branch is followed by a wait on every outstanding load, prefetches included).
struct PairSel {
  int kq;        // z column read (k, k - Q, or 0 when unused)
  bool lin, sq;  // k < Q: zh;  Q <= k < 2Q: -zh^2
  bool pab;      // k == 2Q: P_ab
  bool one;      // k == 2Q+1: 1, beyond: 0
};

__device__ __forceinline__ PairSel pair_sel(int k, int Q) {
  PairSel p;
  p.kq = k < Q ? k : (k < 2 * Q ? k - Q : 0);
  p.lin = k < Q;
  p.sq = k >= Q && k < 2 * Q;
  p.pab = k == 2 * Q;
  p.one = k == 2 * Q + 1;
  return p;
}

// Same value as pair_feature for the same (k, za, zb), with the k = 2Q entry P_ab read
// from Pab (pair_p_kernel) instead of summed over q: the hot loops of the forward and
// per-point passes evaluate features per row block, and that sum was most of their
// feature work.
__device__ __forceinline__ double pair_feature_sel(const PairSel &p, const double *za, const double *zb,
                                                   double pab) {
  const double h = 0.5 * (za[p.kq] + zb[p.kq]);
  const double f = p.lin ? h : -h * h;
  return (p.lin || p.sq) ? f : (p.pab ? pab : (p.one ? 1.0 : 0.0));
}

// P[a][b] = -sum_q (z_aq - z_bq)^2 / (4 l_q), the pair feature k = 2Q, same operations in
// the same order as pair_feature.
__global__ __launch_bounds__(256) void pair_p_kernel(int M, int Q, const double *__restrict__ z,
                                                     const double *__restrict__ ls, double *__restrict__ P) {
  __shared__ double il4[FQ_MAX];
  if (threadIdx.x < Q) il4[threadIdx.x] = 0.25 / exp(2.0 * ls[threadIdx.x]);
  __syncthreads();
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)M * M) return;
  const int a = (int)(e / M), b = (int)(e % M);
  double p = 0.0;
  for (int q = 0; q < Q; ++q) {
    const double dz = z[a * Q + q] - z[b * Q + q];
    p -= dz * dz * il4[q];
  }
  P[e] = p;
}

// v summed over the four 16-lane rows of the wave (lanes l, l^16, l^32, l^48): the
// xor-shuffle steps 16 then 32 as cross-row lane swaps (same bits, no LDS).
__device__ __forceinline__ double sum_rows4(double v) { return gpx_swap_add<32>(gpx_swap_add<16>(v)); }

// Per-point constants: mu, d = 1/(2v+l), c = sum [log(l/(l+2v))/2 - d mu^2].
__device__ __forceinline__ double point_consts(int n, int N, int Q, const double *mx, const double *vx,
                                               const double *lam, double *mu, double *d) {
  double c = 0.0;
  for (int q = 0; q < Q; ++q) {
    const double m = (n < N) ? mx[(long)n * Q + q] : 0.0;
    const double v = (n < N) ? vx[(long)n * Q + q] : 0.0;
    const double dd = 1.0 / (2.0 * v + lam[q]);
    mu[q] = m;
    d[q] = dd;
    c += 0.5 * log(lam[q] / (lam[q] + 2.0 * v)) - dd * m * m;
  }
  return c;
}

// ---------------------------------------------------------------------------
// forward: part[s][n][d] = sum over the tiles of split s of w * Bp[a,b,d] psi2[n,a,b]
// One block = NSUB x 16 points; the Bp loads of a row block are shared by the NSUB
// point tiles.  Bp is the symmetrised Bhat in pair-major layout (M x M x D) so the D
// values of a pair are contiguous.
// ---------------------------------------------------------------------------
// KSC: the number of 4-wide feature slices, (2Q+2+3)/4, fixed at compile time for the
// common shapes (0: read from Q at run time).
template <int DB, int NSUB, int KSC>
__global__ __launch_bounds__(256) void fused_fwd_kernel(int N, int M, int Q, int D,
                                                        const double *__restrict__ mx,
                                                        const double *__restrict__ vx,
                                                        const double *__restrict__ z,
                                                        const double *__restrict__ ls,
                                                        const double *__restrict__ sf,
                                                        const double *__restrict__ Bp,
                                                        const double *__restrict__ Pab, int S,
                                                        double *__restrict__ part) {
  // DB = 0: D <= 4, the contraction with Bp as per-lane FMAs over the lane's four pair
  // columns (summed over the four lane rows once, at the end) instead of MFMAs that would
  // fill 4 of their 16 output rows
  constexpr bool VS = DB == 0;
  constexpr int DBA = VS ? 1 : DB;
  __shared__ double zs[FM_MAX][FQ_MAX + 1];
  __shared__ double lam[FQ_MAX];
  __shared__ double red[4][DBA * 16][17];
  __shared__ double etab[32];
  int pb, split;
  if (!split_block((N + 16 * NSUB - 1) / (16 * NSUB), S, pb, split)) return;
  const int n0 = pb * 16 * NSUB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < M * Q; e += 256) zs[e / Q][e % Q] = z[e];
  if (tid < Q) lam[tid] = exp(2.0 * ls[tid]);
  gpx_exp_table_init(etab);
  __syncthreads();
  constexpr int KSN = KSC ? KSC : FKS_MAX;
  const int KS = KSC ? KSC : (2 * Q + 2 + 3) / 4;
  double bf[NSUB][KSN];
  bool nvalid[NSUB];
#pragma unroll
  for (int u = 0; u < NSUB; ++u) {
    const int n = n0 + 16 * u + (lane & 15);
    nvalid[u] = n < N;
    double mu[FQ_MAX], dq[FQ_MAX];
    const double c = point_consts(n, N, Q, mx, vx, lam, mu, dq);
#pragma unroll
    for (int s = 0; s < KSN; ++s)
      bf[u][s] = (s < KS) ? point_feature((lane >> 4) + 4 * s, Q, mu, dq, c) : 0.0;
  }
  PairSel ps[KSN];
#pragma unroll
  for (int s = 0; s < KSN; ++s) ps[s] = pair_sel((lane >> 4) + 4 * s, Q);
  const double sf2 = exp(2.0 * sf[0]), sf4 = sf2 * sf2;
  gpx_d4 acc[NSUB][DBA];
#pragma unroll
  for (int u = 0; u < NSUB; ++u)
#pragma unroll
    for (int i = 0; i < DBA; ++i) acc[u][i] = gpx_d4{0.0, 0.0, 0.0, 0.0};
  const int T = (M + 15) / 16, ntiles = T * (T + 1) / 2;
  for (int t = split; t < ntiles; t += S) {
    int A, B;
    decode_tile(t, T, A, B);
    const double w = (A == B) ? sf4 : 2.0 * sf4;
    // column of this lane, clamped for the loads: a padding column's exponent is that of
    // a real pair (finite) and its weight is 0 below
    const int jl = min(16 * B + (lane & 15), M - 1);
    for (int i = wave; i < 16; i += 4) {
      const int a = 16 * A + i;
      if (a >= M) break;
      // unconditional loads (column b and row d clamped; rows d >= D of acc are never
      // stored), issued together and waited for once
      const double pab = Pab[(long)a * M + jl];
      // MFMA form: lane (d = 16 db + lane%16, b = lane/16 + 4r); VS: lane (b = lane/16 + 4r,
      // d < 4), the four rows of lane/16 reading the same values
      double bv[DBA][4], bs4[VS ? 4 : 1][4];
      if constexpr (VS) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int b = min(16 * B + (lane >> 4) + 4 * r, M - 1);
#pragma unroll
          for (int d = 0; d < 4; ++d) bs4[r][d] = Bp[((long)a * M + b) * D + min(d, D - 1)];
        }
      } else {
#pragma unroll
        for (int db = 0; db < DB; ++db)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int d = min(16 * db + (lane & 15), D - 1), b = min(16 * B + (lane >> 4) + 4 * r, M - 1);
            bv[db][r] = Bp[((long)a * M + b) * D + d];
          }
      }
      double af[KSN];
#pragma unroll
      for (int s = 0; s < KSN; ++s) af[s] = (s < KS) ? pair_feature_sel(ps[s], zs[a], zs[jl], pab) : 0.0;
#pragma unroll
      for (int u = 0; u < NSUB; ++u) {
        gpx_d4 e = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KSN; ++s)
          if (s < KS) e = gpx_mfma16(af[s], bf[u][s], e);
        double p[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int b = 16 * B + (lane >> 4) + 4 * r;
          // exp on every lane and a 0/1 weight instead of a branch around it (padding
          // rows/columns have finite exponents: zero pair features, or the features of a
          // point at 0 with zero variance); at NSUB = 4 with a run-time slice count the
          // branch-free form needs more registers than two waves per SIMD leave
          if constexpr (NSUB <= 2 || KSC != 0)
            p[r] = w * gpx_exp_tab(e[r], etab) * ((b < M && nvalid[u]) ? 1.0 : 0.0);
          else
            p[r] = (b < M && nvalid[u]) ? w * gpx_exp_tab(e[r], etab) : 0.0;
        }
        if constexpr (VS) {
          // acc[u][0][d]: this lane's partial of part[n = lane%16][d] over its pair columns
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int d = 0; d < 4; ++d) acc[u][0][d] = fma(bs4[r][d], p[r], acc[u][0][d]);
        } else {
#pragma unroll
          for (int db = 0; db < DB; ++db)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[u][db] = gpx_mfma16(bv[db][r], p[r], acc[u][db]);
        }
      }
    }
  }
#pragma unroll
  for (int u = 0; u < NSUB; ++u) {
    if constexpr (VS) {
      // the four lane rows summed (fixed order), then rows d < 4 of the block sum below
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const double t = gpx_swap_add<16>(gpx_swap_add<32>(acc[u][0][d]));
        if ((lane >> 4) == 0) red[wave][d][lane & 15] = t;
      }
    } else {
#pragma unroll
      for (int db = 0; db < DB; ++db)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave][db * 16 + (lane >> 4) + 4 * r][lane & 15] = acc[u][db][r];
    }
    __syncthreads();
    for (int idx = tid; idx < (VS ? 4 : DB * 16) * 16; idx += 256) {
      const int dd = idx >> 4, nn = idx & 15, n = n0 + 16 * u + nn;
      const double s = red[0][dd][nn] + red[1][dd][nn] + red[2][dd][nn] + red[3][dd][nn];
      if (dd < D && n < N) part[((long)split * N + n) * D + dd] = s;
    }
    __syncthreads();
  }
}

// out[(a*M + b)*D + d] = (B[d,a,b] + B[d,b,a]) / 2
__global__ void sym_pairmajor_kernel(int D, int M, const double *__restrict__ B, double *__restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long MM = (long)M * M;
  if (e >= D * MM) return;
  const long ab = e / D;
  const int d = (int)(e % D), a = (int)(ab / M), b = (int)(ab % M);
  out[e] = 0.5 * (B[d * MM + (long)a * M + b] + B[d * MM + (long)b * M + a]);
}


// ---------------------------------------------------------------------------
// backward, per-point pass ([pair x n] orientation): for 16 points per block and the
// column tiles B of split s, visited column-major with A <= B:
// T_na (row sums, plus column sums of the off-diagonal tiles) and V_nq.  Row a is
// owned by wave (a % 16) % 4, so row sums need no barrier; a column tile's sums are
// kept in registers over its A tiles and merged once, in wave order.
// ---------------------------------------------------------------------------
// DFULL: D == 4 DS, so every slice d = (lane>>4) + 4s is a real column and the B-hat loads
// need no clamp (one address per row block, immediate offsets per slice).
template <int DS, int KSC, bool DFULL>
__global__ __launch_bounds__(256) void fused_bwd_points_kernel(
    int N, int M, int Q, int D, const double *__restrict__ mx, const double *__restrict__ vx,
    const double *__restrict__ z, const double *__restrict__ ls, const double *__restrict__ sf,
    const double *__restrict__ Bp, const double *__restrict__ dv, const double *__restrict__ Pab, int S,
    double *__restrict__ Tpart, double *__restrict__ Vpart) {
  // dynamic LDS, sized by the shape (points_lds_bytes): z rows (M x (Q+1)) and the row
  // sums T (16 x (M+1)); at M = 128, Q = 8 that leaves room for three blocks per CU
  extern __shared__ double dyn[];
  const int ZS = Q + 1, TS = M + 1;
  double *zs = dyn;
  double *Tacc = dyn + M * ZS;
  __shared__ double lam[FQ_MAX];
  __shared__ double cst[4][16][17];  // column-sum staging; reused for the V reduction
  int pb, split;
  if (!split_block((N + 15) / 16, S, pb, split)) return;
  const int n0 = pb * 16;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < M * Q; e += 256) zs[(e / Q) * ZS + e % Q] = z[e];
  for (int e = tid; e < 16 * TS; e += 256) Tacc[e] = 0.0;
  if (tid < Q) lam[tid] = exp(2.0 * ls[tid]);
  __syncthreads();
  const int n = n0 + (lane & 15);
  const bool nvalid = n < N;
  constexpr int KSN = KSC ? KSC : FKS_MAX;
  const int KS = KSC ? KSC : (2 * Q + 2 + 3) / 4;
  double bf[KSN];
  PairSel ps[KSN];
  {
    double mu[FQ_MAX], dq[FQ_MAX];
    const double c = point_consts(n, N, Q, mx, vx, lam, mu, dq);
#pragma unroll
    for (int s = 0; s < KSN; ++s) {
      bf[s] = (s < KS) ? point_feature((lane >> 4) + 4 * s, Q, mu, dq, c) : 0.0;
      ps[s] = pair_sel((lane >> 4) + 4 * s, Q);
    }
  }
  // dv[n][d] for d = (lane>>4) + 4s  (B operand of the G contraction); B-hat is read at
  // d clamped to D - 1, which these zeros cancel
  double dvr[DS];
  int dcl[DS];
#pragma unroll
  for (int s = 0; s < DS; ++s) {
    const int d = (lane >> 4) + 4 * s;
    dvr[s] = (nvalid && d < D) ? dv[(long)n * D + d] : 0.0;
    dcl[s] = DFULL ? d : min(d, D - 1);
  }
  // V operand column: q clamped to Q - 1 (rows q >= Q of V are never stored)
  const int qcl = min(lane & 15, Q - 1);
  const double sf2 = exp(2.0 * sf[0]), sf4 = sf2 * sf2;
  gpx_d4 vacc = gpx_d4{0.0, 0.0, 0.0, 0.0};
  const int T = (M + 15) / 16;
  // Column tiles of this split in a reversed snake order: round r takes the tiles
  // T-1-rS ... T-S-rS, split s the s-th of them in even rounds and the (S-1-s)-th in odd
  // ones.  Tile B holds B + 1 row blocks, so the splits get (nearly) equal work, largest
  // tiles first.
  for (int r = 0; r * S < T; ++r) {
    const int B = T - 1 - (r * S + ((r & 1) ? S - 1 - split : split));
    if (B < 0) continue;  // uniform over the workgroup
    double colp[4] = {0.0, 0.0, 0.0, 0.0};
    const int ja = 16 * B + (lane & 15);
    // loads use the column clamped to M - 1: a padding column's exponent is then that of
    // a real pair (finite), and the 0 weight below removes it
    const int jl = min(ja, M - 1);
    const double *zb = zs + jl * ZS;
    // row blocks (A, i) of this column tile for this wave, flattened as it = 16 A + i with
    // it = wave, wave + 4, ...; B-hat and P_ab of the next row block are fetched during
    // this one (at a row clamped to nit - 1, so every load is in bounds and unconditional)
    const int nit = min((B + 1) * 16, M);
    double bcur[DS], bnxt[DS], pcur, pnxt;
    {
      const int it0 = min(wave, nit - 1);
#pragma unroll
      for (int s = 0; s < DS; ++s) bcur[s] = Bp[((long)it0 * M + jl) * D + dcl[s]];
      pcur = Pab[(long)it0 * M + jl];
    }
    for (int it = wave; it < nit; it += 4) {
      const int A = it >> 4, a = it;
      const bool off = (A != B);
      const double w = off ? 2.0 : 1.0;
      const int itn = min(it + 4, nit - 1);
#pragma unroll
      for (int s = 0; s < DS; ++s) bnxt[s] = Bp[((long)itn * M + jl) * D + dcl[s]];
      pnxt = Pab[(long)itn * M + jl];
      const double *za = zs + a * ZS;
      gpx_d4 e = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KSN; ++s)
        if (s < KS) e = gpx_mfma16(pair_feature_sel(ps[s], za, zb, pcur), bf[s], e);
      gpx_d4 g = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < DS; ++s) g = gpx_mfma16(bcur[s], dvr[s], g);
      double L[4];
      double rs = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = 16 * B + (lane >> 4) + 4 * r;
        const double p = sf4 * exp(e[r]) * ((b < M && nvalid) ? 1.0 : 0.0);  // branch-free, as above
        L[r] = g[r] * p;
        rs += L[r];
        if (off) colp[r] += L[r];
      }
      rs = sum_rows4(rs);
      if ((lane >> 4) == 0) Tacc[(lane & 15) * TS + a] += rs;
      // V[q x n] += w z_aq z_bq L[pair x n]  (padding pairs have L = 0)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = min(16 * B + (lane >> 4) + 4 * r, M - 1);
        const double zz = w * za[qcl] * zs[b * ZS + qcl];
        vacc = gpx_mfma16(zz, L[r], vacc);
      }
#pragma unroll
      for (int s = 0; s < DS; ++s) bcur[s] = bnxt[s];
      pcur = pnxt;
    }
    if (B > 0) {
      // merge the column sums of tile column B: stage per wave, then one thread per
      // (point, column) adds the four waves in order
#pragma unroll
      for (int r = 0; r < 4; ++r) cst[wave][lane & 15][(lane >> 4) + 4 * r] = colp[r];
      __syncthreads();
      {
        const int nn = tid >> 4, bb = tid & 15, b = 16 * B + bb;
        if (b < M) Tacc[nn * TS + b] += cst[0][nn][bb] + cst[1][nn][bb] + cst[2][nn][bb] + cst[3][nn][bb];
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // V: lane holds V[q = (lane>>4) + 4r][n = lane & 15]
#pragma unroll
  for (int r = 0; r < 4; ++r) cst[wave][(lane >> 4) + 4 * r][lane & 15] = vacc[r];
  __syncthreads();
  for (int idx = tid; idx < 16 * 16; idx += 256) {
    const int q = idx >> 4, nn = idx & 15;
    if (q < Q && n0 + nn < N)
      Vpart[((long)split * N + n0 + nn) * Q + q] = cst[0][q][nn] + cst[1][q][nn] + cst[2][q][nn] + cst[3][q][nn];
  }
  for (int e = tid; e < 16 * M; e += 256) {
    const int nn = e / M, a = e % M;
    if (n0 + nn < N) Tpart[((long)split * N + n0 + nn) * M + a] = Tacc[nn * TS + a];
  }
}

// Dynamic LDS of fused_bwd_points_kernel.
inline size_t points_lds_bytes(int M, int Q) { return sizeof(double) * ((size_t)M * (Q + 1) + 16 * (size_t)(M + 1)); }

// Per-point features of the per-pair pass, formed once per launch: for each point n,
// pf[n][k] (k < 4*FKS_MAX, the exponent operand) and dq[n][j] = [1 | d_n | 0...] (16).
constexpr int PF_W = 4 * FKS_MAX + 16;

__global__ __launch_bounds__(256) void fused_point_feat_kernel(int N, int Q, const double *__restrict__ mx,
                                                               const double *__restrict__ vx,
                                                               const double *__restrict__ ls,
                                                               double *__restrict__ pf) {
  __shared__ double lam[FQ_MAX];
  if (threadIdx.x < Q) lam[threadIdx.x] = exp(2.0 * ls[threadIdx.x]);
  __syncthreads();
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  double mu[FQ_MAX], dd[FQ_MAX];
  const double c = point_consts(n, N, Q, mx, vx, lam, mu, dd);
  double *o = pf + (long)n * PF_W;
  for (int k = 0; k < 4 * FKS_MAX; ++k) o[k] = point_feature(k, Q, mu, dd, c);
  o[4 * FKS_MAX] = 1.0;
  for (int q = 0; q < 15; ++q) o[4 * FKS_MAX + 1 + q] = (q < Q) ? dd[q] : 0.0;
}

// ---------------------------------------------------------------------------
// backward, per-pair pass ([n x pair] orientation): each wave owns one row block
// (a, B-tile) and sums over the points of its range:
//   dBhat[d, a, b] += sum_n dv_nd psi2_nab,  LW[a, b, q'] += sum_n [1 | d_nq'] L_nab
// Two forms: for D <= 16 the waves read the precomputed point features and dv straight
// from memory (no LDS, no barriers); for D > 16 a workgroup stages each 16-point block
// in LDS once for its four waves (fused_bwd_pairs_staged).
// ---------------------------------------------------------------------------
template <int DB, int DS, int KSC, int NR>
__device__ __forceinline__ void bwd_pairs_body(
    int N, int M, int Q, int D, const double *__restrict__ z, const double *__restrict__ ls,
    const double *__restrict__ sf, const double *__restrict__ Bp, const double *__restrict__ dv,
    const double *__restrict__ pf, int nsplit, int nchunk, double *__restrict__ dBpart,
    double *__restrict__ LWpart) {
  // A block holds 4 NR row blocks of one tile (16 / (4 NR) blocks per tile); wave w owns
  // rows i = 4 NR h + w + 4 u (u < NR): NR independent MFMA chains that share every point
  // operand the wave loads (NR = 2 halves the point loads, NR = 1 halves the registers).
  static_assert(NR == 1 || NR == 2, "rows per wave");
  constexpr int BPT = 4 / NR;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int T = (M + 15) / 16;
  const int t = blockIdx.x / BPT, h = blockIdx.x % BPT;
  int A, B;
  decode_tile(t, T, A, B);
  int a[NR];
  bool act[NR];
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    a[u] = 16 * A + 4 * NR * h + wave + 4 * u;
    act[u] = a[u] < M;
  }
  if (!act[0]) return;                        // whole wave (a[u] > a[0]); no barriers below
  const int ns = blockIdx.y;
  const int nbeg = ns * nchunk, nend = min(N, nbeg + nchunk);
  constexpr int KSN = KSC ? KSC : FKS_MAX;
  const int KS = KSC ? KSC : (2 * Q + 2 + 3) / 4;
  const int jb = 16 * B + (lane & 15);       // this lane's pair column b
  const bool bvalid = jb < M;
  double il4[FQ_MAX], zb[FQ_MAX];
#pragma unroll
  for (int q = 0; q < FQ_MAX; ++q) {
    il4[q] = (q < Q) ? 0.25 * exp(-2.0 * ls[q]) : 0.0;
    zb[q] = (q < Q && bvalid) ? z[jb * Q + q] : 0.0;
  }
  double af[NR][KSN], bs[NR][DS];
#pragma unroll
  for (int h2 = 0; h2 < NR; ++h2) {
    double za[FQ_MAX];
#pragma unroll
    for (int q = 0; q < FQ_MAX; ++q) za[q] = (q < Q && act[h2]) ? z[a[h2] * Q + q] : 0.0;
#pragma unroll
    for (int s = 0; s < KSN; ++s)
      af[h2][s] = (s < KS && bvalid && act[h2]) ? pair_feature((lane >> 4) + 4 * s, Q, za, zb, il4) : 0.0;
#pragma unroll
    for (int s = 0; s < DS; ++s) {
      const int d = (lane >> 4) + 4 * s;
      bs[h2][s] = (bvalid && act[h2] && d < D) ? Bp[((long)a[h2] * M + jb) * D + d] : 0.0;
    }
  }
  const double sf2 = exp(2.0 * sf[0]), sf4 = sf2 * sf2;
  gpx_d4 dacc[NR][DB], lacc[NR];
#pragma unroll
  for (int h2 = 0; h2 < NR; ++h2) {
    lacc[h2] = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int db = 0; db < DB; ++db) dacc[h2][db] = gpx_d4{0.0, 0.0, 0.0, 0.0};
  }
  for (int n0 = nbeg; n0 < nend; n0 += 16) {
    // Every point operand is loaded unconditionally, at a point index clamped to nend - 1
    // and a column clamped into range, so the loads issue together and are waited for
    // once.  Padding points get weight 0 through p; rows d >= D of dacc and q' > Q of lacc
    // are never stored.
    const int nA = min(n0 + (lane & 15), nend - 1);  // point of this lane as an A-operand row
    double pa[KSN], da[DS];
#pragma unroll
    for (int s = 0; s < KSN; ++s) pa[s] = pf[(long)nA * PF_W + (lane >> 4) + 4 * s];
#pragma unroll
    for (int s = 0; s < DS; ++s) da[s] = dv[(long)nA * D + min((lane >> 4) + 4 * s, D - 1)];
    gpx_d4 e[NR], g[NR];
#pragma unroll
    for (int h2 = 0; h2 < NR; ++h2) {
      e[h2] = gpx_d4{0.0, 0.0, 0.0, 0.0};
      g[h2] = gpx_d4{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int s = 0; s < KSN; ++s)
      if (s < KS) {
#pragma unroll
        for (int h2 = 0; h2 < NR; ++h2) e[h2] = gpx_mfma16(pa[s], af[h2][s], e[h2]);
      }
#pragma unroll
    for (int s = 0; s < DS; ++s) {
#pragma unroll
      for (int h2 = 0; h2 < NR; ++h2) g[h2] = gpx_mfma16(da[s], bs[h2][s], g[h2]);
    }
    // operands of the contractions over the points, loaded while the exps run
    double dvv[4][DB], dq[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int nr = min(n0 + (lane >> 4) + 4 * r, nend - 1);  // point of this k-slice
#pragma unroll
      for (int db = 0; db < DB; ++db) dvv[r][db] = dv[(long)nr * D + min(16 * db + (lane & 15), D - 1)];
      dq[r] = pf[(long)nr * PF_W + 4 * FKS_MAX + (lane & 15)];
    }
    double p[NR][4], L[NR][4];
#pragma unroll
    for (int h2 = 0; h2 < NR; ++h2)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int nn = (lane >> 4) + 4 * r;
        // branch-free: padding entries have finite exponents (zero pair features, or a
        // real point's features), times a 0 weight
        p[h2][r] = sf4 * exp(e[h2][r]) * ((n0 + nn < nend && bvalid && act[h2]) ? 1.0 : 0.0);
        L[h2][r] = g[h2][r] * p[h2][r];
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int db = 0; db < DB; ++db)
#pragma unroll
        for (int h2 = 0; h2 < NR; ++h2) dacc[h2][db] = gpx_mfma16(dvv[r][db], p[h2][r], dacc[h2][db]);
#pragma unroll
      for (int h2 = 0; h2 < NR; ++h2) lacc[h2] = gpx_mfma16(dq[r], L[h2][r], lacc[h2]);
    }
  }
#pragma unroll
  for (int h2 = 0; h2 < NR; ++h2) {
    if (!act[h2]) continue;
    // dacc[db]: lane holds dBhat[d = 16db + (lane>>4) + 4r][pair b = jb]
#pragma unroll
    for (int db = 0; db < DB; ++db)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int d = 16 * db + (lane >> 4) + 4 * r;
        if (d < D && jb < M) dBpart[(((long)ns * D + d) * M + a[h2]) * M + jb] = dacc[h2][db][r];
      }
    // lacc: lane holds LW[q' = (lane>>4) + 4r][pair b = jb]
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int qq = (lane >> 4) + 4 * r;
      if (qq <= Q && jb < M) LWpart[(((long)ns * M + a[h2]) * M + jb) * (Q + 1) + qq] = lacc[h2][r];
    }
  }
}

#define GPX_PAIRS_ARGS                                                                                   \
  int N, int M, int Q, int D, const double *__restrict__ z, const double *__restrict__ ls,                \
      const double *__restrict__ sf, const double *__restrict__ Bp, const double *__restrict__ dv,       \
      const double *__restrict__ pf, int nsplit, int nchunk, double *__restrict__ dBpart,                \
      double *__restrict__ LWpart

template <int DB, int DS, int KSC, int NR>
__global__ __launch_bounds__(256) void fused_bwd_pairs_kernel(GPX_PAIRS_ARGS) {
  bwd_pairs_body<DB, DS, KSC, NR>(N, M, Q, D, z, ls, sf, Bp, dv, pf, nsplit, nchunk, dBpart, LWpart);
}

// The same pass held to 128 registers (four waves per SIMD, no accumulation registers):
// faster over long point ranges (>= 16 blocks of 16 points per workgroup), slower over
// short ones, where the lower latency of the unconstrained build wins.
template <int DB, int DS, int KSC, int NR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void fused_bwd_pairs_occ4_kernel(
    GPX_PAIRS_ARGS) {
  bwd_pairs_body<DB, DS, KSC, NR>(N, M, Q, D, z, ls, sf, Bp, dv, pf, nsplit, nchunk, dBpart, LWpart);
}

// Row k of a staged 16 x 32 dv block keeps column c at c ^ dv_swizzle(k).  The two read
// patterns (16 rows x 4 columns, and 4 rows x 16 columns) then fall on distinct LDS banks
// two lanes a bank: rows k and k+8 share their column permutation, and of every four
// consecutive rows two keep the halves of the row in place and two swap them.
__device__ __forceinline__ int dv_swizzle(int k) { return 4 * (((k & 1) << 2) | ((k >> 1) & 3)); }

// Per-pair pass for 16 < D <= 32 with the point operands staged in LDS: each 16-point
// block of pf (16 x PF_W, contiguous) and dv (16 x D) is read from memory once per
// workgroup, coalesced, one block ahead into registers, and written to the other half
// of a double buffer while the waves work on the current one (one barrier per block).
// Rows past the range end are staged as zeros (weight 0 through p).  Held to three waves
// per SIMD, which leaves no accumulation registers and fits 168 VGPRs.
// Same operands, sums and order as bwd_pairs_body with NR = 1, hence the same results.
template <int DB, int DS, int KSC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void fused_bwd_pairs_staged(
    GPX_PAIRS_ARGS) {
  constexpr int PA_W = 4 * FKS_MAX;  // exponent-operand columns of a pf row
  __shared__ double pas[2][16][PA_W];
  __shared__ double pqs[2][16][16];
  __shared__ double dvs[2][16][32];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int T = (M + 15) / 16;
  const int t = blockIdx.x >> 2, h = blockIdx.x & 3;
  int A, B;
  decode_tile(t, T, A, B);
  if (16 * A + 4 * h >= M) return;  // the whole workgroup: no barrier is skipped by part of it
  const int a = 16 * A + 4 * h + wave;
  const bool act = a < M;           // inactive waves still stage and meet the barriers
  const int ns = blockIdx.y;
  const int nbeg = ns * nchunk, nend = min(N, nbeg + nchunk);
  constexpr int KSN = KSC ? KSC : FKS_MAX;
  const int KS = KSC ? KSC : (2 * Q + 2 + 3) / 4;
  const int jb = 16 * B + (lane & 15);
  const bool bvalid = jb < M && act;
  double af[KSN], bs[DS];
  {
    double il4[FQ_MAX], zb[FQ_MAX], za[FQ_MAX];
#pragma unroll
    for (int q = 0; q < FQ_MAX; ++q) {
      il4[q] = (q < Q) ? 0.25 * exp(-2.0 * ls[q]) : 0.0;
      zb[q] = (q < Q && bvalid) ? z[jb * Q + q] : 0.0;
      za[q] = (q < Q && act) ? z[a * Q + q] : 0.0;
    }
#pragma unroll
    for (int s = 0; s < KSN; ++s)
      af[s] = (s < KS && bvalid) ? pair_feature((lane >> 4) + 4 * s, Q, za, zb, il4) : 0.0;
#pragma unroll
    for (int s = 0; s < DS; ++s) {
      const int d = (lane >> 4) + 4 * s;
      bs[s] = (bvalid && d < D) ? Bp[((long)a * M + jb) * D + d] : 0.0;
    }
  }
  const double sf2 = exp(2.0 * sf[0]), sf4 = sf2 * sf2;
  gpx_d4 dacc[DB], lacc = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int db = 0; db < DB; ++db) dacc[db] = gpx_d4{0.0, 0.0, 0.0, 0.0};
  constexpr int NPF = (16 * PF_W + 255) / 256;
  const int sk = tid >> 4, sc = tid & 15;  // staging: dv row and column of this thread
  double rp[NPF], rd[DB];
  auto fetch = [&](int n0) {
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int e = tid + 256 * i;
      rp[i] = (e < 16 * PF_W && n0 + e / PF_W < nend) ? pf[(long)n0 * PF_W + e] : 0.0;
    }
#pragma unroll
    for (int db = 0; db < DB; ++db)
      rd[db] = (n0 + sk < nend && 16 * db + sc < D) ? dv[(long)(n0 + sk) * D + 16 * db + sc] : 0.0;
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int e = tid + 256 * i;
      if (e < 16 * PF_W) {
        const int k = e / PF_W, c = e % PF_W;
        if (c < PA_W) pas[buf][k][c] = rp[i]; else pqs[buf][k][c - PA_W] = rp[i];
      }
    }
#pragma unroll
    for (int db = 0; db < DB; ++db) dvs[buf][sk][(16 * db + sc) ^ dv_swizzle(sk)] = rd[db];
  };
  fetch(nbeg);
  stage(0);
  __syncthreads();
  const int ka = lane & 15, kq = lane >> 4;
  const int swa = dv_swizzle(ka);
  int buf = 0;
  for (int n0 = nbeg; n0 < nend; n0 += 16) {
    const bool more = n0 + 16 < nend;  // uniform over the workgroup
    if (more) fetch(n0 + 16);
    if (act) {
      double pa[KSN], da[DS];
#pragma unroll
      for (int s = 0; s < KSN; ++s) pa[s] = (s < KS) ? pas[buf][ka][kq + 4 * s] : 0.0;
#pragma unroll
      for (int s = 0; s < DS; ++s) da[s] = dvs[buf][ka][(kq + 4 * s) ^ swa];
      gpx_d4 e = gpx_d4{0.0, 0.0, 0.0, 0.0}, g = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KSN; ++s)
        if (s < KS) e = gpx_mfma16(pa[s], af[s], e);
#pragma unroll
      for (int s = 0; s < DS; ++s) g = gpx_mfma16(da[s], bs[s], g);
      double dvv[4][DB], dq[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = kq + 4 * r;
#pragma unroll
        for (int db = 0; db < DB; ++db) dvv[r][db] = dvs[buf][k][(16 * db + ka) ^ dv_swizzle(k)];
        dq[r] = pqs[buf][k][ka];
      }
      double p[4], L[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        p[r] = sf4 * exp(e[r]) * ((n0 + kq + 4 * r < nend && bvalid) ? 1.0 : 0.0);
        L[r] = g[r] * p[r];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int db = 0; db < DB; ++db) dacc[db] = gpx_mfma16(dvv[r][db], p[r], dacc[db]);
        lacc = gpx_mfma16(dq[r], L[r], lacc);
      }
    }
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  if (!act) return;
#pragma unroll
  for (int db = 0; db < DB; ++db)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int d = 16 * db + kq + 4 * r;
      if (d < D && jb < M) dBpart[(((long)ns * D + d) * M + a) * M + jb] = dacc[db][r];
    }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int qq = kq + 4 * r;
    if (qq <= Q && jb < M) LWpart[(((long)ns * M + a) * M + jb) * (Q + 1) + qq] = lacc[r];
  }
}
#undef GPX_PAIRS_ARGS

// Wave-scope ordering of LDS accesses (a wave reading what its other lanes wrote).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// backward, both passes in one: the per-pair sums of fused_bwd_pairs_staged /
// bwd_pairs_body and the per-point sums of fused_bwd_points_kernel from one
// evaluation of p and G per (point, pair).  A workgroup of MG_NW = 8 waves owns the 16
// rows of one tile (A, B) over one point range; wave w holds rows a = 16A + w + 8u
// (u < 2).  Per row and 16-point block the exponent, G, p, L and the dB-hat / LW
// contractions are those of the per-pair kernels (same operands and order).  DB, DS: the
// 16-row and 4-row slices of the D contraction (2, 8 for D <= 32; 1, DS for D <= 16).
// Added per row:
//   U[n][q'] = sum_b L[n][b] [1 | z_b][q']  (L transposed through the wave's LDS, 4 MFMAs)
//   row sums  T[n][a] = U[n][0]                      -> Tpart slot B, column a
//   V[n][q]  += w z_aq U[n][1+q]  (w = 2 off the diagonal tile, 1 on it)
// and per workgroup (A < B) the column sums sum_a L[n][a][b] -> Tpart slot A, column b.
// Column c of tile C thus has exactly T slots (row sums from tiles (C, s), s >= C; column
// sums from tiles (s, C), s < C), each written once; V has one slot per tile.  The
// 8-wave sums of the column sums and V go through LDS in wave order (double-buffered
// by block parity, so one barrier per block serves the staging and both sums); a block's
// sums are read and written out during the next block, beside its first MFMAs.
// ---------------------------------------------------------------------------
constexpr int MG_NW = 8;
template <int DB, int DS, int KSC>
__global__ __launch_bounds__(64 * MG_NW) void fused_bwd_merged_kernel(
    int N, int M, int Q, int D, const double *__restrict__ z, const double *__restrict__ ls,
    const double *__restrict__ sf, const double *__restrict__ Bp, const double *__restrict__ dv,
    const double *__restrict__ pf, int nsplit, int nchunk, double *__restrict__ dBpart,
    double *__restrict__ LWpart, double *__restrict__ Tpart, double *__restrict__ Vpart) {
  constexpr int NR = 2, NT = 64 * MG_NW;
  constexpr int PA_W = 4 * FKS_MAX;
  __shared__ double pas[2][16][PA_W];
  __shared__ double pqs[2][16][16];
  __shared__ double dvs[2][16][32];
  __shared__ double trn[MG_NW][16][17];        // per-wave transpose of L
  __shared__ double csum[2][MG_NW][16][17];    // per-wave column sums  [n][b]
  __shared__ double vsum[2][MG_NW][16][17];    // per-wave V terms      [n][q']
  // pair features of the wave's rows (B operand of the exponent), kept in LDS when the
  // slice count is fixed at compile time: in registers they push the kernel past 256
  constexpr bool AF_LDS = KSC != 0;
  constexpr int KSL = AF_LDS ? KSC : 1;
  __shared__ double afs[MG_NW][2][KSL][64];
  __shared__ double etab[32];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int T = (M + 15) / 16;
  const int tile = blockIdx.x;
  int A, B;
  decode_tile(tile, T, A, B);
  const bool off = A != B;
  const int ns = blockIdx.y;
  const int nbeg = ns * nchunk, nend = min(N, nbeg + nchunk);
  constexpr int KSN = KSC ? KSC : FKS_MAX;
  const int KS = KSC ? KSC : (2 * Q + 2 + 3) / 4;
  const int ka = lane & 15, kq = lane >> 4;
  const int jb = 16 * B + ka;
  int a[NR];
  bool act[NR];
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    a[u] = 16 * A + wave + 8 * u;
    act[u] = a[u] < M;
  }
  double af[AF_LDS ? 1 : NR][AF_LDS ? 1 : KSN], bs[NR][DS], zw[NR];
  {
    double il4[FQ_MAX], zb[FQ_MAX];
#pragma unroll
    for (int q = 0; q < FQ_MAX; ++q) {
      il4[q] = (q < Q) ? 0.25 * exp(-2.0 * ls[q]) : 0.0;
      zb[q] = (q < Q && jb < M) ? z[jb * Q + q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const bool bv = jb < M && act[u];
      double za[FQ_MAX];
#pragma unroll
      for (int q = 0; q < FQ_MAX; ++q) za[q] = (q < Q && act[u]) ? z[a[u] * Q + q] : 0.0;
#pragma unroll
      for (int s = 0; s < KSN; ++s) {
        const double f = (s < KS && bv) ? pair_feature((lane >> 4) + 4 * s, Q, za, zb, il4) : 0.0;
        if constexpr (AF_LDS) afs[wave][u][s][lane] = f; else af[u][s] = f;
      }
#pragma unroll
      for (int s = 0; s < DS; ++s) {
        const int d = kq + 4 * s;
        bs[u][s] = (bv && d < D) ? Bp[((long)a[u] * M + jb) * D + d] : 0.0;
      }
      // V weight of this lane's column q' = ka of U: w z_a[q'-1]
      zw[u] = (act[u] && ka >= 1 && ka <= Q) ? (off ? 2.0 : 1.0) * z[a[u] * Q + ka - 1] : 0.0;
    }
  }
  // B operand of U: [1 | z_b] for b = 16B + kq + 4s (rows), q' = ka (columns); padding
  // columns b >= M give zeros (their L is 0 as well)
  double zx[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int b = 16 * B + kq + 4 * s;
    zx[s] = (b < M && ka <= Q) ? (ka == 0 ? 1.0 : z[b * Q + ka - 1]) : 0.0;
  }
  const double sf2 = exp(2.0 * sf[0]), sf4 = sf2 * sf2;
  gpx_d4 dacc[NR][DB], lacc[NR];
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    lacc[u] = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int db = 0; db < DB; ++db) dacc[u][db] = gpx_d4{0.0, 0.0, 0.0, 0.0};
  }
  // D <= 4 with Q <= 5 (one D slice, three feature slices): the dB-hat and LW contractions
  // as per-lane FMAs over the lane's four points (summed over the four lane rows once, at
  // the end), where the MFMA forms would fill 4 and at most 6 of their 16 output rows
  constexpr bool VSM = DB == 1 && DS == 1 && KSC == 3;
  constexpr int VQ = 6;  // Q + 1 <= 6
  double vd[VSM ? NR : 1][4], vl[VSM ? NR : 1][VQ];
#pragma unroll
  for (int u = 0; u < (VSM ? NR : 1); ++u) {
#pragma unroll
    for (int d = 0; d < 4; ++d) vd[u][d] = 0.0;
#pragma unroll
    for (int q = 0; q < VQ; ++q) vl[u][q] = 0.0;
  }
  constexpr int NPF = (16 * PF_W + NT - 1) / NT;
  double rp[NPF], rd;
  const int sk = tid >> 5, sc = tid & 31;  // dv staging: row sk (< 16), column sc (< 32)
  auto fetch = [&](int n0) {
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int e = tid + NT * i;
      rp[i] = (e < 16 * PF_W && n0 + e / PF_W < nend) ? pf[(long)n0 * PF_W + e] : 0.0;
    }
    rd = (n0 + sk < nend && sc < D) ? dv[(long)(n0 + sk) * D + sc] : 0.0;
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int e = tid + NT * i;
      if (e < 16 * PF_W) {
        const int k = e / PF_W, c = e % PF_W;
        if (c < PA_W) pas[buf][k][c] = rp[i]; else pqs[buf][k][c - PA_W] = rp[i];
      }
    }
    dvs[buf][sk][sc ^ dv_swizzle(sk)] = rd;
  };
  // the 8-wave sums of block n0's column sums and V terms (buffer cb), written out
  auto reduce = [&](int n0, int cb) {
    if (tid < 256) {
      const int nn = tid >> 4, bb = tid & 15, n = n0 + nn, b = 16 * B + bb;
      if (off && n < nend && b < M) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < MG_NW; ++w) t += csum[cb][w][nn][bb];
        Tpart[((long)A * N + n) * M + b] = t;
      }
    } else {
      const int e = tid - 256, nn = e >> 4, qq = e & 15, n = n0 + nn;
      if (qq >= 1 && qq <= Q && n < nend) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < MG_NW; ++w) v += vsum[cb][w][nn][qq];
        Vpart[((long)tile * N + n) * Q + qq - 1] = v;
      }
    }
  };
  fetch(nbeg);
  stage(0);
  gpx_exp_table_init(etab);
  __syncthreads();
  const int swa = dv_swizzle(ka);
  int buf = 0;
  for (int n0 = nbeg; n0 < nend; n0 += 16) {
    const bool more = n0 + 16 < nend;  // uniform over the workgroup
    if (more) fetch(n0 + 16);
    double pa[KSN], da[DS], dvv[4][DB], dq[4];
#pragma unroll
    for (int s = 0; s < KSN; ++s) pa[s] = (s < KS) ? pas[buf][ka][kq + 4 * s] : 0.0;
#pragma unroll
    for (int s = 0; s < DS; ++s) da[s] = dvs[buf][ka][(kq + 4 * s) ^ swa];
    if constexpr (!VSM) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = kq + 4 * r;
#pragma unroll
        for (int db = 0; db < DB; ++db) dvv[r][db] = dvs[buf][k][(16 * db + ka) ^ dv_swizzle(k)];
        dq[r] = pqs[buf][k][ka];
      }
    }
    double cs[4] = {0.0, 0.0, 0.0, 0.0}, vw[4] = {0.0, 0.0, 0.0, 0.0};
    double pp[VSM ? NR : 1][4], LL[VSM ? NR : 1][4];
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      gpx_d4 e = gpx_d4{0.0, 0.0, 0.0, 0.0}, g = gpx_d4{0.0, 0.0, 0.0, 0.0};
      if (act[u]) {  // wave-uniform
#pragma unroll
        for (int s = 0; s < KSN; ++s)
          if (s < KS) {
            double f;
            if constexpr (AF_LDS) f = afs[wave][u][s][lane]; else f = af[u][s];
            e = gpx_mfma16(pa[s], f, e);
          }
#pragma unroll
        for (int s = 0; s < DS; ++s) g = gpx_mfma16(da[s], bs[u][s], g);
      }
      // the previous block's 8-wave sums, while this block's first MFMAs run (its buffer
      // half is not written again before the next barrier)
      if (u == 0 && n0 > nbeg) reduce(n0 - 16, buf ^ 1);
      if (!act[u]) continue;
      double p[4], L[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        p[r] = sf4 * gpx_exp_tab(e[r], etab) * ((n0 + kq + 4 * r < nend && jb < M) ? 1.0 : 0.0);
        L[r] = g[r] * p[r];
      }
      if constexpr (VSM) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          pp[u][r] = p[r];
          LL[u][r] = L[r];
        }
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
          for (int db = 0; db < DB; ++db) dacc[u][db] = gpx_mfma16(dvv[r][db], p[r], dacc[u][db]);
          lacc[u] = gpx_mfma16(dq[r], L[r], lacc[u]);
        }
      }
      // U = L [1 | z_b]: L (lane: n = kq + 4r, b = ka) transposed to the A-operand
      // layout (lane: n = ka, b = kq + 4s) through the wave's LDS
#pragma unroll
      for (int r = 0; r < 4; ++r) trn[wave][kq + 4 * r][ka] = L[r];
      wave_lds_sync();
      double lt[4];
#pragma unroll
      for (int s = 0; s < 4; ++s) lt[s] = trn[wave][ka][kq + 4 * s];
      wave_lds_sync();
      gpx_d4 U = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 4; ++s) U = gpx_mfma16(lt[s], zx[s], U);
      // lane holds U[n = kq + 4r][q' = ka]
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + kq + 4 * r;
        if (ka == 0 && n < nend) Tpart[((long)B * N + n) * M + a[u]] = U[r];
        vw[r] += zw[u] * U[r];
        cs[r] += L[r];
      }
    }
    if constexpr (VSM) {
      // lane: pair column b = ka, points n = kq + 4r (r < 4): dBhat[d][b] += dv[n][d] p[n][b],
      // LW[b][q'] += [1 | d_n][q'] L[n][b]
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = kq + 4 * r;
        double dvl[4], dql[VQ];
#pragma unroll
        for (int d = 0; d < 4; ++d) dvl[d] = dvs[buf][k][d ^ dv_swizzle(k)];
#pragma unroll
        for (int q = 0; q < VQ; ++q) dql[q] = pqs[buf][k][q];
#pragma unroll
        for (int u = 0; u < NR; ++u) {
          if (!act[u]) continue;
#pragma unroll
          for (int d = 0; d < 4; ++d) vd[u][d] = fma(dvl[d], pp[u][r], vd[u][d]);
#pragma unroll
          for (int q = 0; q < VQ; ++q) vl[u][q] = fma(dql[q], LL[u][r], vl[u][q]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      csum[buf][wave][kq + 4 * r][ka] = cs[r];
      vsum[buf][wave][kq + 4 * r][ka] = vw[r];
    }
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  if (nend > nbeg) reduce(nbeg + (nend - nbeg - 1) / 16 * 16, buf ^ 1);  // the last block's
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    if (!act[u]) continue;
    if constexpr (VSM) {
      // sum over the four lane rows (lanes ka + 16 kq), in a fixed order
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const double t = gpx_swap_add<16>(gpx_swap_add<32>(vd[u][d]));
        if (kq == 0 && d < D && jb < M) dBpart[(((long)ns * D + d) * M + a[u]) * M + jb] = t;
      }
#pragma unroll
      for (int q = 0; q < VQ; ++q) {
        const double t = gpx_swap_add<16>(gpx_swap_add<32>(vl[u][q]));
        if (kq == 0 && q <= Q && jb < M) LWpart[(((long)ns * M + a[u]) * M + jb) * (Q + 1) + q] = t;
      }
      continue;
    }
#pragma unroll
    for (int db = 0; db < DB; ++db)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int d = 16 * db + kq + 4 * r;
        if (d < D && jb < M) dBpart[(((long)ns * D + d) * M + a[u]) * M + jb] = dacc[u][db][r];
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int qq = kq + 4 * r;
      if (qq <= Q && jb < M) LWpart[(((long)ns * M + a[u]) * M + jb) * (Q + 1) + qq] = lacc[u][r];
    }
  }
}

// Sum of the per-pair pass's point-range partials.  One workgroup per stored 16 x 16
// tile pair (A <= B, the half the pair kernel writes) and output slice (k < D: dBhat[k],
// k >= D: column q = k - D of LW), 256 entries each; rows are read coalesced and the
// mirrored tile (B, A) is written through an LDS transpose.  dBhat (D x M x M) and LW
// ((M*M) x (Q+1)) come out in the layouts of the materialising path.
__global__ __launch_bounds__(256) void fused_pairs_finish(int M, int Q, int D, int nsplit,
                                                          const double *__restrict__ dBpart,
                                                          const double *__restrict__ LWpart,
                                                          double *__restrict__ dBhat, double *__restrict__ LW) {
  __shared__ double tile[16][17];
  const int T = (M + 15) / 16;
  int t = blockIdx.x, A = 0;
  while (t >= T - A) {
    t -= T - A;
    ++A;
  }
  const int B = A + t, k = blockIdx.y, tid = threadIdx.x, i = tid >> 4, j = tid & 15;
  const long MM = (long)M * M;
  const int a = A * 16 + i, b = B * 16 + j;
  if (k < D) {
    double s = 0.0;
    if (a < M && b < M) {
#pragma unroll 8
      for (int p = 0; p < nsplit; ++p) s += dBpart[(((long)p * D + k) * M + a) * M + b];
      dBhat[k * MM + (long)a * M + b] = s;
    }
    if (A != B) {
      tile[i][j] = s;
      __syncthreads();
      const int a2 = B * 16 + i, b2 = A * 16 + j;
      if (a2 < M && b2 < M) dBhat[k * MM + (long)a2 * M + b2] = tile[j][i];
    }
  } else {
    const int Q1 = Q + 1, q = k - D;
    if (a < M && b < M) {
      double s = 0.0;
#pragma unroll 8
      for (int p = 0; p < nsplit; ++p) s += LWpart[(((long)p * M + a) * M + b) * Q1 + q];
      LW[((long)a * M + b) * Q1 + q] = s;
      if (A != B) LW[((long)b * M + a) * Q1 + q] = s;
    }
  }
}


// Per-point finish of the fused adjoint: T = sum of the S slots of Tpart, V = sum of the
// SV slots of Vpart, Lp = dpsi1 psi1,
// then the same per-point formulas (and outputs) as the materialising path's row pass.
// One wave per point: its lanes stride over the M columns and the sums are wave
// reductions, so no workgroup barrier is needed.
// QC: Q fixed at compile time for the common shapes (0: read at run time).
template <int QC>
__global__ __launch_bounds__(256) void fused_rows_finish(
    int N, int M, int Q, int S, int SV, const double *__restrict__ mx, const double *__restrict__ vx,
    const double *__restrict__ z, const double *__restrict__ ls, const double *__restrict__ Tpart,
    const double *__restrict__ Vpart, const double *__restrict__ dpsi1, const double *__restrict__ psi1,
    double *__restrict__ dmx, double *__restrict__ dvx, double *__restrict__ Tout,
    double *__restrict__ Lpout, double *__restrict__ Hout, double *__restrict__ s0out,
    double *__restrict__ Dn, double *__restrict__ Pn, double *__restrict__ En) {
  constexpr int QN = QC ? QC : FQ_MAX, FRED = 4 * QN + 2;
  if (QC) Q = QC;
  const int lane = threadIdx.x & 63, n = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (n >= N) return;
  double mu[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) mu[q] = (q < Q) ? mx[(long)n * Q + q] : 0.0;
  // [0] S0, [1] sum Lp, [2+q] S1, [2+Q+q] sum T z^2, [2+2Q+q] B1, [2+3Q+q] C1
  double c[FRED];
#pragma unroll
  for (int i = 0; i < FRED; ++i) c[i] = 0.0;
  // T = sum of the S slots of Tpart for this lane's columns m = lane + 64 j (j < 4, M <= 256),
  // each column summed in slot order; the loads of four slots at clamped (always valid)
  // indices issue together, and the clamped lanes' sums are never read
  constexpr int MJ = (FM_MAX + 63) / 64;
  const int nj = (M + 63) >> 6;
  double Tm[MJ];
#pragma unroll
  for (int j = 0; j < MJ; ++j) Tm[j] = 0.0;
  for (int s0 = 0; s0 < S; s0 += 4) {
    double tv[4][MJ];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double *row = Tpart + ((long)min(s0 + u, S - 1) * N + n) * M;
#pragma unroll
      for (int j = 0; j < MJ; ++j) tv[u][j] = j < nj ? row[min(lane + 64 * j, M - 1)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (s0 + u < S) {
#pragma unroll
        for (int j = 0; j < MJ; ++j) Tm[j] += tv[u][j];
      }
  }
#pragma unroll
  for (int j = 0; j < MJ; ++j) {
    const int m = lane + 64 * j;
    if (m >= M) break;
    const double T = Tm[j];
    const double Lp = dpsi1[(long)n * M + m] * psi1[(long)n * M + m];
    Tout[(long)n * M + m] = T;
    Lpout[(long)n * M + m] = Lp;
    c[0] += T;
    c[1] += Lp;
#pragma unroll
    for (int q = 0; q < QN; ++q)
      if (q < Q) {
        const double zq = z[m * Q + q], zm = zq - mu[q];
        c[2 + q] += T * zq;
        c[2 + QN + q] += T * zq * zq;
        c[2 + 2 * QN + q] += Lp * zm;
        c[2 + 3 * QN + q] += Lp * zm * zm;
      }
  }
#pragma unroll
  for (int i = 0; i < FRED; ++i) {
    if (i >= 2 && (i - 2) % QN >= Q) continue;
    c[i] = gpx_wave_sum(c[i]);
  }
  double out_S1 = 0.0, out_TZ2 = 0.0, out_B1 = 0.0, out_C1 = 0.0;
#pragma unroll
  for (int q = 0; q < QN; ++q)
    if (q == lane) {
      out_S1 = c[2 + q];
      out_TZ2 = c[2 + QN + q];
      out_B1 = c[2 + 2 * QN + q];
      out_C1 = c[2 + 3 * QN + q];
    }
  // V = sum of the SV slots of Vpart: lane l sums the Q values of slots l, l + 64, ...
  // (one row of Q contiguous values each), then a wave sum per q
  double Vq[QN];
#pragma unroll
  for (int q = 0; q < QN; ++q) Vq[q] = 0.0;
  for (int s = lane; s < SV; s += 64) {
    const double *row = Vpart + ((long)s * N + n) * Q;
#pragma unroll
    for (int q = 0; q < QN; ++q)
      if (q < Q) Vq[q] += row[q];
  }
  double Vsel = 0.0;
#pragma unroll
  for (int q = 0; q < QN; ++q) {
    if (q >= Q) continue;
    const double t = gpx_wave_sum(Vq[q]);
    if (q == lane) Vsel = t;
  }
  if (lane < Q) {
    const int q = lane;
    const double V = Vsel;
    const double S0 = c[0], A1 = c[1];
    const double S1 = out_S1, TZ2 = out_TZ2, B1 = out_B1, C1 = out_C1;
    const double l = exp(2.0 * ls[q]);
    const double m = mx[(long)n * Q + q], vv = vx[(long)n * Q + q];
    const double ell = exp(ls[q]);
    const double d = 1.0 / (2.0 * vv + l), e = 1.0 / (vv + l);
    const double S2 = 0.5 * (TZ2 + V);
    const double quad = m * m * S0 - 2.0 * m * S1 + S2;
    dmx[(long)n * Q + q] = -2.0 * d * (m * S0 - S1) + e * B1;
    dvx[(long)n * Q + q] = 2.0 * d * d * quad - d * S0 + 0.5 * e * (e * C1 - A1);
    Hout[(long)n * Q + q] = e * ell * (e * C1 + (vv / l) * A1) +
                            2.0 * ell * ((vv / l) * d * S0 + d * d * quad + (TZ2 - V) / (2.0 * l * l));
    Dn[(long)n * (Q + 1) + 1 + q] = d;
    Pn[(long)n * 2 * Q + q] = d * m;
    Pn[(long)n * 2 * Q + Q + q] = d;
    En[(long)n * 2 * Q + q] = e;
    En[(long)n * 2 * Q + Q + q] = e * m;
  }
  if (lane == 0) {
    Dn[(long)n * (Q + 1)] = 1.0;
    s0out[2 * n] = c[0];
    s0out[2 * n + 1] = c[1];
  }
}

// psi1, m = psi1 Ahat^T and v = sf2 + sum_s part - m^2 for FIN_PTS points per block.
// Points per block of the forward finish (8: twice the blocks of 16, so the latency-bound
// psi1 and mean loops have two waves per SIMD at N = 4096).  The split partials' loads
// issue four at a time (summed in split order).
constexpr int FIN_PTS = 8;

__global__ __launch_bounds__(256) void fused_fwd_finish(int N, int M, int Q, int D, const double *__restrict__ mx,
                                                        const double *__restrict__ vx,
                                                        const double *__restrict__ z,
                                                        const double *__restrict__ ls,
                                                        const double *__restrict__ sf,
                                                        const double *__restrict__ Ahat,
                                                        const double *__restrict__ part, int S,
                                                        double *__restrict__ psi1, double *__restrict__ mout,
                                                        double *__restrict__ vout) {
  __shared__ double p1[FIN_PTS][FM_MAX + 1];
  __shared__ double rden[FIN_PTS][FQ_MAX], mus[FIN_PTS][FQ_MAX], lts[FIN_PTS][FQ_MAX], lcs[FIN_PTS];
  const int n0 = blockIdx.x * FIN_PTS, tid = threadIdx.x;
  const double sf2 = exp(2.0 * sf[0]);
  // per point and dimension (one thread each): 1 / (2 (v + l)), mx, and log(l / (l + v)) / 2,
  // whose sum over q (in q order) is the point's constant (independent of the inducing point)
  if (tid < FIN_PTS * Q) {
    const int nn = tid / Q, q = tid % Q, n = n0 + nn;
    const double l = exp(2.0 * ls[q]), v = n < N ? vx[(long)n * Q + q] : 0.0;
    rden[nn][q] = 1.0 / (2.0 * (v + l));
    mus[nn][q] = n < N ? mx[(long)n * Q + q] : 0.0;
    lts[nn][q] = 0.5 * log(l / (l + v));
  }
  __syncthreads();
  if (tid < FIN_PTS) {
    double lc = 0.0;
    for (int q = 0; q < Q; ++q) lc += lts[tid][q];
    lcs[tid] = lc;
  }
  __syncthreads();
  for (int e = tid; e < FIN_PTS * M; e += 256) {
    const int nn = e / M, m = e % M, n = n0 + nn;
    double val = 0.0;
    if (n < N) {
      double acc = 0.0;
      for (int q = 0; q < Q; ++q) {
        const double dlt = mus[nn][q] - z[m * Q + q];
        acc -= dlt * dlt * rden[nn][q];
      }
      val = sf2 * exp(acc + lcs[nn]);
      psi1[(long)n * M + m] = val;
    }
    p1[nn][m] = val;
  }
  __syncthreads();
  for (int e = tid; e < FIN_PTS * D; e += 256) {
    const int nn = e / D, d = e % D, n = n0 + nn;
    if (n >= N) continue;
    double s = 0.0;
    for (int sp0 = 0; sp0 < S; sp0 += 4) {
      double pv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) pv[u] = part[((long)min(sp0 + u, S - 1) * N + n) * D + d];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (sp0 + u < S) s += pv[u];
    }
    // four independent partial sums (k mod 4): a 128-long chain of dependent FMAs was
    // this loop's latency
    const double *arow = Ahat + (long)d * M;
    double m4[4] = {0.0, 0.0, 0.0, 0.0};
    int k = 0;
#pragma unroll 2
    for (; k + 4 <= M; k += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) m4[u] += p1[nn][k + u] * arow[k + u];
    }
    for (int u = 0; k < M; ++k, ++u) m4[u] += p1[nn][k] * arow[k];
    const double m = (m4[0] + m4[1]) + (m4[2] + m4[3]);
    mout[(long)n * D + d] = m;
    vout[(long)n * D + d] = sf2 + s - m * m;
  }
}

}  // namespace

extern "C" long gpx_psi_tail_workspace(int N, int M, int Q);
extern "C" int gpx_gemm_batched(int transA, int transB, int m, int n, int k, double alpha,
                                const double *A, int lda, long strideA, const double *B, int ldb,
                                long strideB, double beta, double *C, int ldc, long strideC,
                                int batch, int splits, double *work, int tile, void *stream);
extern "C" int gpx_psi_bwd_tail(int N, int M, int Q, const double *z, const double *ls, const double *sf,
                                const double *T, const double *Lp, const double *H, const double *s0,
                                const double *Dn, const double *Pn, const double *En, const double *LW,
                                double *TP, double *LE, double *cols0, double *dsf2, double *dell, double *dZ, double *kwork,
                                void *stream);

extern "C" {

long gpx_fused_bwd_workspace(int N, int M, int Q, int D, int splits, int nsplit) {
  const long MM = (long)M * M;
  long s = (long)nsplit * D * MM + (long)nsplit * MM * (Q + 1) + MM * (Q + 1);  // dBpart, LWpart, LW
  s += (long)fused_t_slots(N, M, D, splits) * N * M + (long)fused_v_slots(N, M, D, splits) * N * Q;  // Tpart, Vpart
  s += 3L * N * M + (long)N * Q + 2L * N;                                     // dpsi1, T, Lp, H, s0
  s += (long)N * (Q + 1) + 4L * N * Q + 4L * M * Q + 2;                       // Dn, Pn, En, TP, LE, cols0
  s += (long)N * PF_W;                                                        // point features
  s += gpx_psi_tail_workspace(N, M, Q);                                       // tail split-K partials
  s += MM;                                                                    // pair feature P
  return s;
}

}  // extern "C"

namespace {

// Workspace of the fused adjoint (gpx_fused_bwd_workspace doubles), carved in order.
struct FusedBwdWork {
  double *dBpart, *LWpart, *LW, *Tpart, *Vpart, *dpsi1, *Tn, *Lp, *H, *s0, *Dn, *Pn, *En, *TP, *LE, *cols0, *pf,
      *Pab;
  FusedBwdWork(double *work, int N, int M, int Q, int D, int splits, int nsplit) {
    const long MM = (long)M * M;
    dBpart = work;
    LWpart = dBpart + (long)nsplit * D * MM;
    LW = LWpart + (long)nsplit * MM * (Q + 1);
    Tpart = LW + MM * (Q + 1);
    Vpart = Tpart + (long)fused_t_slots(N, M, D, splits) * N * M;
    dpsi1 = Vpart + (long)fused_v_slots(N, M, D, splits) * N * Q;
    Tn = dpsi1 + (long)N * M;
    Lp = Tn + (long)N * M;
    H = Lp + (long)N * M;
    s0 = H + (long)N * Q;
    Dn = s0 + 2L * N;
    Pn = Dn + (long)N * (Q + 1);
    En = Pn + 2L * N * Q;
    TP = En + 2L * N * Q;
    LE = TP + 2L * M * Q;
    cols0 = LE + 2L * M * Q;
    pf = cols0 + 2;
    Pab = pf + (long)N * PF_W + gpx_psi_tail_workspace(N, M, Q);  // after the tail's partials
  }
};

bool fused_bwd_args_ok(int N, int M, int Q, int D, int splits, int nsplit) {
  // Q <= FQ_MAX - 1: the [1 | d_n] contraction fills the Q+1 rows of one 16-row MFMA tile
  return N >= 1 && M >= 1 && M <= FM_MAX && Q >= 1 && Q <= FQ_MAX - 1 && D >= 1 && D <= 32 && splits >= 1 &&
         nsplit >= 1 && splits <= (M + 15) / 16;  // the per-point pass splits over column tiles
}

}  // namespace

extern "C" {

// Adjoint of the moment-matched layer without materialising psi2 or dL/dpsi2:
// given dm_all = dm - 2 dv m and dv (N x D), Ahat (D x M), Bp (pair-major symmetrised
// Bhat, M x M x D) and psi1 from the forward: dBhat (D x M x M), dsf2, dell, dZ, dmx, dvx.
// dL/dpsi2 = sum_d dv_nd Bsym[d] is symmetric by construction (kernels.py:415-418).
// The work splits into three ordered parts on one `work` buffer: _points (point and pair
// features, the per-point pass and the row finish: dmx and dvx, what the layer below
// needs), _pairs (the per-pair pass: dBhat and LW) and _tail (dsf2, dell, dZ).  _pairs
// and _tail may run on another stream once _points has completed there; _main is _points
// then _pairs.
int gpx_fused_mm_bwd_points(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                            const double *ls, const double *sf, const double *Ahat, const double *Bp,
                            const double *psi1, const double *dm_all, const double *dv, int splits, int nsplit,
                            double *dmx, double *dvx, double *work, void *stream) {
  if (!fused_bwd_args_ok(N, M, Q, D, splits, nsplit)) return GPX_EINVAL;
  if (!mx || !vx || !z || !ls || !sf || !Ahat || !Bp || !psi1 || !dm_all || !dv || !dmx || !dvx || !work)
    return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  FusedBwdWork W(work, N, M, Q, D, splits, nsplit);
  hipLaunchKernelGGL(fused_point_feat_kernel, dim3(gpx_cdiv(N, 256)), dim3(256), 0, st, N, Q, mx, vx, ls, W.pf);
  GPX_CHECK_LAUNCH();
  const int ks = (2 * Q + 2 + 3) / 4;  // compile-time slice counts for Q = 4, 5 (3) and 7, 8 (5)
  if (fused_merged(N, D)) {
    // both passes in one kernel: the per-pair partials (dBpart, LWpart) are written here
    // too, and gpx_fused_mm_bwd_pairs only sums them
    const int T = (M + 15) / 16, ntiles = T * (T + 1) / 2;
    const int nchunk = ((N + nsplit - 1) / nsplit + 15) / 16 * 16;
    dim3 gm(ntiles, nsplit);
#define GPX_MERGED(DBV, DSV, KSV)                                                                             \
  hipLaunchKernelGGL((fused_bwd_merged_kernel<DBV, DSV, KSV>), gm, dim3(64 * MG_NW), 0, st, N, M, Q, D, z, ls, sf,  \
                     Bp, dv, (const double *)W.pf, nsplit, nchunk, W.dBpart, W.LWpart, W.Tpart, W.Vpart)
#define GPX_MERGED_KS(DBV, DSV) \
  if (ks == 3) GPX_MERGED(DBV, DSV, 3); else if (ks == 5) GPX_MERGED(DBV, DSV, 5); else GPX_MERGED(DBV, DSV, 0);
    if (D <= 4) { GPX_MERGED_KS(1, 1) }
    else if (D <= 8) { GPX_MERGED_KS(1, 2) }
    else if (D <= 16) { GPX_MERGED_KS(1, 4) }
    else { GPX_MERGED_KS(2, 8) }
#undef GPX_MERGED_KS
#undef GPX_MERGED
    GPX_CHECK_LAUNCH();
  } else {
  hipLaunchKernelGGL(pair_p_kernel, dim3(gpx_cdiv((long)M * M, 256)), dim3(256), 0, st, M, Q, z, ls, W.Pab);
  GPX_CHECK_LAUNCH();
  // per-point pass over tile splits
  dim3 gq(split_grid(gpx_cdiv(N, 16), splits));
#define GPX_POINTS1(DSV, KSV, FULL)                                                                             \
  hipLaunchKernelGGL((fused_bwd_points_kernel<DSV, KSV, FULL>), gq, dim3(256), points_lds_bytes(M, Q), st, N, M, Q, \
                     D, mx, vx, z, ls, sf, Bp, dv, (const double *)W.Pab, splits, W.Tpart, W.Vpart)
#define GPX_POINTS(DSV, KSV)                                      \
  if (D == 4 * DSV) GPX_POINTS1(DSV, KSV, true);                 \
  else GPX_POINTS1(DSV, KSV, false);
  if (D <= 4) {
    if (ks == 3) { GPX_POINTS(1, 3) } else if (ks == 5) { GPX_POINTS(1, 5) } else { GPX_POINTS(1, 0) }
  } else if (D <= 8) {
    if (ks == 3) { GPX_POINTS(2, 3) } else if (ks == 5) { GPX_POINTS(2, 5) } else { GPX_POINTS(2, 0) }
  } else if (D <= 16) {
    if (ks == 3) { GPX_POINTS(4, 3) } else if (ks == 5) { GPX_POINTS(4, 5) } else { GPX_POINTS(4, 0) }
  } else {
    if (ks == 3) { GPX_POINTS(8, 3) } else if (ks == 5) { GPX_POINTS(8, 5) } else { GPX_POINTS(8, 0) }
  }
#undef GPX_POINTS
#undef GPX_POINTS1
  GPX_CHECK_LAUNCH();
  }
  // dpsi1 = dm_all Ahat
  int rc = gpx_gemm_batched(0, 0, N, M, D, 1.0, dm_all, D, 0, Ahat, M, 0, 0.0, W.dpsi1, M, 0, 1, 1, nullptr, 0,
                            stream);
  if (rc) return rc;
#define GPX_ROWS(QV)                                                                                           \
  hipLaunchKernelGGL((fused_rows_finish<QV>), dim3(gpx_cdiv(N, 4)), dim3(256), 0, st, N, M, Q,                 \
                     fused_t_slots(N, M, D, splits), fused_v_slots(N, M, D, splits), mx, vx, z, ls,              \
                     (const double *)W.Tpart, (const double *)W.Vpart, (const double *)W.dpsi1, psi1, dmx, dvx,   \
                     W.Tn, W.Lp, W.H, W.s0, W.Dn, W.Pn, W.En)
  if (Q == 4) GPX_ROWS(4); else if (Q == 5) GPX_ROWS(5); else if (Q == 8) GPX_ROWS(8); else GPX_ROWS(0);
#undef GPX_ROWS
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_fused_mm_bwd_pairs(int N, int M, int Q, int D, const double *z, const double *ls, const double *sf,
                           const double *Bp, const double *dv, int splits, int nsplit, double *dBhat, double *work,
                           void *stream) {
  if (!fused_bwd_args_ok(N, M, Q, D, splits, nsplit)) return GPX_EINVAL;
  if (!z || !ls || !sf || !Bp || !dv || !dBhat || !work) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  const int T = (M + 15) / 16, ntiles = T * (T + 1) / 2;
  FusedBwdWork W(work, N, M, Q, D, splits, nsplit);
  if (fused_merged(N, D)) {  // the partials came from the merged kernel of the per-point part
    hipLaunchKernelGGL(fused_pairs_finish, dim3(ntiles, D + Q + 1), dim3(256), 0, st, M, Q, D, nsplit,
                       (const double *)W.dBpart, (const double *)W.LWpart, dBhat, W.LW);
    GPX_CHECK_LAUNCH();
    return GPX_OK;
  }
  // per-pair pass over point ranges
  const int nchunk = ((N + nsplit - 1) / nsplit + 15) / 16 * 16;
  constexpr int NR = 1;  // rows per wave (one: three waves per SIMD at D = 32; two: one)
  dim3 gp(ntiles * 4 / NR, nsplit);
  const int ks = (2 * Q + 2 + 3) / 4;
  const bool occ4 = nchunk >= 256;  // long point ranges: the four-waves-per-SIMD build
#define GPX_PAIRS(DBV, DSV, KSV)                                                                                 \
  do {                                                                                                           \
    if (occ4)                                                                                                    \
      hipLaunchKernelGGL((fused_bwd_pairs_occ4_kernel<DBV, DSV, KSV, NR>), gp, dim3(256), 0, st, N, M, Q, D, z, \
                         ls, sf, Bp, dv, (const double *)W.pf, nsplit, nchunk, W.dBpart, W.LWpart);             \
    else                                                                                                         \
      hipLaunchKernelGGL((fused_bwd_pairs_kernel<DBV, DSV, KSV, NR>), gp, dim3(256), 0, st, N, M, Q, D, z, ls,  \
                         sf, Bp, dv, (const double *)W.pf, nsplit, nchunk, W.dBpart, W.LWpart);                 \
  } while (0)
#define GPX_PAIRS_STAGED(KSV)                                                                                   \
  hipLaunchKernelGGL((fused_bwd_pairs_staged<2, 8, KSV>), gp, dim3(256), 0, st, N, M, Q, D, z, ls, sf, Bp, dv, \
                     (const double *)W.pf, nsplit, nchunk, W.dBpart, W.LWpart)
  // DS = ceil(D / 4) rounded up to a power of two: slices of the D contraction
  if (D <= 4) {
    if (ks == 3) GPX_PAIRS(1, 1, 3); else if (ks == 5) GPX_PAIRS(1, 1, 5); else GPX_PAIRS(1, 1, 0);
  } else if (D <= 8) {
    if (ks == 3) GPX_PAIRS(1, 2, 3); else if (ks == 5) GPX_PAIRS(1, 2, 5); else GPX_PAIRS(1, 2, 0);
  } else if (D <= 16) {
    if (ks == 3) GPX_PAIRS(1, 4, 3); else if (ks == 5) GPX_PAIRS(1, 4, 5); else GPX_PAIRS(1, 4, 0);
  } else {
    if (ks == 3) GPX_PAIRS_STAGED(3); else if (ks == 5) GPX_PAIRS_STAGED(5); else GPX_PAIRS_STAGED(0);
  }
#undef GPX_PAIRS
#undef GPX_PAIRS_STAGED
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(fused_pairs_finish, dim3(ntiles, D + Q + 1), dim3(256), 0, st, M, Q, D, nsplit,
                     (const double *)W.dBpart, (const double *)W.LWpart, dBhat, W.LW);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_fused_mm_bwd_main(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                          const double *ls, const double *sf, const double *Ahat, const double *Bp,
                          const double *psi1, const double *dm_all, const double *dv, int splits, int nsplit,
                          double *dBhat, double *dmx, double *dvx, double *work, void *stream) {
  if (!dBhat) return GPX_EINVAL;
  const int rc = gpx_fused_mm_bwd_points(N, M, Q, D, mx, vx, z, ls, sf, Ahat, Bp, psi1, dm_all, dv, splits, nsplit,
                                         dmx, dvx, work, stream);
  if (rc) return rc;
  return gpx_fused_mm_bwd_pairs(N, M, Q, D, z, ls, sf, Bp, dv, splits, nsplit, dBhat, work, stream);
}

int gpx_fused_mm_bwd_tail(int N, int M, int Q, int D, const double *z, const double *ls, const double *sf,
                          int splits, int nsplit, double *dsf2, double *dell, double *dZ, double *work,
                          void *stream) {
  if (!fused_bwd_args_ok(N, M, Q, D, splits, nsplit)) return GPX_EINVAL;
  if (!z || !ls || !sf || !dsf2 || !dell || !dZ || !work) return GPX_EINVAL;
  FusedBwdWork W(work, N, M, Q, D, splits, nsplit);
  return gpx_psi_bwd_tail(N, M, Q, z, ls, sf, W.Tn, W.Lp, W.H, W.s0, W.Dn, W.Pn, W.En, W.LW, W.TP, W.LE, W.cols0,
                          dsf2, dell, dZ, W.pf + (long)N * PF_W, stream);
}

int gpx_fused_mm_bwd(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                     const double *ls, const double *sf, const double *Ahat, const double *Bp,
                     const double *psi1, const double *dm_all, const double *dv, int splits, int nsplit,
                     double *dBhat, double *dsf2, double *dell, double *dZ, double *dmx, double *dvx,
                     double *work, void *stream) {
  if (!dsf2 || !dell || !dZ) return GPX_EINVAL;
  int rc = gpx_fused_mm_bwd_main(N, M, Q, D, mx, vx, z, ls, sf, Ahat, Bp, psi1, dm_all, dv, splits, nsplit, dBhat,
                                 dmx, dvx, work, stream);
  if (rc) return rc;
  return gpx_fused_mm_bwd_tail(N, M, Q, D, z, ls, sf, splits, nsplit, dsf2, dell, dZ, work, stream);
}

}  // extern "C"

extern "C" {


// Symmetrised Bhat in pair-major layout: out[a, b, d] = (B[d,a,b] + B[d,b,a]) / 2.
int gpx_sym_pairmajor(int D, int M, const double *B, double *out, void *stream) {
  if (D < 1 || M < 1 || !B || !out || B == out) return GPX_EINVAL;
  const long n = (long)D * M * M;
  hipLaunchKernelGGL(sym_pairmajor_kernel, dim3(gpx_cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, D, M, B, out);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

long gpx_fused_fwd_workspace(int N, int M, int D, int splits) { return (long)splits * N * D + (long)M * M; }

// Moment-matched layer forward without materialising psi2:
// psi1 (N x M), mout = psi1 Ahat^T, vout = sf2 + <Bsym[d], psi2_n> - mout^2 (N x D).
// Bp is the symmetrised Bhat in pair-major layout (gpx_sym_pairmajor); `splits`
// spreads the pair tiles over grid.y and `nsub` (1, 2 or 4) sets 16*nsub points per block.
static int fused_fwd_check(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                           const double *ls, const double *sf, const double *Ahat, const double *Bp, int splits,
                           int nsub, const double *psi1, const double *mout, const double *vout) {
  if (N < 1 || M < 1 || M > FM_MAX || Q < 1 || Q > FQ_MAX || D < 1 || D > 32 || splits < 1)
    return GPX_EINVAL;
  if (nsub != 1 && nsub != 2 && nsub != 4) return GPX_EINVAL;
  if (!mx || !vx || !z || !ls || !sf || !Ahat || !Bp || !psi1 || !mout || !vout) return GPX_EINVAL;
  const int T = (M + 15) / 16, ntiles = T * (T + 1) / 2;
  if (splits > ntiles) return GPX_EINVAL;
  return GPX_OK;
}

// The launches of gpx_fused_mm_fwd over a checked argument set and a workspace of
// gpx_fused_fwd_workspace(N, M, D, splits) doubles.
static int fused_fwd_launch(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                            const double *ls, const double *sf, const double *Ahat, const double *Bp, int splits,
                            int nsub, double *psi1, double *mout, double *vout, double *work, hipStream_t st) {
  double *Pab = work + (long)splits * N * D;
  hipLaunchKernelGGL(pair_p_kernel, dim3(gpx_cdiv((long)M * M, 256)), dim3(256), 0, st, M, Q, z, ls, Pab);
  GPX_CHECK_LAUNCH();
  dim3 grid(split_grid(gpx_cdiv(N, 16 * nsub), splits));
  const int ks = (2 * Q + 2 + 3) / 4;  // compile-time slice counts for Q = 4, 5 (3) and 7, 8 (5)
#define GPX_FWD1(DBV, NS, KSV)                                                                                 \
  hipLaunchKernelGGL((fused_fwd_kernel<DBV, NS, KSV>), grid, dim3(256), 0, st, N, M, Q, D, mx, vx, z, ls, sf, Bp, \
                     (const double *)Pab, splits, work)
#define GPX_FWD(DBV, NS)                    \
  do {                                      \
    if (ks == 3)                            \
      GPX_FWD1(DBV, NS, 3);                 \
    else if (ks == 5)                       \
      GPX_FWD1(DBV, NS, 5);                 \
    else                                    \
      GPX_FWD1(DBV, NS, 0);                 \
  } while (0)
  if (D <= 4) {
    if (nsub == 1) GPX_FWD(0, 1); else if (nsub == 2) GPX_FWD(0, 2); else GPX_FWD(0, 4);
  } else if (D <= 16) {
    if (nsub == 1) GPX_FWD(1, 1); else if (nsub == 2) GPX_FWD(1, 2); else GPX_FWD(1, 4);
  } else {
    if (nsub == 1) GPX_FWD(2, 1); else if (nsub == 2) GPX_FWD(2, 2); else GPX_FWD(2, 4);
  }
#undef GPX_FWD
#undef GPX_FWD1
  GPX_CHECK_LAUNCH();
  hipLaunchKernelGGL(fused_fwd_finish, dim3(gpx_cdiv(N, FIN_PTS)), dim3(256), 0, st, N, M, Q, D, mx, vx, z, ls, sf,
                     Ahat, (const double *)work, splits, psi1, mout, vout);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

int gpx_fused_mm_fwd(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                     const double *ls, const double *sf, const double *Ahat, const double *Bp,
                     int splits, int nsub, double *psi1, double *mout, double *vout, double *work,
                     void *stream) {
  const int rc = fused_fwd_check(N, M, Q, D, mx, vx, z, ls, sf, Ahat, Bp, splits, nsub, psi1, mout, vout);
  if (rc != GPX_OK) return rc;
  if (!work) return GPX_EINVAL;
  return fused_fwd_launch(N, M, Q, D, mx, vx, z, ls, sf, Ahat, Bp, splits, nsub, psi1, mout, vout, work,
                          (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free forward.  The workspace of gpx_fused_mm_fwd (per-split partial sums
// and the pair constants P_ab) lives in a file-scope device array instead of caller
// memory.  Each GPU holds its own copy of the array from module load to process exit;
// host code never passes the array's name to a kernel (in host code that name is a
// host-side stand-in), only the device address hipGetSymbolAddress returns for the
// current device, looked up once per device and kept.
// ---------------------------------------------------------------------------
namespace {

constexpr long kNoAllocScratch = 4L << 20;  // doubles (32 MiB)
__device__ __align__(256) double scratch[kNoAllocScratch];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (NULL on failure).
double *scratch_on_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

long gpx_fused_fwd_noalloc_max(void) { return kNoAllocScratch; }

int gpx_fused_mm_fwd_noalloc(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                             const double *ls, const double *sf, const double *Ahat, const double *Bp, int splits,
                             int nsub, double *psi1, double *mout, double *vout, void *stream) {
  const int rc = fused_fwd_check(N, M, Q, D, mx, vx, z, ls, sf, Ahat, Bp, splits, nsub, psi1, mout, vout);
  if (rc != GPX_OK) return rc;
  if (gpx_fused_fwd_workspace(N, M, D, splits) > kNoAllocScratch) return GPX_ETOOBIG;
  double *work = scratch_on_current_device();
  if (!work) return GPX_EHIP;
  return fused_fwd_launch(N, M, Q, D, mx, vx, z, ls, sf, Ahat, Bp, splits, nsub, psi1, mout, vout, work,
                          (hipStream_t)stream);
}

}  // extern "C"
