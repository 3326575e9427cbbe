// Batched SPD inverse + log-determinant, one workgroup per M x M matrix.
//
// Replaces the per-output-dim np.linalg.inv / slogdet / solve calls of the layer
// algebra (base_models.py:462-476; aep_models.py:62-95, 523-533).  The matrix stays in
// registers, as MFMA accumulator tiles, for the whole inversion.
//
// Algorithm: Gauss-Jordan elimination without pivoting (stable for SPD input) over
// 16 x 16 pivot blocks, on the FP64 matrix cores (spd_gj_kernel below).
#include "gpx_common.h"

namespace {

// Blocked Gauss-Jordan on the matrix cores, 16 x 16 pivot blocks.  Eliminating block
// K with P = A[K,K] (rows of K: A[K,:], columns of K: c = A[:,K]):
//   A[I,J] -= c_I P^-1 A[K,J]     (I, J != K)
//   A[K,J] <- P^-1 A[K,J],  A[I,K] <- -c_I P^-1,  A[K,K] <- P^-1,
// and after every block the matrix holds A^-1.  All MT x MT tiles are kept, in MFMA
// accumulator layout, dealt round-robin to the NW waves (register r of lane l holds
// tile entry [l/16 + 4r][l%16]): keeping only a symmetric half and mirroring it costs
// two orders of magnitude in accuracy downstream on ill-conditioned input, because the
// two halves of a computed inverse carry errors that cancel in later products.
// Per block: the owners of block row and column K publish them to LDS; the panel waves
// run Gauss-Jordan over the 16 scalar pivots on [P | I | A[K, other columns]], giving
// P^-1 A[K,:] and P^-1 without ever multiplying by an explicit P^-1 (which loses an order
// of magnitude on ill-conditioned blocks); every wave then updates its tiles with four
// 16x16x4 MFMAs per tile.  The scalar pivots are the D of A = L D L^T: their logs sum to
// log det A and the first non-positive one flags a matrix that is not positive definite.
// The matrix is padded with the identity up to a multiple of 16.
template <int MT, int NW>
__global__ __launch_bounds__(NW * 64) void spd_gj_kernel(int M, int batch, const double *__restrict__ A,
                                                         int lda, long sA, double *__restrict__ out, int ldo,
                                                         long sO, double *__restrict__ logdet,
                                                         int *__restrict__ info) {
  constexpr int MP = 16 * MT;
  constexpr int RS = MP + 1;  // LDS row stride (doubles)
  constexpr int T = MT * MT;
  constexpr int TPW = (T + NW - 1) / NW;
  constexpr int NPW = MP > 16 ? (MP - 16 + 31) / 32 : 1;  // waves that eliminate the panel
  static_assert(NPW <= NW, "not enough waves for the panel");
  // rowK: block row K (read by the panel only, so one buffer); colT: block column K,
  // transposed (read by the tile updates while the next block's column is published,
  // so two); rvK, Pi: the panel's results
  __shared__ double rowK[16 * RS];
  __shared__ double colT[2][16 * RS];
  __shared__ double rvK[16 * RS];
  __shared__ double Pi[16 * 17];
  __shared__ double piv[MP];
  __shared__ int failk[2];
  __shared__ double red[16];
  const int b = blockIdx.x;
  if (b >= batch) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int lr = lane >> 4, lc = lane & 15;
  const double *Ab = A + b * sA;
  gpx_d4 acc[TPW];
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int t = w + NW * u, I = t / MT, J = t % MT;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = I * 16 + lr + 4 * r, j = J * 16 + lc;
      acc[u][r] = (t < T && i < M && j < M) ? Ab[(long)i * lda + j] : (i == j ? 1.0 : 0.0);
    }
  }
  int fail = 0;
  for (int K = 0; K < MT; ++K) {
    double *cT = colT[K & 1];
    // (1) block row K and block column K (transposed) to LDS
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int t = w + NW * u, I = t / MT, J = t % MT;
      if (t >= T) continue;
      if (I == K) {
#pragma unroll
        for (int r = 0; r < 4; ++r) rowK[(lr + 4 * r) * RS + J * 16 + lc] = acc[u][r];
      }
      if (J == K) {
#pragma unroll
        for (int r = 0; r < 4; ++r) cT[lc * RS + I * 16 + lr + 4 * r] = acc[u][r];
      }
    }
    __syncthreads();
    // (2) the panel.  In each panel wave lane c < 16 holds pivot column K*16 + c, lane
    //     16 + c identity column c (ending as column c of P^-1) and lanes 32.. hold 32 of
    //     the other columns each (ending as P^-1 A[K, col]).  Every panel wave runs the
    //     same pivots on the same pivot columns, so their copies agree bit for bit.
    if (w < NPW) {
      int col = 0;
      bool live = true;
      if (lane < 16) {
        col = K * 16 + lane;
      } else if (lane >= 32) {
        const int m = 32 * w + lane - 32;
        live = m < MP - 16;
        col = m < K * 16 ? m : m + 16;
      }
      double v[16];
#pragma unroll
      for (int i = 0; i < 16; ++i)
        v[i] = (lane >= 16 && lane < 32) ? (i == lane - 16 ? 1.0 : 0.0) : (live ? rowK[i * RS + col] : 0.0);
      int bad = 0;
#pragma unroll
      for (int p = 0; p < 16; ++p) {
        const double d = gpx_readlane_d(v[p], p);
        if (!bad && !(d > 0.0 && d < INFINITY)) bad = p + 1;
        const double rd = gpx_recip(d);
        double colp[16];  // column p, read from the pivot lane
#pragma unroll
        for (int i = 0; i < 16; ++i)
          if (i != p) colp[i] = gpx_readlane_d(v[i], p);
        const double rowp = v[p] * rd;
#pragma unroll
        for (int i = 0; i < 16; ++i)
          if (i != p) v[i] = fma(-colp[i], rowp, v[i]);
        v[p] = rowp;
        if (w == 0 && lane == 0) piv[K * 16 + p] = d;
      }
      if (lane >= 16 && lane < 32) {
        if (w == 0) {
#pragma unroll
          for (int i = 0; i < 16; ++i) Pi[i * 17 + lane - 16] = v[i];
        }
      } else if (lane >= 32 && live) {
#pragma unroll
        for (int i = 0; i < 16; ++i) rvK[i * RS + col] = v[i];
      }
      if (w == 0 && lane == 0) failk[K & 1] = bad ? K * 16 + bad : 0;
    }
    __syncthreads();
    if (failk[K & 1]) {
      fail = failk[K & 1];
      break;
    }
    // (3) every wave: its tiles
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int t = w + NW * u, I = t / MT, J = t % MT;
      if (t >= T) continue;
      if (I == K) {
        if (J == K) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[u][r] = Pi[(lr + 4 * r) * 17 + lc];
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[u][r] = rvK[(lr + 4 * r) * RS + J * 16 + lc];
        }
      } else if (J == K) {
        gpx_d4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
          c = gpx_mfma16(-cT[(4 * ks + lr) * RS + I * 16 + lc], Pi[(4 * ks + lr) * 17 + lc], c);
        acc[u] = c;
      } else {
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
          acc[u] = gpx_mfma16(-cT[(4 * ks + lr) * RS + I * 16 + lc], rvK[(4 * ks + lr) * RS + J * 16 + lc],
                              acc[u]);
      }
      // keep each tile's operand loads next to its MFMAs: hoisting every tile's loads
      // ahead of the first MFMA needs 16 registers per tile and spills for MT > 8
      if (TPW > 6) __builtin_amdgcn_sched_barrier(0);
    }
  }
  __syncthreads();  // pivots of the last block were written after the last barrier of the loop
  double ld = 0.0;
  for (int k = threadIdx.x; k < (fail ? 0 : M); k += NW * 64) ld += log(piv[k]);
  ld = gpx_block_sum(ld, red);
  if (threadIdx.x == 0) {
    if (logdet) logdet[b] = fail ? NAN : ld;
    if (info) info[b] = fail;
  }
  if (!out) return;
  double *Ob = out + b * sO;
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int t = w + NW * u, I = t / MT, J = t % MT;
    if (t >= T) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = I * 16 + lr + 4 * r, j = J * 16 + lc;
      if (i < M && j < M) Ob[(long)i * ldo + j] = fail ? NAN : acc[u][r];
    }
  }
}

template <int MT>
int launch_sweep(int M, int batch, const double *A, int lda, long sA, double *out, int ldo, long sO,
                 double *logdet, int *info, hipStream_t st) {
  // waves: enough for the panel and for at most ~10 tiles each
  constexpr int NW = MT <= 3 ? 4 : (MT <= 9 ? 8 : 16);
  hipLaunchKernelGGL((spd_gj_kernel<MT, NW>), dim3(batch), dim3(NW * 64), 0, st, M, batch, A, lda, sA, out,
                     ldo, sO, logdet, info);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

typedef int (*sweep_fn)(int, int, const double *, int, long, double *, int, long, double *, int *, hipStream_t);
const sweep_fn kSweep[16] = {launch_sweep<1>,  launch_sweep<2>,  launch_sweep<3>,  launch_sweep<4>,
                             launch_sweep<5>,  launch_sweep<6>,  launch_sweep<7>,  launch_sweep<8>,
                             launch_sweep<9>,  launch_sweep<10>, launch_sweep<11>, launch_sweep<12>,
                             launch_sweep<13>, launch_sweep<14>, launch_sweep<15>, launch_sweep<16>};

}  // namespace

extern "C" {

// Largest M the SPD kit accepts.
int gpx_spd_max_m(void) { return 256; }

// For each b < batch: out[b] = inv(A[b]), logdet[b] = log det A[b], info[b] = 0, or
// info[b] = k+1 (and NaN outputs) when pivot k is not positive.  `out` may alias `A`
// (in-place) or be NULL (log-determinant only); logdet / info may be NULL.
int gpx_spd_inv_batched(int M, int batch, const double *A, int lda, long strideA, double *out,
                        int ldo, long strideOut, double *logdet, int *info, void *stream) {
  if (M < 0 || batch < 0 || M > 256) return GPX_EINVAL;
  if (M == 0 || batch == 0) return GPX_OK;
  if (!A || lda < M || (out && ldo < M)) return GPX_EINVAL;
  return kSweep[(M + 15) / 16 - 1](M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info,
                                   (hipStream_t)stream);
}

}  // extern "C"
