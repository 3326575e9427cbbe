// Batched SPD inverse + log-determinant, one workgroup per M x M matrix.
//
// Replaces the per-output-dim np.linalg.inv / slogdet / solve calls of the layer
// algebra (base_models.py:462-476; aep_models.py:62-95, 523-533).  The matrix stays in
// registers for the whole factorisation: thread (ty, tx) of a G x G block owns the
// strided sub-grid rows ty + G*a, cols tx + G*c (a, c < TB).
//
// Algorithm: Gauss-Jordan elimination without pivoting (stable for SPD input), NB
// pivots per step.  With K the next NB indices and P = A[K,K]:
//   A[i,j] -= A[i,K] P^-1 A[K,j],  A[K,j] <- P^-1 A[K,j],  A[i,K] <- -A[i,K] P^-1,
//   A[K,K] <- P^-1,
// which is NB scalar Gauss-Jordan steps merged, so one barrier serves NB pivots.
// P^-1 comes from a scalar Gauss-Jordan on the NB x NB block, whose pivots are the
// pivots of the full elimination: the D of A = L D L^T.  Hence sum(log pivot) is
// log det A, and the first non-positive pivot flags a matrix that is not positive
// definite.  The matrix is padded with the identity up to a multiple of NB.
#include "gpx_common.h"

namespace {

template <int G, int TB, int NB>
__global__ __launch_bounds__(G *G) void spd_gj_kernel(int M, int batch, const double *__restrict__ A,
                                                      int lda, long sA, double *__restrict__ out,
                                                      int ldo, long sO, double *__restrict__ logdet,
                                                      int *__restrict__ info) {
  constexpr int W = G * TB;
  static_assert(G % NB == 0, "pivot blocks must not straddle thread rows");
  __shared__ double rowk[2][NB][W];
  __shared__ double colk[2][NB][W];
  __shared__ double rvs[2][NB][W];
  __shared__ double Pis[2][NB][NB];
  __shared__ int failk[2];
  __shared__ double piv[W];
  __shared__ double red[16];
  const int b = blockIdx.x;
  if (b >= batch) return;
  const double *Ab = A + b * sA;
  const int tx = threadIdx.x % G, ty = threadIdx.x / G;
  double r[TB][TB];
#pragma unroll
  for (int a = 0; a < TB; ++a)
#pragma unroll
    for (int c = 0; c < TB; ++c) {
      const int i = ty + G * a, j = tx + G * c;
      r[a][c] = (i < M && j < M) ? Ab[(long)i * lda + j] : (i == j ? 1.0 : 0.0);
    }
  int fail = 0;
  const int Mp = (M + NB - 1) / NB * NB;
  for (int k = 0; k < Mp; k += NB) {
    const int buf = (k / NB) & 1, kb = k / G, km = k % G;
    const bool rowK = ty >= km && ty < km + NB, colK = tx >= km && tx < km + NB;
    const int tr = ty - km, uc = tx - km;
    // (1) rows K (owned by ty in [km, km+NB)) and columns K (tx in [km, km+NB)) to LDS
    if (rowK) {
#pragma unroll
      for (int a = 0; a < TB; ++a)
        if (a == kb)
#pragma unroll
          for (int c = 0; c < TB; ++c) rowk[buf][tr][tx + G * c] = r[a][c];
    }
    if (colK) {
#pragma unroll
      for (int c = 0; c < TB; ++c)
        if (c == kb)
#pragma unroll
          for (int a = 0; a < TB; ++a) colk[buf][uc][ty + G * a] = r[a][c];
    }
    __syncthreads();
    // (2) wave 0: P^-1 by scalar Gauss-Jordan on the NB x NB pivot block, its pivots,
    //     and rv = P^-1 A[K, :] for every column, into LDS
    if (threadIdx.x < 64) {
      double Pi[NB][NB];
#pragma unroll
      for (int t = 0; t < NB; ++t)
#pragma unroll
        for (int u = 0; u < NB; ++u) Pi[t][u] = rowk[buf][t][k + u];
      int bad = 0;
#pragma unroll
      for (int s2 = 0; s2 < NB; ++s2) {
        const double p = Pi[s2][s2];
        if (!bad && !(p > 0.0 && p < INFINITY)) bad = s2 + 1;
        const double rp = 1.0 / p;
        if (threadIdx.x == 0) piv[k + s2] = p;
        double rs[NB], cs[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) rs[u] = Pi[s2][u] * rp;
#pragma unroll
        for (int t = 0; t < NB; ++t) cs[t] = Pi[t][s2];
#pragma unroll
        for (int t = 0; t < NB; ++t)
#pragma unroll
          for (int u = 0; u < NB; ++u) Pi[t][u] = fma(-cs[t], rs[u], Pi[t][u]);
#pragma unroll
        for (int u = 0; u < NB; ++u) Pi[s2][u] = rs[u];
#pragma unroll
        for (int t = 0; t < NB; ++t) Pi[t][s2] = -cs[t] * rp;
        Pi[s2][s2] = rp;
      }
      if (threadIdx.x == 0) {
#pragma unroll
        for (int t = 0; t < NB; ++t)
#pragma unroll
          for (int u = 0; u < NB; ++u) Pis[buf][t][u] = Pi[t][u];
        failk[buf] = bad ? k + bad : 0;
      }
      for (int j = threadIdx.x; j < W; j += 64) {
        double rk[NB];
#pragma unroll
        for (int u = 0; u < NB; ++u) rk[u] = rowk[buf][u][j];
#pragma unroll
        for (int t = 0; t < NB; ++t) {
          double v = 0.0;
#pragma unroll
          for (int u = 0; u < NB; ++u) v = fma(Pi[t][u], rk[u], v);
          rvs[buf][t][j] = v;
        }
      }
    }
    __syncthreads();
    // (3) every thread: A[i, j] -= A[i, K] rv[:, j]; rows K <- rv, columns K <- -A[:, K] P^-1,
    //     block K x K <- P^-1
    if (failk[buf]) {
      fail = failk[buf];
      break;
    }
    double cv[TB][NB];
#pragma unroll
    for (int a = 0; a < TB; ++a)
#pragma unroll
      for (int t = 0; t < NB; ++t) cv[a][t] = colk[buf][t][ty + G * a];
#pragma unroll
    for (int c = 0; c < TB; ++c) {
      double rv[NB];
#pragma unroll
      for (int t = 0; t < NB; ++t) rv[t] = rvs[buf][t][tx + G * c];
#pragma unroll
      for (int a = 0; a < TB; ++a) {
        double v = r[a][c];
#pragma unroll
        for (int t = 0; t < NB; ++t) v = fma(-cv[a][t], rv[t], v);
        r[a][c] = v;
      }
      if (rowK) {
        double mine = rv[0];
#pragma unroll
        for (int t = 1; t < NB; ++t)
          if (t == tr) mine = rv[t];
#pragma unroll
        for (int a = 0; a < TB; ++a)
          if (a == kb) r[a][c] = mine;
      }
    }
    if (colK) {
      double pc[NB];
#pragma unroll
      for (int t = 0; t < NB; ++t) pc[t] = Pis[buf][t][uc];
#pragma unroll
      for (int c = 0; c < TB; ++c)
        if (c == kb) {
#pragma unroll
          for (int a = 0; a < TB; ++a) {
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < NB; ++t) v = fma(-cv[a][t], pc[t], v);
            r[a][c] = v;
          }
          if (rowK) {
#pragma unroll
            for (int a = 0; a < TB; ++a)
              if (a == kb) r[a][c] = Pis[buf][tr][uc];
          }
        }
    }
  }
  // log det = sum of log pivots, one log per thread and a fixed-order block sum (the
  // logs are kept out of the elimination loop, which is latency-bound)
  __syncthreads();  // the last pivots were written after the last barrier of the loop
  double ld = 0.0;
  for (int k = threadIdx.x; k < (fail ? 0 : M); k += G * G) ld += log(piv[k]);
  ld = gpx_block_sum(ld, red);
  if (threadIdx.x == 0) {
    if (logdet) logdet[b] = fail ? NAN : ld;
    if (info) info[b] = fail;
  }
  if (!out) return;
  double *Ob = out + b * sO;
#pragma unroll
  for (int a = 0; a < TB; ++a)
#pragma unroll
    for (int c = 0; c < TB; ++c) {
      const int i = ty + G * a, j = tx + G * c;
      if (i < M && j < M) Ob[(long)i * ldo + j] = fail ? NAN : r[a][c];
    }
}

// NB = 4 pivots per step where the registers allow it; the 1024-thread variants for
// M > 128 hold 49-64 matrix entries in their 128 registers per lane and take one pivot
// per step.
template <int G, int TB>
int launch_gj(int M, int batch, const double *A, int lda, long sA, double *out, int ldo, long sO,
              double *logdet, int *info, hipStream_t st) {
  constexpr int NB = (G == 16 || TB <= 4) ? 4 : 1;
  hipLaunchKernelGGL((spd_gj_kernel<G, TB, NB>), dim3(batch), dim3(G * G), 0, st, M, batch, A, lda, sA,
                     out, ldo, sO, logdet, info);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

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
  hipStream_t st = (hipStream_t)stream;
  // 1024 threads from M > 32 on: more waves per matrix hide the per-step latency
  if (M <= 32) return launch_gj<16, 2>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 64) return launch_gj<32, 2>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 96) return launch_gj<32, 3>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 128) return launch_gj<32, 4>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 224) return launch_gj<32, 7>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  return launch_gj<32, 8>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
}

}  // extern "C"
