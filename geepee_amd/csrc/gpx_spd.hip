// Batched SPD inverse + log-determinant, one workgroup per M x M matrix.
//
// Replaces the per-output-dim np.linalg.inv / slogdet / solve calls of the layer
// algebra (base_models.py:462-476; aep_models.py:62-95, 523-533).  The matrix stays in
// registers for the whole factorisation: thread (ty, tx) of a G x G block owns the
// strided sub-grid rows ty + G*a, cols tx + G*c (a, c < TB), so a step touches every
// element once and reads only one row and one column of broadcast data from LDS.
// The algorithm is Gauss-Jordan elimination without pivoting, which is stable for SPD
// input; its pivots are the D of A = L D L^T, so sum(log pivot) is log det A and the
// first non-positive pivot flags a matrix that is not positive definite.
#include "gpx_common.h"

namespace {

template <int G, int TB>
__global__ __launch_bounds__(G *G) void spd_gj_kernel(int M, int batch, const double *__restrict__ A,
                                                      int lda, long sA, double *__restrict__ out,
                                                      int ldo, long sO, double *__restrict__ logdet,
                                                      int *__restrict__ info) {
  constexpr int W = G * TB;
  __shared__ double rowk[2][W];
  __shared__ double colk[2][W];
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
  double ld = 0.0;
  int fail = 0;
  for (int k = 0; k < M; ++k) {
    const int buf = k & 1, kb = k / G, km = k % G;
    if (ty == km) {
#pragma unroll
      for (int a = 0; a < TB; ++a)
        if (a == kb)
#pragma unroll
          for (int c = 0; c < TB; ++c) rowk[buf][tx + G * c] = r[a][c];
    }
    if (tx == km) {
#pragma unroll
      for (int c = 0; c < TB; ++c)
        if (c == kb)
#pragma unroll
          for (int a = 0; a < TB; ++a) colk[buf][ty + G * a] = r[a][c];
    }
    __syncthreads();
    const double p = rowk[buf][k];
    if (!(p > 0.0) || !(p < INFINITY)) {
      fail = k + 1;
      break;
    }
    const double rp = 1.0 / p;
    ld += log(p);
    double cv[TB], rv[TB];
#pragma unroll
    for (int a = 0; a < TB; ++a) cv[a] = colk[buf][ty + G * a];
#pragma unroll
    for (int c = 0; c < TB; ++c) rv[c] = rowk[buf][tx + G * c] * rp;
#pragma unroll
    for (int a = 0; a < TB; ++a)
#pragma unroll
      for (int c = 0; c < TB; ++c) r[a][c] = fma(-cv[a], rv[c], r[a][c]);
    // row k becomes A[k][j] / p, column k becomes -A[i][k] / p, the pivot 1 / p
    if (ty == km) {
#pragma unroll
      for (int a = 0; a < TB; ++a)
        if (a == kb)
#pragma unroll
          for (int c = 0; c < TB; ++c) r[a][c] = rv[c];
    }
    if (tx == km) {
#pragma unroll
      for (int c = 0; c < TB; ++c)
        if (c == kb)
#pragma unroll
          for (int a = 0; a < TB; ++a) r[a][c] = -cv[a] * rp;
    }
    if (ty == km && tx == km) {
#pragma unroll
      for (int a = 0; a < TB; ++a)
        if (a == kb)
#pragma unroll
          for (int c = 0; c < TB; ++c)
            if (c == kb) r[a][c] = rp;
    }
  }
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

template <int G, int TB>
int launch_gj(int M, int batch, const double *A, int lda, long sA, double *out, int ldo, long sO,
              double *logdet, int *info, hipStream_t st) {
  hipLaunchKernelGGL((spd_gj_kernel<G, TB>), dim3(batch), dim3(G * G), 0, st, M, batch, A, lda, sA,
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
  if (M <= 32) return launch_gj<16, 2>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 64) return launch_gj<16, 4>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 112) return launch_gj<16, 7>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 128) return launch_gj<16, 8>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  if (M <= 224) return launch_gj<32, 7>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
  return launch_gj<32, 8>(M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info, st);
}

}  // extern "C"
