// Batched, strided FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4f64).
//
// Row-major C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b].  It carries every dense
// contraction of the layer algebra that the reference writes as np.dot / np.einsum
// over D x M x M stacks (aep_models.py:156, 197, 239-258, 453-493, 525-546;
// base_models.py:470-488, 497), including the N x M^2 projections of psi2.
// Split-K partial slabs (one per z-slice of the grid) are summed by a deterministic
// second pass, so results do not depend on scheduling.
#include "gpx_common.h"

namespace {

constexpr int BK = 16;

template <int BM, int BN>
__global__ __launch_bounds__(256) void gemm_kernel(
    int m, int n, int k, int tA, int tB, double alpha,
    const double *__restrict__ A, int lda, long sA,
    const double *__restrict__ B, int ldb, long sB,
    double beta, double *__restrict__ C, int ldc, long sC,
    int batch, int splits, int kchunk, double *__restrict__ part) {
  // 4 waves as 2x2; each wave owns a (BM/2) x (BN/2) block of 16x16 MFMA tiles.
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  const int z = blockIdx.z;
  const int bidx = z / splits, sidx = z % splits;
  if (bidx >= batch) return;
  A += bidx * sA;
  B += bidx * sB;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kbeg = sidx * kchunk, kend = min(k, kbeg + kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  gpx_d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = gpx_d4{0.0, 0.0, 0.0, 0.0};

  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    // stage A tile as As[kk][mm]
    for (int e = tid; e < BM * BK; e += 256) {
      int mm, kk;
      if (tA) { kk = e / BM; mm = e % BM; } else { mm = e / BK; kk = e % BK; }
      const int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < m && gk < kend) v = tA ? A[(long)gk * lda + gm] : A[(long)gm * lda + gk];
      As[kk][mm] = v;
    }
    for (int e = tid; e < BN * BK; e += 256) {
      int nn, kk;
      if (tB) { nn = e / BK; kk = e % BK; } else { kk = e / BN; nn = e % BN; }
      const int gn = n0 + nn, gk = k0 + kk;
      double v = 0.0;
      if (gn < n && gk < kend) v = tB ? B[(long)gn * ldb + gk] : B[(long)gk * ldb + gn];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kq = 0; kq < BK; kq += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[kq + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[kq + (lane >> 4)][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = gpx_mfma16(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }

  if (part) {
    double *P = part + ((long)sidx * batch + bidx) * (long)m * n;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gm = m0 + wm * WM + i * 16 + (lane >> 4) + 4 * r;
          const int gn = n0 + wn * WN + j * 16 + (lane & 15);
          if (gm < m && gn < n) P[(long)gm * n + gn] = acc[i][j][r];
        }
    return;
  }
  C += bidx * sC;
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * WM + i * 16 + (lane >> 4) + 4 * r;
        const int gn = n0 + wn * WN + j * 16 + (lane & 15);
        if (gm < m && gn < n) {
          double *c = C + (long)gm * ldc + gn;
          const double v = alpha * acc[i][j][r];
          *c = (beta == 0.0) ? v : v + beta * *c;
        }
      }
}

__global__ void splitk_reduce(int m, int n, int batch, int splits, double alpha,
                              const double *__restrict__ part, double beta,
                              double *__restrict__ C, int ldc, long sC) {
  const long total = (long)batch * m * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total;
       e += (long)gridDim.x * blockDim.x) {
    const int b = (int)(e / ((long)m * n));
    const long r = e % ((long)m * n);
    const int i = (int)(r / n), j = (int)(r % n);
    double s = 0.0;
    for (int q = 0; q < splits; ++q) s += part[((long)q * batch + b) * (long)m * n + r];
    double *c = C + b * sC + (long)i * ldc + j;
    const double v = alpha * s;
    *c = (beta == 0.0) ? v : v + beta * *c;
  }
}

}  // namespace

extern "C" {

// Workspace (doubles) that gpx_gemm_batched needs for a given split count.
long gpx_gemm_workspace(int m, int n, int batch, int splits) {
  return splits > 1 ? (long)splits * batch * m * n : 0;
}

int gpx_gemm_batched(int transA, int transB, int m, int n, int k, double alpha,
                     const double *A, int lda, long strideA,
                     const double *B, int ldb, long strideB,
                     double beta, double *C, int ldc, long strideC,
                     int batch, int splits, double *work, int tile, void *stream) {
  if (m < 0 || n < 0 || k < 0 || batch < 0 || splits < 1) return GPX_EINVAL;
  if (m == 0 || n == 0 || batch == 0) return GPX_OK;
  if (!A || !B || !C) return GPX_EINVAL;
  if (splits > 1 && !work) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (k == 0) {  // C = beta * C
    hipLaunchKernelGGL(splitk_reduce, dim3(gpx_cdiv((long)batch * m * n, 256) > 4096 ? 4096 : gpx_cdiv((long)batch * m * n, 256)),
                       dim3(256), 0, st, m, n, batch, 0, 0.0, (const double *)nullptr, beta, C, ldc, strideC);
    GPX_CHECK_LAUNCH();
    return GPX_OK;
  }
  int kchunk = (k + splits - 1) / splits;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  splits = (k + kchunk - 1) / kchunk;
  double *part = splits > 1 ? work : nullptr;
  if ((long)batch * splits > 65535L) return GPX_EINVAL;
  if (tile == 32 || (tile == 0 && (long)gpx_cdiv(m, 64) * gpx_cdiv(n, 64) * batch * splits < 256)) {
    dim3 grid(gpx_cdiv(n, 32), gpx_cdiv(m, 32), batch * splits);
    hipLaunchKernelGGL((gemm_kernel<32, 32>), grid, dim3(256), 0, st, m, n, k, transA, transB, alpha,
                       A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC, batch, splits, kchunk, part);
  } else {
    dim3 grid(gpx_cdiv(n, 64), gpx_cdiv(m, 64), batch * splits);
    hipLaunchKernelGGL((gemm_kernel<64, 64>), grid, dim3(256), 0, st, m, n, k, transA, transB, alpha,
                       A, lda, strideA, B, ldb, strideB, beta, C, ldc, strideC, batch, splits, kchunk, part);
  }
  GPX_CHECK_LAUNCH();
  if (part) {
    long total = (long)batch * m * n;
    unsigned g = gpx_cdiv(total, 256);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(splitk_reduce, dim3(g), dim3(256), 0, st, m, n, batch, splits, alpha,
                       (const double *)part, beta, C, ldc, strideC);
    GPX_CHECK_LAUNCH();
  }
  return GPX_OK;
}

}  // extern "C"
