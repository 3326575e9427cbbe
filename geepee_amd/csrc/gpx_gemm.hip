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

constexpr int BK = 32;

// Output epilogue: C = v + beta C + gamma E (beta == 0 never reads C; E may be NULL; an
// E stride of 0 broadcasts one matrix over the batch).
struct GemmEpi {
  double beta, gamma;
  const double *E;
  int ldE;
  long sE;
};

__device__ __forceinline__ void epi_store(double *c, double v, const GemmEpi &ep, int b, int i, int j) {
  double r = v;
  if (ep.beta != 0.0) r += ep.beta * *c;
  if (ep.E) r += ep.gamma * ep.E[b * ep.sE + (long)i * ep.ldE + j];
  *c = r;
}

// Stage one BK-slice of op(A) (BM x BK) and op(B) (BK x BN) into registers.
template <int BM, int BN>
__device__ __forceinline__ void gemm_load(double *ra, double *rb, const double *__restrict__ A,
                                          const double *__restrict__ B, int tA, int tB, int lda, int ldb,
                                          int m, int n, int m0, int n0, int k0, int kend) {
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < LA; ++u) {
    const int e = tid + 256 * u;
    int mm, kk;
    if (tA) { kk = e / BM; mm = e % BM; } else { mm = e / BK; kk = e % BK; }
    const int gm = m0 + mm, gk = k0 + kk;
    ra[u] = (gm < m && gk < kend) ? (tA ? A[(long)gk * lda + gm] : A[(long)gm * lda + gk]) : 0.0;
  }
#pragma unroll
  for (int u = 0; u < LB; ++u) {
    const int e = tid + 256 * u;
    int nn, kk;
    if (tB) { nn = e / BK; kk = e % BK; } else { kk = e / BN; nn = e % BN; }
    const int gn = n0 + nn, gk = k0 + kk;
    rb[u] = (gn < n && gk < kend) ? (tB ? B[(long)gn * ldb + gk] : B[(long)gk * ldb + gn]) : 0.0;
  }
}

template <int BM, int BN>
__device__ __forceinline__ void gemm_store(const double *ra, const double *rb, double (*As)[BM + 1],
                                           double (*Bs)[BN + 1], int tA, int tB) {
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  const int tid = threadIdx.x;
#pragma unroll
  for (int u = 0; u < LA; ++u) {
    const int e = tid + 256 * u;
    int mm, kk;
    if (tA) { kk = e / BM; mm = e % BM; } else { mm = e / BK; kk = e % BK; }
    As[kk][mm] = ra[u];
  }
#pragma unroll
  for (int u = 0; u < LB; ++u) {
    const int e = tid + 256 * u;
    int nn, kk;
    if (tB) { nn = e / BK; kk = e % BK; } else { kk = e / BN; nn = e % BN; }
    Bs[kk][nn] = rb[u];
  }
}

// Block -> (column tile, row tile, z) of a batched product (z = batch element, or batch
// element and split).  With xcd the launch is 1-D and all the tiles of one z run on one
// XCD: blocks are dealt to the eight XCDs round-robin by id, so block p runs on XCD p % 8;
// z takes XCD z % 8 and its tiles are the blocks p = 8 (tpz (z / 8) + t) + z % 8.  The
// operand panels that a z's tiles re-read then stay in that XCD's L2.
struct TileIdx {
  int x, y, z;
};
__device__ __forceinline__ TileIdx gemm_tile(int xcd, int tx, int ty) {
  if (!xcd) return {(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
  const int tpz = tx * ty, p = blockIdx.x, j = p >> 3;
  const int t = j % tpz;
  return {t % tx, t / tx, (p & 7) + 8 * (j / tpz)};
}

// 4 waves as 2x2; each wave owns (BM/2) x (BN/2) of 16x16 MFMA tiles.  K advances in
// slices of BK through two LDS buffers; the next slice is fetched from global memory into
// registers while the current one feeds the matrix cores (one barrier per slice).
template <int BM, int BN>
__global__ __launch_bounds__(256) void gemm_kernel(
    int m, int n, int k, int tA, int tB, double alpha,
    const double *__restrict__ A, int lda, long sA,
    const double *__restrict__ B, int ldb, long sB,
    GemmEpi ep, double *__restrict__ C, int ldc, long sC,
    int batch, int splits, int kchunk, double *__restrict__ part, int xcd) {
  constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  __shared__ double As[2][BK][BM + 1];
  __shared__ double Bs[2][BK][BN + 1];
  const TileIdx ti = gemm_tile(xcd, (n + BN - 1) / BN, (m + BM - 1) / BM);
  const int z = ti.z;
  const int bidx = z / splits, sidx = z % splits;
  if (bidx >= batch) return;
  A += bidx * sA;
  B += bidx * sB;
  const int m0 = ti.y * BM, n0 = ti.x * BN;
  const int kbeg = sidx * kchunk, kend = min(k, kbeg + kchunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  gpx_d4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = gpx_d4{0.0, 0.0, 0.0, 0.0};
  double ra[LA], rb[LB];
  int cur = 0;
  if (kbeg < kend) {
    gemm_load<BM, BN>(ra, rb, A, B, tA, tB, lda, ldb, m, n, m0, n0, kbeg, kend);
    gemm_store<BM, BN>(ra, rb, As[0], Bs[0], tA, tB);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) gemm_load<BM, BN>(ra, rb, A, B, tA, tB, lda, ldb, m, n, m0, n0, k0 + BK, kend);
#pragma unroll
    for (int kq = 0; kq < BK; kq += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[cur][kq + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) b[j] = Bs[cur][kq + (lane >> 4)][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = gpx_mfma16(a[i], b[j], acc[i][j]);
    }
    if (more) gemm_store<BM, BN>(ra, rb, As[cur ^ 1], Bs[cur ^ 1], tA, tB);
    __syncthreads();
    cur ^= 1;
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
        if (gm < m && gn < n) epi_store(C + (long)gm * ldc + gn, alpha * acc[i][j][r], ep, bidx, gm, gn);
      }
}

// Sum of the split-K partials part[q][b][i][j] (q < splits) and the epilogue.  A workgroup
// takes 32 consecutive outputs; its 8 groups of 32 threads each sum the splits q = g,
// g + 8, ... of them (coalesced rows, loads in flight together), and the 8 group sums are
// added in group order, so every output's order of summation is fixed.  splits = 0 (k = 0)
// stores the epilogue of a zero product.  Grid: ceil(batch*m*n / 32) workgroups of 256.
constexpr int SPLITK_OUT = 32;
__global__ __launch_bounds__(256) void splitk_reduce(int m, int n, int batch, int splits, double alpha,
                                                     const double *__restrict__ part, GemmEpi ep,
                                                     double *__restrict__ C, int ldc, long sC) {
  __shared__ double red[8][SPLITK_OUT + 1];
  const long mn = (long)m * n, total = (long)batch * mn;
  const int o = threadIdx.x & (SPLITK_OUT - 1), g = threadIdx.x / SPLITK_OUT;
  const long e = (long)blockIdx.x * SPLITK_OUT + o;
  double s = 0.0;
  if (e < total) {
#pragma unroll 4
    for (int q = g; q < splits; q += 8) s += part[(long)q * total + e];
  }
  red[g][o] = s;
  __syncthreads();
  if (g == 0 && e < total) {
    double t = red[0][o];
#pragma unroll
    for (int h = 1; h < 8; ++h) t += red[h][o];
    const int b = (int)(e / mn);
    const long r = e % mn;
    const int i = (int)(r / n), j = (int)(r % n);
    epi_store(C + b * sC + (long)i * ldc + j, alpha * t, ep, b, i, j);
  }
}

// Short contractions (k <= FULLK_MAX): every operand element of the block's 32-row
// panel of op(A) and 32-column panel of op(B) is fetched in one round into LDS (one
// memory latency instead of one per K slice), then the 2x2 waves run their MFMAs over
// the whole K.  Same output layout and epilogue as gemm_kernel; no split-K.
constexpr int FULLK_MAX = 128;

__global__ __launch_bounds__(256) void gemm_fullk_kernel(
    int m, int n, int k, int tA, int tB, double alpha,
    const double *__restrict__ A, int lda, long sA,
    const double *__restrict__ B, int ldb, long sB,
    GemmEpi ep, double *__restrict__ C, int ldc, long sC, int batch, int xcd) {
  __shared__ double As[FULLK_MAX][33];
  __shared__ double Bs[FULLK_MAX][33];
  const TileIdx ti = gemm_tile(xcd, (n + 31) / 32, (m + 31) / 32);
  const int bidx = ti.z;
  if (bidx >= batch) return;
  A += bidx * sA;
  B += bidx * sB;
  const int m0 = ti.y * 32, n0 = ti.x * 32;
  const int kp = (k + 3) & ~3;
  const int tid = threadIdx.x;
  // all loads first (16 + 16 per thread in flight), then the LDS stores.  Thread t covers
  // k = k0 + 8j of one row / column (shifts only, no division), reading along the
  // operand's contiguous dimension.
  constexpr int U = FULLK_MAX / 8;
  double ra[U], rb[U];
  // op(A)[mm][kk]: A stored (m x k) reads 8 consecutive k of a row; A^T stored (k x m)
  // reads 32 consecutive m of a k-row
  const int am = tA ? (tid & 31) : (tid >> 3), ak = tA ? (tid >> 5) : (tid & 7);
  const int bn = tB ? (tid >> 3) : (tid & 31), bk = tB ? (tid & 7) : (tid >> 5);
  const int gam = m0 + am, gbn = n0 + bn;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int kka = ak + 8 * u, kkb = bk + 8 * u;
    ra[u] = (gam < m && kka < k) ? (tA ? A[(long)kka * lda + gam] : A[(long)gam * lda + kka]) : 0.0;
    rb[u] = (gbn < n && kkb < k) ? (tB ? B[(long)gbn * ldb + kkb] : B[(long)kkb * ldb + gbn]) : 0.0;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int kka = ak + 8 * u, kkb = bk + 8 * u;
    if (kka < kp) As[kka][am] = ra[u];
    if (kkb < kp) Bs[kkb][bn] = rb[u];
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6, wm = wave >> 1, wn = wave & 1;
  gpx_d4 acc = gpx_d4{0.0, 0.0, 0.0, 0.0};
  // eight 4-wide slices at a time: their LDS reads issue ahead of the MFMAs, which then
  // run back to back on one accumulator (same order as one slice at a time)
  const int ar = wm * 16 + (lane & 15), bc = wn * 16 + (lane & 15), g = lane >> 4;
  int kq = 0;
  for (; kq + 32 <= kp; kq += 32) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a[u] = As[kq + 4 * u + g][ar];
      b[u] = Bs[kq + 4 * u + g][bc];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = gpx_mfma16(a[u], b[u], acc);
  }
  for (; kq < kp; kq += 4) acc = gpx_mfma16(As[kq + g][ar], Bs[kq + g][bc], acc);
  C += bidx * sC;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int gm = m0 + wm * 16 + (lane >> 4) + 4 * r;
    const int gn = n0 + wn * 16 + (lane & 15);
    if (gm < m && gn < n) epi_store(C + (long)gm * ldc + gn, alpha * acc[r], ep, bidx, gm, gn);
  }
}

// k == 1: an outer product.  One thread per output over the flattened (batch, row,
// column) index, with the product and epilogue of the MFMA path, which for k = 1 adds
// exact zeros to round(a b), so the results are the same bits at a fraction of the latency.
__global__ __launch_bounds__(256) void rank1_kernel(int m, int n, int tA, int tB, double alpha,
                                                    const double *__restrict__ A, int lda, long sA,
                                                    const double *__restrict__ B, int ldb, long sB, GemmEpi ep,
                                                    double *__restrict__ C, int ldc, long sC, int batch) {
  const long mn = (long)m * n, e = blockIdx.x * 256L + threadIdx.x;
  if (e >= mn * batch) return;
  const int b = (int)(e / mn), r = (int)(e - b * mn), i = r / n, j = r - i * n;
  const double a = tA ? A[b * sA + i] : A[b * sA + (long)i * lda];
  const double bb = tB ? B[b * sB + (long)j * ldb] : B[b * sB + j];
  epi_store(C + b * sC + (long)i * ldc + j, alpha * (a * bb), ep, b, i, j);
}

// n == 1: matrix-vector products.  One wave per output row: lanes stride over k, then a
// fixed-order wave reduction, so a row's value depends only on its own operands.
__global__ __launch_bounds__(256) void gemv_kernel(int m, int k, int tA, int tB, double alpha,
                                                   const double *__restrict__ A, int lda, long sA,
                                                   const double *__restrict__ B, int ldb, long sB, GemmEpi ep,
                                                   double *__restrict__ C, int ldc, long sC, int batch) {
  const long o = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (o >= (long)batch * m) return;  // whole wave; no barriers below
  const int b = (int)(o / m), i = (int)(o - (long)b * m);
  const double *Ab = A + b * sA, *Bb = B + b * sB;
  double s = 0.0;
  for (int kk = lane; kk < k; kk += 64)
    s = fma(tA ? Ab[(long)kk * lda + i] : Ab[(long)i * lda + kk], tB ? Bb[kk] : Bb[(long)kk * ldb], s);
  s = gpx_wave_sum(s);
  if (lane == 0) epi_store(C + b * sC + (long)i * ldc, alpha * s, ep, b, i, 0);
}

}  // namespace

extern "C" {

// Workspace (doubles) that gpx_gemm_batched needs for a given split count.
long gpx_gemm_workspace(int m, int n, int batch, int splits) {
  return splits > 1 ? (long)splits * batch * m * n : 0;
}

int gpx_gemm_batched_ex(int transA, int transB, int m, int n, int k, double alpha,
                        const double *A, int lda, long strideA,
                        const double *B, int ldb, long strideB,
                        double beta, double *C, int ldc, long strideC,
                        double gamma, const double *E, int ldE, long strideE,
                        int batch, int splits, double *work, int tile, void *stream) {
  if (m < 0 || n < 0 || k < 0 || batch < 0 || splits < 1) return GPX_EINVAL;
  if (E && (ldE < n || strideE < 0)) return GPX_EINVAL;
  GemmEpi ep{beta, gamma, E, ldE, strideE};
  if (m == 0 || n == 0 || batch == 0) return GPX_OK;
  if (!A || !B || !C) return GPX_EINVAL;
  if (splits > 1 && !work) return GPX_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (k == 0) {  // C = beta * C
    hipLaunchKernelGGL(splitk_reduce, dim3(gpx_cdiv((long)batch * m * n, SPLITK_OUT)), dim3(256), 0, st, m, n, batch,
                       0, 0.0, (const double *)nullptr, ep, C, ldc, strideC);
    GPX_CHECK_LAUNCH();
    return GPX_OK;
  }
  // outer products and matrix-vector products (automatic tile choice only)
  if (tile == 0 && splits == 1 && k == 1) {
    hipLaunchKernelGGL(rank1_kernel, dim3(gpx_cdiv((long)batch * m * n, 256)), dim3(256), 0, st, m, n, transA,
                       transB, alpha, A, lda, strideA, B, ldb, strideB, ep, C, ldc, strideC, batch);
    GPX_CHECK_LAUNCH();
    return GPX_OK;
  }
  if (tile == 0 && splits == 1 && n == 1) {
    hipLaunchKernelGGL(gemv_kernel, dim3(gpx_cdiv((long)batch * m, 4)), dim3(256), 0, st, m, k, transA, transB,
                       alpha, A, lda, strideA, B, ldb, strideB, ep, C, ldc, strideC, batch);
    GPX_CHECK_LAUNCH();
    return GPX_OK;
  }
  if (splits == 1 && k <= FULLK_MAX && tile != 64 &&
      (tile == 32 || (long)gpx_cdiv(m, 64) * gpx_cdiv(n, 64) * batch < 256)) {
    if (batch > 65535) return GPX_EINVAL;
    const int xcd = batch >= 8;
    const unsigned tpz = gpx_cdiv(n, 32) * gpx_cdiv(m, 32);
    dim3 grid = xcd ? dim3(8 * gpx_cdiv(batch, 8) * tpz) : dim3(gpx_cdiv(n, 32), gpx_cdiv(m, 32), batch);
    hipLaunchKernelGGL(gemm_fullk_kernel, grid, dim3(256), 0, st, m, n, k, transA, transB, alpha, A, lda, strideA,
                       B, ldb, strideB, ep, C, ldc, strideC, batch, xcd);
    GPX_CHECK_LAUNCH();
    return GPX_OK;
  }
  int kchunk = (k + splits - 1) / splits;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  splits = (k + kchunk - 1) / kchunk;
  double *part = splits > 1 ? work : nullptr;
  if ((long)batch * splits > 65535L) return GPX_EINVAL;
  const int nz = batch * splits, xcd = nz >= 8;
  if (tile == 32 || (tile == 0 && (long)gpx_cdiv(m, 64) * gpx_cdiv(n, 64) * batch * splits < 256)) {
    const unsigned tpz = gpx_cdiv(n, 32) * gpx_cdiv(m, 32);
    dim3 grid = xcd ? dim3(8 * gpx_cdiv(nz, 8) * tpz) : dim3(gpx_cdiv(n, 32), gpx_cdiv(m, 32), nz);
    hipLaunchKernelGGL((gemm_kernel<32, 32>), grid, dim3(256), 0, st, m, n, k, transA, transB, alpha,
                       A, lda, strideA, B, ldb, strideB, ep, C, ldc, strideC, batch, splits, kchunk, part, xcd);
  } else {
    const unsigned tpz = gpx_cdiv(n, 64) * gpx_cdiv(m, 64);
    dim3 grid = xcd ? dim3(8 * gpx_cdiv(nz, 8) * tpz) : dim3(gpx_cdiv(n, 64), gpx_cdiv(m, 64), nz);
    hipLaunchKernelGGL((gemm_kernel<64, 64>), grid, dim3(256), 0, st, m, n, k, transA, transB, alpha,
                       A, lda, strideA, B, ldb, strideB, ep, C, ldc, strideC, batch, splits, kchunk, part, xcd);
  }
  GPX_CHECK_LAUNCH();
  if (part) {
    const long total = (long)batch * m * n;
    hipLaunchKernelGGL(splitk_reduce, dim3(gpx_cdiv(total, SPLITK_OUT)), dim3(256), 0, st, m, n, batch, splits, alpha,
                       (const double *)part, ep, C, ldc, strideC);
    GPX_CHECK_LAUNCH();
  }
  return GPX_OK;
}

int gpx_gemm_batched(int transA, int transB, int m, int n, int k, double alpha,
                     const double *A, int lda, long strideA,
                     const double *B, int ldb, long strideB,
                     double beta, double *C, int ldc, long strideC,
                     int batch, int splits, double *work, int tile, void *stream) {
  return gpx_gemm_batched_ex(transA, transB, m, n, k, alpha, A, lda, strideA, B, ldb, strideB, beta, C, ldc,
                             strideC, 0.0, nullptr, 0, 0, batch, splits, work, tile, stream);
}

}  // extern "C"
