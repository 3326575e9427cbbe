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

template <int MT>
struct SpdLds {
  static constexpr int MP = 16 * MT, RS = MP + 1;  // RS: LDS row stride (doubles)
  // rowK: block row K (read by the panel only, so one buffer); colT: block column K,
  // transposed (read by the tile updates while the next block's column is published,
  // so two); rvK, Pi: the panel's results
  double rowK[16 * RS];
  double colT[2][16 * RS];
  double rvK[16 * RS];
  double Pi[16 * 17];
  double piv[MP];
  int failk[2];
  double red[16];
};

// Phase 2 for panel wave pw of block K.  Lane c < 16 holds pivot column K*16 + c, lane
// 16 + c identity column c (ending as column c of P^-1) and lanes 32.. hold CPL of the
// other columns each (ending as P^-1 A[K, col]).  Every panel wave runs the same pivots
// on the same pivot columns, so their copies agree bit for bit.
template <int MT, int CPL = 1>
__device__ __forceinline__ void spd_panel(SpdLds<MT> &sm, int K, int pw, int lane) {
  constexpr int MP = SpdLds<MT>::MP, RS = SpdLds<MT>::RS;
  int col[CPL];
  bool live[CPL];
#pragma unroll
  for (int c = 0; c < CPL; ++c) {
    live[c] = c == 0 || lane >= 32;
    col[c] = K * 16 + lane;
    if (lane >= 32) {
      const int m = 32 * CPL * pw + 32 * c + lane - 32;
      live[c] = m < MP - 16;
      col[c] = m < K * 16 ? m : m + 16;
    }
  }
  const bool ident = lane >= 16 && lane < 32;
  double v[CPL][16];
#pragma unroll
  for (int c = 0; c < CPL; ++c)
#pragma unroll
    for (int i = 0; i < 16; ++i)
      v[c][i] = ident ? (c == 0 && i == lane - 16 ? 1.0 : 0.0) : (live[c] ? sm.rowK[i * RS + col[c]] : 0.0);
  int bad = 0;
#pragma unroll
  for (int p = 0; p < 16; ++p) {
    const double d = gpx_readlane_d(v[0][p], p);
    if (!bad && !(d > 0.0 && d < INFINITY)) bad = p + 1;
    const double rd = gpx_recip(d);
    double colp[16];  // column p, read from the pivot lane
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (i != p) colp[i] = gpx_readlane_d(v[0][i], p);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const double rowp = v[c][p] * rd;
#pragma unroll
      for (int i = 0; i < 16; ++i)
        if (i != p) v[c][i] = fma(-colp[i], rowp, v[c][i]);
      v[c][p] = rowp;
    }
    if (pw == 0 && lane == 0) sm.piv[K * 16 + p] = d;
  }
  if (ident) {
    if (pw == 0) {
#pragma unroll
      for (int i = 0; i < 16; ++i) sm.Pi[i * 17 + lane - 16] = v[0][i];
    }
  } else if (lane >= 32) {
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      if (live[c]) {
#pragma unroll
        for (int i = 0; i < 16; ++i) sm.rvK[i * RS + col[c]] = v[c][i];
      }
  }
  if (pw == 0 && lane == 0) sm.failk[K & 1] = bad ? K * 16 + bad : 0;
}

// Phase 1 for one tile: block row K and block column K (transposed) to LDS.
template <int MT>
__device__ __forceinline__ void spd_publish(SpdLds<MT> &sm, int K, int I, int J, const gpx_d4 &acc, int lr,
                                            int lc) {
  constexpr int RS = SpdLds<MT>::RS;
  double *cT = sm.colT[K & 1];
  if (I == K) {
#pragma unroll
    for (int r = 0; r < 4; ++r) sm.rowK[(lr + 4 * r) * RS + J * 16 + lc] = acc[r];
  }
  if (J == K) {
#pragma unroll
    for (int r = 0; r < 4; ++r) cT[lc * RS + I * 16 + lr + 4 * r] = acc[r];
  }
}

// Phase 3 for one tile.
template <int MT>
__device__ __forceinline__ void spd_update(SpdLds<MT> &sm, int K, int I, int J, gpx_d4 &acc, int lr, int lc) {
  constexpr int RS = SpdLds<MT>::RS;
  const double *cT = sm.colT[K & 1];
  if (I == K) {
    if (J == K) {
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = sm.Pi[(lr + 4 * r) * 17 + lc];
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] = sm.rvK[(lr + 4 * r) * RS + J * 16 + lc];
    }
  } else if (J == K) {
    gpx_d4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      c = gpx_mfma16(-cT[(4 * ks + lr) * RS + I * 16 + lc], sm.Pi[(4 * ks + lr) * 17 + lc], c);
    acc = c;
  } else {
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      acc = gpx_mfma16(-cT[(4 * ks + lr) * RS + I * 16 + lc], sm.rvK[(4 * ks + lr) * RS + J * 16 + lc], acc);
  }
}

template <int MT, int NW>
__device__ __forceinline__ void spd_finish(SpdLds<MT> &sm, int M, int b, int fail, double *logdet, int *info) {
  __syncthreads();  // pivots of the last block were written after the last barrier of the loop
  double ld = 0.0;
  for (int k = threadIdx.x; k < (fail ? 0 : M); k += NW * 64) ld += log(sm.piv[k]);
  ld = gpx_block_sum(ld, sm.red);
  if (threadIdx.x == 0) {
    if (logdet) logdet[b] = fail ? NAN : ld;
    if (info) info[b] = fail;
  }
}

// Blocked Gauss-Jordan on the matrix cores, 16 x 16 pivot blocks.  Eliminating block
// K with P = A[K,K] (rows of K: A[K,:], columns of K: c = A[:,K]):
//   A[I,J] -= c_I P^-1 A[K,J]     (I, J != K)
//   A[K,J] <- P^-1 A[K,J],  A[I,K] <- -c_I P^-1,  A[K,K] <- P^-1,
// and after every block the matrix holds A^-1.  All MT x MT tiles are kept, in MFMA
// accumulator layout, dealt round-robin to the tile waves (register r of lane l holds
// tile entry [l/16 + 4r][l%16]): keeping only a symmetric half and mirroring it costs
// two orders of magnitude in accuracy downstream on ill-conditioned input, because the
// two halves of a computed inverse carry errors that cancel in later products.
// Per block: the owners of block row and column K publish them to LDS; the panel waves
// run Gauss-Jordan over the 16 scalar pivots on [P | I | A[K, other columns]], giving
// P^-1 A[K,:] and P^-1 without ever multiplying by an explicit P^-1 (which loses an order
// of magnitude on ill-conditioned blocks); the tile waves then update their tiles with
// four 16x16x4 MFMAs per tile.  The scalar pivots are the D of A = L D L^T: their logs
// sum to log det A and the first non-positive one flags a matrix that is not positive
// definite.  The matrix is padded with the identity up to a multiple of 16.  Every wave
// holds tiles and the first NPW waves also run the panel (giving the panel waves of its
// own, two columns per lane, measured slower at every size on MI355X).
template <int MT, int NW>
__global__ __launch_bounds__(NW * 64) void spd_gj_kernel(int M, int batch, const double *__restrict__ A,
                                                         int lda, long sA, double *__restrict__ out, int ldo,
                                                         long sO, double *__restrict__ logdet,
                                                         int *__restrict__ info) {
  constexpr int MP = 16 * MT;
  constexpr int T = MT * MT;
  constexpr int NPW = MP > 16 ? (MP - 16 + 31) / 32 : 1;  // panel waves
  constexpr int TPW = (T + NW - 1) / NW;
  static_assert(NPW <= NW, "not enough waves for the panel");
  __shared__ SpdLds<MT> sm;
  const int b = blockIdx.x;
  if (b >= batch) return;
  // w in a scalar register: the tile indices and their LDS offsets are then wave-uniform
  // scalars rather than a few vector registers per tile
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane >> 4, lc = lane & 15;
  int fail = 0;
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
#pragma unroll 1
  for (int K = 0; K < MT; ++K) {
    // (1) block row and column K to LDS
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int t = w + NW * u;
      if (t < T) spd_publish<MT>(sm, K, t / MT, t % MT, acc[u], lr, lc);
    }
    __syncthreads();
    // (2) the panel
    if (w < NPW) spd_panel<MT>(sm, K, w, lane);
    __syncthreads();
    if (sm.failk[K & 1]) {
      fail = sm.failk[K & 1];
      break;
    }
    // (3) the tiles, one tile's operands at a time (hoisting the LDS loads of later tiles
    // above this one's MFMAs would hold all of them in registers at once)
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
      const int t = w + NW * u;
      if (t < T) spd_update<MT>(sm, K, t / MT, t % MT, acc[u], lr, lc);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  spd_finish<MT, NW>(sm, M, b, fail, logdet, info);
  if (!out) return;
  double *Ob = out + b * sO;
  // the store indices are formed afresh from an opaque copy of the lane: otherwise the
  // compiler keeps the load's per-tile indices live through the whole elimination
  int lane2 = lane;
  asm volatile("" : "+v"(lane2));
  const int lr2 = lane2 >> 4, lc2 = lane2 & 15;
#pragma unroll
  for (int u = 0; u < TPW; ++u) {
    const int t = w + NW * u, I = t / MT, J = t % MT;
    if (t >= T) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = I * 16 + lr2 + 4 * r, j = J * 16 + lc2;
      if (i < M && j < M) Ob[(long)i * ldo + j] = fail ? NAN : acc[u][r];
    }
  }
}

template <int MT>
int launch_gj(int M, int batch, const double *A, int lda, long sA, double *out, int ldo, long sO,
                 double *logdet, int *info, hipStream_t st) {
  // 8 waves from M > 48 on (measured on MI355X against 16: equal at M = 200, faster
  // everywhere else)
  constexpr int NW = MT <= 3 ? 4 : 8;
  hipLaunchKernelGGL((spd_gj_kernel<MT, NW>), dim3(batch), dim3(NW * 64), 0, st, M, batch, A, lda, sA, out, ldo,
                     sO, logdet, info);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

typedef int (*gj_fn)(int, int, const double *, int, long, double *, int, long, double *, int *, hipStream_t);
const gj_fn kGJ[16] = {launch_gj<1>,  launch_gj<2>,  launch_gj<3>,  launch_gj<4>,
                             launch_gj<5>,  launch_gj<6>,  launch_gj<7>,  launch_gj<8>,
                             launch_gj<9>,  launch_gj<10>, launch_gj<11>, launch_gj<12>,
                             launch_gj<13>, launch_gj<14>, launch_gj<15>, launch_gj<16>};

// Pivot-failure flags of several SPD launches folded into one status word pair:
// status[0] = the largest info entry (0 = every matrix was positive definite),
// status[1] = max(status[1], status[0]) (sticky across evaluations until the host
// clears it).  One wave; the list descriptors travel by value in the kernel arguments,
// so a captured launch needs no host memory.
constexpr int kStatusLists = 16;
struct StatusLists {
  const int *ptr[kStatusLists];
  int count[kStatusLists];
};

__global__ void __launch_bounds__(64) status_fold_kernel(int nlists, StatusLists L, int *status) {
  int m = 0;
  for (int j = 0; j < nlists; ++j) {
    const int *p = L.ptr[j];
    for (int i = threadIdx.x; i < L.count[j]; i += 64) m = max(m, p[i]);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if (threadIdx.x == 0) {
    status[0] = m;
    status[1] = max(status[1], m);
  }
}

}  // namespace

extern "C" {

int gpx_status_fold(int nlists, const int *const *lists, const int *counts, int *status,
                    void *stream) {
  if (nlists < 0 || nlists > kStatusLists || !status || (nlists > 0 && (!lists || !counts)))
    return GPX_EINVAL;
  StatusLists L = {};
  for (int j = 0; j < nlists; ++j) {
    if (counts[j] < 0 || (counts[j] > 0 && !lists[j])) return GPX_EINVAL;
    L.ptr[j] = lists[j];
    L.count[j] = counts[j];
  }
  hipLaunchKernelGGL(status_fold_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, nlists, L, status);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

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
  return kGJ[(M + 15) / 16 - 1](M, batch, A, lda, strideA, out, ldo, strideOut, logdet, info,
                                   (hipStream_t)stream);
}

}  // extern "C"
