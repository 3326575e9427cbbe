// Grouped predictive mean / variance for many resident models in one launch.
//
// One workgroup per work item = up to SERVE_TILE query points of one request against one
// model.  Each model is a chain of at most two sparse-GP layers in posterior form:
//   deterministic inputs (first layer): kfu = SE(x, Z), m = kfu A^T,
//       v = sf2 + kfu B_det kfu^T                      (base_models.py:265-284)
//   Gaussian inputs (second layer):      m = psi1 A^T,
//       v = sf2 + <B_sto, psi2> - m^2                   (base_models.py:286-307)
// psi2 is never stored: every entry is recomputed from the point's input moments and
// folded into the B_sto contraction.  Every per-point reduction runs in a fixed order
// over thread indices, independent of the other points of the launch, so a query's
// result is bitwise the same alone or batched with any other requests.
#include "gpx_common.h"

#define SERVE_TILE 16
#define SERVE_MMAX 256
#define SERVE_QMAX 16
#define SERVE_HMAX 8

struct GpxServeLayer {
  int type;  // 0: deterministic inputs, 1: Gaussian inputs
  int M, Din, Dout;
  const double *zu, *ls, *sf;
  const double *A;  // Dout x M
  const double *B;  // Dout x M x M
};

struct GpxServeModel {
  int nlayers;
  int Din, Dout;
  int pad_;
  const double *sn;  // log noise std for predict_y (may be NULL)
  GpxServeLayer layer[2];
};

struct GpxServeItem {
  int model;
  int rows;
  long in_off;   // doubles into X (rows x Din)
  long out_off;  // doubles into mean / var (rows x Dout)
};

namespace {

// Sum TILE per-thread values over the block (fixed order); result in tot[0..TILE).
__device__ __forceinline__ void tile_sum(const double *v, double (*red)[SERVE_TILE], double *tot) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < SERVE_TILE; ++p) {
    const double s = gpx_wave_sum(v[p]);
    if (lane == 0) red[w][p] = s;
  }
  __syncthreads();
  if (threadIdx.x < SERVE_TILE)
    tot[threadIdx.x] = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
  __syncthreads();
}

__global__ __launch_bounds__(256) void serve_kernel(const GpxServeModel *__restrict__ models,
                                                    const GpxServeItem *__restrict__ items,
                                                    const double *__restrict__ X,
                                                    double *__restrict__ mean, double *__restrict__ var,
                                                    int noisy) {
  __shared__ double xin[SERVE_TILE][SERVE_QMAX];
  __shared__ double hm[SERVE_TILE][SERVE_HMAX], hv[SERVE_TILE][SERVE_HMAX];
  __shared__ double feat[SERVE_TILE][SERVE_MMAX];
  __shared__ double zs[SERVE_MMAX][SERVE_QMAX];
  __shared__ double red[4][SERVE_TILE];
  __shared__ double tm[SERVE_TILE], tv[SERVE_TILE];
  __shared__ double pc[SERVE_TILE], pd[SERVE_TILE][SERVE_QMAX];
  __shared__ double red2[4][16][17];
  __shared__ double vqs[SERVE_TILE][16];
  const GpxServeItem it = items[blockIdx.x];
  const GpxServeModel &mod = models[it.model];
  const int tid = threadIdx.x, rows = it.rows;
  for (int e = tid; e < SERVE_TILE * mod.Din; e += 256) {
    const int p = e / mod.Din, q = e % mod.Din;
    xin[p][q] = (p < rows) ? X[it.in_off + (long)p * mod.Din + q] : 0.0;
  }
  __syncthreads();
  for (int l = 0; l < mod.nlayers; ++l) {
    const GpxServeLayer &L = mod.layer[l];
    const int M = L.M, Q = L.Din;
    const double sf2 = exp(2.0 * L.sf[0]);
    const bool last = (l == mod.nlayers - 1);
    for (int e = tid; e < M * Q; e += 256) zs[e / Q][e % Q] = L.zu[e];
    __syncthreads();
    // features: kfu (type 0) or psi1 (type 1), SERVE_TILE x M
    for (int e = tid; e < SERVE_TILE * M; e += 256) {
      const int p = e / M, m = e % M;
      double acc = 0.0, lc = 0.0;
      for (int q = 0; q < Q; ++q) {
        const double l2 = exp(2.0 * L.ls[q]);
        if (L.type == 0) {
          const double d = xin[p][q] - zs[m][q];
          acc += d * d / l2;
        } else {
          const double v = hv[p][q], d = hm[p][q] - zs[m][q];
          acc += d * d / (v + l2);
          lc += 0.5 * log(l2 / (l2 + v));
        }
      }
      feat[p][m] = sf2 * exp(-0.5 * acc + lc);
    }
    if (L.type == 1 && tid < SERVE_TILE) {
      double c = 0.0;
      for (int q = 0; q < Q; ++q) {
        const double l2 = exp(2.0 * L.ls[q]);
        c += 0.5 * log(l2 / (l2 + 2.0 * hv[tid][q]));
        pd[tid][q] = 1.0 / (2.0 * hv[tid][q] + l2);
      }
      pc[tid] = c;
    }
    __syncthreads();
    if (L.type == 1) {
      // <B_d, psi2_p> for all d at once on the matrix cores: psi2 exponents of a 16-pair x
      // 16-point block as a K = 2Q+2 feature product, one exp per entry, then the
      // contraction with the symmetrised B_d; pairs (a, b) visited as tiles A <= B
      // (off-diagonal weight 2), row blocks split over the 4 waves, partial sums
      // combined in wave order.
      const int lane = tid & 63, wave = tid >> 6, pp = lane & 15;
      const int KS = (2 * Q + 2 + 3) / 4;
      double il4[SERVE_HMAX], lam[SERVE_HMAX];
#pragma unroll
      for (int q = 0; q < SERVE_HMAX; ++q) {
        lam[q] = (q < Q) ? exp(2.0 * L.ls[q]) : 1.0;
        il4[q] = 0.25 / lam[q];
      }
      double bf[5];
      {
        double c = 0.0;
        for (int q = 0; q < Q; ++q) {
          const double v = hv[pp][q], mu = hm[pp][q], dd = 1.0 / (2.0 * v + lam[q]);
          c += 0.5 * log(lam[q] / (lam[q] + 2.0 * v)) - dd * mu * mu;
        }
#pragma unroll
        for (int s2 = 0; s2 < 5; ++s2) {
          const int k = (lane >> 4) + 4 * s2;
          double f = 0.0;
          if (s2 < KS) {
            if (k < Q) f = 2.0 * hm[pp][k] / (2.0 * hv[pp][k] + lam[k]);
            else if (k < 2 * Q) f = 1.0 / (2.0 * hv[pp][k - Q] + lam[k - Q]);
            else if (k == 2 * Q) f = 1.0;
            else if (k == 2 * Q + 1) f = c;
          }
          bf[s2] = f;
        }
      }
      const double sf4 = sf2 * sf2;
      const int T = (M + 15) / 16;
      gpx_d4 acc = gpx_d4{0.0, 0.0, 0.0, 0.0};
      for (int B = 0; B < T; ++B)
        for (int A = 0; A <= B; ++A) {
          const double wgt = (A == B) ? sf4 : 2.0 * sf4;
          for (int i = wave; i < 16; i += 4) {
            const int a = 16 * A + i;
            if (a >= M) break;
            const int ja = 16 * B + (lane & 15);
            gpx_d4 e = gpx_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s2 = 0; s2 < 5; ++s2) {
              if (s2 < KS) {
                const int k = (lane >> 4) + 4 * s2;
                double af = 0.0;
                if (ja < M) {
                  if (k < Q) af = 0.5 * (zs[a][k] + zs[ja][k]);
                  else if (k < 2 * Q) {
                    const double hh = 0.5 * (zs[a][k - Q] + zs[ja][k - Q]);
                    af = -hh * hh;
                  } else if (k == 2 * Q) {
                    for (int q = 0; q < Q; ++q) {
                      const double dz = zs[a][q] - zs[ja][q];
                      af -= dz * dz * il4[q];
                    }
                  } else if (k == 2 * Q + 1) af = 1.0;
                }
                e = gpx_mfma16(af, bf[s2], e);
              }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int b = 16 * B + (lane >> 4) + 4 * r, d = lane & 15;
              const double pv = (b < M) ? wgt * exp(e[r]) : 0.0;
              const double bd = (b < M && d < L.Dout)
                                    ? 0.5 * (L.B[((long)d * M + a) * M + b] + L.B[((long)d * M + b) * M + a])
                                    : 0.0;
              acc = gpx_mfma16(bd, pv, acc);
            }
          }
        }
      // acc register r of lane: (d = (lane>>4) + 4r, point = lane & 15)
#pragma unroll
      for (int r = 0; r < 4; ++r) red2[wave][(lane >> 4) + 4 * r][lane & 15] = acc[r];
      __syncthreads();
      {
        const int d = tid >> 4, p = tid & 15;
        vqs[p][d] = red2[0][d][p] + red2[1][d][p] + red2[2][d][p] + red2[3][d][p];
      }
      __syncthreads();
    }
    for (int d = 0; d < L.Dout; ++d) {
      const double *Ad = L.A + (long)d * M, *Bd = L.B + (long)d * M * M;
      double vm[SERVE_TILE], vq[SERVE_TILE];
#pragma unroll
      for (int p = 0; p < SERVE_TILE; ++p) vm[p] = vq[p] = 0.0;
      const int b = tid;
      if (b < M) {
        const double ab = Ad[b];
#pragma unroll
        for (int p = 0; p < SERVE_TILE; ++p) vm[p] = feat[p][b] * ab;
        if (L.type == 0) {
          double w[SERVE_TILE];
#pragma unroll
          for (int p = 0; p < SERVE_TILE; ++p) w[p] = 0.0;
          for (int a = 0; a < M; ++a) {
            const double bab = Bd[(long)a * M + b];
#pragma unroll
            for (int p = 0; p < SERVE_TILE; ++p) w[p] = fma(feat[p][a], bab, w[p]);
          }
#pragma unroll
          for (int p = 0; p < SERVE_TILE; ++p) vq[p] = feat[p][b] * w[p];
        } else if (tid == 0) {
#pragma unroll
          for (int p = 0; p < SERVE_TILE; ++p) vq[p] = vqs[p][d];
        }
      }
      tile_sum(vm, red, tm);
      tile_sum(vq, red, tv);
      if (tid < SERVE_TILE) {
        const int p = tid;
        const double m = tm[p];
        const double v = (L.type == 0) ? sf2 + tv[p] : sf2 + tv[p] - m * m;
        if (last) {
          if (p < rows) {
            const double sn2 = (noisy && mod.sn) ? exp(2.0 * mod.sn[0]) : 0.0;
            mean[it.out_off + (long)p * L.Dout + d] = m;
            var[it.out_off + (long)p * L.Dout + d] = v + sn2;
          }
        } else {
          hm[p][d] = m;
          hv[p][d] = v;
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace

extern "C" {

int gpx_serve_tile(void) { return SERVE_TILE; }
int gpx_serve_desc_sizes(int *model_bytes, int *item_bytes) {
  if (!model_bytes || !item_bytes) return GPX_EINVAL;
  *model_bytes = (int)sizeof(GpxServeModel);
  *item_bytes = (int)sizeof(GpxServeItem);
  return GPX_OK;
}

// One launch over `nitems` work items (each <= SERVE_TILE rows of one request).  Models
// need M <= 256, layer inputs <= 16 dims, hidden width <= 8; the caller validates the
// descriptor tables (gpx_serve_validate) when it builds them.
int gpx_predict_grouped(int nitems, const void *models, const void *items, const double *X,
                        double *mean, double *var, int noisy, void *stream) {
  if (nitems < 0 || !models || !items || !X || !mean || !var) return GPX_EINVAL;
  if (nitems == 0) return GPX_OK;
  hipLaunchKernelGGL(serve_kernel, dim3(nitems), dim3(256), 0, (hipStream_t)stream,
                     (const GpxServeModel *)models, (const GpxServeItem *)items, X, mean, var, noisy);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

// Host-side check of one model descriptor (host copy) against the kernel's limits.
int gpx_serve_validate(const void *model) {
  const GpxServeModel *m = (const GpxServeModel *)model;
  if (!m || m->nlayers < 1 || m->nlayers > 2) return GPX_EINVAL;
  if (m->Din < 1 || m->Din > SERVE_QMAX) return GPX_EINVAL;
  int din = m->Din;
  for (int l = 0; l < m->nlayers; ++l) {
    const GpxServeLayer &L = m->layer[l];
    if (L.type != (l == 0 ? 0 : 1)) return GPX_EINVAL;
    if (L.M < 1 || L.M > SERVE_MMAX || L.Din != din || L.Dout < 1) return GPX_EINVAL;
    if (l < m->nlayers - 1 && L.Dout > SERVE_HMAX) return GPX_EINVAL;
    if (L.type == 1 && L.Dout > 16) return GPX_EINVAL;  // one 16-row MFMA tile of outputs
    if (!L.zu || !L.ls || !L.sf || !L.A || !L.B) return GPX_EINVAL;
    din = L.Dout;
  }
  if (din != m->Dout) return GPX_EINVAL;
  return GPX_OK;
}

}  // extern "C"
