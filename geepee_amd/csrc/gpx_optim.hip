// On-device Adam step over the flat parameter vector (sorted-key order).
#include "gpx_common.h"

namespace {

// Contraction is off so every operation rounds as numpy's does: the update is the
// same sequence of IEEE operations as geepee/utils.py:108-112 (sqrt and division are
// correctly rounded on the device).
#pragma clang fp contract(off)
__global__ void adam_kernel(long n, double *__restrict__ x, const double *__restrict__ g,
                            double *__restrict__ m, double *__restrict__ v, double lr, double b1, double b2,
                            double eps, double c1, double c2, const double *__restrict__ energy,
                            double *__restrict__ cost) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
    double gi = g[i];
    if (!isfinite(gi)) gi = 0.0;  // non-finite gradient entries become 0 (utils.py:46-53)
    const double mi = (1.0 - b1) * gi + b1 * m[i];
    const double vi = (1.0 - b2) * (gi * gi) + b2 * v[i];
    m[i] = mi;
    v[i] = vi;
    const double mhat = mi / c1, vhat = vi / c2;
    x[i] = x[i] - lr * mhat / (sqrt(vhat) + eps);
  }
  if (cost && blockIdx.x == 0 && threadIdx.x == 0) cost[0] = energy[0];
}
#pragma clang fp contract(on)

}  // namespace

extern "C" {

// One Adam iteration t (1-based) in place on x, m, v (n doubles each) from gradient g
// (utils.py:93-117: b1, b2, eps, bias correction 1 - b^t).  When cost is given,
// cost[0] = energy[0] (the objective value of this iteration) is recorded too.
int gpx_adam_step(long n, double *x, const double *g, double *m, double *v, double lr, double b1, double b2,
                  double eps, int t, const double *energy, double *cost, void *stream) {
  if (n < 1 || !x || !g || !m || !v || t < 1 || (cost && !energy)) return GPX_EINVAL;
  const double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
  unsigned grid = gpx_cdiv(n, 256);
  if (grid > 2048) grid = 2048;
  hipLaunchKernelGGL(adam_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, n, x, g, m, v, lr, b1, b2, eps,
                     c1, c2, energy, cost);
  GPX_CHECK_LAUNCH();
  return GPX_OK;
}

}  // extern "C"
