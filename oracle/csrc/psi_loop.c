/* Scalar CPU restatement of the psi1/psi2 loop of geepee's compute_psi_weave
 * (geepee/kernels.py:201-234).  TEST INFRASTRUCTURE ONLY: it is the timed CPU
 * baseline of the oracle's faithful-cost mode, never part of the product.
 *
 * Same cost structure as the reference loop: for every point n and pair
 * (a, b <= a) the exponent is accumulated over q, one exp per pair, the lower
 * triangle is mirrored, and psi1 is accumulated on the b == 0 pass.
 * Layout: mx, vx, ld1, ld2 are N x Q row-major; z is M x Q; ell2 is Q;
 * psi1 is N x M; psi2 is N x M x M.  ld1/ld2 are 0.5*log(l/(l+v)) and
 * 0.5*log(l/(l+2v)) per (n, q).
 */
#include <math.h>

void psi_loop(int N, int M, int Q, double sf2,
              const double *mx, const double *vx, const double *z,
              const double *ell2, const double *ld1, const double *ld2,
              double *psi1, double *psi2)
{
    for (int n = 0; n < N; ++n) {
        const double *mu = mx + (long)n * Q;
        const double *v = vx + (long)n * Q;
        double *p2 = psi2 + (long)n * M * M;
        for (int a = 0; a < M; ++a) {
            double e1 = 0.0;
            for (int b = 0; b <= a; ++b) {
                double e2 = 0.0;
                for (int q = 0; q < Q; ++q) {
                    double l = ell2[q];
                    double za = z[a * Q + q], zb = z[b * Q + q];
                    if (b == 0) {
                        double d = mu[q] - za;
                        e1 += -d * d / 2.0 / (v[q] + l) + ld1[n * Q + q];
                    }
                    double h = mu[q] - 0.5 * (za + zb);
                    double dz = za - zb;
                    e2 += -dz * dz / (4.0 * l) - h * h / (2.0 * v[q] + l) + ld2[n * Q + q];
                }
                double val = sf2 * sf2 * exp(e2);
                p2[(long)a * M + b] = val;
                p2[(long)b * M + a] = val;
            }
            psi1[(long)n * M + a] = sf2 * exp(e1);
        }
    }
}
