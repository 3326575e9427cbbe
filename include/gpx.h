/* libgpx: MI355X (gfx950) FP64 kernels for the sparse-GP approximate-EP hot path.
 *
 * C-ABI over caller-owned device memory.  Conventions for every entry point:
 *   - arrays are float64, row-major, on the current HIP device;
 *   - log hyper-parameters (ls, sf, sn, R) are read from device pointers, so no
 *     host synchronisation is needed and calls can be captured in a hipGraph;
 *   - `stream` is a hipStream_t (NULL = default stream);
 *   - the return value is 0 on success, GPX_EINVAL (-1) for an invalid argument
 *     and GPX_EHIP (-2) when a launch fails; no exceptions cross the ABI;
 *   - functions with a `work` argument need a device workspace of the size the
 *     matching *_workspace() query returns (in doubles);
 *   - nothing allocates device memory and nothing keeps state between calls.
 *
 * Each declaration names the reference code (thangbui/geepee, file:line) whose
 * arithmetic it replaces.  The reference's only native boundary is the weave
 * inline C++ of compute_psi_weave (geepee/kernels.py:181-240); the other entry
 * points replace numpy / scipy calls on the same path.
 */
#ifndef GPX_H
#define GPX_H

#ifdef __cplusplus
extern "C" {
#endif

#define GPX_OK 0
#define GPX_EINVAL (-1)
#define GPX_EHIP (-2)
#define GPX_ETOOBIG (-3) /* the allocation-free entry point's scratch is too small */

/* ---- dense algebra -------------------------------------------------------- */

/* C[b] = alpha op(A[b]) op(B[b]) + beta C[b] on the FP64 matrix cores; strides of 0
 * broadcast an operand over the batch; splits > 1 splits K into partial slabs summed in a
 * fixed order (needs gpx_gemm_workspace doubles); tile = 0 (auto), 32 or 64.
 * Replaces the np.dot / np.einsum contractions of geepee/aep_models.py:156, 197, 239-258,
 * 453-493, 525-546 and geepee/base_models.py:470-488, 497. */
long gpx_gemm_workspace(int m, int n, int batch, int splits);
int gpx_gemm_batched(int transA, int transB, int m, int n, int k, double alpha,
                     const double *A, int lda, long strideA, const double *B, int ldb,
                     long strideB, double beta, double *C, int ldc, long strideC, int batch,
                     int splits, double *work, int tile, void *stream);
/* gpx_gemm_batched with a bias term in the same pass:
 * C[b] = alpha op(A[b]) op(B[b]) + beta C[b] + gamma E[b] (E row-major with leading
 * dimension ldE; strideE = 0 broadcasts one E over the batch; E = NULL drops the term).
 * Carries the A K^-1 B - K^-1 and A + B C forms of geepee/aep_models.py:62-95, 453-493
 * in one launch instead of a product and a separate axpy. */
int gpx_gemm_batched_ex(int transA, int transB, int m, int n, int k, double alpha,
                        const double *A, int lda, long strideA, const double *B, int ldb,
                        long strideB, double beta, double *C, int ldc, long strideC,
                        double gamma, const double *E, int ldE, long strideE, int batch,
                        int splits, double *work, int tile, void *stream);

/* Batched SPD inverse and log-determinant, one workgroup per matrix (M <= 256):
 * out[b] = inv(A[b]) (out may alias A or be NULL), logdet[b] = log det A[b],
 * info[b] = 0 or k+1 when pivot k is not positive (outputs NaN).  A[b] is read whole and
 * taken to be symmetric; out[b] is symmetric to rounding (both halves are computed, since
 * their rounding errors cancel in later products).
 * Replaces np.linalg.inv / slogdet / solve at geepee/base_models.py:462-476 and
 * geepee/aep_models.py:62-95, 523-533. */
int gpx_spd_max_m(void);
int gpx_spd_inv_batched(int M, int batch, const double *A, int lda, long strideA, double *out,
                        int ldo, long strideOut, double *logdet, int *info, void *stream);

/* Folds the `info` arrays of up to 16 gpx_spd_inv_batched calls (host arrays of nlists
 * device pointers and their lengths) into a device status pair: status[0] = the largest
 * entry (0 when every matrix was positive definite), status[1] = max(status[1],
 * status[0]), a flag that stays set across evaluations until the caller clears it.
 * The host raises numpy.linalg.LinAlgError from it, as np.linalg.inv does at
 * geepee/base_models.py:464, 471 and geepee/aep_models.py:525, without synchronising
 * inside a captured evaluation. */
int gpx_status_fold(int nlists, const int *const *lists, const int *counts, int *status,
                    void *stream);

/* ---- SE kernel and psi statistics ------------------------------------------ */

/* K[n,m] = exp(2 sf) exp(-sum_q (x_nq - z_mq)^2 / (2 exp(2 ls_q))) (+ jitter if n == m).
 * Replaces compute_kernel (geepee/kernels.py:10-22) as called at base_models.py:277, 462
 * and aep_models.py:154. */
int gpx_se_kernel(int N, int M, int Q, const double *x, int ldx, const double *z,
                  const double *ls, const double *sf, double jitter, double *K, void *stream);

/* psi1 (N x M) and psi2 (N x M x M, NULL to skip) for x_n ~ N(mx_n, diag vx_n); psi2 is
 * exactly symmetric.  Q <= gpx_psi_qmax().
 * Replaces compute_psi_weave (geepee/kernels.py:181-240), the reference's weave FFI. */
int gpx_psi_qmax(void);
int gpx_psi_fwd(int N, int M, int Q, const double *mx, const double *vx, const double *z,
                const double *ls, const double *sf, double *psi1, double *psi2, void *stream);

/* Adjoint of gpx_psi_fwd given dL/dpsi1 and a symmetric dL/dpsi2: dsf2 (w.r.t. exp(2 sf)),
 * dell (Q, w.r.t. exp(ls)), dZ (M x Q), dmx, dvx (N x Q).
 * Replaces compute_psi_derivatives -> psi1compDer / psi2compDer
 * (geepee/kernels.py:302-309, 355-444). */
long gpx_psi_bwd_workspace(int N, int M, int Q);
int gpx_psi_bwd(int N, int M, int Q, const double *mx, const double *vx, const double *z,
                const double *ls, const double *sf, const double *dpsi1, const double *psi1,
                const double *dpsi2, const double *psi2, double *dsf2, double *dell, double *dZ,
                double *dmx, double *dvx, double *work, void *stream);

/* Adjoint of K(x, Z): dsf2, dell, dZ and optionally dx (N x Q); M <= 256, Q <= 16.
 * Replaces compute_kfu_derivatives -> kfucompDer (geepee/kernels.py:312-313, 381-399). */
long gpx_kfu_bwd_workspace(int N, int M, int Q);
int gpx_kfu_bwd(int N, int M, int Q, const double *dK, const double *K, const double *x,
                const double *z, const double *ls, const double *sf, double *dsf2,
                double *dell, double *dZ, double *dx, double *work, void *stream);

/* d tr(Mm Kzz) / d(2 sf, 2 ls, z) for Kzz = SE(z, z) without jitter: gsf, gls (Q), gz (M x Q);
 * M <= 256, Q <= 16.
 * Replaces d_trace_MKzz_dhypers (geepee/kernels.py:447-475). */
long gpx_dtrace_workspace(int M, int Q);
int gpx_dtrace_kzz(int M, int Q, const double *Mm, const double *z, const double *ls,
                   const double *sf, double *gsf, double *gls, double *gz, double *work,
                   void *stream);

/* ---- q(u) plumbing and layer helpers --------------------------------------- */

/* R[d] = upper-triangular from packed row-major triu values with exp'd diagonal.
 * Replaces the loop of Base_SGP_Layer.update_hypers (geepee/base_models.py:645-653). */
int gpx_eta_to_R(int D, int M, const double *eta, double *R, void *stream);

/* Kuu of nl (<= 16) layers of one M in one launch: K (nl x M x M), K[l] = the squared
 * exponential kernel of layer l's inducing inputs z[l] (M x Q[l]) with its ls[l] (Q[l]) and
 * sf[l] (1), plus jitter on the diagonal; the entries of gpx_se_kernel(z[l], z[l]).  The
 * compute_kernel(zu, zu) + jitter of Base_SGP_Layer.update_hypers
 * (geepee/base_models.py:456-461) batched over a deep GP's layers. */
int gpx_kuu_batched(int nl, int M, const double *const *z, const int *Q, const double *const *ls,
                    const double *const *sf, double jitter, double *K, void *stream);

/* The posterior and cavity precisions of nl (<= 16) layers of one M in one launch: from
 * X = R^T R of the layers' output dims stacked (sum D[l] x M x M) and their Kuu^-1 (nl x M x M),
 * prec holds per layer [X_l + Kuu^-1_l (D[l]); beta X_l + Kuu^-1_l (D[l])].  The
 * Kuu^-1 + theta1 and Kuu^-1 + beta theta1 of compute_cavity and update_posterior
 * (geepee/aep_models.py:523-533, base_models.py:462-470) batched over a deep GP's layers. */
int gpx_prec_form(int nl, const int *D, int M, const double *X, const double *Ki, double beta, double *prec,
                  void *stream);

/* out[d, triu(i,j)] = (R (G + G^T))[i,j] * (i == j ? R_ii : 1).
 * Replaces the deta1_R packing (geepee/aep_models.py:267-280, 573-584;
 * base_models.py:504-514). */
int gpx_pack_dR(int D, int M, const double *R, const double *G, double *out, void *stream);

/* out[i] = a x[i % nx] + b y[i % ny] + c w[i % nw] (y, w may be NULL): the broadcast
 * additions of the layer algebra (e.g. Kuuinv + beta theta1, aep_models.py:524). */
int gpx_lincomb(long n, double a, const double *x, long nx, double b, const double *y,
                long ny, double c, const double *w, long nw, double *out, void *stream);

/* out[b] = a (X[b] + X[b]^T) (the dL/dpsi2 symmetrisation of geepee/kernels.py:415-418). */
int gpx_sym(int batch, int M, double a, const double *X, double *out, void *stream);

/* out[j] = alpha sum_i P[i, j] + beta out[j] (np.sum(..., axis=0) on the path). */
int gpx_colsum(int rows, int cols, const double *P, double alpha, double beta, double *out,
               void *stream);

/* Two-stage column sum for tall P (many rows): chunked partial sums into `work`
 * (gpx_colsum_workspace doubles) and their fixed-order sum; same contract as
 * gpx_colsum (cols <= 65535). */
long gpx_colsum_workspace(int rows, int cols);
int gpx_colsum_tall(int rows, int cols, const double *P, double alpha, double beta, double *out, double *work,
                    void *stream);

/* v[n,d] = exp(2 sf) + sum_a T[d,n,a] kfu[n,a] with T[d] = kfu Bhat_det[d]
 * (geepee/aep_models.py:153-157, base_models.py:280-282). */
int gpx_det_var(int N, int D, int M, const double *T, const double *kfu, const double *sf,
                double *v, void *stream);

/* v = exp(2 sf) + Bpsi2 - m^2 (geepee/aep_models.py:196-198, base_models.py:302-304). */
int gpx_mm_var(long n, const double *m, const double *Bp, const double *sf, double *v,
               void *stream);

/* out[b] = sum_i x[b,i] y[b,i] (the mu^T Sigma^-1 mu terms of aep_models.py:80, 93). */
int gpx_batched_dot(int batch, int n, const double *x, const double *y, double *out,
                    void *stream);

/* out[d][n][m] = w[n,d] X[n][m]: the dv-weighted rows behind the syrk-like einsums
 * 'nd,na,nb->dab' (geepee/aep_models.py:464, 490). */
int gpx_rowscale(int N, int D, int M, const double *w, const double *X, double *out, void *stream);

/* out[n][m] = alpha sum_d w[n,d] T[d][n][m] + beta out[n][m]: the einsum
 * 'nd,dab,na->nb' (geepee/aep_models.py:455) with T[d] = kfu Bhat_det[d]. */
int gpx_wsum(int N, int D, int M, double alpha, const double *w, const double *T, double beta,
             double *out, void *stream);

/* out = dm - 2 dv m (geepee/aep_models.py:235). */
int gpx_dmall(long n, const double *dm, const double *dv, const double *m, double *out, void *stream);

/* dls = dell exp(ls) + 2 gls, dsf = 2 exp(2 sf) (dsf2 + sumdv) + 2 gsf: the hyper-parameter
 * chain of geepee/aep_models.py:244-247, 296-298, 456-460, 502-504. */
int gpx_hyper_grads(int Q, const double *ls, const double *sf, const double *dell, const double *gls,
                    const double *dsf2, const double *sumdv, const double *gsf, double *dls, double *dsf,
                    void *stream);

/* phi = (D/2) ldK + sp/2 sum_d (sgnP ldP_d + dotP_d) + sc/2 sum_d (sgnC ldC_d + dotC_d):
 * SGP_Layer.compute_phi (geepee/aep_models.py:62-114). */
int gpx_phi(int D, const double *ldK, const double *ldP, double sgnP, const double *dotP,
            const double *ldC, double sgnC, const double *dotC, double sp, double sc, double *out,
            void *stream);
/* gpx_phi for the natural parameterisation in one launch: ldP = log det Sigma^-1,
 * ldC = log det Sigma-hat^-1 (D each), mu, muhat, theta2 (D x M); the quadratic terms
 * mu^T Sigma^-1 mu and mu-hat^T Sigma-hat^-1 mu-hat are mu . theta2 and
 * beta mu-hat . theta2 (Sigma^-1 mu = theta2, Sigma-hat^-1 mu-hat = beta theta2), in
 * place of the np.linalg.solve calls of geepee/aep_models.py:72-95. */
int gpx_phi_nat(int D, int M, const double *ldK, const double *ldP, const double *ldC, const double *mu,
                const double *muhat, const double *theta2, double beta, double sp, double sc, double *out,
                void *stream);

/* ---- likelihoods ----------------------------------------------------------- */

/* Gaussian tilted log-normaliser and its adjoint: scal[0] = logZ (incl. the alpha
 * constant), scal[1] = dsn * scale (scal holds 130 doubles: the rest are block partials); dm, dv = scale * dlogZ/dm, dlogZ/dv; dmall (optional)
 * = dm - 2 dv m.  Replaces Gauss_Layer.compute_log_Z + backprop_grads
 * (geepee/lik_layers.py:104-181). */
int gpx_gauss_tilted(int B, int D, const double *m, const double *v, const double *y,
                     const double *sn, double alpha, double scale, double *dm, double *dv,
                     double *dmall, double *scal, void *stream);

/* The Gaussian tilted log-normaliser over S Monte Carlo samples: m, v, dm, dv are
 * (S, B, D), y is (B, D); per (n, d) the log-mean-exp over the samples, dm, dv the
 * softmax-weighted per-sample gradients times scale; scal as gpx_gauss_tilted (130 doubles).
 * Replaces the 3-D branch of Gauss_Layer.compute_log_Z + backprop_grads
 * (geepee/lik_layers.py:139-152, 154-181). */
int gpx_gauss_tilted_mc(int S, int B, int D, const double *m, const double *v, const double *y,
                        const double *sn, double alpha, double scale, double *dm, double *dv,
                        double *scal, void *stream);

/* Probit tilted log-normaliser for y in {-1, +1}: closed form at alpha = 1, 10-point
 * Gauss-Hermite quadrature otherwise, with the reference's small additive constants.
 * three = 0: (B, D) arrays (S must be 1); three = 1: (S, B, D) Monte Carlo samples, averaged
 * by log-mean-exp.  scal[0] = logZ (scal holds 130 doubles); dm, dv scaled; dmall optional.
 * Replaces Probit_Layer.compute_log_Z (geepee/lik_layers.py:285-411). */
int gpx_probit_tilted(int S, int three, int B, int D, const double *m, const double *v,
                      const double *y, double alpha, double scale, double *dm, double *dv,
                      double *dmall, double *scal, void *stream);

/* Monte Carlo inputs x[s] = eps[s] * sqrt(vx) + mx for draws eps (S, N, Q), and the adjoint
 * dmx = sum_s dx[s], dvx = sum_s dx[s] eps[s] / (2 sqrt(vx)).
 * Replace geepee/aep_models.py:170-173 (base_models.py:320-322) and backprop_grads_reparam
 * (geepee/base_models.py:373-388). */
int gpx_mc_inputs(int S, int N, int Q, const double *eps, const double *mx, const double *vx,
                  double *x, void *stream);
int gpx_mc_reparam_bwd(int S, int N, int Q, const double *dx, const double *eps, const double *vx,
                       double *dmx, double *dvx, void *stream);

/* GPSSM transition normaliser over S Monte Carlo propagations mp, vp (S, B, D) against the
 * next-state cavities mt, vt (B, D): scal[0] = scale * logZ, scal[1] = dsn (130 doubles);
 * dmp, dvp (S, B, D) and dmt, dvt (B, D) scaled.  Replaces the 3-D branch of
 * SGPSSM.compute_transition_tilted (geepee/aep_models.py:1344-1372). */
int gpx_transition_tilted_mc(int S, int B, int D, const double *mp, const double *vp,
                             const double *mt, const double *vt, const double *sn, double alpha,
                             double scale, double *dmp, double *dvp, double *dmt, double *dvt,
                             double *scal, void *stream);

/* GPSSM transition tilted normaliser (scaled): scal[0] = scale * logZ, scal[1] = dsn (scal
 * holds 130 doubles: the rest are block partials).
 * Replaces SGPSSM.compute_transition_tilted (geepee/aep_models.py:1317-1343). */
int gpx_transition_tilted(int B, int D, const double *mp, const double *vp, const double *mt,
                          const double *vt, const double *sn, double alpha, double scale,
                          double *dmall, double *dvp, double *dmt, double *dvt, double *scal,
                          void *stream);

/* Linear-Gaussian emission tilted normaliser (Dout <= 8): scal[0] = scale * logZ, dC, dR,
 * dmx, dvx (all scaled).  Replaces Gauss_Emis.compute_emission_tilted
 * (geepee/lik_layers.py:573-627). */
long gpx_emission_workspace(int Nb, int Dout, int Din);
int gpx_emission_tilted(int Nb, int Dout, int Din, const double *C, const double *Rp,
                        const double *y, int ldy, const double *mx, int ldm, const double *vx,
                        int ldv, double alpha, double scale, double *dmx, double *dvx, double *dC,
                        double *dR, double *scal, double *work, void *stream);

/* Emission predictive mean C mf and variance diag(C diag(vf) C^T) + R.
 * Replaces Gauss_Emis.output_probabilistic (geepee/lik_layers.py:558-571), diagonal only as
 * returned by Base_SGPSSM.predict_y (geepee/base_models.py:1544-1560). */
int gpx_emission_predict(int Nb, int Dout, int Din, const double *C, const double *Rp, const double *mf,
                         const double *vf, double *my, double *vy, void *stream);

/* ---- variational free energy (VFE) family -------------------------------------- */

/* KL(q(u) || p(u)) = (D logdet Kuu - sum_d logdet Su_d - D M + sum_d <Kuu^-1, Su_d + mu_d mu_d^T>) / 2
 * into out[0], with logdet Su_d = sgnS * ldS[d] (ldS = logdet Su^-1, sgnS = -1, in the natural
 * parameterisation).  Replaces SGP_Layer.compute_KL (geepee/vfe_models.py:309-326). */
int gpx_vfe_kl(int D, int M, const double *ldK, const double *ldS, double sgnS, const double *Kuuinv,
               const double *Splusmm, double *out, void *stream);

/* Expected Gaussian log-likelihood E log N(y | f, exp(2 sn)) of (B, D) propagations
 * (three = 0, S = 1) or the mean over S Monte Carlo samples (S, B, D) (three = 1):
 * scal[0] = scale * E, scal[1] = dsn * scale, dm and dv times scale (scal holds 130 doubles:
 * the rest are block partials).  Replaces Gauss_Layer.compute_log_lik_exp and
 * backprop_grads_log_lik_exp (geepee/lik_layers.py:183-236). */
int gpx_gauss_lik_exp(int S, int three, int B, int D, const double *m, const double *v, const double *y,
                      const double *sn, double scale, double *dm, double *dv, double *scal, void *stream);

/* Expected probit log-likelihood E log Phi(y f) by 10-point Gauss-Hermite quadrature, (B, D) or
 * averaged over (S, B, D) samples: scal[0] = scale * E; dm, dv times scale (130 doubles).
 * Replaces Probit_Layer.compute_log_lik_exp (geepee/lik_layers.py:412-455). */
int gpx_probit_lik_exp(int S, int three, int B, int D, const double *m, const double *v, const double *y,
                       double scale, double *dm, double *dv, double *scal, void *stream);

/* GPSSM expected transition log-likelihood E log N(x_t | f(x_{t-1}), exp(2 sn)) of the
 * propagations mp, vp ((B, D), or (S, B, D) averaged) against the posterior of x_t (mt, vt):
 * scal[0] = scale * E, scal[1] = dsn; dmp, dvp, dmt, dvt scaled (130 doubles).
 * Replaces SGPSSM.compute_transition_log_lik_exp (geepee/vfe_models.py:1080-1120). */
int gpx_transition_lik_exp(int S, int three, int B, int D, const double *mp, const double *vp, const double *mt,
                           const double *vt, const double *sn, double scale, double *dmp, double *dvp,
                           double *dmt, double *dvt, double *scal, void *stream);

/* Linear-Gaussian emission expected log-likelihood E log N(y | C x, diag R), R = exp(2 Rp), under
 * x ~ N(mx, diag vx) (Dout, Din <= 16): scal[0] = scale * E; dmx, dvx, dC, dR scaled; `work`
 * holds gpx_emission_lik_exp_workspace(Dout, Din) doubles.  Replaces
 * Gauss_Emis.compute_emission_log_lik_exp (geepee/lik_layers.py:629-676). */
long gpx_emission_lik_exp_workspace(int Dout, int Din);
int gpx_emission_lik_exp(int Nb, int Dout, int Din, const double *C, const double *Rp, const double *y, int ldy,
                         const double *mx, int ldm, const double *vx, int ldv, double scale, double *dmx,
                         double *dvx, double *dC, double *dR, double *scal, double *work, void *stream);

/* SGPLVM posterior q(x) of the rows idx (int64, NULL = rows 0..B-1) from the factors x1, x2:
 * natural parameters prior + (x1, exp(2 x2)) when nat_param, else mean x1 / exp(2 x2) and
 * variance 1 / exp(2 x2).  Replaces Base_SGPLVM.get_posterior_x (geepee/base_models.py:895-911)
 * as used by the VFE SGPLVM (geepee/vfe_models.py:772-780). */
int gpx_lvm_x_post(int B, int Q, const long *idx, const double *x1, const double *x2, double prior_mean,
                   double prior_var, int nat_param, double *mx, double *vx, void *stream);

/* SGPLVM: scal[0] = scale_x * KL(q(x_idx) || N(prior_mean, prior_var)) and the gradients of
 * the rows idx w.r.t. x1, x2 from dmx, dvx plus scale_x times the KL's (rows outside idx are
 * not written).  Replaces SGPLVM.compute_KL_x (geepee/vfe_models.py:847-864) and
 * Base_SGPLVM.compute_posterior_grad_x (geepee/base_models.py:913-929); 130 doubles in scal. */
int gpx_lvm_vfe_x_grads(int B, int Q, const long *idx, const double *x1, const double *x2, const double *dmx,
                        const double *dvx, double prior_mean, double prior_var, int nat_param, double scale_x,
                        double *g1, double *g2, double *scal, void *stream);

/* SGPSSM: entropy term scal[0] = scale_ent (Be Q (1 + log 2 pi) / 2 + sum log(var) / 2) of the
 * window [s, e) of the latent path and the gradients w.r.t. its factors (rows s..e-1 written):
 * dup / vup the emission-input grads (Be rows, leading dim ldu), dmt / dvt the
 * transition-target grads (Be-1 rows, leading dim Q), dmn / dvn the dynamics-input grads
 * (Be-1 rows, leading dim ldn).  Replaces geepee/vfe_models.py:1030-1060 with
 * Base_SGPSSM.compute_posterior_grad_x (geepee/base_models.py:1730-1752); 130 doubles in scal. */
int gpx_ssm_vfe_x_grads(int N, int Q, int s, int e, const double *xf1, const double *x2, double prior_mean,
                        double prior_var, const double *dup, const double *vup, int ldu, const double *dmt,
                        const double *dvt, const double *dmn, const double *dvn, int ldn, double scale_ent,
                        double *g1, double *g2, double *scal, void *stream);

/* ---- fused moment-matched layer (psi2 never written to HBM) -------------------- */

/* Symmetrised Bhat in pair-major layout, out[a, b, d] = (B[d,a,b] + B[d,b,a]) / 2: the operand
 * layout of the fused kernels (the symmetrisation of geepee/kernels.py:415-418). */
int gpx_sym_pairmajor(int D, int M, const double *B, double *out, void *stream);

/* Forward of a Gaussian-input layer: psi1 (N x M), mout = psi1 Ahat^T and
 * vout = sf2 + <Bhat[d], psi2_n> - mout^2 with psi2 recomputed per 16x16 tile (FP64 MFMA for
 * the exponent GEMM and the Bhat contraction).  splits (grid.y over pair tiles) and nsub
 * (1, 2, 4: 16*nsub points per block) are tuning knobs; `work` holds
 * gpx_fused_fwd_workspace(N, M, D, splits) doubles.  Replaces compute_psi_weave followed by
 * the einsums of geepee/aep_models.py:193-198 (and base_models.py:298-304). */
long gpx_fused_fwd_workspace(int N, int M, int D, int splits);
int gpx_fused_mm_fwd(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                     const double *ls, const double *sf, const double *Ahat, const double *Bp,
                     int splits, int nsub, double *psi1, double *mout, double *vout, double *work,
                     void *stream);

/* gpx_fused_mm_fwd without the `work` argument and without any allocation: the workspace
 * is a file-scope device array of gpx_fused_fwd_noalloc_max() doubles (32 MiB), one copy
 * per GPU, held until the process ends.  No counterpart in the reference: this entry is for
 * callers that cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Returns GPX_ETOOBIG, before any launch, when
 * gpx_fused_fwd_workspace(N, M, D, splits) exceeds the array.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (concurrent streams, host threads) race;
 * the kernels write every workspace entry they read, so no clearing is needed. */
long gpx_fused_fwd_noalloc_max(void);
int gpx_fused_mm_fwd_noalloc(int N, int M, int Q, int D, const double *mx, const double *vx,
                             const double *z, const double *ls, const double *sf, const double *Ahat,
                             const double *Bp, int splits, int nsub, double *psi1, double *mout,
                             double *vout, void *stream);

/* Adjoint of the same layer without psi2 or dL/dpsi2 in HBM: from dm_all = dm - 2 dv m and dv
 * (N x D) produces dBhat = sum_n dv_nd psi2_n (D x M x M), and the psi adjoints dsf2, dell, dZ,
 * dmx, dvx.  Replaces geepee/aep_models.py:235-243 with compute_psi_derivatives
 * (geepee/kernels.py:302-309, 355-444). */
long gpx_fused_bwd_workspace(int N, int M, int Q, int D, int splits, int nsplit);
int gpx_fused_mm_bwd(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                     const double *ls, const double *sf, const double *Ahat, const double *Bp,
                     const double *psi1, const double *dm_all, const double *dv, int splits, int nsplit,
                     double *dBhat, double *dsf2, double *dell, double *dZ, double *dmx, double *dvx,
                     double *work, void *stream);
/* gpx_fused_mm_bwd in parts over the same workspace, run in this order: _points gives dmx,
 * dvx (the input gradients the layer below needs first); _pairs gives dBhat; _tail gives
 * dsf2, dell, dZ.  _pairs and _tail may be queued on another stream that waits for
 * _points, so a deep GP's backward can continue down the layers while the
 * hyper-parameter adjoint finishes.  _main is _points then _pairs.  Same reference code as
 * gpx_fused_mm_bwd. */
int gpx_fused_mm_bwd_points(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                            const double *ls, const double *sf, const double *Ahat, const double *Bp,
                            const double *psi1, const double *dm_all, const double *dv, int splits, int nsplit,
                            double *dmx, double *dvx, double *work, void *stream);
int gpx_fused_mm_bwd_pairs(int N, int M, int Q, int D, const double *z, const double *ls, const double *sf,
                           const double *Bp, const double *dv, int splits, int nsplit, double *dBhat, double *work,
                           void *stream);
int gpx_fused_mm_bwd_main(int N, int M, int Q, int D, const double *mx, const double *vx, const double *z,
                          const double *ls, const double *sf, const double *Ahat, const double *Bp,
                          const double *psi1, const double *dm_all, const double *dv, int splits, int nsplit,
                          double *dBhat, double *dmx, double *dvx, double *work, void *stream);
int gpx_fused_mm_bwd_tail(int N, int M, int Q, int D, const double *z, const double *ls, const double *sf,
                          int splits, int nsplit, double *dsf2, double *dell, double *dZ, double *work,
                          void *stream);

/* Shared final stage of both psi adjoints (pair sums LW and per-point statistics in, dsf2,
 * dell, dZ out); the tail of compute_psi_derivatives (geepee/kernels.py:430-442, 371-376). */
int gpx_psi_bwd_tail(int N, int M, int Q, const double *z, const double *ls, const double *sf,
                     const double *T, const double *Lp, const double *H, const double *s0,
                     const double *Dn, const double *Pn, const double *En, const double *LW,
                     double *TP, double *LE, double *cols0, double *dsf2, double *dell, double *dZ,
                     void *stream);

/* ---- serving ------------------------------------------------------------------ */

/* One launch of predictive mean / variance for many (request, model) work items of at
 * most gpx_serve_tile() points each, against resident models given by a device table of
 * GpxServeModel descriptors (layout in geepee_amd/csrc/gpx_serve.hip; sizes from
 * gpx_serve_desc_sizes).  psi2 is recomputed per point, never stored; per-point results do
 * not depend on the other points in the launch.  noisy adds exp(2 sn) (predict_y).
 * Replaces Base_SGPR.predict_f / predict_y and Base_SDGPR.predict_f_mm / predict_y
 * (geepee/base_models.py:985-1034, 1140-1158, 1264-1275) over many models at once. */
int gpx_serve_tile(void);
int gpx_serve_desc_sizes(int *model_bytes, int *item_bytes);
int gpx_serve_validate(const void *model);
int gpx_predict_grouped(int nitems, const void *models, const void *items, const double *X,
                        double *mean, double *var, int noisy, void *stream);

/* ---- latent-variable (x-factor) terms --------------------------------------- */

/* SGPLVM cavity inputs of rows idx (NULL = all rows): get_cavity_x
 * (geepee/aep_models.py:840-861) with factor_x2 = exp(2 x2) (base_models.py:895-911). */
int gpx_lvm_x_cavity(int B, int Q, const long *idx, const double *x1, const double *x2, double alpha,
                     double prior_mean, double prior_var, int nat_param, double *mcav, double *vcav,
                     void *stream);

/* SGPLVM x gradients (rows idx of the N x Q outputs) and scal[0] = the x log-partition
 * energy (scal holds 130 doubles: the rest are block partials):
 * aep_models.py:784-805, 817-838, 863-867 and base_models.py:913-929. */
int gpx_lvm_x_grads(int N, int B, int Q, const long *idx, const double *x1, const double *x2,
                    const double *gmx, const double *gvx, double alpha, double prior_mean,
                    double prior_var, int nat_param, double *g1, double *g2, double *scal,
                    void *stream);

/* SGPSSM cavity means / variances of every x_t with the 3/2/prior factor ties:
 * compute_cavity_x (geepee/aep_models.py:1376-1387) over base_models.py:1719-1725. */
int gpx_ssm_x_cavity(int N, int Q, const double *xf1, const double *x2, double alpha, double prior_mean,
                     double prior_var, double *cm, double *cv, void *stream);

/* Posterior means / variances of every latent state x_t from the tied factors (three inside
 * the chain, two at its ends, the prior at t = 0).  Replaces Base_SGPSSM.get_posterior_x
 * (geepee/base_models.py:1562-1576) over update_hypers (:1705-1728). */
int gpx_ssm_x_posterior(int N, int Q, const double *xf1, const double *x2, double prior_mean,
                        double prior_var, double *pm, double *pv, void *stream);

/* SDGPR_H hidden-value factors (h1, h2 with precision factor exp(2 h2), two per value):
 * cavity means / variances, and the gradients through posterior, cavity and cavity inputs
 * (dmc, dvc) with scal[0] = the posterior and cavity log-partition terms (scal holds 66
 * doubles).  Replace compute_cavity_h, compute_grads_hidden and compute_phi_posterior_h /
 * compute_phi_cavity_h (geepee/aep_models.py:1620-1700, 1737-1752). */
int gpx_hidden_cavity(int N, int D, const double *h1, const double *h2, double alpha, double *cm,
                      double *cv, void *stream);
int gpx_hidden_grads(int N, int D, const double *h1, const double *h2, double alpha,
                     const double *dmc, const double *dvc, double *g1, double *g2, double *scal,
                     void *stream);

/* SGPSSM x-factor gradients and scal[0] = phi_prior_x + phi_post_x + phi_cav_x for the
 * window [s, e) (scal holds 130 doubles: the rest are block partials):
 * aep_models.py:1170-1196, 1208-1315, 1389-1437. */
int gpx_ssm_x_grads(int N, int Q, int s, int e, const double *xf1, const double *x2, double alpha,
                    double prior_mean, double prior_var, const double *dup, const double *vup, int ldu,
                    const double *dmt, const double *dvt, const double *dmn, const double *dvn, int ldn,
                    double *g1, double *g2, double *scal, void *stream);

/* ---- optimiser ------------------------------------------------------------------ */

/* One Adam iteration t (1-based) in place on the flat parameter vector x with moments
 * m, v from the flat gradient g, all n doubles in sorted-key order: the update of
 * adam (geepee/utils.py:93-117; b1, b2, eps, bias correction 1 - b^t), with non-finite
 * gradient entries taken as 0 as ObjectiveWrapper does (utils.py:46-53).  The same
 * IEEE operation sequence as the reference (no contraction).  cost (optional)
 * receives energy[0]. */
int gpx_adam_step(long n, double *x, const double *g, double *m, double *v, double lr, double b1, double b2,
                  double eps, int t, const double *energy, double *cost, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GPX_H */
