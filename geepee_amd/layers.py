"""Sparse-GP layer on the GPU: q(u), cavity, moment propagation and adjoints.

API of the reference's ``SGP_Layer`` (geepee/aep_models.py:26-586 over
geepee/base_models.py:168-658).  State lives in device tensors; every arithmetic
step is a libgpx kernel launch (``dops``).  Shapes: D = output dims, M = inducing
points, Q = input dims; matrices are (D, M, M) stacks, vectors (D, M).
"""
import torch

from . import dops
from .config import JITTER, PROP_MM, PROP_MC, PROP_LIN
from .dops import mm, lincomb, colsum


def _col(v):
    """(D, M) -> (D, M, 1) view for batched mat-vecs."""
    return v.reshape(v.shape[0], v.shape[1], 1)


def _add_outer(C, u, v, alpha):
    """C[d] += alpha u_d v_d^T (batched rank-1)."""
    mm(_col(u), _col(v), C, tB=True, alpha=alpha, beta=1.0)
    return C


def _plus_outer(X, u, v, alpha=1.0, gamma=1.0, out=None):
    """gamma X[d] + alpha u_d v_d^T into a new (or given) stack, in one launch."""
    if out is None:
        out = torch.empty_like(X)
    return mm(_col(u), _col(v), out, tB=True, alpha=alpha, E=X, gamma=gamma)


def _sandwich(L, X, R, alpha=1.0, out=None, beta=1.0, E=None, gamma=0.0):
    """alpha L X R (+ beta out, + gamma E; both added in the second product's epilogue)
    with L, R (M, M) or (D, M, M) and X (D, M, M)."""
    T = mm(X, R)
    if out is None:
        return mm(L, T, alpha=alpha, E=E, gamma=gamma)
    return mm(L, T, out, alpha=alpha, beta=beta, E=E, gamma=gamma)


_eyes = {}


def _eye(M, like):
    key = (M, like.device)
    if key not in _eyes:
        _eyes[key] = torch.eye(M, dtype=like.dtype, device=like.device)
    return _eyes[key]


class SGP_Layer(object):

    def __init__(self, no_train, input_size, output_size, no_pseudo, nat_param=True):
        self.N = no_train
        self.Din = input_size
        self.Dout = output_size
        self.M = no_pseudo
        self.nat_param = nat_param
        self._pred_ready = False
        # fused=True: moment-matched propagation never writes psi2 to HBM; the forward
        # returns None in psi2's slot and the adjoint recomputes it.
        self.fused = False

    # ------------------------------------------------------------------
    # q(u) and Kuu (base_models.py:454-488, 630-658)
    # ------------------------------------------------------------------
    def update_hypers(self, params, key_suffix=''):
        """params: dict of device tensors (see models._DeviceParams)."""
        self.ls = params['ls' + key_suffix]
        self.sf = params['sf' + key_suffix]
        self.zu = params['zu' + key_suffix]
        self.theta_2 = params['eta2' + key_suffix]
        self.theta_1_R = dops.eta_to_R(params['eta1_R' + key_suffix], self.M)
        # theta1 = R^T R; the natural parameterisation only ever needs Kuu^-1 + c theta1,
        # formed by the product's epilogue where it is used
        self.theta_1 = None if self.nat_param else mm(self.theta_1_R, self.theta_1_R, tA=True)
        self._adj = None
        self.compute_kuu()
        self.update_posterior()

    def prepare(self, params, alpha, key_suffix=''):
        """update_hypers followed by compute_cavity(alpha).  In the natural
        parameterisation both precisions, Kuu^-1 + theta1 and Kuu^-1 + beta theta1, are
        known once Kuu^-1 is, so their 2D inverses run as one batched SPD launch."""
        if not self.nat_param:
            self.update_hypers(params, key_suffix)
            self.compute_cavity(alpha)
            return
        self.prepare_params(params, key_suffix)
        self.compute_kuu()
        prec = dops.empty(2 * self.Dout, self.M, self.M, like=self.Kuuinv)
        self.precisions_into(prec, alpha)
        self.finish_prepare(prec, *dops.spd_inv(prec), alpha=alpha)

    # the parts of prepare (natural parameterisation), so that a model with several layers
    # of one M can invert all their Kuu in one batched SPD launch and all their precisions
    # in another (SDGPR._prepare_layers)
    def prepare_params(self, params, key_suffix='', R=None):
        """R: theta1's factor already unpacked (a view of a batch over layers), if given."""
        self.ls = params['ls' + key_suffix]
        self.sf = params['sf' + key_suffix]
        self.zu = params['zu' + key_suffix]
        self.theta_2 = params['eta2' + key_suffix]
        self.theta_1_R = dops.eta_to_R(params['eta1_R' + key_suffix], self.M) if R is None else R
        self.theta_1 = None
        self._adj = None

    def kuu_into(self, out):
        """Kuu (with the jitter) into out (M, M)."""
        self.Kuu = dops.se(self.zu, self.zu, self.ls, self.sf, jitter=JITTER, out=out)

    def set_kuu_inverse(self, Kuuinv, ld, info):
        self.Kuuinv, self.ld_Kuu, self.info_Kuu = Kuuinv, ld, info

    def precisions_into(self, prec, alpha):
        """[Kuu^-1 + theta1; Kuu^-1 + beta theta1] into prec (2D, M, M)."""
        D = self.Dout
        beta = (self.N - alpha) * 1.0 / self.N
        R, Ki = self.theta_1_R, self.Kuuinv
        mm(R, R, prec[:D], tA=True, E=Ki, gamma=1.0)             # Kuu^-1 + theta1
        mm(R, R, prec[D:], tA=True, alpha=beta, E=Ki, gamma=1.0)  # Kuu^-1 + beta theta1

    def finish_prepare(self, prec, cov, ld, info, alpha):
        """The rest of prepare from the inverted precisions (views of batched buffers
        are fine)."""
        D, M = self.Dout, self.M
        beta = (self.N - alpha) * 1.0 / self.N
        self.Suinv, self.Suhatinv = prec[:D], prec[D:]
        self.Su, self.Suhat = cov[:D], cov[D:]
        self.ld_Suinv, self.ld_Suhatinv = ld[:D], ld[D:]
        self.info_Su, self.info_Suhat = info[:D], info[D:]
        self.mu = mm(self.Su, _col(self.theta_2)).reshape(D, M)
        self._ld_post, self._sgn_post = self.ld_Suinv, -1.0
        self.Splusmm = _plus_outer(self.Su, self.mu, self.mu)
        self._pred_ready = False
        self.muhat = mm(self.Suhat, _col(self.theta_2), alpha=beta).reshape(D, M)
        self._beta, self._alpha = beta, alpha
        self._cavity_tail()

    def compute_kuu(self):
        self.Kuu = dops.se(self.zu, self.zu, self.ls, self.sf, jitter=JITTER)
        self.Kuuinv, self.ld_Kuu, self.info_Kuu = dops.spd_inv(self.Kuu)

    def update_posterior(self):
        D, M = self.Dout, self.M
        if self.nat_param:
            R = self.theta_1_R
            self.Suinv = mm(R, R, dops.empty(D, M, M, like=self.Kuu), tA=True, E=self.Kuuinv, gamma=1.0)
            self.Su, self.ld_Suinv, self.info_Su = dops.spd_inv(self.Suinv)
            self.mu = mm(self.Su, _col(self.theta_2)).reshape(D, M)
            self._ld_post, self._sgn_post = self.ld_Suinv, -1.0
        else:
            self.Su = self.theta_1
            self.Suinv, self.ld_Su, self.info_Su = dops.spd_inv(self.Su)
            self.mu = self.theta_2
            self._ld_post, self._sgn_post = self.ld_Su, 1.0
        self.Splusmm = _plus_outer(self.Su, self.mu, self.mu)
        self._pred_ready = False

    def _posterior_predictive(self):
        """A, B_sto, B_det of update_posterior (base_models.py:477-488)."""
        if self._pred_ready:
            return
        Ki = self.Kuuinv
        self.A = mm(self.mu, Ki, tB=True)
        self.B_sto = _sandwich(Ki, self.Splusmm, Ki, E=Ki, gamma=-1.0)
        self.B_det = _sandwich(Ki, self.Su, Ki, E=Ki, gamma=-1.0)
        self._pred_ready = True

    # ------------------------------------------------------------------
    # cavity and phi (aep_models.py:62-114, 513-546)
    # ------------------------------------------------------------------
    def compute_cavity(self, alpha):
        D, M = self.Dout, self.M
        beta = (self.N - alpha) * 1.0 / self.N
        Ki = self.Kuuinv
        self._beta, self._alpha = beta, alpha
        if self.nat_param:
            R = self.theta_1_R
            self.Suhatinv = mm(R, R, dops.empty(D, M, M, like=Ki), tA=True, alpha=beta, E=Ki, gamma=1.0)
            self.Suhat, self.ld_Suhatinv, self.info_Suhat = dops.spd_inv(self.Suhatinv)
            self.muhat = mm(self.Suhat, _col(self.theta_2), alpha=beta).reshape(D, M)
        else:
            # Kuuinv + beta (Suinv - Kuuinv);  muhat = Suhat beta Suinv mu
            self.Suhatinv = lincomb(dops.empty(D, M, M, like=Ki), 1.0 - beta, Ki, beta, self.Suinv)
            self.Suhat, self.ld_Suhatinv, self.info_Suhat = dops.spd_inv(self.Suhatinv)
            f2 = mm(self.Suinv, _col(self.mu))
            self.muhat = mm(self.Suhat, f2, alpha=beta).reshape(D, M)
        self._cavity_tail()

    def adjoint_constants(self, kind, alpha=None):
        """Parts of the adjoint that depend on q(u) and Kuu only, formed once per q(u)
        (after prepare / compute_cavity, so they can run beside the forward pass):
          'reg' (natural parameters): the posterior chain dSuinv_post = -Su dSu Su with
            dSu = sp/2 (Sigma^-1 + theta2 theta2^T), Kc = sum_d dSuinv_post - (D/2) Kuu,
            KiS2 = 2 Ki Suhat - I, and the cavity's normalising-factor terms
            Ecav = sc/2 (Sigma-hat^-1 - beta^2 theta2 theta2^T);
          'mm': Mc = sp sum_d Splusmm + sc sum_d Splusmmhat, McKi = Mc Ki, T1 = -sp/2 Splusmm
            - sc beta/2 Splusmmhat, c2 = sp mu + sc beta muhat, KiS2 = 2 Ki Splusmmhat - I
        (sp = N/alpha - 1, sc = -N/alpha; aep_models.py:245-304, 412-511).  The backward
        passes use them when present."""
        N, D, M = self.N, self.Dout, self.M
        alpha = self._alpha if alpha is None else alpha
        beta = (N - alpha) * 1.0 / N
        sp, sc = N * 1.0 / alpha - 1.0, -N * 1.0 / alpha
        Ki = self.Kuuinv
        c = {'kind': kind, 'alpha': alpha}
        if kind == 'reg':
            if not self.nat_param:
                return
            th = self.theta_2
            dSu = _plus_outer(self.Suinv, th, th, alpha=0.5 * sp, gamma=0.5 * sp)
            c['dS_post'] = _sandwich(self.Su, dSu, self.Su, alpha=-1.0)
            c['Kc'] = lincomb(torch.empty_like(Ki), -0.5 * D, self.Kuu)
            colsum(c['dS_post'], D, M * M, out=c['Kc'].reshape(-1), beta=1.0)
            c['KiS2'] = mm(Ki, self.Suhat, alpha=2.0, E=_eye(M, Ki), gamma=-1.0)
            c['Ecav'] = _plus_outer(self.Suhatinv, th, th, alpha=-0.5 * sc * beta * beta, gamma=0.5 * sc)
        else:
            c['Mc'] = colsum(self.Splusmm, D, M * M, alpha=sp).reshape(M, M)
            colsum(self.Splusmmhat, D, M * M, out=c['Mc'].reshape(-1), alpha=sc, beta=1.0)
            c['McKi'] = mm(c['Mc'], Ki)
            c['T1'] = lincomb(torch.empty_like(self.Splusmm), -0.5 * sp, self.Splusmm, -0.5 * sc * beta,
                              self.Splusmmhat)
            c['c2'] = lincomb(torch.empty_like(self.mu), sp, self.mu, sc * beta, self.muhat)
            c['KiS2'] = mm(Ki, self.Splusmmhat, alpha=2.0, E=_eye(M, Ki), gamma=-1.0)
        self._adj = c

    def _adjoint(self, kind, alpha):
        c = getattr(self, '_adj', None)
        if c is None or c['kind'] != kind or c['alpha'] != alpha:
            self.adjoint_constants(kind, alpha)
            c = self._adj
        return c

    def _cavity_tail(self):
        """Ahat, Bhat_sto, Bhat_det from muhat, Suhat (aep_models.py:535-546):
        Bhat_det = Ki Suhat Ki - Ki (the -Ki in the second product's epilogue) and
        Bhat_sto = Ki (Suhat + muhat muhat^T) Ki - Ki = Bhat_det + Ahat Ahat^T, since
        Ahat = Ki muhat."""
        Ki = self.Kuuinv
        self.Ahat = mm(self.muhat, Ki, tB=True)
        self.Splusmmhat = _plus_outer(self.Suhat, self.muhat, self.muhat)
        self.Bhat_det = _sandwich(Ki, self.Suhat, Ki, E=Ki, gamma=-1.0)
        self.Bhat_sto = _plus_outer(self.Bhat_det, self.Ahat, self.Ahat)
        # the fused kernels' operand layout, formed with the rest of the q(u) set-up (off the
        # forward pass's critical path in a deep GP)
        self._Bp = None
        if self.fused and dops.fused_fits(self.M, self.Din, self.Dout):
            self._Bp = dops.sym_pairmajor(self.Bhat_sto)

    def compute_phi(self, alpha=1.0):
        """Device scalar (1,) of the weighted log-partitions (aep_models.py:97-114)."""
        N = self.N
        if self.nat_param:
            # Sigma^-1 mu = theta2 and Sigma-hat^-1 mu-hat = beta theta2 (aep_models.py:72-95
            # solve for them); one launch for the whole phi
            return dops.phi_nat(self.Dout, self.M, self.ld_Kuu, self._ld_post, self.ld_Suhatinv,
                                self.mu, self.muhat, self.theta_2, self._beta,
                                N * 1.0 / alpha - 1.0, -N * 1.0 / alpha)
        dotP = dops.batched_dot(self.mu, mm(self.Suinv, _col(self.mu)).reshape(self.Dout, self.M))
        dotC = dops.batched_dot(self.muhat,
                                mm(self.Suhatinv, _col(self.muhat)).reshape(self.Dout, self.M))
        return dops.phi(self.Dout, self.ld_Kuu, self._ld_post, self._sgn_post, dotP,
                        self.ld_Suhatinv, -1.0, dotC, N * 1.0 / alpha - 1.0, -N * 1.0 / alpha)

    # ------------------------------------------------------------------
    # forward propagation (aep_models.py:116-199; base_models.py:235-307)
    # ------------------------------------------------------------------
    def forward_prop_thru_cav(self, mx, vx=None, mode=PROP_MM, eps=None):
        """aep_models.py:116-140.  mode=PROP_MC takes the standard-normal draws eps
        (S, B, Din) as a device tensor (the reference draws them inside, with
        np.random.randn(MC_NO_SAMPLES, B, Din)) and returns ((S, B, Dout) means and
        variances, the stacked kfu (S*B, M) and inputs (S*B, Din))."""
        if vx is None:
            return self._forward_prop_deterministic_thru_cav(mx)
        if mode == PROP_MM:
            return self._forward_prop_random_thru_cav_mm(mx, vx)
        if mode == PROP_MC:
            return self._forward_prop_random_thru_cav_mc(mx, vx, eps)
        if mode == PROP_LIN:
            raise NotImplementedError('propagation mode %s is not on the GPU path' % mode)
        raise NotImplementedError('unknown propagation mode')

    def _forward_prop_random_thru_cav_mc(self, mx, vx, eps):
        """_forward_prop_random_thru_cav_mc (aep_models.py:160-180)."""
        S, B, Q = eps.shape
        x = dops.mc_inputs(eps, mx, vx).reshape(S * B, Q)
        m, v, kfu = self._forward_prop_deterministic_thru_cav(x)
        return m.reshape(S, B, self.Dout), v.reshape(S, B, self.Dout), kfu, x

    def _forward_prop_deterministic_thru_cav(self, x):
        kfu = dops.se(x, self.zu, self.ls, self.sf)
        mout = mm(kfu, self.Ahat, tB=True)
        self._T_det = mm(kfu, self.Bhat_det)          # (D, N, M): kfu Bhat_det[d]
        vout = dops.det_var(self._T_det, kfu, self.sf)
        return mout, vout, kfu

    def _forward_prop_random_thru_cav_mm(self, mx, vx):
        N = mx.shape[0]
        M, D = self.M, self.Dout
        if self.fused and dops.fused_fits(M, mx.shape[1], D):
            if getattr(self, '_Bp', None) is None:
                self._Bp = dops.sym_pairmajor(self.Bhat_sto)
            psi1, mout, vout = dops.fused_mm_fwd(mx, vx, self.zu, self.ls, self.sf, self.Ahat, self._Bp)
            return mout, vout, psi1, None
        psi1, psi2 = dops.psi_fwd(mx, vx, self.zu, self.ls, self.sf)
        mout = mm(psi1, self.Ahat, tB=True)
        Bp = mm(psi2.reshape(N, M * M), self.Bhat_sto.reshape(D, M * M), tB=True)
        vout = dops.mm_var(mout, Bp, self.sf)
        return mout, vout, psi1, psi2

    def forward_prop_thru_post_mc(self, mx, vx, eps):
        """_forward_prop_random_thru_post_mc (base_models.py:309-332) for draws eps
        (S, B, Din): (S, B, Dout) means and variances."""
        S, B, Q = eps.shape
        x = dops.mc_inputs(eps, mx, vx).reshape(S * B, Q)
        m, v = self.forward_prop_thru_post(x)
        return m.reshape(S, B, self.Dout), v.reshape(S, B, self.Dout)

    def forward_prop_thru_post(self, mx, vx=None, mode=PROP_MM, return_info=False):
        self._posterior_predictive()
        if vx is None:
            kfu = dops.se(mx, self.zu, self.ls, self.sf)
            mout = mm(kfu, self.A, tB=True)
            T = mm(kfu, self.B_det)
            vout = dops.det_var(T, kfu, self.sf)
            return (mout, vout, kfu) if return_info else (mout, vout)
        if mode != PROP_MM:
            raise NotImplementedError('propagation mode %s is not on the GPU path' % mode)
        N = mx.shape[0]
        M, D = self.M, self.Dout
        psi1, psi2 = dops.psi_fwd(mx, vx, self.zu, self.ls, self.sf)
        mout = mm(psi1, self.A, tB=True)
        Bp = mm(psi2.reshape(N, M * M), self.B_sto.reshape(D, M * M), tB=True)
        vout = dops.mm_var(mout, Bp, self.sf)
        return (mout, vout, psi1, psi2) if return_info else (mout, vout)

    # ------------------------------------------------------------------
    # adjoints
    # ------------------------------------------------------------------
    def _cav_grad_u(self, dmu, dSu, alpha):
        """compute_cav_grad_u (aep_models.py:548-586) -> deta1_R, deta2, dKuuinv."""
        D, M = self.Dout, self.M
        beta = (self.N - alpha) * 1.0 / self.N
        if self.nat_param:
            _add_outer(dSu, dmu, self.theta_2, beta)
            dSuinv = _sandwich(self.Suhat, dSu, self.Suhat, alpha=-1.0)
            dKuuinv = colsum(dSuinv, D, M * M).reshape(M, M)
            deta1 = dops.pack_dR(self.theta_1_R, lincomb(torch.empty_like(dSuinv), beta, dSuinv))
            deta2 = mm(self.Suhat, _col(dmu), alpha=beta).reshape(D, M)
        else:
            f2 = mm(self.Suinv, _col(self.mu)).reshape(D, M)
            _add_outer(dSu, dmu, f2, beta)
            dSuhatinv = _sandwich(self.Suhat, dSu, self.Suhat, alpha=-1.0)
            Suhatdmu = mm(self.Suhat, _col(dmu)).reshape(D, M)
            dSuinv = lincomb(torch.empty_like(dSuhatinv), beta, dSuhatinv)
            _add_outer(dSuinv, Suhatdmu, self.mu, beta)
            dtheta1 = _sandwich(self.Suinv, dSuinv, self.Suinv, alpha=-1.0)
            deta1 = dops.pack_dR(self.theta_1_R, dtheta1)
            deta2 = mm(self.Suinv, _col(Suhatdmu), alpha=beta).reshape(D, M)
            dKuuinv = colsum(dSuhatinv, D, M * M, alpha=(1 - beta)).reshape(M, M)
        return deta1, deta2, dKuuinv

    def _post_grad_u(self, dmu, dSu):
        """compute_posterior_grad_u (base_models.py:490-516)."""
        D, M = self.Dout, self.M
        if self.nat_param:
            _add_outer(dSu, dmu, self.theta_2, 1.0)
            dSuinv = _sandwich(self.Su, dSu, self.Su, alpha=-1.0)
            dKuuinv = colsum(dSuinv, D, M * M).reshape(M, M)
            deta1 = dops.pack_dR(self.theta_1_R, dSuinv)
            deta2 = mm(self.Su, _col(dmu)).reshape(D, M)
            return deta1, deta2, dKuuinv
        deta1 = dops.pack_dR(self.theta_1_R, dSu)
        return deta1, dmu, None

    def backprop_grads_reg(self, m, v, dm, dv, kfu, x, alpha=1.0):
        """backprop_grads_reg (aep_models.py:412-511): dm, dv already scaled."""
        if self.nat_param:
            return self._backprop_grads_reg_nat(dm, dv, kfu, x, alpha)
        N = self.N
        D, M = self.Dout, self.M
        B = kfu.shape[0]
        Ki = self.Kuuinv
        scale_post = N * 1.0 / alpha - 1.0
        scale_cav = -N * 1.0 / alpha
        kfuKi = mm(kfu, Ki)
        # grads wrt kfu: dm Ahat + 2 sum_d dv_nd (kfu Bhat_det[d])_n
        dkfu = mm(dm, self.Ahat)
        dops.wsum(dv, self._T_det, dkfu, alpha=2.0, beta=1.0)
        dsf2, dell, dzu, _ = dops.kfu_bwd(dkfu, kfu, x, self.zu, self.ls, self.sf)
        sumdv = colsum(dv, B * D, 1)
        # cavity mean / covariance
        dmucav = mm(dm, kfuKi, tA=True)
        dSucav = mm(dops.rowscale(dv, kfuKi), kfuKi, tA=True)
        Suinvmuhat = mm(self.Suhatinv, _col(self.muhat)).reshape(D, M)
        lincomb(dmucav, 1.0, dmucav, scale_cav, Suinvmuhat)
        lincomb(dSucav, 1.0, dSucav, 0.5 * scale_cav, self.Suhatinv)
        _add_outer(dSucav, Suinvmuhat, Suinvmuhat, -0.5 * scale_cav)
        deta1_cav, deta2_cav, dKi_cav = self._cav_grad_u(dmucav, dSucav, alpha)
        # posterior
        Suinvmu = mm(self.Suinv, _col(self.mu)).reshape(D, M)
        dmu = lincomb(torch.empty_like(Suinvmu), scale_post, Suinvmu)
        dSu = lincomb(torch.empty_like(self.Suinv), 0.5 * scale_post, self.Suinv)
        _add_outer(dSu, Suinvmu, Suinvmu, -0.5 * scale_post)
        deta1_post, deta2_post, dKi_post = self._post_grad_u(dmu, dSu)
        # Kuuinv: phi prior, Ahat, Bhat_det, cavity, posterior
        dAhat = mm(dm, kfu, tA=True)
        dKi = mm(dAhat, self.muhat, tA=True)
        lincomb(dKi, 1.0, dKi, -0.5 * D, self.Kuu)
        dBhat = mm(dops.rowscale(dv, kfu), kfu, tA=True)
        KiS = mm(Ki, self.Suhat)
        mm(KiS.reshape(D * M, M), dBhat.reshape(D * M, M), dKi, tA=True, alpha=2.0, beta=1.0)
        colsum(dBhat, D, M * M, out=dKi.reshape(-1), alpha=-1.0, beta=1.0)
        lincomb(dKi, 1.0, dKi, 1.0, dKi_cav)
        if dKi_post is not None:
            lincomb(dKi, 1.0, dKi, 1.0, dKi_post)
        M_inner = _sandwich(Ki, dKi, Ki, alpha=-1.0)
        gsf, gls, gz = dops.dtrace(M_inner, self.zu, self.ls, self.sf)
        dls, dsf = dops.hyper(self.ls, self.sf, dell, gls, dsf2, sumdv, gsf)
        lincomb(dzu, 1.0, dzu, 1.0, gz)
        deta1 = lincomb(deta1_cav, 1.0, deta1_cav, 1.0, deta1_post)
        deta2 = lincomb(deta2_cav, 1.0, deta2_cav, 1.0, deta2_post)
        return {'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': deta2}

    def _backprop_grads_reg_nat(self, dm, dv, kfu, x, alpha):
        """backprop_grads_reg in the natural parameterisation, with the same terms as the
        general path (compute_cav_grad_u, compute_posterior_grad_u; aep_models.py:412-511,
        548-586, base_models.py:490-516) regrouped into fewer launches:
          Sigma-hat^-1 mu-hat = beta theta2 and Sigma^-1 mu = theta2 (no mat-vecs),
          the cavity's normalising-factor terms ride in the products' epilogues,
          dSu (posterior) = sp/2 (Sigma^-1 + theta2 theta2^T)  (its two rank-1 terms
            -sp/2 theta2 theta2^T and dmu theta2^T = sp theta2 theta2^T combined),
          deta2 = beta Sigma-hat dmu_cav + sp mu  (Sigma dmu_post = sp Sigma theta2 = sp mu),
          deta1 = pack(R, beta dSuinv_cav + dSuinv_post) (pack is linear),
          2 KiS^T dBhat - sum_d dBhat = (2 KiS - I)^T dBhat in one product,
          the cavity's dm^T (kfu Ki) = dAhat Ki and sum_n dv_nd (kfu Ki)_n^T (kfu Ki)_n =
            Ki^T dBhat Ki (M x M products in place of two contractions over the batch)."""
        N = self.N
        D, M = self.Dout, self.M
        B = kfu.shape[0]
        Ki = self.Kuuinv
        th = self.theta_2
        beta = self._beta
        scale_post = N * 1.0 / alpha - 1.0
        scale_cav = -N * 1.0 / alpha
        # grads wrt kfu: dm Ahat + 2 sum_d dv_nd (kfu Bhat_det[d])_n
        dkfu = mm(dm, self.Ahat)
        dops.wsum(dv, self._T_det, dkfu, alpha=2.0, beta=1.0)
        dsf2, dell, dzu, _ = dops.kfu_bwd(dkfu, kfu, x, self.zu, self.ls, self.sf)
        sumdv = colsum(dv, B * D, 1)
        dAhat = mm(dm, kfu, tA=True)
        dBhat = mm(dops.rowscale(dv, kfu), kfu, tA=True)
        # cavity: dmu_cav, dSu_cav (+ the normalising factor), then dSuinv_cav
        c = self._adjoint('reg', alpha)
        dmucav = mm(dAhat, Ki, E=th, gamma=scale_cav * beta)
        dSucav = mm(Ki, mm(dBhat, Ki), tA=True, E=c['Ecav'], gamma=1.0)
        _add_outer(dSucav, dmucav, th, beta)
        dS_cav = _sandwich(self.Suhat, dSucav, self.Suhat, alpha=-1.0)
        # posterior chain (data independent, from adjoint_constants)
        deta1 = dops.pack_dR(self.theta_1_R, lincomb(dSucav, beta, dS_cav, 1.0, c['dS_post']))
        deta2 = mm(self.Suhat, _col(dmucav), alpha=beta, E=_col(self.mu), gamma=scale_post).reshape(D, M)
        # Kuuinv: phi prior, Ahat, Bhat_det, cavity, posterior
        dKi = mm(dAhat, self.muhat, tA=True, E=c['Kc'], gamma=1.0)
        mm(c['KiS2'].reshape(D * M, M), dBhat.reshape(D * M, M), dKi, tA=True, beta=1.0)
        colsum(dS_cav, D, M * M, out=dKi.reshape(-1), beta=1.0)
        M_inner = _sandwich(Ki, dKi, Ki, alpha=-1.0)
        gsf, gls, gz = dops.dtrace(M_inner, self.zu, self.ls, self.sf)
        dls, dsf = dops.hyper(self.ls, self.sf, dell, gls, dsf2, sumdv, gsf)
        lincomb(dzu, 1.0, dzu, 1.0, gz)
        return {'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': deta2}

    def backprop_grads_lvm_mc(self, m, v, dm, dv, kfu, x, alpha=1.0):
        """backprop_grads_lvm_mc (aep_models.py:307-410): hyper grads of the layer at the
        stacked Monte Carlo inputs x (S*B rows) and the grads w.r.t. x; dm, dv already
        scaled, S*B*Dout entries in sample-major order.  The cavity terms are formed as
        -Suhat (sum_n dv kfuKi kfuKi^T) Suhat - (Suhat sum_n dm kfuKi) muhat^T, the
        regrouping of the reference's per-point products Suhat Kuu^-1 kfu_n."""
        N = self.N
        D, M = self.Dout, self.M
        B = kfu.shape[0]
        Ki = self.Kuuinv
        beta = (N - alpha) * 1.0 / N
        dm = dm.reshape(B, D)
        dv = dv.reshape(B, D)
        kfuKi = mm(kfu, Ki)
        dkfu = mm(dm, self.Ahat)
        dops.wsum(dv, self._T_det, dkfu, alpha=2.0, beta=1.0)
        dsf2, dell, dzu, dx = dops.kfu_bwd(dkfu, kfu, x, self.zu, self.ls, self.sf, want_dx=True)
        sumdv = colsum(dv, B * D, 1)
        dmucav = mm(dm, kfuKi, tA=True)
        dSucav = mm(dops.rowscale(dv, kfuKi), kfuKi, tA=True)
        dSinv = _sandwich(self.Suhat, dSucav, self.Suhat, alpha=-1.0)
        dSinvM = mm(self.Suhat, _col(dmucav)).reshape(D, M)
        _add_outer(dSinv, dSinvM, self.muhat, -1.0)
        c = self._adjoint('mm', alpha)
        dtheta1 = lincomb(torch.empty_like(dSinv), beta, dSinv, 1.0, c['T1'])
        dtheta2 = lincomb(torch.empty_like(dSinvM), beta, dSinvM, 1.0, c['c2'])
        deta1 = dops.pack_dR(self.theta_1_R, dtheta1)
        dAhat = mm(dm, kfu, tA=True)
        dKi = colsum(dSinv, D, M * M).reshape(M, M)
        mm(dAhat, self.muhat, dKi, tA=True, beta=1.0)
        dBhat = mm(dops.rowscale(dv, kfu), kfu, tA=True)
        KiS2 = mm(Ki, self.Suhat, alpha=2.0, E=_eye(M, Ki), gamma=-1.0)
        mm(KiS2.reshape(D * M, M), dBhat.reshape(D * M, M), dKi, tA=True, beta=1.0)
        Minner = lincomb(torch.empty_like(dKi), 1.0, c['Mc'], -2.0, dKi)
        M_all = _sandwich(Ki, Minner, Ki, alpha=0.5, E=Ki, gamma=0.5 * D)
        gsf, gls, gz = dops.dtrace(M_all, self.zu, self.ls, self.sf)
        dls, dsf = dops.hyper(self.ls, self.sf, dell, gls, dsf2, sumdv, gsf)
        lincomb(dzu, 1.0, dzu, 1.0, gz)
        return {'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': dtheta2}, dx

    def backprop_grads_reparam(self, dx, mx, vx, eps):
        """backprop_grads_reparam (base_models.py:373-388): {'mx', 'vx'} of
        x = eps sqrt(vx) + mx."""
        dmx, dvx = dops.mc_reparam_bwd(dx, eps, vx)
        return {'mx': dmx, 'vx': dvx}

    def backprop_grads_lvm_mm(self, m, v, dm, dv, psi1, psi2, mx, vx, alpha=1.0, dm_all=None,
                              side=None, keep=None):
        """backprop_grads_lvm_mm (aep_models.py:201-304): dm, dv already scaled.
        Returns (hyper grads, {'mx', 'vx'}).

        With a `side` stream, only the input gradients (the psi adjoint) are on the
        current stream; the hyper-parameter chain through Kuu^-1 and q(u) is forked onto
        `side`, so the caller can carry on with the next layer's backward and must join
        `side` before reading the hyper grads.  Tensors of the current stream that the
        fork reads are appended to `keep`, which the caller holds until the join."""
        N = self.N
        D, M = self.Dout, self.M
        B = psi1.shape[0]
        Ki = self.Kuuinv
        beta = (N - alpha) * 1.0 / N
        scale_post = N * 1.0 / alpha - 1.0
        scale_cav = -N * 1.0 / alpha
        if dm_all is None:
            dm_all = dops.dmall(dm, dv, m)
        dAhat = mm(dm_all, psi1, tA=True)
        rest = None
        if psi2 is None and side is not None:
            # the psi adjoint's per-pair pass and hyper-parameter tail join the forked chain
            # below: the layer below needs only dmx, dvx
            dmx, dvx, rest = dops.fused_mm_bwd_split(
                mx, vx, self.zu, self.ls, self.sf, self.Ahat, self._Bp, psi1, dm_all, dv)
        elif psi2 is None:
            dBhat, dsf2, dell, dzu, dmx, dvx = dops.fused_mm_bwd(
                mx, vx, self.zu, self.ls, self.sf, self.Ahat, self._Bp, psi1, dm_all, dv)
        else:
            dBhat = mm(dv, psi2.reshape(B, M * M), tA=True).reshape(D, M, M)
            dpsi1 = mm(dm_all, self.Ahat)
            Bsym = dops.sym(self.Bhat_sto, 0.5)
            dpsi2 = mm(dv, Bsym.reshape(D, M * M)).reshape(B, M, M)
            dsf2, dell, dzu, dmx, dvx = dops.psi_bwd(mx, vx, self.zu, self.ls, self.sf,
                                                     dpsi1, psi1, dpsi2, psi2)
            del dpsi2
        if side is None:
            grad = self._lvm_hyper_chain(dv, dAhat, dBhat, dsf2, dell, dzu, B, alpha)
        else:
            side.wait_stream(torch.cuda.current_stream())
            if rest is None:
                keep.extend([dv, dAhat, dBhat, dsf2, dell, dzu])
            else:
                keep.extend([dv, dAhat] + rest.keep)
            with torch.cuda.stream(side):
                if rest is not None:
                    dBhat, dsf2, dell, dzu = rest()
                grad = self._lvm_hyper_chain(dv, dAhat, dBhat, dsf2, dell, dzu, B, alpha)
        return grad, {'mx': dmx, 'vx': dvx}

    def _lvm_hyper_chain(self, dv, dAhat, dBhat, dsf2, dell, dzu, B, alpha):
        """The q(u) / Kuu^-1 / hyper-parameter part of backprop_grads_lvm_mm
        (aep_models.py:245-304) from dAhat, dBhat and the psi adjoint."""
        N = self.N
        D, M = self.Dout, self.M
        Ki = self.Kuuinv
        beta = (N - alpha) * 1.0 / N
        scale_post = N * 1.0 / alpha - 1.0
        scale_cav = -N * 1.0 / alpha
        sumdv = colsum(dv, B * D, 1)
        dvcav = _sandwich(Ki, dBhat, Ki)
        dmcav = mm(dvcav, _col(self.muhat), alpha=2.0).reshape(D, M)
        mm(dAhat, Ki, dmcav, tB=True, beta=1.0)
        _add_outer(dvcav, dmcav, self.theta_2, beta)
        dvcavinv = _sandwich(self.Suhat, dvcav, self.Suhat, alpha=-1.0)
        c = self._adjoint('mm', alpha)
        dtheta2 = mm(self.Suhat, _col(dmcav), alpha=beta, E=_col(c['c2']), gamma=1.0).reshape(D, M)
        dKi = colsum(dvcavinv, D, M * M).reshape(M, M)
        mm(dAhat, self.muhat, dKi, tA=True, beta=1.0)
        # 2 sum_d KiS_d^T dBhat_d - sum_d dBhat_d as one product with KiS2 = 2 KiS - I
        mm(c['KiS2'].reshape(D * M, M), dBhat.reshape(D * M, M), dKi, tA=True, beta=1.0)
        dtheta1 = lincomb(torch.empty_like(dvcavinv), beta, dvcavinv, 1.0, c['T1'])
        deta1 = dops.pack_dR(self.theta_1_R, dtheta1)
        # M_all = Ki (Mc - 2 dKi) Ki / 2 + (D/2) Ki, with Mc Ki from the adjoint constants
        M_all = mm(Ki, mm(dKi, Ki, alpha=-2.0, E=c['McKi'], gamma=1.0), alpha=0.5, E=Ki, gamma=0.5 * D)
        gsf, gls, gz = dops.dtrace(M_all, self.zu, self.ls, self.sf)
        dls, dsf = dops.hyper(self.ls, self.sf, dell, gls, dsf2, sumdv, gsf)
        dzu = lincomb(torch.empty_like(dzu), 1.0, dzu, 1.0, gz)
        return {'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': dtheta2}
