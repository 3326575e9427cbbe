"""Variational free energy (VFE) sparse GP models on the GPU: SGPR, SGPLVM, SGPSSM.

Drop-in for geepee/vfe_models.py:292-1120: the same constructors and
``objective_function(params, mb_size, alpha='not_used', prop_mode) -> (energy, grad_dict)``
(the uncollapsed bound, divided by N), with the posterior q(u) in place of the AEP
cavity.  They share the AEP models' parameter transfer, prediction, optimisers and
checkpointing (geepee_amd.aep_models), and the layer's q(u) / Kuu algebra
(geepee_amd.layers).  Minibatches and Monte Carlo draws come from numpy's legacy global
RNG in the reference's order, so a common seed gives the same samples.
The collapsed bound (SGPR_collapsed, vfe_models.py:15-290) is not on the GPU path.
"""
import numpy as np
import torch

from . import _lib, aep_models as AEP, config, dops, init
from ._lib import P, call, query
from .config import PROP_MM, PROP_MC
from .dops import mm, lincomb, colsum
from .layers import SGP_Layer as _Layer, _col, _sandwich, _eye

F64 = torch.float64


def _scal(like):
    """[result, second result, 2 x 64 block partials] of the reduction entry points."""
    return torch.empty(2 + 2 * 64, dtype=F64, device=like.device)


def _dims(m):
    three = m.dim() == 3
    return (m.shape[0] if three else 1), int(three), m.shape[-2], m.shape[-1]


class Gauss_Layer(AEP.Gauss_Layer):
    """Gaussian likelihood's expected log-likelihood (geepee/lik_layers.py:183-236)."""

    def compute_log_lik_exp(self, m, v, y, scale):
        """(dm, dv, scal) for (B, D) propagations or (S, B, D) Monte Carlo samples:
        scal[0] = scale * E log N(y | f, sn2), scal[1] = scale * dE/dsn, dm and dv scaled."""
        Sn, three, B, D = _dims(m)
        dm, dv = torch.empty_like(m), torch.empty_like(m)
        scal = _scal(m)
        call('gpx_gauss_lik_exp', Sn, three, B, D, P(m), P(v), P(y), P(self.sn), float(scale), P(dm),
             P(dv), P(scal), _lib.S())
        return dm, dv, scal


class Probit_Layer(AEP.Probit_Layer):
    """Probit likelihood's expected log-likelihood by Gauss-Hermite quadrature
    (geepee/lik_layers.py:412-455)."""

    def compute_log_lik_exp(self, m, v, y, scale):
        Sn, three, B, D = _dims(m)
        dm, dv = torch.empty_like(m), torch.empty_like(m)
        scal = _scal(m)
        call('gpx_probit_lik_exp', Sn, three, B, D, P(m), P(v), P(y), float(scale), P(dm), P(dv),
             P(scal), _lib.S())
        return dm, dv, scal


def _make_lik(lik, N, D):
    if lik.lower() == 'gaussian':
        return Gauss_Layer(N, D)
    if lik.lower() == 'probit':
        return Probit_Layer(N, D)
    raise NotImplementedError('likelihood not implemented')


def transition_lik_exp(mp, vp, mt, vt, sn, scale):
    """SGPSSM.compute_transition_log_lik_exp (vfe_models.py:1080-1120) for (B, D)
    propagations or (S, B, D) samples: (dmp, dvp, dmt, dvt, scal), scal[0] = the scaled
    expectation, scal[1] = dsn."""
    Sn, three, B, D = _dims(mp)
    dmp, dvp = torch.empty_like(mp), torch.empty_like(mp)
    dmt, dvt = torch.empty_like(mt), torch.empty_like(mt)
    scal = _scal(mp)
    call('gpx_transition_lik_exp', Sn, three, B, D, P(mp), P(vp), P(mt), P(vt), P(sn), float(scale),
         P(dmp), P(dvp), P(dmt), P(dvt), P(scal), _lib.S())
    return dmp, dvp, dmt, dvt, scal


def emission_lik_exp(C, Rp, y, mx, vx, scale):
    """Gauss_Emis.compute_emission_log_lik_exp (lik_layers.py:629-676), scaled:
    (dmx, dvx, dC, dR, scal) with scal[0] = scale * E log N(y | C x, R)."""
    Nb, Din = mx.shape
    Dout = C.shape[0]
    outs = [torch.empty_like(mx), torch.empty_like(mx), torch.empty_like(C), torch.empty_like(Rp),
            _scal(mx)]
    work = dops.empty(query('gpx_emission_lik_exp_workspace', Dout, Din), like=mx)
    call('gpx_emission_lik_exp', Nb, Dout, Din, P(C), P(Rp), P(y), y.stride(0), P(mx), mx.stride(0), P(vx),
         vx.stride(0), float(scale), *[P(o) for o in outs], P(work), _lib.S())
    return outs


class SGP_Layer(_Layer):
    """SGP_Layer of geepee/vfe_models.py:292-550: propagation through the posterior q(u)
    (base_models.py:265-332) and the adjoints of the expected log-likelihood plus
    KL(q(u) || p(u)).  The q(u)-only parts of the adjoints are formed once per q(u) in
    prepare_post:  dSu's constant 0.5 (Kuu^-1 - Su^-1), the KL's
    dKuu^-1 = sum_d Splusmm_d / 2 - D Kuu / 2, and 2 Kuu^-1 S - I for S = Su (deterministic
    inputs) or Splusmm (moment matching), which turns 2 sum_d (Kuu^-1 S_d)^T dB_d - sum_d dB_d
    into one product."""

    def prepare_post(self, params, key_suffix=''):
        self.update_hypers(params, key_suffix)
        self._posterior_predictive()
        D, M = self.Dout, self.M
        self._C0 = lincomb(torch.empty_like(self.Su), 0.5, self.Kuuinv, -0.5, self.Suinv)
        self._Kc = lincomb(torch.empty_like(self.Kuu), -0.5 * D, self.Kuu)
        colsum(self.Splusmm, D, M * M, out=self._Kc.reshape(-1), alpha=0.5, beta=1.0)
        self._KiS2 = {}

    def _kis2(self, kind):
        if kind not in self._KiS2:
            S = self.Su if kind == 'det' else self.Splusmm
            self._KiS2[kind] = mm(self.Kuuinv, S, alpha=2.0, E=_eye(self.M, S), gamma=-1.0)
        return self._KiS2[kind]

    def compute_KL(self):
        """SGP_Layer.compute_KL (vfe_models.py:309-326) as a device scalar (1,)."""
        out = dops.empty(1, like=self.Kuu)
        call('gpx_vfe_kl', self.Dout, self.M, P(self.ld_Kuu), P(self._ld_post), float(self._sgn_post),
             P(self.Kuuinv), P(self.Splusmm), P(out), _lib.S())
        return out

    def forward_post_det(self, x):
        """forward_prop_thru_post of deterministic inputs with return_info
        (base_models.py:265-284): (mout, vout, kfu); keeps kfu B_det for the adjoint."""
        self._posterior_predictive()
        kfu = dops.se(x, self.zu, self.ls, self.sf)
        mout = mm(kfu, self.A, tB=True)
        self._T_post = mm(kfu, self.B_det)
        vout = dops.det_var(self._T_post, kfu, self.sf)
        return mout, vout, kfu

    def forward_post_mc(self, mx, vx, eps):
        """_forward_prop_random_thru_post_mc with return_info (base_models.py:309-332) for
        draws eps (S, B, Din): (S, B, Dout) means and variances, kfu (S*B, M), inputs x."""
        Sn, B, Q = eps.shape
        x = dops.mc_inputs(eps, mx, vx).reshape(Sn * B, Q)
        m, v, kfu = self.forward_post_det(x)
        return m.reshape(Sn, B, self.Dout), v.reshape(Sn, B, self.Dout), kfu, x

    def _kuu_tail(self, dKi, dsf2, dell, dzu, sumdv):
        """Kuu^-1's adjoint through Kuu into the hyper-parameters."""
        Ki = self.Kuuinv
        M_inner = _sandwich(Ki, dKi, Ki, alpha=-1.0)
        gsf, gls, gz = dops.dtrace(M_inner, self.zu, self.ls, self.sf)
        dls, dsf = dops.hyper(self.ls, self.sf, dell, gls, dsf2, sumdv, gsf)
        lincomb(dzu, 1.0, dzu, 1.0, gz)
        return dsf, dls, dzu

    def backprop_grads_reg(self, m, v, dm, dv, kfu, x, want_dx=False):
        """backprop_grads_reg (vfe_models.py:479-548); with want_dx also the grads w.r.t.
        the inputs x, which is backprop_grads_lvm_mc (:405-477) at the stacked Monte Carlo
        inputs.  dm, dv already scaled; any shape with kfu.shape[0] * Dout entries."""
        D, M = self.Dout, self.M
        B = kfu.shape[0]
        Ki = self.Kuuinv
        dm = dm.reshape(B, D)
        dv = dv.reshape(B, D)
        # grads wrt kfu: dm A + 2 sum_d dv_nd (kfu B_det[d])_n
        dkfu = mm(dm, self.A)
        dops.wsum(dv, self._T_post, dkfu, alpha=2.0, beta=1.0)
        dsf2, dell, dzu, dx = dops.kfu_bwd(dkfu, kfu, x, self.zu, self.ls, self.sf, want_dx=want_dx)
        sumdv = colsum(dv, B * D, 1)
        kfuKi = mm(kfu, Ki)
        # dmu = dm^T kfu Kuu^-1 + Kuu^-1 mu (A = mu Kuu^-1), dSu = sum_n dv kfuKi kfuKi^T + C0
        dmu = mm(dm, kfuKi, tA=True, E=self.A, gamma=1.0)
        dSu = mm(dops.rowscale(dv, kfuKi), kfuKi, tA=True, E=self._C0, gamma=1.0)
        deta1, deta2, dKi_u = self._post_grad_u(dmu, dSu)
        dA = mm(dm, kfu, tA=True)
        dKi = mm(dA, self.mu, tA=True, E=self._Kc, gamma=1.0)
        dB = mm(dops.rowscale(dv, kfu), kfu, tA=True)
        mm(self._kis2('det').reshape(D * M, M), dB.reshape(D * M, M), dKi, tA=True, beta=1.0)
        if dKi_u is not None:
            lincomb(dKi, 1.0, dKi, 1.0, dKi_u)
        dsf, dls, dzu = self._kuu_tail(dKi, dsf2, dell, dzu, sumdv)
        grad = {'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': deta2}
        return (grad, dx) if want_dx else grad

    def backprop_grads_lvm_mm(self, m, v, dm, dv, psi1, psi2, mx, vx):
        """backprop_grads_lvm_mm (vfe_models.py:328-403): dm, dv already scaled.
        Returns (hyper grads, {'mx', 'vx'})."""
        D, M = self.Dout, self.M
        B = psi1.shape[0]
        Ki = self.Kuuinv
        dm_all = dops.dmall(dm, dv, m)
        dpsi1 = mm(dm_all, self.A)
        Bsym = dops.sym(self.B_sto, 0.5)
        dpsi2 = mm(dv, Bsym.reshape(D, M * M)).reshape(B, M, M)
        dsf2, dell, dzu, dmx, dvx = dops.psi_bwd(mx, vx, self.zu, self.ls, self.sf, dpsi1, psi1, dpsi2,
                                                 psi2)
        del dpsi2
        sumdv = colsum(dv, B * D, 1)
        dA = mm(dm_all, psi1, tA=True)
        dB = mm(dv, psi2.reshape(B, M * M), tA=True).reshape(D, M, M)
        # dSu = Kuu^-1 dB Kuu^-1; dmu = 2 dSu mu + Kuu^-1 dA + Kuu^-1 mu
        dSu = _sandwich(Ki, dB, Ki)
        dmu = mm(dSu, _col(self.mu), alpha=2.0, E=_col(self.A), gamma=1.0).reshape(D, M)
        mm(dA, Ki, dmu, tB=True, beta=1.0)
        lincomb(dSu, 1.0, dSu, 1.0, self._C0)
        deta1, deta2, dKi_u = self._post_grad_u(dmu, dSu)
        dKi = mm(dA, self.mu, tA=True, E=self._Kc, gamma=1.0)
        mm(self._kis2('mm').reshape(D * M, M), dB.reshape(D * M, M), dKi, tA=True, beta=1.0)
        if dKi_u is not None:
            lincomb(dKi, 1.0, dKi, 1.0, dKi_u)
        dsf, dls, dzu = self._kuu_tail(dKi, dsf2, dell, dzu, sumdv)
        return ({'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': deta2},
                {'mx': dmx, 'vx': dvx})


def _energy(*terms):
    out = torch.empty(1, dtype=F64, device='cuda')
    if len(terms) == 2:
        return lincomb(out, 1.0, terms[0], 1.0, terms[1])
    return lincomb(out, 1.0, terms[0], 1.0, terms[1], 1.0, terms[2])


class SGPR(AEP.SGPR):
    """vfe_models.py:551-720: minus the uncollapsed bound (expected log-likelihood of the
    minibatch scaled to N, plus KL(q(u))), divided by N.  alpha and prop_mode are not used."""

    def __init__(self, x_train, y_train, no_pseudo, lik='Gaussian', nat_param=True):
        super(SGPR, self).__init__(x_train, y_train, no_pseudo, lik, nat_param)
        self.lik_layer = _make_lik(lik, self.N, self.Dout)
        self.sgp_layer = SGP_Layer(self.N, self.Din, self.Dout, self.M, nat_param)

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        N = self.N
        if idx is None and mb_size < N:
            idx = torch.as_tensor(np.random.choice(N, mb_size, replace=False), device='cuda')
        if idx is None:
            xb, yb = self.x_dev, self.y_dev
        else:
            xb, yb = self.x_dev.index_select(0, idx), self.y_dev.index_select(0, idx)
        scale = -N * 1.0 / yb.shape[0]
        dp = self._device_params(params)
        layer = self.sgp_layer
        layer.prepare_post(dp)
        self.lik_layer.update_hypers(dp)
        m, v, kfu = layer.forward_post_det(xb)
        dm, dv, scal = self.lik_layer.compute_log_lik_exp(m, v, yb, scale)
        grads = layer.backprop_grads_reg(m, v, dm, dv, kfu, xb)
        grads.update(self.lik_layer.hyper_grads(scal))
        return _energy(scal[0:1], layer.compute_KL()), grads, N


class SGPLVM(AEP.SGPLVM):
    """vfe_models.py:722-864: the expected log-likelihood under q(x) (moment matching or
    Monte Carlo), minus KL(q(x)) of the minibatch (both scaled to N), minus KL(q(u)); the
    whole divided by N (vfe_models.py:838-845)."""

    def __init__(self, y_train, hidden_size, no_pseudo, lik='Gaussian', prior_mean=0, prior_var=1,
                 nat_param=True):
        super(SGPLVM, self).__init__(y_train, hidden_size, no_pseudo, lik, prior_mean, prior_var,
                                     nat_param)
        self.lik_layer = _make_lik(lik, self.N, self.Dout)
        self.sgp_layer = SGP_Layer(self.N, self.Din, self.Dout, self.M, nat_param)

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        N, Din = self.N, self.Din
        if idx is None and mb_size != N:
            idx = torch.as_tensor(np.random.choice(N, mb_size, replace=False), dtype=torch.int64,
                                  device='cuda')
        yb = self.y_dev if idx is None else self.y_dev.index_select(0, idx)
        B = yb.shape[0]
        scale = -N * 1.0 / B
        dp = self._device_params(params)
        layer = self.sgp_layer
        layer.prepare_post(dp)
        self.lik_layer.update_hypers(dp)
        nat = int(bool(self.nat_param))
        x1, x2 = dp['x1'], dp['x2']
        mx = torch.empty(B, Din, dtype=F64, device='cuda')
        vx = torch.empty_like(mx)
        call('gpx_lvm_x_post', B, Din, P(idx), P(x1), P(x2), float(self.prior_mean),
             float(self.prior_var), nat, P(mx), P(vx), _lib.S())
        if prop_mode == PROP_MC:
            # vfe_models.py:795 (base_models.py:320): the draws follow the minibatch draw
            eps = AEP._mc_draws(config.MC_NO_SAMPLES, B, Din)
            m3, v3, kfu, x = layer.forward_post_mc(mx, vx, eps)
            dm, dv, scal = self.lik_layer.compute_log_lik_exp(m3, v3, yb, scale)
            grads, dx = layer.backprop_grads_reg(m3, v3, dm, dv, kfu, x, want_dx=True)
            dmx, dvx = dops.mc_reparam_bwd(dx, eps, vx)
        else:
            m, v, psi1, psi2 = layer.forward_prop_thru_post(mx, vx, return_info=True)
            dm, dv, scal = self.lik_layer.compute_log_lik_exp(m, v, yb, scale)
            grads, gin = layer.backprop_grads_lvm_mm(m, v, dm, dv, psi1, psi2, mx, vx)
            dmx, dvx = gin['mx'], gin['vx']
            del psi2
        grads.update(self.lik_layer.hyper_grads(scal))
        g1 = torch.zeros_like(x1) if idx is not None else torch.empty_like(x1)
        g2 = torch.zeros_like(x2) if idx is not None else torch.empty_like(x2)
        xs = _scal(x1)
        call('gpx_lvm_vfe_x_grads', B, Din, P(idx), P(x1), P(x2), P(dmx), P(dvx), float(self.prior_mean),
             float(self.prior_var), nat, N * 1.0 / B, P(g1), P(g2), P(xs), _lib.S())
        grads['x1'] = g1
        grads['x2'] = g2
        return _energy(scal[0:1], xs[0:1], layer.compute_KL()), grads, N


class SGPSSM(AEP.SGPSSM):
    """vfe_models.py:866-1120: expected transition and emission log-likelihoods of a
    window of the latent path (scaled to the whole sequence), plus the entropy of q(x) on
    the window, minus the KL terms of q(u); divided by N.  The prior term of x_0 is left
    out, as in the reference (vfe_models.py:1045-1048)."""

    def __init__(self, y_train, hidden_size, no_pseudo, lik='Gaussian', prior_mean=0, prior_var=1,
                 x_control=None, gp_emi=False, control_to_emi=True, nat_param=True):
        super(SGPSSM, self).__init__(y_train, hidden_size, no_pseudo, lik, prior_mean, prior_var,
                                     x_control, gp_emi, control_to_emi)
        self.nat_param = nat_param
        self.dyn_layer = SGP_Layer(self.N - 1, self.Din + self.Dcon_dyn, self.Din, self.M, nat_param)
        if gp_emi:
            self.emi_layer = SGP_Layer(self.N, self.Din + self.Dcon_emi, self.Dout, self.M, nat_param)
            self.lik_layer = _make_lik(lik, self.N, self.Dout)

    def init_hypers(self, y_train=None):
        y = self.y_train if y_train is None else np.asarray(y_train, dtype=np.float64)
        return init.sgpssm(y, self.Din, self.M, self.x_control, self.gp_emi, self.Dcon_emi > 0,
                           nat_param=self.nat_param)

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        if idx is not None:
            raise NotImplementedError('SGPSSM draws contiguous windows, not index sets')
        N, Din = self.N, self.Din
        if mb_size >= N:
            s, e = 0, N
        else:
            s = int(np.random.randint(0, N - mb_size))
            e = s + mb_size
        Bd, Be = e - 1 - s, e - s
        sc_dyn = -(N - 1) * 1.0 / Bd
        sc_emi = -N * 1.0 / Be
        dp = self._device_params(params)
        dyn = self.dyn_layer
        dyn.prepare_post(dp, key_suffix='_dynamic')
        if self.gp_emi:
            self.emi_layer.prepare_post(dp, key_suffix='_emission')
            self.lik_layer.update_hypers(dp, key_suffix='_emission')
        else:
            self.emi_layer.update_hypers(dp, key_suffix='_emission')
        self.sn = dp['sn']
        xf1, x2 = dp['x_factor_1'], dp['x_factor_2']
        pm, pv = dops.ssm_posterior(xf1, x2, self.prior_mean, self.prior_var)
        tm1_m, tm1_v = self._with_control(pm[s:e - 1], pv[s:e - 1], s, e - 1, self.Dcon_dyn)
        up_m, up_v = self._with_control(pm[s:e], pv[s:e], s, e, self.Dcon_emi)
        tm1_m, tm1_v, up_m, up_v = (t.contiguous() for t in (tm1_m, tm1_v, up_m, up_v))
        t_m, t_v = pm[s + 1:e].contiguous(), pv[s + 1:e].contiguous()
        mc = prop_mode == PROP_MC
        if mc:
            # vfe_models.py:986-999: the dynamics draws come first
            eps = AEP._mc_draws(config.MC_NO_SAMPLES, Bd, Din + self.Dcon_dyn)
            m3, v3, kfu, x = dyn.forward_post_mc(tm1_m, tm1_v, eps)
            dmp, dvp, dmt, dvt, scal_t = transition_lik_exp(m3, v3, t_m, t_v, self.sn, sc_dyn)
            g_dyn, dx = dyn.backprop_grads_reg(m3, v3, dmp, dvp, kfu, x, want_dx=True)
            dmn, dvn = dops.mc_reparam_bwd(dx, eps, tm1_v)
        else:
            mp, vp, psi1, psi2 = dyn.forward_prop_thru_post(tm1_m, tm1_v, return_info=True)
            dmp, dvp, dmt, dvt, scal_t = transition_lik_exp(mp, vp, t_m, t_v, self.sn, sc_dyn)
            g_dyn, gin = dyn.backprop_grads_lvm_mm(mp, vp, dmp, dvp, psi1, psi2, tm1_m, tm1_v)
            dmn, dvn = gin['mx'], gin['vx']
            del psi2
        grads = {'sn': scal_t[1:2]}
        for k, v in g_dyn.items():
            grads[k + '_dynamic'] = v
        energy = _energy(scal_t[0:1], dyn.compute_KL())
        if self.gp_emi:
            emi = self.emi_layer
            ye = self.y_dev[s:e]
            if mc:
                eps = AEP._mc_draws(config.MC_NO_SAMPLES, Be, Din + self.Dcon_emi)
                m3, v3, kfu, x = emi.forward_post_mc(up_m, up_v, eps)
                dm, dv, scal_e = self.lik_layer.compute_log_lik_exp(m3, v3, ye, sc_emi)
                g_emi, dx = emi.backprop_grads_reg(m3, v3, dm, dv, kfu, x, want_dx=True)
                dup, vup = dops.mc_reparam_bwd(dx, eps, up_v)
            else:
                m, v, e1, e2 = emi.forward_prop_thru_post(up_m, up_v, return_info=True)
                dm, dv, scal_e = self.lik_layer.compute_log_lik_exp(m, v, ye, sc_emi)
                g_emi, gin = emi.backprop_grads_lvm_mm(m, v, dm, dv, e1, e2, up_m, up_v)
                dup, vup = gin['mx'], gin['vx']
                del e2
            for k, v in g_emi.items():
                grads[k + '_emission'] = v
            grads.update(self.lik_layer.hyper_grads(scal_e, '_emission'))
            lincomb(energy, 1.0, energy, 1.0, scal_e[0:1], 1.0, emi.compute_KL())
        else:
            dup, vup, dC, dR, lz = emission_lik_exp(self.emi_layer.C, self.emi_layer.Rp, self.y_dev[s:e],
                                                    up_m, up_v, sc_emi)
            grads['C_emission'] = dC
            grads['R_emission'] = dR
            lincomb(energy, 1.0, energy, 1.0, lz[0:1])
        g1, g2 = torch.zeros_like(xf1), torch.zeros_like(x2)
        xs = _scal(xf1)
        call('gpx_ssm_vfe_x_grads', N, Din, s, e, P(xf1), P(x2), float(self.prior_mean), float(self.prior_var),
             P(dup), P(vup), dup.stride(0), P(dmt), P(dvt), P(dmn), P(dvn), dmn.stride(0), -N * 1.0 / Be,
             P(g1), P(g2), P(xs), _lib.S())
        grads['x_factor_1'] = g1
        grads['x_factor_2'] = g2
        lincomb(energy, 1.0, energy, 1.0, xs[0:1])
        return energy, grads, N
