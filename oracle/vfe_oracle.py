"""CPU restatement (numpy float64) of geepee's variational free energy (VFE) family.

TEST INFRASTRUCTURE ONLY, like aep_oracle.py (whose layer state, kernels and adjoints it
reuses): only ``tests/`` may import it.  Each function cites the reference lines it
follows in geepee/vfe_models.py and geepee/lik_layers.py.

Pins (tests/test_oracle_pins_vfe.py): the SGPR energy at the optimal q(u) equals the
collapsed Titsias bound and in general the closed-form uncollapsed bound; the SGPLVM
energy at the optimal q(u) equals the Titsias-Lawrence bound minus KL(q(x)); the SGPSSM
energy equals an independently written free energy evaluated by quadrature; the
collapsed SGPR equals the Titsias bound as alpha -> 0 and the FITC marginal likelihood
at alpha = 1; every gradient against finite differences; the expected log-likelihoods
against quadrature.
"""
import numpy as np

from oracle import aep_oracle as O
from oracle.aep_oracle import ein, layer_update, pack_dR, dtrace_MKzz, kfu_grads, psi_grads, \
    psi_stats, se_kernel, _post_grad_u, _finish, JITTER, GH_DEGREE, MC_NO_SAMPLES, PROP_MC, PROP_MM


# ----------------------------------------------------------------------------
# SGP_Layer (vfe_models.py:292-550): the layer at its posterior
# ----------------------------------------------------------------------------

def kl_u(L):
    """SGP_Layer.compute_KL (vfe_models.py:303-324): KL(q(u) || p(u)) summed over outputs."""
    logdet_prior = L.Dout * np.linalg.slogdet(L.Kuu)[1]
    logdet_post = np.sum(np.linalg.slogdet(L.Su)[1])
    trace_term = np.sum(L.Kuuinv * L.Splusmm)
    return 0.5 * (logdet_prior - logdet_post - L.Dout * L.M + trace_term)


def fwd_det(L, x):
    """Posterior propagation of deterministic inputs, returning kfu as well
    (base_models.py:265-284 with return_info)."""
    kfu = se_kernel(L.ls, L.sf, x, L.zu)
    mout = ein('nm,dm->nd', kfu, L.A)
    vout = np.exp(2 * L.sf) + ein('dab,na,nb->nd', L.B_det, kfu, kfu)
    return mout, vout, kfu


def fwd_mm(L, mx, vx):
    """base_models.py:286-307 with return_info: (mout, vout, psi1, psi2)."""
    psi1, psi2 = psi_stats(L.ls, L.sf, mx, vx, L.zu)
    mout = ein('nm,dm->nd', psi1, L.A)
    vout = np.exp(2.0 * L.sf) + ein('dab,nab->nd', L.B_sto, psi2) - mout ** 2
    return mout, vout, psi1, psi2


def fwd_mc(L, mx, vx, eps):
    """base_models.py:309-332 with return_info for draws eps (S, B, Din)."""
    S, B, Din = eps.shape
    x = (eps * np.sqrt(vx) + mx).reshape(S * B, Din)
    m, v, kfu = fwd_det(L, x)
    return m.reshape(S, B, -1), v.reshape(S, B, -1), kfu, x


def _kuu_tail(L, dKuuinv, dsf, dls, dzu):
    M_inner = -np.dot(L.Kuuinv, np.dot(dKuuinv, L.Kuuinv))
    g_sf, g_ls, g_z = dtrace_MKzz(L.ls, L.sf, L.zu, M_inner, L.Kuu - np.diag(JITTER * np.ones(L.M)))
    return dsf + 2 * g_sf, dls + 2 * g_ls, dzu + g_z


def bwd_det(L, dm, dv, kfu, x, grad_x=False):
    """SGP_Layer.backprop_grads_reg (vfe_models.py:472-548); with grad_x the Monte Carlo
    form backprop_grads_lvm_mc (:406-470), which also returns the grads w.r.t. x."""
    ell = np.exp(L.ls)
    sf2 = np.exp(2 * L.sf)
    dm = dm.reshape(kfu.shape[0], L.Dout)
    dv = dv.reshape(kfu.shape[0], L.Dout)
    kfuKuuinv = np.dot(kfu, L.Kuuinv)
    dkfu = ein('nd,dm->nm', dm, L.A) + 2 * ein('nd,dab,na->nb', dv, L.B_det, kfu)
    out = kfu_grads(dkfu, kfu, sf2, ell, L.zu, x, grad_x=grad_x)
    dsf2, dls, dzu = out[:3]
    dls = dls * ell
    dsf = 2 * sf2 * (dsf2 + np.sum(dv))
    dmu = ein('nd,nm->dm', dm, kfuKuuinv) + ein('ab,db->da', L.Kuuinv, L.mu)
    dSu = ein('na,nd,nb->dab', kfuKuuinv, dv, kfuKuuinv) + 0.5 * (L.Kuuinv - L.Suinv)
    deta1, deta2, dKi_u = _post_grad_u(L, dmu, dSu)
    dA = ein('nd,nm->dm', dm, kfu)
    dKi_m = ein('da,db->ab', dA, L.mu)
    KuuinvS = ein('ab,dbc->dac', L.Kuuinv, L.Su)
    dB = ein('nd,na,nb->dab', dv, kfu, kfu)
    dKi_v = 2 * ein('dab,dac->bc', KuuinvS, dB) - np.sum(dB, axis=0)
    dKi_KL = -0.5 * L.Dout * L.Kuu + 0.5 * np.sum(L.Splusmm, axis=0)
    dsf, dls, dzu = _kuu_tail(L, dKi_m + dKi_v + dKi_u + dKi_KL, dsf, dls, dzu)
    grad = {'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': deta2}
    return (grad, out[3]) if grad_x else grad


def bwd_mm(L, m, dm, dv, psi1, psi2, mx, vx):
    """SGP_Layer.backprop_grads_lvm_mm (vfe_models.py:326-404)."""
    ell = np.exp(L.ls)
    sf2 = np.exp(2 * L.sf)
    dm_all = dm - 2 * dv * m
    dpsi1 = ein('nd,dm->nm', dm_all, L.A)
    dpsi2 = ein('nd,dab->nab', dv, L.B_sto)
    dsf2, dls, dzu, dmx, dvx = psi_grads(dpsi1, psi1, dpsi2, psi2, ell, sf2, mx, vx, L.zu)
    dls = dls * ell
    dsf = 2 * sf2 * (dsf2 + np.sum(dv))
    dA = ein('nd,nm->dm', dm_all, psi1)
    dB = ein('nd,nab->dab', dv, psi2)
    dSu = ein('ab,dbc,ce->dae', L.Kuuinv, dB, L.Kuuinv)
    dmu = 2 * ein('dab,db->da', dSu, L.mu) + ein('ab,db->da', L.Kuuinv, dA)
    dmu = dmu + ein('ab,db->da', L.Kuuinv, L.mu)
    dSu = dSu + 0.5 * (L.Kuuinv - L.Suinv)
    deta1, deta2, dKi_u = _post_grad_u(L, dmu, dSu)
    dKi_KL = -0.5 * L.Dout * L.Kuu + 0.5 * np.sum(L.Splusmm, axis=0)
    dKi_A = ein('da,db->ab', dA, L.mu)
    KuuinvSmmd = ein('ab,dbc->dac', L.Kuuinv, L.Splusmm)
    dKi_B = 2 * ein('dab,dac->bc', KuuinvSmmd, dB) - np.sum(dB, axis=0)
    dsf, dls, dzu = _kuu_tail(L, dKi_A + dKi_B + dKi_u + dKi_KL, dsf, dls, dzu)
    return {'sf': dsf, 'ls': dls, 'zu': dzu, 'eta1_R': deta1, 'eta2': deta2}, {'mx': dmx, 'vx': dvx}


# ----------------------------------------------------------------------------
# expected log-likelihoods (lik_layers.py:183-236, 412-455, 629-676)
# ----------------------------------------------------------------------------

def gauss_lik_exp(sn, m, v, y):
    """Gauss_Layer.compute_log_lik_exp (lik_layers.py:183-215): (E log N(y|f, sn2), d/dm,
    d/dv), the 3-D branch averaging over Monte Carlo samples."""
    sn2 = np.exp(2.0 * sn)
    if m.ndim == 2:
        t1 = -0.5 * np.log(2 * np.pi * sn2)
        t2 = -0.5 / sn2 * (y ** 2 - 2 * y * m + m ** 2 + v)
        return float(np.sum(t1 + t2)), 1.0 / sn2 * (y - m), -0.5 / sn2 * np.ones_like(v)
    t = -0.5 * np.log(2 * np.pi * sn2) - 0.5 * ((y - m) ** 2 + v) / sn2
    S = m.shape[0]
    return float(np.sum(np.mean(t, axis=0))), (y - m) / sn2 / S, -0.5 / sn2 * np.ones_like(v) / S


def gauss_lik_exp_dsn(sn, m, v, y, scale):
    """Gauss_Layer.backprop_grads_log_lik_exp (lik_layers.py:217-236)."""
    sn2 = np.exp(2.0 * sn)
    if m.ndim == 2:
        return scale * np.sum(-1 + 1 / sn2 * (y ** 2 - 2 * y * m + m ** 2 + v))
    return scale * np.sum(-1 + ((y - m) ** 2 + v) / sn2) / m.shape[0]


def probit_lik_exp(m, v, y):
    """Probit_Layer.compute_log_lik_exp (lik_layers.py:412-455): Gauss-Hermite quadrature
    of E log Phi(y f) of degree GH_DEGREE."""
    from scipy.stats import norm
    gx, gw = np.polynomial.hermite.hermgauss(GH_DEGREE)
    shape = (-1,) + (1,) * m.ndim
    gx, gw = gx.reshape(shape), gw.reshape(shape) / np.sqrt(np.pi)
    ts = gx * np.sqrt(2 * v[None]) + m[None]
    ll = np.sum(gw * norm.logcdf(ts * y))
    g = y * gw * norm.pdf(ts * y) / norm.cdf(ts * y)
    dm = np.sum(g, axis=0)
    dv = np.sum(g * 0.5 * gx * np.sqrt(2 / v[None]), axis=0)
    if m.ndim == 3:
        S = m.shape[0]
        return float(ll / S), dm / S, dv / S
    return float(ll), dm, dv


def emission_lik_exp(C, Rpar, y, mx, vx, scale):
    """Gauss_Emis.compute_emission_log_lik_exp (lik_layers.py:629-676): E log N(y | C x, R)
    under x ~ N(mx, diag vx); returns (logZ*scale, {'mx','vx'}, {'C','R'})."""
    R = np.exp(2 * Rpar)
    Nb, Dout = y.shape
    Din = C.shape[1]
    Cmx = ein('ab,nb->na', C, mx)
    CRC = np.dot(C.T, np.dot(np.diag(1 / R), C))
    logZ = -0.5 * Nb * Dout * np.log(2 * np.pi) - 0.5 * Nb * np.sum(np.log(R)) \
        - 0.5 * np.sum(np.sum((y - Cmx) ** 2, axis=0) / R) \
        - 0.5 * np.sum(np.sum(vx, axis=0) * np.diag(CRC))
    dR = -0.5 * Nb / R + 0.5 * np.sum((y - Cmx) ** 2, axis=0) / R ** 2 \
        + 0.5 * np.diag(np.dot(C, np.dot(np.diag(np.sum(vx, axis=0)), C.T))) / R ** 2
    dR = dR * 2 * R
    dC = np.dot(np.diag(1 / R), ein('na,nb->ab', y - Cmx, mx)) \
        - np.dot(np.diag(1 / R), np.dot(C, np.diag(np.sum(vx, axis=0))))
    dmx = ein('ba,na->nb', np.dot(C.T, np.diag(1 / R)), y - Cmx)
    dvx = np.tile(np.reshape(-0.5 * np.diag(CRC), [1, Din]), [Nb, 1])
    return logZ * scale, {'mx': dmx * scale, 'vx': dvx * scale}, {'C': dC * scale, 'R': dR * scale}


def transition_lik_exp(sn, m_prop, v_prop, m_t, v_t, scale):
    """SGPSSM.compute_transition_log_lik_exp (vfe_models.py:1085-1120): E log N(x_t | f, sn2)
    for propagations (B, D) or Monte Carlo (S, B, D); (logZ, dmprop, dvprop, dmt, dvt, dsn)."""
    sn2 = np.exp(2 * sn)
    t1 = -0.5 * np.log(2 * np.pi * sn2)
    t2 = -0.5 / sn2 * (m_t ** 2 + v_t - 2 * m_t * m_prop + m_prop ** 2 + v_prop)
    if m_prop.ndim == 2:
        logZ = scale * np.sum(t1 + t2)
        dmt = -scale / sn2 * (m_t - m_prop)
        dvt = -scale * 0.5 / sn2 * np.ones_like(v_t)
        return logZ, -dmt, dvt, dmt, dvt, scale * np.sum(-1 - 2 * t2)
    K = m_prop.shape[0]
    logZ = scale * np.sum(t1 + t2) / K
    dmprop = scale / sn2 * (m_t - m_prop) / K
    dvprop = -scale * 0.5 / sn2 * np.ones_like(v_prop) / K
    return logZ, dmprop, dvprop, -np.sum(dmprop, axis=0), np.sum(dvprop, axis=0), \
        scale * np.sum(-1 - 2 * t2) / K


def _lik_exp(lik, params, m, v, y, scale, sfx=''):
    if lik.lower() == 'probit':
        ll, dm, dv = probit_lik_exp(m, v, y)
        return ll, dm, dv, {}
    ll, dm, dv = gauss_lik_exp(params['sn' + sfx], m, v, y)
    dsn = gauss_lik_exp_dsn(params['sn' + sfx], m, v, y, scale)
    return ll, dm, dv, {'sn' + sfx: np.reshape(dsn, np.shape(params['sn' + sfx]))}


# ----------------------------------------------------------------------------
# models (vfe_models.py:551-1120)
# ----------------------------------------------------------------------------

def sgpr_objective(x, y, params, M, mb_size, nat_param=True, fixed=(), idxs=None, lik='Gaussian'):
    """SGPR.objective_function (vfe_models.py:573-636): minus the uncollapsed bound,
    divided by N."""
    N = y.shape[0]
    if idxs is None and mb_size < N:
        idxs = np.random.choice(N, mb_size, replace=False)
    xb, yb = (x, y) if idxs is None else (x[idxs], y[idxs])
    scale = -N * 1.0 / yb.shape[0]
    L = layer_update(params, '', M, N, nat_param)
    m, v, kfu = fwd_det(L, xb)
    ll, dm, dv, glik = _lik_exp(lik, params, m, v, yb, scale)
    grads = bwd_det(L, scale * dm, scale * dv, kfu, xb)
    grads.update(glik)
    return _finish(scale * ll + kl_u(L), grads, fixed, N)


def _lvm_post_x(params, prior_mean, prior_var, nat_param):
    f1 = np.asarray(params['x1'])
    f2 = np.exp(2 * np.asarray(params['x2']))
    if nat_param:
        return prior_mean / prior_var + f1, 1.0 / prior_var + f2, f2
    return f1 / f2, 1.0 / f2, f2


def sgplvm_objective(y, params, M, mb_size, prior_mean=0.0, prior_var=1.0, nat_param=True,
                     fixed=(), idxs=None, prop_mode=PROP_MM, eps=None, lik='Gaussian'):
    """SGPLVM.objective_function (vfe_models.py:734-845) with compute_KL_x (:847-855) and
    Base_SGPLVM.compute_posterior_grad_x (base_models.py:913-929)."""
    N = y.shape[0]
    if idxs is None:
        idxs = np.arange(N) if mb_size == N else np.random.choice(N, mb_size, replace=False)
    yb = y[idxs]
    B = yb.shape[0]
    Din = params['x1'].shape[1]
    scale = -N * 1.0 / B
    L = layer_update(params, '', M, N, nat_param)
    p1, p2, f2 = _lvm_post_x(params, prior_mean, prior_var, nat_param)
    mx = p1[idxs] / p2[idxs]
    vx = 1.0 / p2[idxs]
    if prop_mode == PROP_MC:
        if eps is None:
            eps = np.random.randn(MC_NO_SAMPLES, B, Din)
        m3, v3, kfu, xs = fwd_mc(L, mx, vx, eps)
        ll, dm, dv, glik = _lik_exp(lik, params, m3, v3, yb, scale)
        grads, dx = bwd_det(L, scale * dm, scale * dv, kfu, xs, grad_x=True)
        dmx, dvx = O.reparam_bwd(dx, vx, eps)
    else:
        m, v, psi1, psi2 = fwd_mm(L, mx, vx)
        ll, dm, dv, glik = _lik_exp(lik, params, m, v, yb, scale)
        grads, gin = bwd_mm(L, m, scale * dm, scale * dv, psi1, psi2, mx, vx)
        dmx, dvx = gin['mx'], gin['vx']
    grads.update(glik)
    m0, v0 = prior_mean, prior_var
    kl_x = np.sum(0.5 * (np.log(v0) - np.log(vx) + (vx + (mx - m0) ** 2) / v0 - 1))
    scale_x = N * 1.0 / B
    dmx = dmx + scale_x * (mx - m0) / v0
    dvx = dvx + scale_x * (-0.5 / vx + 0.5 / v0)
    g1 = np.zeros_like(p1)
    g2 = np.zeros_like(p1)
    if nat_param:
        g1[idxs] = dmx / p2[idxs]
        g2[idxs] = (-dmx * p1[idxs] / p2[idxs] ** 2 - dvx / p2[idxs] ** 2) * 2 * f2[idxs]
    else:
        g1[idxs] = dmx
        g2[idxs] = dvx * 2 * f2[idxs]
    grads['x1'] = g1
    grads['x2'] = g2
    return _finish(scale * ll + scale_x * kl_x + kl_u(L), grads, fixed, N)


def sgpssm_objective(y, params, M, Din, mb_size, prior_mean=0.0, prior_var=1.0, x_control=None,
                     gp_emi=False, control_to_emi=True, fixed=(), start_idx=None, prop_mode=PROP_MM,
                     eps=None, lik='Gaussian'):
    """SGPSSM.objective_function (vfe_models.py:893-1083) with the tied posterior of
    Base_SGPSSM.update_hypers (base_models.py:1705-1728) and compute_posterior_grad_x
    (:1730-1752); the x prior term is left out as in the reference (:1045)."""
    N, Dout = y.shape
    Dcon = 0 if x_control is None else x_control.shape[1]
    Dcon_emi = Dcon if control_to_emi else 0
    if mb_size >= N:
        dyn_idxs, emi_idxs = np.arange(0, N - 1), np.arange(0, N)
    else:
        if start_idx is None:
            start_idx = np.random.randint(0, N - mb_size)
        emi_idxs = np.arange(start_idx, start_idx + mb_size)
        dyn_idxs = np.arange(start_idx, start_idx + mb_size - 1)
    yb = y[emi_idxs]
    Bd, Be = dyn_idxs.shape[0], emi_idxs.shape[0]
    sc_dyn = -(N - 1) * 1.0 / Bd
    sc_emi = -N * 1.0 / Be
    dyn = layer_update(params, '_dynamic', M, N - 1, True)
    emi = layer_update(params, '_emission', M, N, True) if gp_emi else None
    xf1 = np.asarray(params['x_factor_1'])
    xf2 = np.exp(2 * np.asarray(params['x_factor_2']))
    p1, p2 = 3 * xf1, 3 * xf2
    p1[[0, -1], :] = 2 * xf1[[0, -1], :]
    p2[[0, -1], :] = 2 * xf2[[0, -1], :]
    p1[0, :] += prior_mean / prior_var
    p2[0, :] += 1.0 / prior_var
    pm, pv = p1[emi_idxs] / p2[emi_idxs], 1.0 / p2[emi_idxs]
    t_m, t_v = pm[1:], pv[1:]
    tm1_m, tm1_v = pm[:-1], pv[:-1]
    if Dcon > 0:
        tm1_m = np.hstack((tm1_m, x_control[dyn_idxs]))
        tm1_v = np.hstack((tm1_v, np.zeros((Bd, Dcon))))
    up_m, up_v = pm, pv
    if Dcon_emi > 0:
        up_m = np.hstack((up_m, x_control[emi_idxs]))
        up_v = np.hstack((up_v, np.zeros((Be, Dcon_emi))))
    sn = params['sn']
    mc = prop_mode == PROP_MC
    if mc:
        e_dyn = np.random.randn(MC_NO_SAMPLES, Bd, Din + Dcon) if eps is None else eps[0]
        m3, v3, kfu, xs = fwd_mc(dyn, tm1_m, tm1_v, e_dyn)
        lz_dyn, dmp, dvp, dmt, dvt, dsn = transition_lik_exp(sn, m3, v3, t_m, t_v, sc_dyn)
        g_dyn, dx = bwd_det(dyn, dmp, dvp, kfu, xs, grad_x=True)
        gm, gv = O.reparam_bwd(dx, tm1_v, e_dyn)
        gin_dyn = {'mx': gm, 'vx': gv}
    else:
        mp, vp, a1, a2 = fwd_mm(dyn, tm1_m, tm1_v)
        lz_dyn, dmp, dvp, dmt, dvt, dsn = transition_lik_exp(sn, mp, vp, t_m, t_v, sc_dyn)
        g_dyn, gin_dyn = bwd_mm(dyn, mp, dmp, dvp, a1, a2, tm1_m, tm1_v)
    grads = {'sn': np.reshape(dsn, np.shape(sn))}
    for k, val in g_dyn.items():
        grads[k + '_dynamic'] = val
    if gp_emi:
        if mc:
            e_emi = np.random.randn(MC_NO_SAMPLES, Be, Din + Dcon_emi) if eps is None else eps[1]
            m3, v3, kfu, xs = fwd_mc(emi, up_m, up_v, e_emi)
            ll, dm, dv, glik = _lik_exp(lik, params, m3, v3, yb, sc_emi, '_emission')
            g_emi, dx = bwd_det(emi, sc_emi * dm, sc_emi * dv, kfu, xs, grad_x=True)
            gm, gv = O.reparam_bwd(dx, up_v, e_emi)
            gin_emi = {'mx': gm, 'vx': gv}
        else:
            m, v, a1, a2 = fwd_mm(emi, up_m, up_v)
            ll, dm, dv, glik = _lik_exp(lik, params, m, v, yb, sc_emi, '_emission')
            g_emi, gin_emi = bwd_mm(emi, m, sc_emi * dm, sc_emi * dv, a1, a2, up_m, up_v)
        lz_emi = sc_emi * ll
        grads.update(glik)
    else:
        lz_emi, gin_emi, g_emi = emission_lik_exp(params['C_emission'], params['R_emission'], yb,
                                                  up_m, up_v, sc_emi)
    for k, val in g_emi.items():
        grads[k + '_emission'] = val
    sc_ent = -N * 1.0 / Be
    x_ent = sc_ent * (Be * Din * (0.5 + 0.5 * np.log(2 * np.pi)) + np.sum(0.5 * np.log(pv)))
    dm = gin_emi['mx'][:, :Din].copy()
    dv = gin_emi['vx'][:, :Din] + sc_ent * 0.5 / pv
    dm[1:] += dmt
    dm[:-1] += gin_dyn['mx'][:, :Din]
    dv[1:] += dvt
    dv[:-1] += gin_dyn['vx'][:, :Din]
    g1 = np.zeros_like(xf1)
    g2 = np.zeros_like(xf1)
    q1, q2 = p1[emi_idxs], p2[emi_idxs]
    sx = 3.0 * np.ones((Be, 1))
    sx[np.where(emi_idxs == 0)[0]] = 2
    sx[np.where(emi_idxs == N - 1)[0]] = 2
    g1[emi_idxs] = dm / q2 * sx
    g2[emi_idxs] = (-dm * q1 / q2 ** 2 - dv / q2 ** 2) * sx * 2 * xf2[emi_idxs]
    grads['x_factor_1'] = g1
    grads['x_factor_2'] = g2
    energy = lz_dyn + lz_emi + x_ent + kl_u(dyn) + (kl_u(emi) if gp_emi else 0.0)
    return _finish(energy, grads, fixed, N)


def sgpr_collapsed_objective(x, y, params, M, alpha=1.0):
    """SGPR_collapsed.objective_function (vfe_models.py:39-131): the collapsed power-EP
    bound of the sparse GP (the Titsias bound as alpha -> 0, FITC at alpha = 1) and its
    analytic gradients, divided by N."""
    N, Dout = y.shape
    if Dout != 1:
        # the reference's gradient broadcasts a (Dout, N) array against (M, N) (:101-103)
        raise ValueError('the collapsed SGPR gradient is defined for one output')
    ls, sf, zu, sn = (np.asarray(params[k], dtype=np.float64) for k in ('ls', 'sf', 'zu', 'sn'))
    sf2, sn2 = np.exp(2 * sf), np.exp(2 * sn)
    Kuu_noiseless = se_kernel(ls, sf, zu, zu)
    Kuu = Kuu_noiseless + np.diag(JITTER * np.ones(M))
    Lu = np.linalg.cholesky(Kuu)
    Kuf = se_kernel(ls, sf, zu, x)
    V = np.linalg.solve(Lu, Kuf)
    r = sf2 - np.sum(V * V, axis=0)
    G = alpha * r + sn2
    iG = 1.0 / G
    A = np.eye(M) + np.dot(V * iG, V.T)
    La = np.linalg.cholesky(A)
    B = np.linalg.solve(La, V)
    iGy = iG[:, None] * y
    z = iGy - np.dot(B.T, np.dot(B, iGy)) * iG[:, None]
    energy = 0.5 * np.sum(z * y) + Dout * np.sum(np.log(np.diag(La))) + 0.5 * Dout * np.sum(np.log(G)) \
        + 0.5 * N * Dout * np.log(2 * np.pi) \
        + 0.5 * Dout * (1 - alpha) / alpha * np.sum(np.log(1 + alpha * r / sn2))
    zus = zu / np.exp(ls)
    xs = x / np.exp(ls)
    R = np.linalg.solve(Lu.T, V)
    RiG = R * iG
    RdQ = -np.dot(np.dot(R, z), z.T) + RiG - np.dot(np.dot(RiG, B.T), B) * iG
    dG = z ** 2 + (-iG + iG ** 2 * np.sum(B * B, axis=0))[:, None]
    tmp = alpha * dG - (1 - alpha) / (1 + alpha * r[:, None] / sn2) / sn2
    RdQ2 = RdQ + R * tmp.T
    KW = Kuf * RdQ2
    KWR = Kuu_noiseless * (np.dot(RdQ2, R.T))
    P = (np.dot(KW, xs) - np.dot(KWR, zus) + (np.sum(KWR, axis=1) - np.sum(KW, axis=1))[:, None] * zus)
    dzu = P / np.exp(ls)
    dls = (-np.sum(P * zus, axis=0)
           - np.sum((np.dot(KW.T, zus) - np.sum(KW, axis=0)[:, None] * xs) * xs, axis=0))
    dsn = -np.sum(dG) * sn2 - (1 - alpha) * np.sum(r / (1 + alpha * r / sn2)) / sn2
    dsf = (np.sum(Kuf * RdQ) - alpha * np.sum(r[:, None] * dG)
           + (1 - alpha) * np.sum(r / (1 + alpha * r / sn2)) / sn2)
    grads = {'zu': dzu, 'ls': dls, 'sn': np.reshape(dsn, np.shape(params['sn'])),
             'sf': np.reshape(dsf, np.shape(params['sf']))}
    return _finish(energy, grads, (), N)


def sgpr_collapsed_predict(x, y, params, M, xs, alpha=0.01, marginal=True, noisy=False):
    """SGPR_collapsed.predict_f / predict_y (vfe_models.py:133-185, 198-216)."""
    ls, sf, zu, sn = (np.asarray(params[k], dtype=np.float64) for k in ('ls', 'sf', 'zu', 'sn'))
    sf2, sn2 = np.exp(2 * sf), np.exp(2 * sn)
    Kuu = se_kernel(ls, sf, zu, zu) + np.diag(JITTER * np.ones(M))
    Lu = np.linalg.cholesky(Kuu)
    Kuf = se_kernel(ls, sf, zu, x)
    Kut = se_kernel(ls, sf, zu, xs)
    V = np.linalg.solve(Lu, Kuf)
    r = sf2 - np.sum(V * V, axis=0)
    iG = 1.0 / (alpha * r + sn2)
    A = np.eye(M) + np.dot(V * iG, V.T)
    La = np.linalg.cholesky(A)
    B = np.linalg.solve(La, V)
    yhatB = np.dot((y * iG[:, None]).T, B.T)
    beta = np.linalg.solve(Lu.T, np.linalg.solve(La.T, yhatB.T))
    W1 = np.eye(M) - np.linalg.solve(La.T, np.linalg.solve(La, np.eye(M))).T
    W = np.linalg.solve(Lu.T, np.linalg.solve(Lu.T, W1).T)
    KtuW = np.dot(Kut.T, W)
    mf = np.dot(Kut.T, beta)
    if marginal:
        vf = sf2 - np.sum(KtuW * Kut.T, axis=1)
        return mf, (vf + sn2 if noisy else vf)
    Ktt = se_kernel(ls, sf, xs, xs) + np.diag(JITTER * np.ones(xs.shape[0]))
    vf = Ktt - np.dot(KtuW, Kut)
    return mf, (vf + sn2 * np.eye(xs.shape[0]) if noisy else vf)
