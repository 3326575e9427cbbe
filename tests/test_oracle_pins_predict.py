"""Pins for the oracle's prediction functions and the GPSSM x-factor energy terms.

Each check is fixed by mathematics that the oracle's code does not restate, so that a
swapped B_det / B_sto, a dropped noise term, a wrong tie count or a wrong constant
fails it:
  * the deterministic-input posterior predictive (base_models.py:265-284) with the
    inducing inputs at the training inputs and q(u) set to the exact posterior equals
    the exact GP predictive of Rasmussen & Williams (2006) eqs. 2.25-2.26, computed
    here by a Cholesky solve; predict_y adds the noise variance;
  * the moment-matched posterior predictive (base_models.py:286-307) equals, by the
    laws of total expectation and variance, a Gauss-Hermite quadrature of the
    deterministic predictive over x ~ N(mx, diag vx);
  * the linear-Gaussian emission predictive (lik_layers.py:558-571) equals a
    quadrature of y = C x + eps over x;
  * the GPSSM energy (aep_models.py:1024-1206, x-factor terms :1389-1437, tied
    factors base_models.py:1705-1728) tends, as alpha -> 0, to minus the
    variational free energy of the same q(x), q(u) -- an expectation evaluated here
    by quadrature of the GP conditional at each node -- plus (T-1) Din log(2 pi)/2:
    the x-factor log-partitions of the reference carry no log(2 pi) terms, and the
    T posterior terms, the 3 or 2 cavity terms per time step and the one prior term
    leave T - 1 of them uncancelled.  A wrong tie count, scale or sign changes the
    limit by a parameter-dependent amount.
"""
import numpy as np
import pytest
from numpy.polynomial.hermite import hermgauss
from scipy.linalg import cho_factor, cho_solve

from oracle import aep_oracle as O
from geepee_amd import workloads as W


def _exact_gp_problem(rs, N=9, Q=2, D=2):
    x = rs.uniform(-2, 2, (N, Q))
    y = rs.randn(N, D)
    ls = np.log(rs.uniform(0.8, 1.4, Q))
    sf = np.array([np.log(0.9)])
    sn2 = 0.15
    return x, y, ls, sf, sn2


def _exact_gp_params(x, y, ls, sf, sn2, nat_param):
    """q(u) = p(u | y) for u = f(x) with Z = x: precision Kuu^-1 + I/sn2 and
    Sigma^-1 mu = y/sn2."""
    N = x.shape[0]
    D = y.shape[1]
    Kuu = O.se_kernel(ls, sf, x, x) + O.JITTER * np.eye(N)
    p = {'ls': ls, 'sf': sf, 'zu': x.copy(), 'sn': np.array([0.5 * np.log(sn2)])}
    iu = np.triu_indices(N)
    if nat_param:
        R = np.eye(N) / np.sqrt(sn2)
        p['eta2'] = (y / sn2).T.copy()
    else:
        S = Kuu - Kuu.dot(np.linalg.solve(Kuu + sn2 * np.eye(N), Kuu))
        S = 0.5 * (S + S.T)
        R = np.linalg.cholesky(S).T
        p['eta2'] = Kuu.dot(np.linalg.solve(Kuu + sn2 * np.eye(N), y)).T.copy()
    R = R.copy()
    R[np.diag_indices(N)] = np.log(R[np.diag_indices(N)])
    p['eta1_R'] = np.tile(R[iu], (D, 1))
    return p, Kuu


def _rw_predict(Kuu, kxs, kss, y, sn2):
    """Rasmussen & Williams eq. 2.25-2.26 with K = Kuu (jitter as part of the prior)."""
    c = cho_factor(Kuu + sn2 * np.eye(Kuu.shape[0]), lower=True)
    mean = kxs.dot(cho_solve(c, y))
    var = kss - np.sum(kxs * cho_solve(c, kxs.T).T, axis=1)
    return mean, np.repeat(var[:, None], y.shape[1], axis=1)


@pytest.mark.parametrize('nat', [True, False])
def test_det_post_predict_equals_exact_gp(nat):
    rs = np.random.RandomState(30 + nat)
    x, y, ls, sf, sn2 = _exact_gp_problem(rs)
    p, Kuu = _exact_gp_params(x, y, ls, sf, sn2, nat)
    xs = rs.uniform(-2.5, 2.5, (13, 2))
    kxs = O.se_kernel(ls, sf, xs, x)
    mean, var = _rw_predict(Kuu, kxs, np.exp(2 * sf[0]) * np.ones(13), y, sn2)
    mf, vf = O.sgpr_predict(p, x.shape[0], x.shape[0], xs, nat_param=nat)
    np.testing.assert_allclose(mf, mean, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(vf, var, rtol=1e-8, atol=1e-11)
    my, vy = O.sgpr_predict(p, x.shape[0], x.shape[0], xs, nat_param=nat, noisy=True)
    np.testing.assert_allclose(my, mean, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(vy, var + sn2, rtol=1e-9)
    if nat:
        # the SGPLVM predictive is the same posterior at deterministic inputs
        mf2, vf2 = O.sgplvm_predict(p, x.shape[0], x.shape[0], xs, noisy=True)
        np.testing.assert_allclose(mf2, mean, rtol=1e-9, atol=1e-11)
        np.testing.assert_allclose(vf2, var + sn2, rtol=1e-9)


def _gh_nodes(mx, vx, deg):
    """Tensor Gauss-Hermite nodes and weights for x ~ N(mx, diag vx)."""
    t, w = hermgauss(deg)
    Q = mx.shape[0]
    grids = np.meshgrid(*[mx[q] + np.sqrt(2 * vx[q]) * t for q in range(Q)], indexing='ij')
    pts = np.stack([g.ravel() for g in grids], axis=1)
    wts = np.prod(np.meshgrid(*([w] * Q), indexing='ij'), axis=0).ravel() / np.pi ** (Q / 2)
    return pts, wts


@pytest.mark.parametrize('Q', [1, 2])
def test_mm_post_predict_is_total_variance_of_det_post(Q):
    rs = np.random.RandomState(40 + Q)
    M, D, N = 6, 3, 50
    p = W.layer_params(rs, Q, D, M)
    p['ls'] = np.log(rs.uniform(0.8, 1.5, Q))
    L = O.layer_update(p, '', M, N, True)
    mx = rs.randn(5, Q)
    vx = rs.uniform(0.05, 0.6, (5, Q))
    mf, vf = O.fwd_mm_post(L, mx, vx)
    for n in range(5):
        pts, wts = _gh_nodes(mx[n], vx[n], 60 if Q == 1 else 40)
        md, vd = O.fwd_det_post(L, pts)
        mean = wts.dot(md)
        var = wts.dot(vd) + wts.dot(md ** 2) - mean ** 2
        np.testing.assert_allclose(mf[n], mean, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(vf[n], var, rtol=1e-9, atol=1e-12)
    # at zero input variance the two predictives coincide
    mf0, vf0 = O.fwd_mm_post(L, mx, np.zeros_like(vx))
    md0, vd0 = O.fwd_det_post(L, mx)
    np.testing.assert_allclose(mf0, md0, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(vf0, vd0, rtol=1e-9, atol=1e-12)


def test_emission_predict_quadrature():
    rs = np.random.RandomState(50)
    Din, Dout = 2, 3
    C = rs.randn(Dout, Din)
    Rpar = np.log(rs.uniform(0.1, 0.5, Dout))
    mf = rs.randn(4, Din)
    vf = rs.uniform(0.1, 0.8, (4, Din))
    my, vy0, vy = O.emission_predict(C, Rpar, mf, vf)
    for n in range(4):
        pts, wts = _gh_nodes(mf[n], vf[n], 20)
        ys = pts.dot(C.T)
        mean = wts.dot(ys)
        cov = np.einsum('p,pa,pb->ab', wts, ys, ys) - np.outer(mean, mean)
        np.testing.assert_allclose(my[n], mean, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(vy0[n], cov, rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(vy[n], cov + np.diag(np.exp(2 * Rpar)), rtol=1e-11, atol=1e-13)


def _gpssm_free_energy(y, p, M, Din, prior_mean, prior_var, deg):
    """Variational free energy (minus the ELBO) of the GPSSM with q(x_t) = N(m_t, v_t)
    built from the tied factors (3 per interior step, 2 at the ends, the prior at t=0)
    and q(u_d) = N(mu_d, Sigma_d) from the dynamics layer's natural parameters."""
    T, Dout = y.shape
    dyn = O.layer_update(p, '_dynamic', M, T - 1, True)
    n = np.full((T, 1), 3.0)
    n[0] = n[-1] = 2.0
    xf1, xf2 = p['x_factor_1'], np.exp(2 * p['x_factor_2'])
    t1, t2 = n * xf1, n * xf2
    t1[0] += prior_mean / prior_var
    t2[0] += 1.0 / prior_var
    v = 1.0 / t2
    m = t1 / t2
    C, R = p['C_emission'], np.exp(2 * p['R_emission'])
    sn2 = np.exp(2 * p['sn'][0])
    elbo = 0.0
    for t in range(T):
        mu_y, var_y = C.dot(m[t]), (C ** 2).dot(v[t])
        elbo += np.sum(-0.5 * np.log(2 * np.pi * R) - 0.5 * ((y[t] - mu_y) ** 2 + var_y) / R)
    Ki = np.linalg.inv(dyn.Kuu)
    kss = np.exp(2 * dyn.sf[0])
    for t in range(1, T):
        pts, wts = _gh_nodes(m[t - 1], v[t - 1], deg)
        k = O.se_kernel(dyn.ls, dyn.sf, pts, dyn.zu)
        kKi = k.dot(Ki)
        for d in range(Din):
            Ef = wts.dot(kKi.dot(dyn.mu[d]))
            Ef2 = wts.dot(kss - np.sum(kKi * k, 1) + np.sum(kKi.dot(dyn.Splusmm[d]) * kKi, 1))
            Ed2 = m[t, d] ** 2 + v[t, d] - 2 * m[t, d] * Ef + Ef2
            elbo += -0.5 * np.log(2 * np.pi * sn2) - 0.5 * Ed2 / sn2
    ldK = np.linalg.slogdet(dyn.Kuu)[1]
    for d in range(Din):
        S, mu = dyn.Su[d], dyn.mu[d]
        elbo -= 0.5 * (np.trace(Ki.dot(S)) + mu.dot(Ki).dot(mu) - M + ldK - np.linalg.slogdet(S)[1])
    elbo -= np.sum(0.5 * (v[0] / prior_var + (m[0] - prior_mean) ** 2 / prior_var - 1
                          + np.log(prior_var / v[0])))
    elbo += np.sum(0.5 * np.log(2 * np.pi * np.e * v[1:]))
    return -elbo


@pytest.mark.parametrize('T,Din', [(3, 1), (4, 1), (6, 1), (4, 2)])
def test_sgpssm_alpha_to_zero_matches_free_energy(T, Din):
    rs = np.random.RandomState(60 + T + 10 * Din)
    M, Dout = 4, 2
    y, p = W.sgpssm_problem(T=T, Din=Din, Dout=Dout, M=M, seed=T + Din)
    # a lengthscale the quadrature resolves, inducing inputs spread over the path
    p['ls_dynamic'] = np.log(0.9) * np.ones(Din)
    p['zu_dynamic'] = rs.randn(M, Din)
    p['x_factor_1'] = rs.randn(T, Din)
    p['x_factor_2'] = np.log(rs.uniform(0.5, 2, (T, Din))) / 2
    p['sn'] = np.array([np.log(rs.uniform(0.2, 0.5))])
    pm, pv = 0.3, 1.7
    F = _gpssm_free_energy(y, p, M, Din, pm, pv, 80 if Din == 1 else 40)

    def e(a):   # the oracle divides by N = T
        return T * O.sgpssm_objective(y, p, M, Din, T, alpha=a, prior_mean=pm, prior_var=pv)[0]
    # Richardson extrapolation over alpha, 2 alpha, 4 alpha removes the O(alpha) and
    # O(alpha^2) terms; a smaller alpha loses more to the cancellation of the N/alpha
    # scaled log-partitions than it gains.  A wrong constant moves the limit by at least
    # log(2 pi)/2 = 0.92 and a wrong tie or scale by a parameter-dependent O(0.1-1).
    a = 1e-5
    limit = (8 * e(a) - 6 * e(2 * a) + e(4 * a)) / 3
    const = (T - 1) * Din * 0.5 * np.log(2 * np.pi)
    assert abs(limit - (F + const)) < 1e-6 * abs(F), (limit, F + const)
    # the limit is not reached at a finite alpha: the check sees the alpha terms
    assert abs(e(0.5) - (F + const)) > 1e-3 * abs(F)
