"""Pins for the oracle's Monte Carlo propagation and probit likelihood.

  * The 3-D (Monte Carlo) branches of the Gaussian and probit tilted normalisers and of
    the GPSSM transition (lik_layers.py:139-152, 360-404; aep_models.py:1344-1372) are a
    log-mean-exp over samples: with S identical samples they must equal the 2-D branch
    exactly (the per-sample gradients summing to the 2-D gradient).
  * The probit normaliser at alpha = 1 is the closed-form Gaussian-probit integral
    Phi(y m / sqrt(1 + v)); for alpha != 1 the reference's 10-point Gauss-Hermite rule
    must agree with adaptive quadrature of the integral of Phi(y f)^alpha N(f; m, v) to
    that rule's accuracy.
  * The reparameterisation adjoint (base_models.py:373-388) against finite differences
    of a function of x = eps sqrt(v) + m.
  * The Monte Carlo layer adjoint (aep_models.py:307-410) against the deterministic-input
    adjoint (aep_models.py:412-511): both are the gradient of the same energy at the
    stacked samples.
  * Whole objectives with Monte Carlo propagation (fixed draws, as the reference's
    tests/test_utils.py:58-60 re-seeds) and with the probit likelihood against finite
    differences.
"""
import numpy as np
import pytest
from scipy import integrate
from scipy.stats import norm

from oracle import aep_oracle as O
from geepee_amd import workloads as W
from tests.fdutil import check_grads


def test_mc_branches_reduce_to_2d_with_identical_samples():
    rs = np.random.RandomState(0)
    B, D, S = 7, 3, 5
    m, v, y = rs.randn(B, D), rs.uniform(0.1, 1.0, (B, D)), rs.randn(B, D)
    sn = np.array([np.log(0.4)])
    m3, v3 = np.repeat(m[None], S, 0), np.repeat(v[None], S, 0)
    for alpha in (1.0, 0.5):
        l2, dm2, dv2 = O.gauss_logZ(sn, m, v, y, alpha)
        l3, dm3, dv3 = O.gauss_logZ_mc(sn, m3, v3, y, alpha)
        assert abs(l3 - l2) <= 1e-12 * abs(l2)
        np.testing.assert_allclose(dm3.sum(0), dm2, rtol=1e-12)
        np.testing.assert_allclose(dv3.sum(0), dv2, rtol=1e-12)
    yb = np.sign(rs.randn(B, D))
    for alpha in (1.0, 0.5):
        l2, dm2, dv2 = O.probit_logZ(m, v, yb, alpha)
        l3, dm3, dv3 = O.probit_logZ(m3, v3, yb, alpha)
        # the 2-D quadrature branch adds 1e-8 to each probability and gradient, the 3-D
        # branch 1e-16 (lik_layers.py:338, 392)
        tol = 1e-12 if alpha == 1.0 else 2e-5
        assert abs(l3 - l2) <= tol * abs(l2)
        np.testing.assert_allclose(dm3.sum(0), dm2, rtol=tol, atol=tol)
        np.testing.assert_allclose(dv3.sum(0), dv2, rtol=tol, atol=tol)
    mt, vt = rs.randn(B, D), rs.uniform(0.1, 1.0, (B, D))
    two = O.transition_tilted(sn, m, v, mt, vt, 0.5, -3.0, D)
    three = O.transition_tilted_mc(sn, m3, v3, mt, vt, 0.5, -3.0, D)
    assert abs(three[0] - two[0]) <= 1e-12 * abs(two[0])
    np.testing.assert_allclose(three[1].sum(0), two[1], rtol=1e-12)   # dm_prop
    np.testing.assert_allclose(three[2].sum(0), two[2], rtol=1e-12)   # dv_prop
    np.testing.assert_allclose(three[3], two[3], rtol=1e-12)          # dm_t
    np.testing.assert_allclose(three[4], two[4], rtol=1e-12)          # dv_t
    assert abs(three[5] - two[5]) <= 1e-12 * abs(two[5])             # dsn


def test_probit_normaliser_against_quadrature():
    rs = np.random.RandomState(1)
    m, v = rs.randn(3, 2), rs.uniform(0.1, 0.8, (3, 2))
    y = np.array([[1.0, -1.0], [-1.0, 1.0], [1.0, 1.0]])
    for alpha, tol in ((1.0, 1e-10), (0.5, 2e-4), (0.2, 2e-4)):
        lz, _, _ = O.probit_logZ(m, v, y, alpha)
        ref = 0.0
        for i in range(3):
            for d in range(2):
                f = lambda t: norm.cdf(y[i, d] * t) ** alpha * norm.pdf(t, m[i, d], np.sqrt(v[i, d]))
                val, _ = integrate.quad(f, -np.inf, np.inf, epsabs=1e-14, epsrel=1e-13)
                ref += np.log(val)
        assert abs(lz - ref) <= tol * abs(ref), (alpha, lz, ref)


@pytest.mark.parametrize('alpha', [1.0, 0.5])
def test_probit_gradients_fd(alpha):
    rs = np.random.RandomState(2)
    m, v = rs.randn(4, 2), rs.uniform(0.2, 0.9, (4, 2))
    y = np.sign(rs.randn(4, 2))

    def f(p):
        lz, dm, dv = O.probit_logZ(p['m'], p['v'], y, alpha)
        return lz, {'m': dm, 'v': dv}
    # the reference's 2-D quadrature gradients carry a +1e-8 offset
    check_grads(f, {'m': m, 'v': v}, h=1e-3, rtol=1e-7 if alpha == 1.0 else 2e-7)


def test_reparam_adjoint_fd():
    rs = np.random.RandomState(3)
    S, B, Q = 4, 3, 2
    eps = rs.randn(S, B, Q)
    mx, vx = rs.randn(B, Q), rs.uniform(0.2, 1.0, (B, Q))
    A = rs.randn(Q, Q)

    def g(x):                       # a smooth function of the samples and its gradient
        return np.sum(np.sin(x.dot(A))), np.cos(x.dot(A)).dot(A.T)

    def f(p):
        x = eps * np.sqrt(p['vx']) + p['mx']
        val, dx = g(x)
        dm, dv = O.reparam_bwd(dx, p['vx'], eps)
        return val, {'mx': dm, 'vx': dv}
    check_grads(f, {'mx': mx, 'vx': vx}, h=1e-3, rtol=1e-9)


def test_mc_layer_adjoint_equals_deterministic_adjoint():
    """backprop_grads_lvm_mc and backprop_grads_reg differentiate the same energy of a
    layer evaluated at fixed inputs, so their hyper-parameter gradients agree."""
    rs = np.random.RandomState(4)
    Q, D, M, N = 2, 3, 6, 40
    p = W.layer_params(rs, Q, D, M)
    L = O.layer_update(p, '', M, N, True)
    O.layer_cavity(L, 0.5)
    x = rs.randn(15, Q)
    m, v, kfu = O.fwd_det_cav(L, x)
    dm, dv = rs.randn(15, D), rs.randn(15, D)
    g_det = O.bwd_det(L, dm, dv, kfu, x, 0.5)
    g_mc, dx = O.bwd_mc(L, dm, dv, kfu, x, 0.5)
    for k in g_det:
        np.testing.assert_allclose(g_mc[k], g_det[k], rtol=1e-9, atol=1e-9 * np.max(np.abs(g_det[k])))
    assert dx.shape == x.shape


def _fixed_eps(shape, seed):
    return np.random.RandomState(seed).randn(*shape)


def test_model_gradients_fd_sgplvm_mc():
    y, p = W.sgplvm_problem(N=12, Q=2, D=3, M=5, seed=0)
    eps = _fixed_eps((O.MC_NO_SAMPLES, 12, 2), 5)
    check_grads(lambda q: O.sgplvm_objective(y, q, 5, 12, alpha=0.5, prop_mode='MC', eps=eps), p)
    idx = np.random.RandomState(4).choice(12, 5, replace=False)
    eps = _fixed_eps((O.MC_NO_SAMPLES, 5, 2), 6)
    check_grads(lambda q: O.sgplvm_objective(y, q, 5, 5, alpha=0.5, idxs=idx, prop_mode='MC',
                                             eps=eps), p)


def test_model_gradients_fd_sgpssm_mc():
    y, p = W.sgpssm_problem(T=12, Din=2, Dout=3, M=5, seed=0)
    eps = (_fixed_eps((O.MC_NO_SAMPLES, 11, 2), 7), None)
    check_grads(lambda q: O.sgpssm_objective(y, q, 5, 2, 100, alpha=0.5, prop_mode='MC', eps=eps), p)
    rs = np.random.RandomState(12)
    T, Din, Dout, M, Dc = 10, 2, 2, 4, 1
    y = rs.randn(T, Dout)
    u = rs.randn(T, Dc)
    p = W.layer_params(rs, Din + Dc, Din, M, x=rs.randn(T, Din + Dc), sfx='_dynamic')
    p.update(W.layer_params(rs, Din + Dc, Dout, M, x=rs.randn(T, Din + Dc), sfx='_emission'))
    p['x_factor_1'] = rs.randn(T, Din)
    p['x_factor_2'] = np.log(rs.uniform(2, 4, (T, Din))) / 2
    p['sn'] = np.array([np.log(0.3)])
    p['sn_emission'] = np.array([np.log(0.2)])
    eps = (_fixed_eps((O.MC_NO_SAMPLES, T - 1, Din + Dc), 8), _fixed_eps((O.MC_NO_SAMPLES, T, Din + Dc), 9))
    check_grads(lambda q: O.sgpssm_objective(y, q, M, Din, 100, alpha=0.5, x_control=u, gp_emi=True,
                                             prop_mode='MC', eps=eps), p, rtol=1e-5)


def test_model_gradients_fd_probit():
    rs = np.random.RandomState(5)
    x = rs.randn(20, 2)
    y = np.sign(np.sin(2 * x[:, :1]) + 0.3 * rs.randn(20, 1))
    p = W.layer_params(rs, 2, 1, 6, x=x)
    for alpha in (1.0, 0.5):
        check_grads(lambda q: O.sgpr_objective(x, y, q, 6, 20, alpha=alpha, lik='probit'), p, rtol=5e-6)
    xd, _, pd, Ms, _ = W.sdgpr_problem(N=15, widths=(3, 2, 1), M=5, seed=0)
    pd.pop('sn')
    yd = np.sign(rs.randn(15, 1))
    check_grads(lambda q: O.sdgpr_objective(xd, yd, q, Ms, 15, alpha=0.5, lik='probit'), pd, rtol=2e-5)


def test_model_gradients_fd_sgplvm_probit():
    y, p = W.sgplvm_problem(N=10, Q=2, D=3, M=5, seed=2)
    yb = np.sign(y)
    p.pop('sn')
    check_grads(lambda q: O.sgplvm_objective(yb, q, 5, 10, alpha=0.5, lik='probit'), p, rtol=1e-5)
    eps = _fixed_eps((O.MC_NO_SAMPLES, 10, 2), 11)
    check_grads(lambda q: O.sgplvm_objective(yb, q, 5, 10, alpha=0.5, lik='probit', prop_mode='MC',
                                             eps=eps), p, rtol=1e-5)
