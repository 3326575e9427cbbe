"""Pins for the oracle's SDGPR_H objective and GPSSM rollouts.

  * SDGPR_H (aep_models.py:1440-1863): every gradient against finite differences of
    the energy; its transition normaliser (:1702-1735) written with log(2 pi v_sum) equals
    the GPSSM transition normaliser (:1317-1343, pinned by quadrature in
    test_oracle_pins.py) for the same arguments; the hidden-factor log-partition
    0.5 (t1^2 / t2 - log t2) against quadrature of exp(t1 h - t2 h^2 / 2).
  * Rollouts (base_models.py:1463-1542): the moment-matched rollout is the iteration of
    the moment-matched posterior predictive (pinned by total-variance quadrature in
    test_oracle_pins_predict.py) from the last latent state, and the Monte Carlo rollout
    at one sample and zero variance steps is the deterministic map.
"""
import numpy as np
import pytest
from scipy import integrate

from oracle import aep_oracle as O
from geepee_amd import workloads as W
from tests.fdutil import check_grads


def dgph_problem(N=15, widths=(2, 2, 1), M=5, seed=0):
    rs = np.random.RandomState(seed)
    x, y, p, Ms, hidden = W.sdgpr_problem(N=N, widths=widths, M=M, seed=seed)
    L = len(widths) - 1
    p['sn_hidden'] = np.log(rs.uniform(0.1, 0.3, L - 1))
    for i in range(L - 1):
        p['h_factor_1_%d' % i] = rs.randn(N, widths[i + 1])
        p['h_factor_2_%d' % i] = np.log(rs.uniform(1.0, 3.0, (N, widths[i + 1]))) / 2
    return x, y, p, Ms, hidden


@pytest.mark.parametrize('widths', [(2, 2, 1), (3, 2, 2, 1)])
def test_sdgpr_h_gradients_fd(widths):
    x, y, p, Ms, _ = dgph_problem(widths=widths)
    check_grads(lambda q: O.sdgpr_h_objective(x, y, q, Ms, 15, alpha=0.5), p, rtol=2e-5)


def test_sdgpr_h_transition_equals_gpssm_transition():
    rs = np.random.RandomState(1)
    mp, vp, mt, vt = rs.randn(6, 2), rs.uniform(0.1, 1, (6, 2)), rs.randn(6, 2), rs.uniform(0.1, 1, (6, 2))
    sn = np.log(0.3)
    for alpha in (1.0, 0.5, 0.01):
        a = O.transition_tilted_h(sn, mp, vp, mt, vt, alpha, -2.0, 2)
        b = O.transition_tilted(np.array([sn]), mp, vp, mt, vt, alpha, -2.0, 2)
        # (logZ, dm_prop, dv_prop, dm_t, dv_t, dsn) against GPSSM's (logZ, dmall-free terms)
        assert abs(a[0] - b[0]) <= 1e-12 * abs(b[0])
        np.testing.assert_allclose(a[2], b[2], rtol=1e-13)
        np.testing.assert_allclose(a[3], b[3], rtol=1e-13)
        np.testing.assert_allclose(a[4], b[4], rtol=1e-13)
        np.testing.assert_allclose(a[1], -b[3], rtol=1e-13)
        assert abs(a[5] - b[5]) <= 1e-12 * abs(b[5])


def test_gaussian_log_partition_quadrature():
    for t1, t2 in ((0.3, 2.0), (-1.5, 0.7), (2.0, 5.0)):
        val, _ = integrate.quad(lambda h: np.exp(t1 * h - 0.5 * t2 * h * h), -np.inf, np.inf, epsrel=1e-13)
        assert abs(np.log(val) - 0.5 * np.log(2 * np.pi) - 0.5 * (t1 ** 2 / t2 - np.log(t2))) < 1e-12


def test_rollouts_are_iterated_predictives():
    y, p = W.sgpssm_problem(T=40, Din=2, Dout=3, M=6, seed=2)
    mx, vx, my, vf, vy = O.sgpssm_predict_forward(p, 6, 40, 2, 5)
    dyn = O.layer_update(p, '_dynamic', 6, 39, True)
    pm, pv = O.sgpssm_posterior_x(p)
    m, v = pm[[-1]], pv[[-1]]
    for t in range(5):
        m, v = O.fwd_mm_post(dyn, m, v)
        np.testing.assert_allclose(mx[t], m[0], rtol=1e-13)
        np.testing.assert_allclose(vx[t], v[0], rtol=1e-13)
        a, b, c = O.emission_predict(p['C_emission'], p['R_emission'], m, v)
        np.testing.assert_allclose(my[t], a[0], rtol=1e-13)
        np.testing.assert_allclose(vy[t], np.diag(c[0]), rtol=1e-13)
    # the posterior of the x factors: three tied factors inside, two at the ends, prior at 0
    f1, f2 = p['x_factor_1'], np.exp(2 * p['x_factor_2'])
    np.testing.assert_allclose(1 / pv[5], 3 * f2[5], rtol=1e-14)
    np.testing.assert_allclose(1 / pv[-1], 2 * f2[-1], rtol=1e-14)
    np.testing.assert_allclose(pm[0] / pv[0], 2 * f1[0], rtol=1e-14)
    # Monte Carlo rollout: x_t is mt + sqrt(vt) eps_t of the deterministic predictive at x_{t-1}
    np.random.seed(3)
    xs, mys, vys = O.sgpssm_predict_forward_mc(p, 6, 40, 2, 3, 4)
    np.random.seed(3)
    e0 = np.random.randn(4, 2)
    x0 = e0 * np.sqrt(pv[[-1]]) + pm[[-1]]
    mt, vt = O.fwd_det_post(dyn, x0)
    e1 = np.random.randn(4, 2)
    np.testing.assert_allclose(xs[0], e1 * np.sqrt(vt) + mt, rtol=1e-13)
