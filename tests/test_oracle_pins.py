"""Pins for the CPU oracle (oracle/aep_oracle.py).

The reference ships no golden vectors or assertions (its tests only print), so the
oracle is pinned by what the mathematics fixes, each check chosen so that a dropped
term, wrong sign or transposed operand fails it:
  * SE kernel vs a brute-force double loop;
  * psi1 / psi2 vs Gauss-Hermite quadrature of E[k(x,z)] and E[k(x,za) k(x,zb)],
    and their exact vx -> 0 limits;
  * every adjoint (psi, kfu, dtrace, emission) and every model gradient vs
    fourth-order central differences of the energy (the reference's own strategy,
    tests/test_utils.py:49-128, made to assert);
  * the Gaussian tilted normaliser (lik_layers.py:118-126) and the GPSSM transition
    normaliser (aep_models.py:1334-1343) vs quadrature of their defining integrals;
  * the linear-Gaussian emission normaliser (lik_layers.py:573-608) vs the
    closed-form Gaussian marginal;
  * whole energies at N=1, alpha=1 (cavity = prior, the energy is minus the exact
    log marginal likelihood) for SGPR, SDGPR and SGPLVM;
  * SGPR and SGPLVM energies at alpha -> 0 with the optimal q(u) vs the collapsed
    Titsias (2009) and Titsias-Lawrence (2010) bounds
    (tests/test_aep_vfe_limits.py:17-72 in the reference only prints this gap).
"""
import numpy as np
import pytest
from numpy.polynomial.hermite import hermgauss
from scipy.stats import multivariate_normal, norm

from oracle import aep_oracle as O
from geepee_amd import workloads as W
from tests.fdutil import check_grads


def _rand_psi_inputs(rs, N, M, Q):
    mx = rs.randn(N, Q)
    vx = rs.uniform(0.05, 0.6, (N, Q))
    z = rs.randn(M, Q)
    ls = np.log(rs.uniform(0.7, 1.6, Q))
    sf = np.array([np.log(rs.uniform(0.5, 1.5))])
    return ls, sf, mx, vx, z


def test_se_kernel_bruteforce():
    rs = np.random.RandomState(0)
    x, z = rs.randn(7, 3), rs.randn(5, 3)
    ls, sf = np.log(rs.uniform(0.5, 2, 3)), np.array([0.3])
    K = O.se_kernel(ls, sf, x, z)
    ref = np.zeros((7, 5))
    for i in range(7):
        for j in range(5):
            r2 = sum((x[i, q] - z[j, q]) ** 2 / np.exp(2 * ls[q]) for q in range(3))
            ref[i, j] = np.exp(2 * sf[0]) * np.exp(-0.5 * r2)
    np.testing.assert_allclose(K, ref, rtol=1e-14)


@pytest.mark.parametrize('Q', [1, 2])
def test_psi_vs_gauss_hermite(Q):
    rs = np.random.RandomState(Q)
    ls, sf, mx, vx, z = _rand_psi_inputs(rs, 3, 4, Q)
    psi1, psi2 = O.psi_stats(ls, sf, mx, vx, z)
    t, w = hermgauss(60)
    for n in range(3):
        grids = np.meshgrid(*[mx[n, q] + np.sqrt(2 * vx[n, q]) * t for q in range(Q)],
                            indexing='ij')
        pts = np.stack([g.ravel() for g in grids], axis=1)
        wts = np.prod(np.meshgrid(*([w] * Q), indexing='ij'), axis=0).ravel() / np.pi ** (Q / 2)
        k = O.se_kernel(ls, sf, pts, z)
        np.testing.assert_allclose(psi1[n], wts.dot(k), rtol=1e-12)
        np.testing.assert_allclose(psi2[n], np.einsum('p,pa,pb->ab', wts, k, k), rtol=1e-11)


def test_psi_zero_variance_limit_and_symmetry():
    rs = np.random.RandomState(3)
    ls, sf, mx, _, z = _rand_psi_inputs(rs, 5, 6, 3)
    psi1, psi2 = O.psi_stats(ls, sf, mx, np.zeros_like(mx), z)
    K = O.se_kernel(ls, sf, mx, z)
    np.testing.assert_allclose(psi1, K, rtol=1e-13)
    np.testing.assert_allclose(psi2, K[:, :, None] * K[:, None, :], rtol=1e-12)
    _, _, mx, vx, _ = _rand_psi_inputs(rs, 5, 6, 3)
    psi1, psi2 = O.psi_stats(ls, sf, mx, vx, z)
    assert np.array_equal(psi2, np.swapaxes(psi2, 1, 2))
    sf2 = np.exp(2 * sf[0])
    assert np.all(psi1 <= sf2) and np.all(np.diagonal(psi2, axis1=1, axis2=2) <= sf2 ** 2)


def test_psi_faithful_c_loop_matches():
    rs = np.random.RandomState(5)
    ls, sf, mx, vx, z = _rand_psi_inputs(rs, 9, 7, 3)
    a1, a2 = O.psi_stats(ls, sf, mx, vx, z)
    O.FAITHFUL = True
    try:
        b1, b2 = O.psi_stats(ls, sf, mx, vx, z)
    finally:
        O.FAITHFUL = False
    np.testing.assert_allclose(b1, a1, rtol=1e-13)
    np.testing.assert_allclose(b2, a2, rtol=1e-13)


def test_psi_adjoint_fd():
    rs = np.random.RandomState(7)
    ls, sf, mx, vx, z = _rand_psi_inputs(rs, 4, 5, 2)
    W1, W2 = rs.randn(4, 5), rs.randn(4, 5, 5)

    def f(p):
        psi1, psi2 = O.psi_stats(p['ls'], p['sf'], p['mx'], p['vx'], p['z'])
        ell, sf2 = np.exp(p['ls']), np.exp(2 * p['sf'])
        dsf2, dl, dz, dmx, dvx = O.psi_grads(W1, psi1, W2, psi2, ell, sf2, p['mx'], p['vx'], p['z'])
        e = np.sum(W1 * psi1) + np.sum(W2 * psi2)
        return e, {'ls': dl * ell, 'sf': 2 * sf2 * dsf2, 'mx': dmx, 'vx': dvx, 'z': dz}
    check_grads(f, dict(ls=ls, sf=sf, mx=mx, vx=vx, z=z), h=1e-3, rtol=1e-8)


def test_kfu_and_dtrace_adjoints_fd():
    rs = np.random.RandomState(8)
    x, z = rs.randn(6, 2), rs.randn(4, 2)
    ls, sf = np.log(rs.uniform(0.7, 1.5, 2)), np.array([0.1])
    G = rs.randn(6, 4)
    Mm = rs.randn(4, 4)

    def f(p):
        K = O.se_kernel(p['ls'], p['sf'], x, p['z'])
        ell, sf2 = np.exp(p['ls']), np.exp(2 * p['sf'])
        dvar, dl, dz, dx = O.kfu_grads(G, K, sf2, ell, p['z'], x, grad_x=True)
        Kzz = O.se_kernel(p['ls'], p['sf'], p['z'], p['z'])
        gs, gl, gz = O.dtrace_MKzz(p['ls'], p['sf'], p['z'], Mm, Kzz)
        e = np.sum(G * K) + np.sum(Mm * Kzz)
        return e, {'ls': dl * ell + 2 * gl, 'sf': 2 * sf2 * dvar + 2 * gs, 'z': dz + gz}
    check_grads(f, dict(ls=ls, sf=sf, z=z), h=1e-3, rtol=1e-8)


def test_gauss_tilted_normaliser_quadrature():
    rs = np.random.RandomState(9)
    t, w = hermgauss(120)
    for alpha in (0.5, 1.0, 0.3):
        m, v, y, sn = rs.randn(), rs.uniform(0.2, 1.0), rs.randn(), np.log(0.4)
        logZ, dm, dv = O.gauss_logZ(np.array([sn]), np.array([[m]]), np.array([[v]]),
                                    np.array([[y]]), alpha)
        f = m + np.sqrt(2 * v) * t
        quad = np.sum(w * norm.pdf(y, f, np.exp(sn)) ** alpha) / np.sqrt(np.pi)
        assert abs(logZ - np.log(quad)) < 1e-12


def test_transition_tilted_normaliser_quadrature():
    rs = np.random.RandomState(10)
    t, w = hermgauss(80)
    for alpha in (0.5, 1.0, 0.2):
        mp, vp, mt, vt, sn = rs.randn(), rs.uniform(0.1, 0.8), rs.randn(), rs.uniform(0.1, 0.8), np.log(0.5)
        logZ = O.transition_tilted(sn, np.array([[mp]]), np.array([[vp]]), np.array([[mt]]),
                                   np.array([[vt]]), alpha, 1.0, 1)[0]
        f = mp + np.sqrt(2 * vp) * t
        x = mt + np.sqrt(2 * vt) * t
        integrand = norm.pdf(x[None, :], f[:, None], np.exp(sn)) ** alpha
        quad = np.einsum('i,j,ij->', w, w, integrand) / np.pi
        assert abs(logZ - np.log(quad)) < 1e-12


def test_emission_tilted_closed_form_and_fd():
    rs = np.random.RandomState(11)
    Nb, Din, Dout, alpha = 5, 2, 3, 0.5
    C, Rp = rs.randn(Dout, Din), np.log(rs.uniform(0.3, 0.8, Dout))
    y, mx, vx = rs.randn(Nb, Dout), rs.randn(Nb, Din), rs.uniform(0.1, 0.5, (Nb, Din))
    logZ, _, _ = O.emission_tilted(C, Rp, y, mx, vx, alpha, 1.0)
    R = np.exp(2 * Rp)
    ref = 0.0
    for n in range(Nb):
        S = np.diag(R / alpha) + C.dot(np.diag(vx[n])).dot(C.T)
        ref += multivariate_normal(C.dot(mx[n]), S).logpdf(y[n])
        ref += 0.5 * Dout * (1 - alpha) * np.log(2 * np.pi) + 0.5 * (1 - alpha) * np.sum(np.log(R)) \
            - 0.5 * Dout * np.log(alpha)
    assert abs(logZ - ref) < 1e-11 * abs(ref)

    def f(p):
        e, gin, gh = O.emission_tilted(p['C'], p['R'], y, p['mx'], p['vx'], alpha, 1.0)
        return e, {'C': gh['C'], 'R': gh['R'], 'mx': gin['mx'], 'vx': gin['vx']}
    check_grads(f, dict(C=C, R=Rp, mx=mx, vx=vx), h=1e-3, rtol=1e-8)


def test_model_gradients_fd_sgpr():
    x, y, p = W.sgpr_problem(N=20, M=6, Din=2, Dout=2, seed=0)
    check_grads(lambda q: O.sgpr_objective(x, y, q, 6, 20, alpha=0.5), p)


def test_model_gradients_fd_sgpr_minibatch_and_natparam_false():
    x, y, p = W.sgpr_problem(N=20, M=5, Din=2, Dout=1, seed=1)
    idx = np.random.RandomState(3).choice(20, 8, replace=False)
    check_grads(lambda q: O.sgpr_objective(x, y, q, 5, 8, alpha=0.3, idxs=idx), p)
    # nat_param=False stores (Su, mu) directly: theta1 = R^T R is then the covariance
    check_grads(lambda r: O.sgpr_objective(x, y, r, 5, 20, alpha=0.5, nat_param=False), p)


def test_model_gradients_fd_sdgpr():
    x, y, p, Ms, _ = W.sdgpr_problem(N=15, widths=(3, 2, 2, 1), M=5, seed=0)
    check_grads(lambda q: O.sdgpr_objective(x, y, q, Ms, 15, alpha=0.5), p, rtol=2e-5)


def test_model_gradients_fd_sgplvm():
    y, p = W.sgplvm_problem(N=12, Q=2, D=3, M=5, seed=0)
    check_grads(lambda q: O.sgplvm_objective(y, q, 5, 12, alpha=0.5), p)
    idx = np.random.RandomState(4).choice(12, 5, replace=False)
    check_grads(lambda q: O.sgplvm_objective(y, q, 5, 5, alpha=0.5, idxs=idx), p)


def test_model_gradients_fd_sgpssm():
    y, p = W.sgpssm_problem(T=12, Din=2, Dout=3, M=5, seed=0)
    check_grads(lambda q: O.sgpssm_objective(y, q, 5, 2, 100, alpha=0.5), p)
    check_grads(lambda q: O.sgpssm_objective(y, q, 5, 2, 6, alpha=0.5, start_idx=3), p)


def test_model_gradients_fd_sgpssm_gp_emission_and_control():
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
    check_grads(lambda q: O.sgpssm_objective(y, q, M, Din, 100, alpha=0.5, x_control=u,
                                             gp_emi=True), p, rtol=1e-5)


def _log_gauss(y, var):
    return -0.5 * np.log(2 * np.pi * var) - 0.5 * y ** 2 / var


def test_energy_known_answer_N1_alpha1():
    """At N=1, alpha=1 the cavity equals the prior, the phi terms cancel and the
    energy is -log N(y | 0, sf2_last + sn2) exactly, whatever q(u)."""
    rs = np.random.RandomState(13)
    x, y, p = W.sgpr_problem(N=1, M=4, Din=2, Dout=2, seed=5)
    p['zu'] = rs.randn(4, 2)
    e, _ = O.sgpr_objective(x, y, p, 4, 1, alpha=1.0)
    ref = -np.sum(_log_gauss(y, np.exp(2 * p['sf'][0]) + np.exp(2 * p['sn'][0])))
    assert abs(e - ref) < 1e-9 * abs(ref)
    x, y, p, Ms, _ = W.sdgpr_problem(N=1, widths=(3, 2, 2, 1), M=4, seed=6)
    p['zu_0'] = rs.randn(4, 3)
    e, _ = O.sdgpr_objective(x, y, p, Ms, 1, alpha=1.0)
    ref = -np.sum(_log_gauss(y, np.exp(2 * p['sf_2'][0]) + np.exp(2 * p['sn'][0])))
    assert abs(e - ref) < 1e-9 * abs(ref)
    y, p = W.sgplvm_problem(N=1, Q=2, D=3, M=4, seed=7)
    p['zu'] = rs.randn(4, 2)
    e, _ = O.sgplvm_objective(y, p, 4, 1, alpha=1.0)
    ref = -np.sum(_log_gauss(y, np.exp(2 * p['sf'][0]) + np.exp(2 * p['sn'][0])))
    assert abs(e - ref) < 1e-9 * abs(ref)


def _optimal_qu(Kuu, Psi1, Psi2, y, sn2, M, D):
    Ki = np.linalg.inv(Kuu)
    theta1 = Ki.dot(Psi2).dot(Ki) / sn2
    R = np.linalg.cholesky(theta1).T
    R[np.diag_indices(M)] = np.log(R[np.diag_indices(M)])
    eta1 = np.tile(R[np.triu_indices(M)], (D, 1))
    eta2 = (Ki.dot(Psi1.T).dot(y) / sn2).T
    return eta1, eta2


def _collapsed_bound(Kuu, Psi0, Psi1, Psi2, y, sn2):
    N = y.shape[0]
    A = Kuu + Psi2 / sn2
    LA = np.linalg.slogdet(A)[1]
    LK = np.linalg.slogdet(Kuu)[1]
    F = 0.0
    for d in range(y.shape[1]):
        yd = y[:, d]
        b = Psi1.T.dot(yd)
        F += -0.5 * N * np.log(2 * np.pi * sn2) - 0.5 * yd.dot(yd) / sn2 \
            + 0.5 * b.dot(np.linalg.solve(A, b)) / sn2 ** 2 - 0.5 * LA + 0.5 * LK \
            - 0.5 * Psi0 / sn2 + 0.5 * np.trace(np.linalg.solve(Kuu, Psi2)) / sn2
    return F


def test_sgpr_alpha_to_zero_matches_titsias_bound():
    x, y, p = W.sgpr_problem(N=30, M=6, Din=2, Dout=2, seed=8)
    N, M, D = 30, 6, 2
    sn2 = np.exp(2 * p['sn'][0])
    Kuu = O.se_kernel(p['ls'], p['sf'], p['zu'], p['zu']) + O.JITTER * np.eye(M)
    Kfu = O.se_kernel(p['ls'], p['sf'], x, p['zu'])
    p['eta1_R'], p['eta2'] = _optimal_qu(Kuu, Kfu, Kfu.T.dot(Kfu), y, sn2, M, D)
    F = _collapsed_bound(Kuu, N * np.exp(2 * p['sf'][0]), Kfu, Kfu.T.dot(Kfu), y, sn2)
    e, _ = O.sgpr_objective(x, y, p, M, N, alpha=1e-6)
    assert abs(e * N + F) < 2e-5 * abs(F)


def test_sgplvm_alpha_to_zero_matches_titsias_lawrence_bound():
    y, p = W.sgplvm_problem(N=30, Q=2, D=3, M=6, seed=9)
    N, Q, M, D = 30, 2, 6, 3
    sn2 = np.exp(2 * p['sn'][0])
    vpost = 1.0 / (1.0 + np.exp(2 * p['x2']))
    mpost = p['x1'] * vpost
    Kuu = O.se_kernel(p['ls'], p['sf'], p['zu'], p['zu']) + O.JITTER * np.eye(M)
    psi1, psi2 = O.psi_stats(p['ls'], p['sf'], mpost, vpost, p['zu'])
    Psi2 = psi2.sum(0)
    p['eta1_R'], p['eta2'] = _optimal_qu(Kuu, psi1, Psi2, y, sn2, M, D)
    F = _collapsed_bound(Kuu, N * np.exp(2 * p['sf'][0]), psi1, Psi2, y, sn2)
    KLx = 0.5 * np.sum(vpost + mpost ** 2 - 1.0 - np.log(vpost))
    e, _ = O.sgplvm_objective(y, p, M, N, alpha=1e-6)
    assert abs(e + (F - KLx)) < 2e-5 * abs(F - KLx)


def test_flatten_roundtrip_and_adam():
    rs = np.random.RandomState(14)
    params = {'b': rs.randn(2, 3), 'a': rs.randn(4), 'c': np.array([1.5])}
    vec, args = O.flatten_dict(params)
    assert np.array_equal(vec[:4], params['a'])
    back = O.unflatten_dict(vec, args)
    for k in params:
        assert np.array_equal(back[k], params[k])
    # Adam on a quadratic converges to the minimiser
    xs, _ = O.adam(lambda v: (np.sum((v - 3) ** 2), 2 * (v - 3)), np.zeros(3), 3000, step_size=1e-2)
    np.testing.assert_allclose(xs, 3.0, atol=1e-3)
