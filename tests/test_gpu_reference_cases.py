"""The reference's own test cases, run on the GPU models (GPU only).

* Gradient checks of tests/test_grads_aep.py:18-30, 124-135, 230-243, 370-410: the same
  model sizes, alpha values and init_hypers starting points (np.random.seed(42)), with
  every analytic gradient coordinate (up to a cap per key) against fourth-order
  central differences of the GPU energy; the reference prints failures at 1e-4
  relative, here they assert at the stated normwise tolerance.
* Known answers that fix the energy independently of any restatement: at N=1 and
  alpha=1 the energy is -log N(y | 0, sf2 + sn2); as alpha -> 0 with the optimal q(u)
  the SGPR energy tends to the collapsed bound of Titsias (2009) and the SGPLVM energy
  to that of Titsias & Lawrence (2010) (tests/test_aep_vfe_limits.py:17-34, 56-72).
"""
import sys

import numpy as np
import pytest

from geepee_amd import workloads as W
from tests.fdutil import check_grads
from oracle import aep_oracle as O
from tests.test_gpu_models import compare, oracle_ref
from tests.test_oracle_pins import _collapsed_bound, _log_gauss, _optimal_qu

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fd(model, params, mb, alpha, rtol, h=1e-4, oracle=None):
    """FD check of the GPU gradient (errors against the largest gradient entry: the
    reference's init_hypers makes q(u) sharp, sn = log 0.01, so the energy carries large
    cancelling terms and the finite differences of tiny coordinates sit in rounding
    noise), then the analytic gradient against the oracle's at the parity bar."""
    def f(p):
        np.random.seed(42)
        return model.objective_function(p, mb, alpha=alpha)
    worst = check_grads(f, params, h=h, rtol=rtol, max_per_key=40, seed=3, scale='global')
    say('  worst FD error %.2e' % worst)
    if oracle is not None:
        want, floor = oracle_ref(lambda: (np.random.seed(42), oracle(params))[1])
        compare(f(params), want, floor=floor)


def _kink(T, process_noise, obs_noise, seed):
    """x_{t+1} = f(x_t) + N(0, pn^2) with the kink map f(x) = 0.8 + (x + 0.2)(1 - 5 /
    (1 + exp(-2x))), y_t = x_t + N(0, on^2)."""
    rs = np.random.RandomState(seed)
    x = np.zeros(T)
    x[0] = rs.randn()
    for t in range(T - 1):
        x[t + 1] = 0.8 + (x[t] + 0.2) * (1 - 5 / (1 + np.exp(-2 * x[t]))) + process_noise * rs.randn()
    return (x + obs_noise * rs.randn(T)).reshape(T, 1)


@pytest.mark.parametrize('nat', [True, False])
def test_gplvm_aep_gaussian(nat):
    from geepee_amd import aep_models as G
    say('gplvm fd nat=%s' % nat)
    np.random.seed(42)
    y = np.random.randn(10, 3)
    model = G.SGPLVM(y, 2, 5, nat_param=nat)
    p = model.init_hypers(y)
    oracle = lambda q: O.sgplvm_objective(y, q, 5, 10, alpha=0.5, nat_param=nat)
    if nat:
        _fd(model, p, 10, 0.5, rtol=1e-5, oracle=oracle)
        return
    # The reference's moment-matched adjoint (aep_models.py:201-304) has no branch for
    # nat_param=False: it differentiates the natural-parameter cavity, so this model's
    # gradient is not the gradient of its energy (finite differences disagree at O(1)
    # in the oracle too).  The drop-in reproduces the reference; parity only.
    np.random.seed(42)
    got = model.objective_function(p, 10, alpha=0.5)
    want, floor = oracle_ref(lambda: (np.random.seed(42), oracle(p))[1])
    compare(got, want, floor=floor)


@pytest.mark.parametrize('nat', [True, False])
def test_gpr_aep_gaussian(nat):
    from geepee_amd import aep_models as G
    say('gpr fd nat=%s' % nat)
    np.random.seed(42)
    y = np.random.randn(20, 3)
    x = np.random.randn(20, 2)
    model = G.SGPR(x, y, 10, nat_param=nat)
    # alpha = 1e-4 scales the q(u) terms by N / alpha ~ 2e5; their cancellation costs
    # about five digits of the finite differences
    _fd(model, model.init_hypers(y), 20, 1e-4, rtol=1e-4, h=1e-3,
        oracle=lambda p: O.sgpr_objective(x, y, p, 10, 20, alpha=1e-4, nat_param=nat))
    _fd(model, model.init_hypers(y), 20, 0.5, rtol=1e-5,
        oracle=lambda p: O.sgpr_objective(x, y, p, 10, 20, alpha=0.5, nat_param=nat))


def test_dgpr_aep_gaussian():
    from geepee_amd import aep_models as G
    say('dgpr fd')
    np.random.seed(42)
    y = np.random.randn(10, 3)
    x = np.random.randn(10, 2)
    model = G.SDGPR(x, y, 5, [3, 2])
    _fd(model, model.init_hypers(y), 10, 1.0, rtol=1e-5,
        oracle=lambda p: O.sdgpr_objective(x, y, p, [5, 5, 5], 10, alpha=1.0))


@pytest.mark.parametrize('gp_emi,T', [(False, 50), (True, 10)])
def test_gpssm_aep_gaussian_kink(gp_emi, T):
    from geepee_amd import aep_models as G
    say('gpssm kink fd gp_emi=%s' % gp_emi)
    y = _kink(T, 0.2, 0.1, seed=7)
    model = G.SGPSSM(y, 1, 4, gp_emi=gp_emi)
    np.random.seed(42)
    _fd(model, model.init_hypers(y), T, 0.5, rtol=1e-5,
        oracle=lambda p: O.sgpssm_objective(y, p, 4, 1, T, alpha=0.5, gp_emi=gp_emi))


def test_energy_known_answer_N1_alpha1():
    from geepee_amd import aep_models as G
    say('N=1 alpha=1 known answer')
    rs = np.random.RandomState(13)
    x, y, p = W.sgpr_problem(N=1, M=4, Din=2, Dout=2, seed=5)
    p['zu'] = rs.randn(4, 2)
    e, _ = G.SGPR(x, y, 4).objective_function(p, 1, alpha=1.0)
    ref = -np.sum(_log_gauss(y, np.exp(2 * p['sf'][0]) + np.exp(2 * p['sn'][0])))
    assert abs(e - ref) < 1e-9 * abs(ref)
    x, y, p, Ms, hidden = W.sdgpr_problem(N=1, widths=(3, 2, 2, 1), M=4, seed=6)
    p['zu_0'] = rs.randn(4, 3)
    e, _ = G.SDGPR(x, y, Ms, hidden).objective_function(p, 1, alpha=1.0)
    ref = -np.sum(_log_gauss(y, np.exp(2 * p['sf_2'][0]) + np.exp(2 * p['sn'][0])))
    assert abs(e - ref) < 1e-9 * abs(ref)


def test_alpha_to_zero_limits():
    from geepee_amd import aep_models as G
    from oracle import aep_oracle as O
    say('alpha -> 0 limits')
    x, y, p = W.sgpr_problem(N=30, M=6, Din=2, Dout=2, seed=8)
    N, M, D = 30, 6, 2
    sn2 = np.exp(2 * p['sn'][0])
    Kuu = O.se_kernel(p['ls'], p['sf'], p['zu'], p['zu']) + O.JITTER * np.eye(M)
    Kfu = O.se_kernel(p['ls'], p['sf'], x, p['zu'])
    p['eta1_R'], p['eta2'] = _optimal_qu(Kuu, Kfu, Kfu.T.dot(Kfu), y, sn2, M, D)
    F = _collapsed_bound(Kuu, N * np.exp(2 * p['sf'][0]), Kfu, Kfu.T.dot(Kfu), y, sn2)
    e, _ = G.SGPR(x, y, M).objective_function(p, N, alpha=1e-6)
    assert abs(e * N + F) < 2e-5 * abs(F)
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
    e, _ = G.SGPLVM(y, Q, M).objective_function(p, N, alpha=1e-6)
    assert abs(e + (F - KLx)) < 2e-5 * abs(F - KLx)
