"""The variational free energy (VFE) family on the GPU against the oracle
(oracle/vfe_oracle.py) on the same seeded inputs, minibatches and Monte Carlo draws
(GPU only).  The draws come from numpy's legacy global RNG in the reference's order on
both sides (minibatch first, then the dynamics layer's, then the emission layer's).

Bar: normwise relative error <= max(1e-9, 10 x the oracle's LU-vs-Cholesky spread) for the
energy and every gradient; 1e-12 for the element-wise likelihood kernels.
"""
import sys

import numpy as np
import pytest
import torch

from oracle import aep_oracle as O
from oracle import vfe_oracle as V
from geepee_amd import workloads as W
from tests.test_gpu_models import compare, compare_pred, oracle_ref, pred_ref, nrel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device='cuda')


def _seeded(seed, fn):
    return lambda: (np.random.seed(seed), fn())[1]


def _moment_qu(p, M, D, var=0.1, seed=1):
    """q(u) in the moment parameterisation: Su = var I, mu ~ N(0, 1/4) (the problems'
    natural parameters read as moments give Su with eigenvalues in the hundreds, where the
    probit's Phi underflows in the reference's float64 arithmetic)."""
    q = dict(p)
    iu = np.triu_indices(M)
    eta = np.zeros((D, len(iu[0])))
    eta[:, iu[0] == iu[1]] = 0.5 * np.log(var)
    q['eta1_R'] = eta
    q['eta2'] = 0.5 * np.random.RandomState(seed).randn(D, M)
    return q


def test_vfe_likelihood_kernels_match_oracle():
    from geepee_amd import vfe_models as G
    say('VFE expected log-likelihood kernels')
    rs = np.random.RandomState(0)
    S, B, D = 5, 301, 3
    m3, v3 = rs.randn(S, B, D), rs.uniform(0.05, 1.5, (S, B, D))
    y = rs.randn(B, D)
    yb = np.sign(rs.randn(B, D))
    sn = np.array([np.log(0.3)])
    sc = -2.5
    lik = G.Gauss_Layer(B, D)
    lik.sn = _dev(sn)
    for m, v in ((m3[0], v3[0]), (m3, v3)):
        dm, dv, scal = lik.compute_log_lik_exp(_dev(m), _dev(v), _dev(y), sc)
        ll, dmo, dvo = V.gauss_lik_exp(sn, m, v, y)
        assert abs(scal[0].item() - sc * ll) <= 1e-12 * abs(sc * ll), m.ndim
        assert nrel(dm.cpu().numpy(), sc * dmo) <= 1e-12 and nrel(dv.cpu().numpy(), sc * dvo) <= 1e-12
        dsn = V.gauss_lik_exp_dsn(sn, m, v, y, sc)
        assert abs(scal[1].item() - dsn) <= 1e-11 * abs(dsn), m.ndim
    plik = G.Probit_Layer(B, D)
    for m, v in ((m3[0], v3[0]), (m3, v3)):
        dm, dv, scal = plik.compute_log_lik_exp(_dev(m), _dev(v), _dev(yb), sc)
        ll, dmo, dvo = V.probit_lik_exp(m, v, yb)
        assert abs(scal[0].item() - sc * ll) <= 1e-12 * abs(sc * ll), m.ndim
        assert nrel(dm.cpu().numpy(), sc * dmo) <= 1e-11 and nrel(dv.cpu().numpy(), sc * dvo) <= 1e-11
    # the far lower tail, where Phi underflows in float64 but log Phi and phi / Phi do not
    from scipy.stats import norm
    mt_, vt_ = np.linspace(-60.0, 8.0, 3 * 64).reshape(64, 3), np.full((64, 3), 0.3)
    yt = np.ones((64, 3))
    dm, dv, scal = plik.compute_log_lik_exp(_dev(mt_), _dev(vt_), _dev(yt), 1.0)
    gx, gw = np.polynomial.hermite.hermgauss(O.GH_DEGREE)
    ts = gx[:, None, None] * np.sqrt(2 * vt_)[None] + mt_[None]
    w = (gw / np.sqrt(np.pi))[:, None, None]
    ll = np.sum(w * norm.logcdf(ts))
    ratio = np.exp(norm.logpdf(ts) - norm.logcdf(ts))
    assert abs(scal[0].item() - ll) <= 1e-12 * abs(ll)
    assert nrel(dm.cpu().numpy(), np.sum(w * ratio, axis=0)) <= 1e-12
    # dv sums nodes of both signs whose terms are ~100x the result here: 1e-10
    assert nrel(dv.cpu().numpy(), np.sum(w * ratio * 0.5 * gx[:, None, None] * np.sqrt(2 / vt_)[None],
                                         axis=0)) <= 1e-10
    mt, vt = rs.randn(B, D), rs.uniform(0.05, 1.0, (B, D))
    for m, v in ((m3[0], v3[0]), (m3, v3)):
        outs = G.transition_lik_exp(_dev(m), _dev(v), _dev(mt), _dev(vt), _dev(sn), sc)
        want = V.transition_lik_exp(sn, m, v, mt, vt, sc)
        assert abs(outs[4][0].item() - want[0]) <= 1e-12 * abs(want[0]), m.ndim
        for got, w in zip(outs[:4], want[1:5]):
            assert nrel(got.cpu().numpy(), w) <= 1e-12, m.ndim
        assert abs(outs[4][1].item() - want[5]) <= 1e-11 * abs(want[5]), m.ndim
    Dout, Din, Nb = 4, 3, 257
    C, Rp = rs.randn(Dout, Din), np.log(rs.uniform(0.2, 1.0, Dout))
    ye, mx, vx = rs.randn(Nb, Dout), rs.randn(Nb, Din), rs.uniform(0.1, 1.0, (Nb, Din))
    dmx, dvx, dC, dR, scal = G.emission_lik_exp(_dev(C), _dev(Rp), _dev(ye), _dev(mx), _dev(vx), sc)
    lz, gin, gh = V.emission_lik_exp(C, Rp, ye, mx, vx, sc)
    assert abs(scal[0].item() - lz) <= 1e-12 * abs(lz)
    for got, w in ((dmx, gin['mx']), (dvx, gin['vx']), (dC, gh['C']), (dR, gh['R'])):
        assert nrel(got.cpu().numpy(), w) <= 1e-12


@pytest.mark.parametrize('nat', [True, False])
@pytest.mark.parametrize('lik', ['Gaussian', 'Probit'])
def test_vfe_sgpr_parity(nat, lik):
    from geepee_amd import vfe_models as G
    say('VFE sgpr nat=%s lik=%s' % (nat, lik))
    x, y, p = W.sgpr_problem(N=200, M=20, Din=2, Dout=2, seed=0)
    if not nat:
        p = _moment_qu(p, 20, 2)
    if lik == 'Probit':
        y = np.sign(y)
        p = {k: v for k, v in p.items() if k != 'sn'}
    model = G.SGPR(x, y, 20, lik=lik, nat_param=nat)
    got = model.objective_function(p, 200)
    want, floor = oracle_ref(lambda: V.sgpr_objective(x, y, p, 20, 200, nat_param=nat, lik=lik))
    compare(got, want, floor=floor)
    np.random.seed(11)
    got = model.objective_function(p, 64)
    want, floor = oracle_ref(_seeded(11, lambda: V.sgpr_objective(x, y, p, 20, 64, nat_param=nat, lik=lik)))
    compare(got, want, floor=floor)
    # predictions through the posterior (vfe_models.py:634-683)
    xs = np.random.RandomState(2).randn(17, 2)
    got = model.predict_f(xs)
    want, floors = pred_ref(lambda: V.fwd_det(O.layer_update(p, '', 20, 200, nat), xs)[:2])
    compare_pred(got, want, floors)


@pytest.mark.parametrize('mode', ['MM', 'MC'])
@pytest.mark.parametrize('nat', [True, False])
def test_vfe_sgplvm_parity(mode, nat):
    from geepee_amd import vfe_models as G
    say('VFE sgplvm mode=%s nat=%s' % (mode, nat))
    y, p = W.sgplvm_problem(N=150, Q=3, D=4, M=12, seed=3)
    if not nat:
        p = dict(p)
        p['x1'] = p['x1'] * 0.1
        p['x2'] = np.log(np.random.RandomState(9).uniform(0.05, 0.5, p['x2'].shape)) / 2
    model = G.SGPLVM(y, 3, 12, nat_param=nat)
    for mb, seed in ((150, 31), (40, 32)):
        np.random.seed(seed)
        got = model.objective_function(p, mb, prop_mode=mode)
        want, floor = oracle_ref(_seeded(seed, lambda: V.sgplvm_objective(y, p, 12, mb, nat_param=nat,
                                                                          prop_mode=mode)))
        compare(got, want, floor=floor)


def test_vfe_sgplvm_probit_mc():
    from geepee_amd import vfe_models as G
    say('VFE sgplvm probit MC')
    y, p = W.sgplvm_problem(N=150, Q=3, D=4, M=12, seed=3)
    y = np.sign(y)
    p = {k: v for k, v in p.items() if k != 'sn'}
    model = G.SGPLVM(y, 3, 12, lik='Probit')
    for mode, seed in (('MC', 33), ('MM', 34)):
        np.random.seed(seed)
        got = model.objective_function(p, 40, prop_mode=mode)
        want, floor = oracle_ref(_seeded(seed, lambda: V.sgplvm_objective(y, p, 12, 40, prop_mode=mode,
                                                                          lik='probit')))
        compare(got, want, floor=floor)


@pytest.mark.parametrize('mode', ['MM', 'MC'])
def test_vfe_sgpssm_parity(mode):
    from geepee_amd import vfe_models as G
    say('VFE sgpssm linear emission mode=%s' % mode)
    y, p = W.sgpssm_problem(T=120, Din=2, Dout=3, M=10, seed=4)
    model = G.SGPSSM(y, 2, 10)
    for mb, seed in ((500, 41), (40, 42)):
        np.random.seed(seed)
        got = model.objective_function(p, mb, prop_mode=mode)
        want, floor = oracle_ref(_seeded(seed, lambda: V.sgpssm_objective(y, p, 10, 2, mb, prop_mode=mode)))
        compare(got, want, floor=floor)
    say('VFE sgpssm GP emission with control mode=%s' % mode)
    rs = np.random.RandomState(12)
    T, Din, Dout, M, Dc = 60, 2, 2, 8, 1
    y = rs.randn(T, Dout)
    u = rs.randn(T, Dc)
    p = W.layer_params(rs, Din + Dc, Din, M, x=rs.randn(T, Din + Dc), sfx='_dynamic')
    p.update(W.layer_params(rs, Din + Dc, Dout, M, x=rs.randn(T, Din + Dc), sfx='_emission'))
    p['x_factor_1'] = rs.randn(T, Din)
    p['x_factor_2'] = np.log(rs.uniform(2, 4, (T, Din))) / 2
    p['sn'] = np.array([np.log(0.3)])
    p['sn_emission'] = np.array([np.log(0.2)])
    model = G.SGPSSM(y, Din, M, x_control=u, gp_emi=True)
    for mb, seed in ((500, 43), (25, 44)):
        np.random.seed(seed)
        got = model.objective_function(p, mb, prop_mode=mode)
        want, floor = oracle_ref(_seeded(seed, lambda: V.sgpssm_objective(y, p, M, Din, mb, x_control=u,
                                                                          gp_emi=True, prop_mode=mode)))
        compare(got, want, floor=floor)
