"""Monte Carlo propagation and the probit likelihood on the GPU against the oracle
(GPU only).  The Monte Carlo draws come from numpy's legacy global RNG in the
reference's order on both sides (minibatch first, then the dynamics layer's draws, then
the emission layer's), so a common seed gives both the same samples.

Bar: normwise relative error <= max(1e-9, 10 x the oracle's LU-vs-Cholesky spread).
"""
import sys

import numpy as np
import pytest
import torch

from oracle import aep_oracle as O
from geepee_amd import workloads as W
from tests.test_gpu_models import compare, oracle_ref, nrel

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _seeded(seed, fn):
    return lambda: (np.random.seed(seed), fn())[1]


def test_likelihood_kernels_match_oracle():
    from geepee_amd import dops
    say('MC / probit likelihood kernels')
    rs = np.random.RandomState(0)
    S, B, D = 5, 301, 3
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device='cuda')
    m3, v3 = rs.randn(S, B, D), rs.uniform(0.05, 1.5, (S, B, D))
    y = rs.randn(B, D)
    yb = np.sign(rs.randn(B, D))
    sn = np.array([np.log(0.3)])
    for alpha in (1.0, 0.5):
        dm, dv, scal = dops.gauss_tilted_mc(dev(m3), dev(v3), dev(y), dev(sn), alpha, -2.0)
        lz, dmo, dvo = O.gauss_logZ_mc(sn, m3, v3, y, alpha)
        assert abs(scal[0].item() - lz) <= 1e-12 * abs(lz)
        assert nrel(dm.cpu().numpy(), -2.0 * dmo) <= 1e-12 and nrel(dv.cpu().numpy(), -2.0 * dvo) <= 1e-12
        dsn = O.gauss_dsn(sn, dvo, B, D, alpha, -2.0)
        assert abs(scal[1].item() - dsn) <= 1e-11 * abs(dsn)
        for m, v in ((m3[0], v3[0]), (m3, v3)):
            dm, dv, dma, scal = dops.probit_tilted(dev(m), dev(v), dev(yb), alpha, -2.0, want_dmall=True)
            lz, dmo, dvo = O.probit_logZ(m, v, yb, alpha)
            assert abs(scal[0].item() - lz) <= 1e-12 * abs(lz), (alpha, m.ndim)
            assert nrel(dm.cpu().numpy(), -2.0 * dmo) <= 1e-11, (alpha, m.ndim)
            assert nrel(dv.cpu().numpy(), -2.0 * dvo) <= 1e-11, (alpha, m.ndim)
            assert nrel(dma.cpu().numpy(), -2.0 * (dmo - 2 * dvo * m)) <= 1e-11
    mt, vt = rs.randn(B, D), rs.uniform(0.05, 1.0, (B, D))
    outs = dops.transition_mc(dev(m3), dev(v3), dev(mt), dev(vt), dev(sn), 0.5, -3.0)
    want = O.transition_tilted_mc(sn, m3, v3, mt, vt, 0.5, -3.0, D)
    assert abs(outs[4][0].item() - want[0]) <= 1e-12 * abs(want[0])
    for got, w in zip(outs[:4], want[1:5]):
        assert nrel(got.cpu().numpy(), w) <= 1e-12
    assert abs(outs[4][1].item() - want[5]) <= 1e-11 * abs(want[5])
    eps, mx, vx = rs.randn(S, B, D), rs.randn(B, D), rs.uniform(0.1, 1.0, (B, D))
    x = dops.mc_inputs(dev(eps), dev(mx), dev(vx))
    assert nrel(x.cpu().numpy(), eps * np.sqrt(vx) + mx) <= 1e-15
    dx = rs.randn(S, B, D)
    dmx, dvx = dops.mc_reparam_bwd(dev(dx), dev(eps), dev(vx))
    wm, wv = O.reparam_bwd(dx, vx, eps)
    assert nrel(dmx.cpu().numpy(), wm) <= 1e-15 and nrel(dvx.cpu().numpy(), wv) <= 1e-14


@pytest.mark.parametrize('lik', ['Gaussian', 'Probit'])
def test_sgplvm_mc_parity(lik):
    from geepee_amd import aep_models as G
    say('sgplvm MC %s' % lik)
    y, p = W.sgplvm_problem(N=150, Q=3, D=4, M=12, seed=3)
    if lik == 'Probit':
        y = np.sign(y)
        p = {k: v for k, v in p.items() if k != 'sn'}
    model = G.SGPLVM(y, 3, 12, lik=lik)
    for mb, seed in ((150, 31), (40, 32)):
        np.random.seed(seed)
        got = model.objective_function(p, mb, alpha=0.5, prop_mode='MC')
        want, floor = oracle_ref(_seeded(seed, lambda: O.sgplvm_objective(y, p, 12, mb, alpha=0.5,
                                                                          prop_mode='MC', lik=lik)))
        compare(got, want, floor=floor)
    if lik == 'Probit':
        got = model.objective_function(p, 150, alpha=0.5)
        want, floor = oracle_ref(lambda: O.sgplvm_objective(y, p, 12, 150, alpha=0.5, lik=lik))
        compare(got, want, floor=floor)


def test_sgpssm_mc_parity():
    from geepee_amd import aep_models as G
    say('sgpssm MC')
    y, p = W.sgpssm_problem(T=120, Din=2, Dout=3, M=10, seed=4)
    model = G.SGPSSM(y, 2, 10)
    for mb, seed in ((500, 41), (40, 42)):
        np.random.seed(seed)
        got = model.objective_function(p, mb, alpha=0.5, prop_mode='MC')
        np.random.seed(seed)
        want = O.sgpssm_objective(y, p, 10, 2, mb, alpha=0.5, prop_mode='MC')
        compare(got, want)
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
    np.random.seed(43)
    got = model.objective_function(p, 500, alpha=0.5, prop_mode='MC')
    np.random.seed(43)
    want = O.sgpssm_objective(y, p, M, Din, 500, alpha=0.5, x_control=u, gp_emi=True, prop_mode='MC')
    compare(got, want)


@pytest.mark.parametrize('alpha', [1.0, 0.5])
def test_probit_regression_parity(alpha):
    from geepee_amd import aep_models as G
    say('probit SGPR / SDGPR alpha=%g' % alpha)
    rs = np.random.RandomState(5)
    x = rs.uniform(-3, 3, (200, 2))
    y = np.sign(np.sin(2 * x[:, :1]) + 0.3 * rs.randn(200, 1))
    p = W.layer_params(rs, 2, 1, 15, x=x)
    model = G.SGPR(x, y, 15, lik='Probit')
    got = model.objective_function(p, 200, alpha=alpha)
    want, floor = oracle_ref(lambda: O.sgpr_objective(x, y, p, 15, 200, alpha=alpha, lik='probit'))
    compare(got, want, floor=floor)
    xd, _, pd, Ms, hidden = W.sdgpr_problem(N=300, widths=(4, 3, 2, 1), M=16, seed=2)
    pd.pop('sn')
    yd = np.sign(np.sin(xd[:, :1]) + 0.2 * rs.randn(300, 1))
    model = G.SDGPR(xd, yd, Ms, hidden, lik='Probit')
    np.random.seed(6)
    got = model.objective_function(pd, 100, alpha=alpha)
    want, floor = oracle_ref(_seeded(6, lambda: O.sdgpr_objective(xd, yd, pd, Ms, 100, alpha=alpha,
                                                                  lik='probit')))
    compare(got, want, floor=floor)
    with pytest.raises(NotImplementedError):
        model.predict_y(xd[:3])


def test_sdgpr_predict_f_mc():
    from geepee_amd import aep_models as G
    say('sdgpr predict_f MC')
    x, y, p, Ms, hidden = W.sdgpr_problem(N=300, widths=(4, 3, 2, 1), M=16, seed=2)
    model = G.SDGPR(x, y, Ms, hidden)
    model.update_hypers(p)
    xs = np.random.RandomState(4).randn(33, 4)
    np.random.seed(51)
    got = model.predict_f(xs, prop_mode='MC', no_samples=20)
    np.random.seed(51)
    want = O.sdgpr_predict_mc(p, Ms, 300, xs, 20)
    for g, w in zip(got, want):
        assert g.shape == w.shape
        assert nrel(g, w) <= 1e-9, nrel(g, w)
