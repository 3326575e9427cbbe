"""SDGPR_H (deep GP with hidden-variable factors) and the GPSSM posterior and rollouts on
the GPU against the oracle (GPU only).

Bar: normwise relative error <= max(1e-9, 10 x the oracle's LU-vs-Cholesky spread).
"""
import sys

import numpy as np
import pytest

from oracle import aep_oracle as O
from geepee_amd import workloads as W
from tests.test_gpu_models import compare, oracle_ref, pred_ref, compare_pred, nrel
from tests.test_oracle_pins_dgph_rollout import dgph_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('widths,fused', [((2, 2, 1), False), ((3, 2, 2, 1), False), ((3, 2, 2, 1), True)])
def test_sdgpr_h_parity(widths, fused):
    from geepee_amd import aep_models as G
    say('SDGPR_H', widths, 'fused' if fused else 'plain')
    x, y, p, Ms, hidden = dgph_problem(N=120, widths=widths, M=10, seed=1)
    model = G.SDGPR_H(x, y, Ms, hidden)
    if fused:
        model.set_fused()
    got = model.objective_function(p, 120, alpha=0.5)
    want, floor = oracle_ref(lambda: O.sdgpr_h_objective(x, y, p, Ms, 120, alpha=0.5))
    compare(got, want, floor=floor)
    xs = np.random.RandomState(3).randn(21, widths[0])
    compare_pred(model.predict_f(xs), *pred_ref(lambda: O.sdgpr_h_predict(p, Ms, 120, xs)))
    compare_pred(model.predict_y(xs), *pred_ref(lambda: O.sdgpr_h_predict(p, Ms, 120, xs, noisy=True)))
    with pytest.raises(NotImplementedError):
        model.objective_function(p, 60, alpha=0.5)


def test_sdgpr_h_probit():
    from geepee_amd import aep_models as G
    say('SDGPR_H probit')
    x, y, p, Ms, hidden = dgph_problem(N=80, widths=(2, 2, 1), M=8, seed=2)
    y = np.sign(y - np.median(y))
    p.pop('sn')
    model = G.SDGPR_H(x, y, Ms, hidden, lik='Probit')
    got = model.objective_function(p, 80, alpha=0.5)
    want, floor = oracle_ref(lambda: O.sdgpr_h_objective(x, y, p, Ms, 80, alpha=0.5, lik='probit'))
    compare(got, want, floor=floor)


def test_sgpssm_posterior_and_rollouts_linear_emission():
    from geepee_amd import aep_models as G
    say('GPSSM rollout (linear emission)')
    y, p = W.sgpssm_problem(T=120, Din=2, Dout=3, M=10, seed=4)
    model = G.SGPSSM(y, 2, 10)
    model.update_hypers(p)
    compare_pred(model.get_posterior_x(), *pred_ref(lambda: O.sgpssm_posterior_x(p)))
    compare_pred(model.get_posterior_y(), *pred_ref(lambda: O.sgpssm_posterior_y(p, 10, 120)))
    got = model.predict_forward(30)
    compare_pred(got, *pred_ref(lambda: O.sgpssm_predict_forward(p, 10, 120, 2, 30)))
    np.random.seed(7)
    got = model.predict_forward(12, prop_mode='MC', no_samples=16)
    np.random.seed(7)
    want = O.sgpssm_predict_forward_mc(p, 10, 120, 2, 12, 16)
    for g, w in zip(got, want):
        assert g.shape == w.shape and nrel(g, w) <= 1e-9, nrel(g, w)


def test_sgpssm_rollouts_gp_emission_with_control():
    from geepee_amd import aep_models as G
    say('GPSSM rollout (GP emission, control)')
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
    model.update_hypers(p)
    uf = rs.randn(15, Dc)
    compare_pred(model.get_posterior_y(),
                 *pred_ref(lambda: O.sgpssm_posterior_y(p, M, T, x_control=u, gp_emi=True)))
    compare_pred(model.predict_forward(15, x_control=uf),
                 *pred_ref(lambda: O.sgpssm_predict_forward(p, M, T, Din, 15, x_control=uf, gp_emi=True)))
    with pytest.raises(NotImplementedError):
        model.predict_forward(3, x_control=uf, prop_mode='MC', no_samples=4)
    # control inputs to the dynamics only: the emission layer reads the latent state alone
    p2 = {k: v for k, v in p.items() if not k.endswith('_emission') or k == 'sn_emission'}
    p2.update(W.layer_params(rs, Din, Dout, M, x=rs.randn(T, Din), sfx='_emission'))
    m2 = G.SGPSSM(y, Din, M, x_control=u, gp_emi=True, control_to_emi=False)
    m2.update_hypers(p2)
    np.random.seed(8)
    got = m2.predict_forward(10, x_control=uf, prop_mode='MC', no_samples=6)
    np.random.seed(8)
    want = O.sgpssm_predict_forward_mc(p2, M, T, Din, 10, 6, x_control=uf, control_to_emi=False, gp_emi=True)
    for g, w in zip(got, want):
        assert g.shape == w.shape and nrel(g, w) <= 1e-9, nrel(g, w)
