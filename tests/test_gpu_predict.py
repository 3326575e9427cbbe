"""Prediction paths of every model against the oracle (GPU only), and the SPD failure
status on the device-resident paths.

The oracle's predictions are pinned by the exact GP predictive, total-variance
quadrature and emission quadrature (tests/test_oracle_pins_predict.py).  Bar: normwise
relative error <= max(1e-9, 10 x the oracle's LU-vs-Cholesky spread) per output.
"""
import sys

import numpy as np
import pytest

from oracle import aep_oracle as O
from geepee_amd import workloads as W
from tests.test_gpu_models import compare_pred, pred_ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_sgplvm_predict():
    from geepee_amd import aep_models as G
    say('sgplvm predict')
    y, p = W.sgplvm_problem(N=150, Q=3, D=4, M=12, seed=3)
    model = G.SGPLVM(y, 3, 12)
    model.update_hypers(p)
    xs = np.random.RandomState(5).randn(41, 3)
    compare_pred(model.predict_f(xs), *pred_ref(lambda: O.sgplvm_predict(p, 12, 150, xs)))
    compare_pred(model.predict_y(xs), *pred_ref(lambda: O.sgplvm_predict(p, 12, 150, xs, noisy=True)))


def test_sgpssm_predict_linear_emission():
    from geepee_amd import aep_models as G
    say('sgpssm predict (linear emission)')
    y, p = W.sgpssm_problem(T=120, Din=2, Dout=3, M=10, seed=4)
    model = G.SGPSSM(y, 2, 10)
    model.update_hypers(p)
    xs = np.random.RandomState(6).randn(37, 2)
    compare_pred(model.predict_f(xs), *pred_ref(lambda: O.sgpssm_predict(p, 10, 120, xs)))
    compare_pred(model.predict_y(xs), *pred_ref(lambda: O.sgpssm_predict(p, 10, 120, xs, noisy=True)))


def test_sgpssm_predict_gp_emission():
    from geepee_amd import aep_models as G
    say('sgpssm predict (GP emission)')
    rs = np.random.RandomState(12)
    T, Din, Dout, M = 60, 2, 2, 8
    y = rs.randn(T, Dout)
    p = W.layer_params(rs, Din, Din, M, x=rs.randn(T, Din), sfx='_dynamic')
    p.update(W.layer_params(rs, Din, Dout, M, x=rs.randn(T, Din), sfx='_emission'))
    p['x_factor_1'] = rs.randn(T, Din)
    p['x_factor_2'] = np.log(rs.uniform(2, 4, (T, Din))) / 2
    p['sn'] = np.array([np.log(0.3)])
    p['sn_emission'] = np.array([np.log(0.2)])
    model = G.SGPSSM(y, Din, M, gp_emi=True)
    model.update_hypers(p)
    xs = rs.randn(29, Din)
    compare_pred(model.predict_f(xs), *pred_ref(lambda: O.sgpssm_predict(p, M, T, xs)))
    compare_pred(model.predict_y(xs),
                 *pred_ref(lambda: O.sgpssm_predict(p, M, T, xs, noisy=True, gp_emi=True)))


def test_sdgpr_predict_y_and_prop_mode():
    from geepee_amd import aep_models as G
    say('sdgpr predict_y')
    x, y, p, Ms, hidden = W.sdgpr_problem(N=300, widths=(4, 3, 2, 1), M=16, seed=2)
    model = G.SDGPR(x, y, Ms, hidden)
    model.update_hypers(p)
    xs = np.random.RandomState(4).randn(33, 4)
    want = pred_ref(lambda: O.sdgpr_predict(p, Ms, 300, xs))
    compare_pred(model.predict_f(xs, prop_mode='MM'), *want)
    compare_pred(model.predict_f(xs, prop_mode='MM', no_samples=50), *want)
    compare_pred(model.predict_y(xs), *pred_ref(lambda: O.sdgpr_predict(p, Ms, 300, xs, noisy=True)))
    with pytest.raises(NotImplementedError):
        model.predict_f(xs, prop_mode='LIN')


def _bad_sgpr():
    """SGPR whose q(u) precision Kuu^-1 + R^T R overflows: a diagonal log entry of R of
    1000 makes R_33 = inf, so the SPD inverse of output dimension 1 meets a pivot that is
    not a positive finite number (the kernel reports it in info without faulting)."""
    from geepee_amd import aep_models as G
    x, y, p = W.sgpr_problem(N=120, M=10, Din=1, Dout=2, seed=7)
    bad = {k: np.array(v, copy=True) for k, v in p.items()}
    iu = np.triu_indices(10)
    k = int(np.where((iu[0] == 3) & (iu[1] == 3))[0][0])
    bad['eta1_R'][1, k] = 1000.0
    return G.SGPR(x, y, 10), p, bad


def test_spd_failure_raises_on_every_path():
    from geepee_amd.aep_models import DeviceParams, _shapes
    from geepee_amd.graphs import GraphedObjective
    say('spd failure status')
    model, good, bad = _bad_sgpr()
    model.objective_function(good, 120, alpha=0.5)
    with pytest.raises(np.linalg.LinAlgError):
        model.objective_function(bad, 120, alpha=0.5)
    # numpy API recovers on the next good evaluation
    e, _ = model.objective_function(good, 120, alpha=0.5)
    assert np.isfinite(e)
    # device-resident evaluation: no synchronisation inside; the status word says it
    dp = DeviceParams(_shapes(bad))
    dp.upload(bad)
    model.objective_function_device(dp.views, 120, alpha=0.5)
    with pytest.raises(np.linalg.LinAlgError):
        model.check_status()
    dp.upload(good)
    model.objective_function_device(dp.views, 120, alpha=0.5)
    model.check_status()
    # graph replay: the status launch is part of the captured evaluation
    go = GraphedObjective(model, good, 120, alpha=0.5)
    for _ in range(2):
        go.objective_function(good)
        with pytest.raises(np.linalg.LinAlgError):
            go.objective_function(bad)
    go.objective_function(good)
    # device Adam: the sticky flag is read once after the loop
    with pytest.raises(np.linalg.LinAlgError):
        model.optimise_device(bad, maxiter=3, graph=True)
    model.optimise_device(good, maxiter=3, graph=True)
