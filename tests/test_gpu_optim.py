"""On-device Adam (GPU only): the update kernel against the reference's numpy update,
and the SGPR workload of N=1000 points, M=50 (one eval then 200 Adam steps) against
the oracle's objective driven by the oracle's Adam."""
import sys

import numpy as np
import pytest
import torch

from oracle import aep_oracle as O
from geepee_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def test_adam_step_matches_numpy_update():
    from geepee_amd import _lib
    say('adam step')
    rs = np.random.RandomState(0)
    n = 10007
    x, m, v = rs.randn(n), 0.1 * rs.randn(n), rs.uniform(0, 0.1, n)
    g = rs.randn(n)
    g[::97] = np.nan
    g[5] = np.inf
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device='cuda')
    xd, md, vd, gd = dev(x), dev(m), dev(v), dev(g)
    e = dev([3.5])
    cost = dev([0.0])
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    for t in (1, 7):
        _lib.call('gpx_adam_step', n, _lib.P(xd), _lib.P(gd), _lib.P(md), _lib.P(vd), lr, b1, b2, eps, t,
                  _lib.P(e), _lib.P(cost), _lib.S())
        gg = np.where(np.isfinite(g), g, 0.0)
        m = (1 - b1) * gg + b1 * m
        v = (1 - b2) * (gg ** 2) + b2 * v
        mhat = m / (1 - b1 ** t)
        vhat = v / (1 - b2 ** t)
        x = x - lr * mhat / (np.sqrt(vhat) + eps)
        assert np.max(np.abs(xd.cpu().numpy() - x)) <= 1e-15 * np.max(np.abs(x))
        assert np.max(np.abs(md.cpu().numpy() - m)) <= 1e-15 * np.max(np.abs(m))
        assert cost.cpu().numpy()[0] == 3.5


@pytest.mark.parametrize('graph', [True, False])
def test_sgpr_eval_then_200_adam_steps(graph):
    from geepee_amd import aep_models as G
    say('sgpr 200 adam steps graph=%s' % graph)
    x, y, p = W.sgpr_problem(N=1000, M=50, seed=0)
    model = G.SGPR(x, y, 50)
    e0, g0 = model.objective_function(p, 1000, alpha=0.5)
    want0 = O.sgpr_objective(x, y, p, 50, 1000, alpha=0.5)
    assert abs(e0 - want0[0]) <= 1e-9 * abs(want0[0])
    final, costs = model.optimise_device(p, maxiter=200, adam_lr=1e-3, alpha=0.5, graph=graph)
    vec0, args = O.flatten_dict(p)
    xo, fs, floor_f, floor_x = _oracle_adam(x, y, p, 50, 1000, 200)
    got = O.flatten_dict(final)[0]
    # Adam's update m / sqrt(v) is scale-free, so coordinates whose gradient is at the
    # rounding level take steps of either sign; two valid float64 evaluations (LU vs
    # Cholesky inverses in the oracle) drift apart the same way.  That drift is the bar.
    assert nrel(costs, fs) <= max(1e-9, 10 * floor_f), (nrel(costs, fs), floor_f)
    assert nrel(got - vec0, xo - vec0) <= max(1e-7, 10 * floor_x), (nrel(got - vec0, xo - vec0), floor_x)
    assert fs[-1] < fs[0]


def _oracle_adam(x, y, p, M, mb, steps):
    """Oracle Adam trajectory (LU inverses) and the drift between LU and Cholesky runs
    of the same trajectory (cost sequence, displacement)."""
    vec0, args = O.flatten_dict(p)

    def f(vec):
        e, g = O.sgpr_objective(x, y, O.unflatten_dict(vec, args), M, mb, alpha=0.5)
        return e, O.flatten_dict(g)[0]

    runs = {}
    try:
        for inv in ('lu', 'chol'):
            O.INV = inv
            runs[inv] = O.adam(f, vec0, steps, step_size=1e-3)
    finally:
        O.INV = 'lu'
    (xo, fs), (xc, fc) = runs['lu'], runs['chol']
    return xo, fs, nrel(fc, fs), nrel(xc - vec0, xo - vec0)


def test_optimise_from_init_hypers_matches_oracle():
    """optimise() with the reference defaults (reinit_hypers=True: init_hypers, then the
    host Adam loop over objective_function) against the oracle from the same start."""
    from geepee_amd import aep_models as G
    say('optimise from init_hypers')
    x, y, _ = W.sgpr_problem(N=200, M=15, Din=2, Dout=1, seed=4)
    model = G.SGPR(x, y, 15)
    np.random.seed(1)
    p0 = model.init_hypers(y)
    np.random.seed(1)
    final = model.optimise(method='adam', maxiter=20, alpha=0.5, adam_lr=1e-3, disp=False)
    vec0, args = O.flatten_dict(p0)
    xo, fs, floor_f, floor_x = _oracle_adam(x, y, p0, 15, 200, 20)
    got = O.flatten_dict(final)[0]
    assert nrel(got - vec0, xo - vec0) <= max(1e-7, 10 * floor_x), (nrel(got - vec0, xo - vec0), floor_x)
    assert fs[-1] < fs[0]


def _oracle_lbfgsb(x, y, p, M, maxiter):
    """scipy's L-BFGS-B over the oracle objective with the options of
    base_models.py:110-121, under LU and Cholesky inverses (the drift between the two is
    the noise floor of the trajectory)."""
    from scipy.optimize import minimize
    vec0, args = O.flatten_dict(p)

    def f(vec):
        e, g = O.sgpr_objective(x, y, O.unflatten_dict(vec, args), M, x.shape[0], alpha=0.5)
        return e, O.flatten_dict(g)[0]

    opts = {'maxfun': 100000, 'maxiter': maxiter, 'disp': False, 'gtol': 1e-6, 'ftol': 1e-6}
    runs = {}
    try:
        for inv in ('lu', 'chol'):
            O.INV = inv
            runs[inv] = minimize(f, vec0, method='L-BFGS-B', jac=True, options=opts)
    finally:
        O.INV = 'lu'
    lu, ch = runs['lu'], runs['chol']
    return lu, nrel(ch.x - vec0, lu.x - vec0), abs(ch.fun - lu.fun) / abs(lu.fun)


def test_optimise_lbfgsb_matches_oracle():
    """optimise(method='L-BFGS-B') (base_models.py:110-121: scipy.optimize.minimize with
    jac=True over the flattened params, always full batch) against the same scipy call on
    the oracle objective."""
    from geepee_amd import aep_models as G
    say('optimise L-BFGS-B')
    x, y, p = W.sgpr_problem(N=200, M=15, Din=2, Dout=1, seed=4)
    model = G.SGPR(x, y, 15)
    model.update_hypers(p)
    final = model.optimise(method='L-BFGS-B', reinit_hypers=False, maxiter=15, alpha=0.5,
                           disp=False)
    lu, floor_x, floor_f = _oracle_lbfgsb(x, y, p, 15, 15)
    vec0 = O.flatten_dict(p)[0]
    got = O.flatten_dict(final)[0]
    err = nrel(got - vec0, lu.x - vec0)
    assert err <= max(1e-7, 10 * floor_x), (err, floor_x)
    e, _ = model.objective_function(final, 200, alpha=0.5)
    assert abs(e - lu.fun) <= max(1e-9, 10 * floor_f) * abs(lu.fun), (e, lu.fun, floor_f)
    assert lu.fun < O.sgpr_objective(x, y, p, 15, 200, alpha=0.5)[0]


def test_save_load_model_and_serving_from_checkpoints(tmp_path):
    """save_model / load_model round trip (base_models.py:149-165) on the GPU models, and
    resident serving models loaded from those files."""
    from geepee_amd import aep_models as G
    from geepee_amd.serving import ModelStore
    say('save / load model')
    x, y, p = W.sgpr_problem(N=120, M=12, Din=2, Dout=1, seed=9)
    m1 = G.SGPR(x, y, 12)
    m1.update_hypers(p)
    f = str(tmp_path / 'sgpr.pickle')
    m1.save_model(f)
    m2 = G.SGPR(x, y, 12)
    m2.load_model(f)
    xs = np.random.RandomState(1).randn(17, 2)
    a, b = m1.predict_y(xs), m2.predict_y(xs)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert m1.objective_function(p, 120, alpha=0.5)[0] == m2.objective_function(m2.get_hypers(), 120,
                                                                                 alpha=0.5)[0]
    xd, yd, pd, Ms, hidden = W.sdgpr_problem(N=200, widths=(3, 2, 1), M=16, seed=5)
    m3 = G.SDGPR(xd, yd, Ms, hidden)
    m3.update_hypers(pd)
    f3 = str(tmp_path / 'sdgpr.pickle')
    m3.save_model(f3)
    specs = [dict(kind='sgpr', Din=2, Ms=[12], N=120, hidden=[], checkpoint=f),
             dict(kind='sdgpr', Din=3, Ms=Ms, N=200, hidden=hidden, checkpoint=f3)]
    store = ModelStore(specs)
    ref = ModelStore([dict(s, params=pp) for s, pp in zip(specs, (p, pd))])
    xs3 = np.random.RandomState(2).randn(9, 3)
    for got, want in zip(store.predict_many([(0, xs), (1, xs3)], noisy=True),
                         ref.predict_many([(0, xs), (1, xs3)], noisy=True)):
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
