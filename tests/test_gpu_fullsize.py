"""Full-size workloads (GPU only), checked through properties that hold at any size:
the analytic gradient against fourth-order central differences of the GPU energy on
sampled coordinates, the fused path against the materialising path, and graph replay
against eager evaluation.

Sizes: SGPLVM N=4096, Q=8, D=32, M=128; SDGPR 10->5->5->1, N=20000, M=100, minibatch
1000; GPSSM T=10000, latent 4, Dout=8, M=200.
"""
import sys

import numpy as np
import pytest

from geepee_amd import workloads as W
from tests.fdutil import check_grads

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _model(name):
    from geepee_amd import aep_models as G
    if name == 'sgplvm':
        y, p = W.sgplvm_problem(N=4096, Q=8, D=32, M=128, seed=1)
        return G.SGPLVM(y, 8, 128), p, 4096
    if name == 'sdgpr':
        x, y, p, Ms, hidden = W.sdgpr_problem(N=20000, widths=(10, 5, 5, 1), M=100, seed=2)
        return G.SDGPR(x, y, Ms, hidden), p, 1000
    y, p = W.sgpssm_problem(T=10000, Din=4, Dout=8, M=200, seed=3)
    return G.SGPSSM(y, 4, 200), p, 10000


def _f(model, mb):
    def f(p):
        np.random.seed(11)
        return model.objective_function(p, mb, alpha=0.5)
    return f


@pytest.mark.parametrize('name', ['sgplvm', 'sdgpr', 'sgpssm'])
def test_fullsize_gradient_vs_finite_differences(name):
    say('full-size FD', name)
    model, p, mb = _model(name)
    model.set_fused()
    # At these sizes the energy's own float64 rounding (terms of ~1e5 scaled by N/alpha
    # that cancel) puts ~1e-6 normwise noise on h = 1e-3 differences: the oracle's
    # finite differences of the SGPLVM energy show the same scatter around the analytic
    # gradient, which the GPU reproduces to ~1e-11.  The bar is 1e-5; the small-size
    # checks in test_gpu_reference_cases hold 1e-6.
    worst = check_grads(_f(model, mb), p, h=1e-3, rtol=1e-5, max_per_key=6, seed=1)
    say('  worst normwise FD error %.2e' % worst)


def _alt_plan(orig):
    """A different, equally valid launch shape of the fused kernels (other split
    counts, hence another fixed order of the partial sums)."""
    def plan(N, M, Q=None, D=None, tuned=True):
        p = dict(orig(N, M, Q, D, tuned=False))
        T = (M + 15) // 16
        nt = T * (T + 1) // 2
        p['splits_f'] = p['splits_f'] + 1 if p['splits_f'] < nt else p['splits_f'] - 1
        p['splits_b'] = p['splits_b'] + 1 if p['splits_b'] < T else max(1, p['splits_b'] - 1)
        p['nsplit'] = p['nsplit'] + 1 if p['nsplit'] < -(-N // 16) else max(1, p['nsplit'] - 1)
        return p
    return plan


@pytest.mark.parametrize('name', ['sgplvm', 'sdgpr', 'sgpssm'])
def test_fullsize_fused_equals_plain_and_graph_equals_eager(name, monkeypatch):
    from geepee_amd import dops
    from geepee_amd.graphs import GraphedObjective
    say('full-size fused/plain/graph', name)
    model, p, mb = _model(name)
    e0, g0 = _f(model, mb)(p)
    model.set_fused()
    e1, g1 = _f(model, mb)(p)
    # noise floor of this problem: the fused path under another launch shape (a
    # different fixed summation order).  Kuu^-1 chains with condition numbers up to
    # ~1e6 (the GPSSM dynamics layer, inducing points drawn from a compact trajectory)
    # amplify last-bit differences far beyond 1e-9 there; the bar is
    # max(1e-9, 100 x floor) per tensor.
    with monkeypatch.context() as mp:
        mp.setattr(dops, 'fused_plan', _alt_plan(dops.fused_plan))
        e1b, g1b = _f(model, mb)(p)
    floor = {k: nrel(g1b[k], g1[k]) for k in g1}
    floor['energy'] = abs(e1b - e1) / abs(e1)
    say('  floor max %.2e' % max(floor.values()))
    assert abs(e1 - e0) <= max(1e-9, 100 * floor['energy']) * abs(e0)
    for k in g0:
        assert nrel(g1[k], g0[k]) < max(1e-9, 100 * floor[k]), (k, nrel(g1[k], g0[k]), floor[k])
    go = GraphedObjective(model, p, mb, alpha=0.5)
    for _ in range(3):
        np.random.seed(11)
        e2, g2 = go.objective_function(p)
        assert abs(e2 - e1) <= 1e-13 * abs(e1)
        for k in g1:
            assert nrel(g2[k], g1[k]) < 1e-13, k
