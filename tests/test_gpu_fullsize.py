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
    worst = check_grads(_f(model, mb), p, h=1e-3, rtol=1e-6, max_per_key=6, seed=1)
    say('  worst normwise FD error %.2e' % worst)


@pytest.mark.parametrize('name', ['sgplvm', 'sdgpr', 'sgpssm'])
def test_fullsize_fused_equals_plain_and_graph_equals_eager(name):
    from geepee_amd.graphs import GraphedObjective
    say('full-size fused/plain/graph', name)
    model, p, mb = _model(name)
    e0, g0 = _f(model, mb)(p)
    model.set_fused()
    e1, g1 = _f(model, mb)(p)
    # the two paths differ in psi2's last bits; the Kuu^-1 chain (condition number up to
    # ~1e6 for the GPSSM dynamics) amplifies that, so the bar is the float64 parity bar
    assert abs(e1 - e0) <= 1e-9 * abs(e0)
    for k in g0:
        assert nrel(g1[k], g0[k]) < 1e-9, k
    go = GraphedObjective(model, p, mb, alpha=0.5)
    for _ in range(3):
        np.random.seed(11)
        e2, g2 = go.objective_function(p)
        assert abs(e2 - e1) <= 1e-13 * abs(e1)
        for k in g1:
            assert nrel(g2[k], g1[k]) < 1e-13, k
