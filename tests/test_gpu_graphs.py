"""hipGraph replay of whole evaluations equals the eager path, replay after replay, with
eager work and synchronisations in between (GPU only)."""
import sys

import numpy as np
import pytest

from geepee_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(a, b):
    ea, ga = a
    eb, gb = b
    assert abs(ea - eb) <= 1e-13 * abs(eb)
    for k in gb:
        assert np.linalg.norm(ga[k] - gb[k]) <= 1e-13 * max(np.linalg.norm(gb[k]), 1e-300), k


def test_graph_replays_sdgpr_minibatch():
    from geepee_amd import aep_models as G
    from geepee_amd.graphs import GraphedObjective
    say('graph sdgpr')
    x, y, p, Ms, hidden = W.sdgpr_problem(N=2000, widths=(5, 3, 3, 1), M=24, seed=5)
    model = G.SDGPR(x, y, Ms, hidden)
    g = GraphedObjective(model, p, 200, alpha=0.5)
    rs = np.random.RandomState(0)
    for seed in (1, 2, 3):
        q = {k: v + 0.01 * rs.randn(*np.shape(v)) for k, v in p.items()}
        np.random.seed(seed)
        eager = model.objective_function(q, 200, alpha=0.5)
        np.random.seed(seed)
        replay = g.objective_function(q)
        _same(replay, eager)


@pytest.mark.parametrize('kind', ['sgplvm', 'sgpssm', 'sgpr'])
def test_graph_replays_full_batch(kind):
    from geepee_amd import aep_models as G
    from geepee_amd.graphs import GraphedObjective
    say('graph', kind)
    if kind == 'sgplvm':
        y, p = W.sgplvm_problem(N=256, Q=3, D=5, M=20, seed=6)
        model, mb = G.SGPLVM(y, 3, 20), 256
    elif kind == 'sgpssm':
        y, p = W.sgpssm_problem(T=200, Din=2, Dout=3, M=12, seed=7)
        model, mb = G.SGPSSM(y, 2, 12), 200
    else:
        x, y, p = W.sgpr_problem(N=300, M=25, Din=2, Dout=1, seed=8)
        model, mb = G.SGPR(x, y, 25), 300
    g = GraphedObjective(model, p, mb, alpha=0.5)
    rs = np.random.RandomState(1)
    for _ in range(3):
        q = {k: v + 0.01 * rs.randn(*np.shape(v)) for k, v in p.items()}
        eager = model.objective_function(q, mb, alpha=0.5)
        _same(g.objective_function(q), eager)
