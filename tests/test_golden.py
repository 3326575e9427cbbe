"""The oracle still reproduces the committed golden fixtures (guards oracle drift)."""
import numpy as np
import pytest

from oracle import aep_oracle as O
from tests import golden_util as G


def _run(name, f):
    m = f['meta']
    i, p = f['in'], f['p']
    if name == 'sgpr':
        return O.sgpr_objective(i['x'], i['y'], p, int(m['M']), int(m['mb']), alpha=float(m['alpha']),
                                idxs=i['idx'])
    if name == 'sdgpr':
        return O.sdgpr_objective(i['x'], i['y'], p, [int(v) for v in m['Ms']], int(m['mb']),
                                 alpha=float(m['alpha']))
    if name == 'sgplvm':
        return O.sgplvm_objective(i['y'], p, int(m['M']), int(m['mb']), alpha=float(m['alpha']))
    return O.sgpssm_objective(i['y'], p, int(m['M']), int(m['Din']), int(m['mb']),
                              alpha=float(m['alpha']))


@pytest.mark.parametrize('name', G.NAMES)
def test_oracle_matches_golden(name):
    f = G.load(name)
    e, g = _run(name, f)
    assert abs(e - f['energy']) <= 1e-12 * abs(f['energy'])
    assert sorted(g) == sorted(f['g'])
    for k in g:
        ref = f['g'][k]
        assert np.linalg.norm(g[k] - ref) <= 1e-12 * max(np.linalg.norm(ref), 1e-300), k
