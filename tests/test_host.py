"""Host-side logic that needs no GPU: launch plans, the autotuner cache, request
generation."""
import json

import numpy as np
import pytest

from geepee_amd import autotune, dops
from geepee_amd.serving import make_requests


def _ntiles(M):
    T = (M + 15) // 16
    return T * (T + 1) // 2


@pytest.mark.parametrize('N', [1, 15, 16, 17, 1000, 4096, 20000])
@pytest.mark.parametrize('M', [1, 16, 33, 100, 128, 200, 256])
def test_fused_plan_within_kernel_limits(N, M, monkeypatch, tmp_path):
    monkeypatch.setenv('GEEPEE_AMD_TUNE_CACHE', str(tmp_path / 'none.json'))
    autotune.clear()
    p = dops.fused_plan(N, M, 4, 8)
    assert p['nsub'] in (1, 2, 4)
    assert 1 <= p['splits_f'] <= _ntiles(M)
    assert 1 <= p['splits_b'] <= _ntiles(M)
    assert 1 <= p['nsplit'] <= -(-N // 16)


@pytest.mark.parametrize('N,M', [(16, 16), (1000, 100), (4096, 128), (10000, 200)])
def test_fused_candidates_within_kernel_limits(N, M):
    fwd, bpt, bpr = autotune.fused_candidates(N, M)
    assert fwd and bpt and bpr
    for c in fwd:
        assert c['nsub'] in (1, 2, 4) and 1 <= c['splits_f'] <= _ntiles(M)
    assert all(1 <= s <= _ntiles(M) for s in bpt)
    assert all(1 <= s <= -(-N // 16) for s in bpr)


def test_gemm_candidates():
    c = autotune.gemm_candidates(128, 128, 4096, 1)
    assert {'tile': 64, 'splits': 1} in c and {'tile': 32, 'splits': 1} in c
    assert all(4096 // x['splits'] >= 64 for x in c)


def test_autotune_cache_round_trip(monkeypatch, tmp_path):
    path = tmp_path / 'tune.json'
    monkeypatch.setenv('GEEPEE_AMD_TUNE_CACHE', str(path))
    autotune.clear()
    assert autotune.lookup('fused', (100, 20, 3, 4)) is None
    plan = {'nsub': 2, 'splits_f': 3, 'splits_b': 5, 'nsplit': 7}
    autotune.store('fused', (100, 20, 3, 4), plan)
    assert json.loads(path.read_text())
    autotune.clear()
    assert autotune.lookup('fused', (100, 20, 3, 4)) == plan
    assert dops.fused_plan(100, 20, 3, 4) == plan
    assert dops.fused_plan(100, 20, 3, 4, tuned=False) != plan
    autotune.clear()


def test_make_requests_shape():
    reqs = make_requests(8, 500, 1000.0, [1 + i for i in range(8)], seed=3)
    t = np.array([r[0] for r in reqs])
    assert np.all(np.diff(t) >= 0)
    assert abs(len(reqs) / t[-1] - 1000.0) < 200.0
    for _, mid, X in reqs:
        assert 0 <= mid < 8 and X.shape[1] == 1 + mid and 1 <= X.shape[0] <= 4096
