"""Grouped predict and the continuous-batching server (GPU only): parity with the oracle
(<= 1e-9 normwise) and batch invariance (a request's result alone and in any mix is
bitwise identical, tolerance 0)."""
import sys

import numpy as np
import pytest

from oracle import aep_oracle as O
from geepee_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nrel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _specs():
    specs = W.serving_models(n_models=8, seed=7, Ms=(8, 16, 32, 24))
    return specs


def _oracle(sp, X, noisy):
    if sp['kind'] == 'sgpr':
        return O.sgpr_predict(sp['params'], sp['Ms'][0], sp['N'], X, noisy=noisy)
    return O.sdgpr_predict(sp['params'], sp['Ms'], sp['N'], X, noisy=noisy)


@pytest.mark.parametrize('noisy', [False, True])
def test_grouped_predict_parity(noisy):
    from geepee_amd.serving import ModelStore
    say('grouped predict parity noisy=%s' % noisy)
    specs = _specs()
    store = ModelStore(specs)
    rs = np.random.RandomState(0)
    reqs = [(i % len(specs), rs.randn(int(rs.randint(1, 70)), specs[i % len(specs)]['Din']))
            for i in range(20)]
    outs = store.predict_many(reqs, noisy=noisy)
    for (mid, X), (m, v) in zip(reqs, outs):
        rm, rv = _oracle(specs[mid], X, noisy)
        assert nrel(m, rm) < 1e-9 and nrel(v, rv) < 1e-9, (mid, nrel(m, rm), nrel(v, rv))
    # the per-request plain path agrees too
    for mid, X in reqs[:4]:
        m, v = store.predict_plain(mid, X, noisy)
        rm, rv = _oracle(specs[mid], X, noisy)
        assert nrel(m, rm) < 1e-9 and nrel(v, rv) < 1e-9


def test_batch_invariance_bitwise():
    from geepee_amd.serving import ModelStore
    say('batch invariance')
    specs = _specs()
    store = ModelStore(specs)
    rs = np.random.RandomState(1)
    probe = (3, rs.randn(37, specs[3]['Din']))
    alone = store.predict_many([probe])[0]
    for trial in range(3):
        mix = [(int(rs.randint(len(specs))), None) for _ in range(6)]
        mix = [(m, rs.randn(int(rs.randint(1, 50)), specs[m]['Din'])) for m, _ in mix]
        pos = int(rs.randint(len(mix) + 1))
        batch = mix[:pos] + [probe] + mix[pos:]
        got = store.predict_many(batch)[pos]
        assert np.array_equal(got[0], alone[0]) and np.array_equal(got[1], alone[1])


def test_server_continuous_batching_matches_alone():
    from geepee_amd.serving import ModelStore, PredictServer, make_requests
    say('server')
    specs = _specs()
    store = ModelStore(specs)
    reqs = make_requests(len(specs), 60, rate=2000.0, dins=[s['Din'] for s in specs], seed=3,
                         max_batch=256)
    server = PredictServer(store, budget_bytes=1 << 20)
    lat, outs, wall = server.run(reqs)
    assert np.all(lat > 0) and wall > 0
    for (t, mid, X), (m, v) in zip(reqs, outs):
        am, av = store.predict_many([(mid, X)])[0]
        assert np.array_equal(m, am) and np.array_equal(v, av)
    lat_p, outs_p, _ = server.run_plain(reqs[:10])
    for (t, mid, X), (m, v), (pm, pv) in zip(reqs[:10], outs[:10], outs_p):
        assert nrel(pm, m) < 1e-9 and nrel(pv, v) < 1e-9
