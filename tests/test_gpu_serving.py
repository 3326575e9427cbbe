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


def _oracle1(sp, X, noisy):
    if sp['kind'] == 'sgpr':
        return O.sgpr_predict(sp['params'], sp['Ms'][0], sp['N'], X, noisy=noisy)
    return O.sdgpr_predict(sp['params'], sp['Ms'], sp['N'], X, noisy=noisy)


def _oracle(sp, X, noisy):
    """Oracle (mean, var) with LU inverses, each with its bar: 1e-9 or ten times the
    oracle's own LU-vs-Cholesky spread, whichever is larger."""
    O.INV = 'lu'
    want = _oracle1(sp, X, noisy)
    O.INV = 'chol'
    try:
        alt = _oracle1(sp, X, noisy)
    finally:
        O.INV = 'lu'
    return [(w, max(1e-9, 10 * nrel(a, w))) for w, a in zip(want, alt)]


def _close(got, ref):
    return all(nrel(g, w) < bar for g, (w, bar) in zip(got, ref))


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
        ref = _oracle(specs[mid], X, noisy)
        assert _close((m, v), ref), (mid, nrel(m, ref[0][0]), nrel(v, ref[1][0]))
    # the per-request plain path agrees too
    for mid, X in reqs[:4]:
        assert _close(store.predict_plain(mid, X, noisy), _oracle(specs[mid], X, noisy))


def test_grouped_predict_parity_served_sizes():
    """The served model set's sizes (M in {32, 64, 128, 256}, Din 1-10, SGPR and two-layer
    SDGPR), requests up to 600 points, against the oracle."""
    from geepee_amd.serving import ModelStore
    say('grouped predict parity, served sizes')
    specs = W.serving_models(n_models=8, seed=4)
    store = ModelStore(specs)
    rs = np.random.RandomState(2)
    reqs = [(i, rs.randn(int(rs.choice([1, 37, 600])), specs[i]['Din'])) for i in range(8)]
    outs = store.predict_many(reqs, noisy=True)
    for (mid, X), (m, v) in zip(reqs, outs):
        ref = _oracle(specs[mid], X, True)
        assert _close((m, v), ref), (mid, specs[mid]['Ms'], nrel(m, ref[0][0]), nrel(v, ref[1][0]))


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
