"""Finite-difference helpers shared by the oracle and GPU gradient tests.

The reference checks gradients coordinate by coordinate with a central difference
(tests/test_utils.py:49-128, eps=1e-5) and only prints failures at 1e-4 relative
(:42-47).  Here a fourth-order central difference is used and failures assert.
"""
import copy

import numpy as np


def fd4(f, params, key, idx, h):
    vals = []
    for s in (2, 1, -1, -2):
        p = copy.deepcopy(params)
        p[key] = np.array(p[key], dtype=np.float64)
        p[key][idx] += s * h
        vals.append(f(p)[0])
    return (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * h)


def check_grads(f, params, h=2e-3, rtol=2e-6, max_per_key=12, seed=0, skip=(), scale='key'):
    """Compare the analytic gradient dict of ``f(params) -> (energy, grads)`` with
    fourth-order finite differences on up to ``max_per_key`` coordinates per key.
    The error of each coordinate is measured against the largest gradient entry of its
    key (scale='key') or of the whole gradient (scale='global'), so cancelling tiny
    coordinates do not fail on rounding."""
    rs = np.random.RandomState(seed)
    _, g = f(params)
    assert sorted(g.keys()) == sorted(params.keys())
    gmax = max(float(np.max(np.abs(np.asarray(v)))) for v in g.values())
    report = []
    for key in sorted(params):
        if key in skip:
            continue
        shape = np.shape(params[key])
        assert np.shape(g[key]) == shape, (key, np.shape(g[key]), shape)
        idxs = list(np.ndindex(shape)) if shape else [()]
        if len(idxs) > max_per_key:
            idxs = [idxs[i] for i in rs.choice(len(idxs), max_per_key, replace=False)]
        gk = np.asarray(g[key], dtype=np.float64)
        den = max(np.max(np.abs(gk)) if scale == 'key' else gmax, 1e-8)
        for idx in idxs:
            num = fd4(f, params, key, idx, h)
            err = abs(num - gk[idx]) / den
            report.append((err, key, idx, gk[idx], num))
    report.sort(key=lambda t: -t[0])
    worst = report[0]
    assert worst[0] < rtol, 'gradient mismatch: %r' % (report[:5],)
    return worst[0]
