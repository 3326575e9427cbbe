"""Synthetic, seeded inputs and parameters for the five benchmark workloads.

Parameters are built directly (not by ``init_hypers``, which needs k-means / PCA /
pre-fits), following the recipe of ``geepee/base_models.py:534-588``:
  zu   random subset of the inputs (or N(0,1) for hidden layers),
  ls   log(median pairwise distance / 2),
  sf   log 0.5, sn log 0.01 unless stated,
  eta1_R  packed upper R with diagonal log(1/sqrt(0.5 U(0,1))) and off-diagonal
          N(0, 0.05^2) (so the q(u) covariances are full),
  eta2 ~ N(0, 1).
All arrays are float64; every function takes an explicit seed and uses a private
``np.random.RandomState`` so workloads are reproducible and independent of the
legacy global RNG that minibatch draws use.
"""
import numpy as np
from scipy.spatial.distance import cdist


def _median_ls(x, rs, cap=1000):
    X1 = x if x.shape[0] <= cap else x[rs.permutation(x.shape[0])[:cap]]
    d = cdist(X1, X1, 'euclidean')
    med = np.median(d[np.triu_indices(X1.shape[0])])
    return np.log(med / 2 + 1e-16)


def layer_params(rs, Din, Dout, M, x=None, sfx='', ls=None, sf=np.log(0.5), off=0.05):
    p = {}
    if x is not None:
        zu = x[rs.permutation(x.shape[0])[:M]].copy()
        lsv = _median_ls(x, rs) if ls is None else ls
    else:
        zu = rs.randn(M, Din)
        lsv = 0.0 if ls is None else ls
    p['zu' + sfx] = zu
    p['ls' + sfx] = lsv * np.ones(Din)
    p['sf' + sfx] = np.array([sf], dtype=np.float64)
    K = M * (M + 1) // 2
    eta1 = np.zeros((Dout, K))
    iu = np.triu_indices(M)
    diag = iu[0] == iu[1]
    for d in range(Dout):
        e = off * rs.randn(K)
        e[diag] = np.log(1.0 / np.sqrt(0.5 * rs.rand(M) + 0.05))
        eta1[d] = e
    p['eta1_R' + sfx] = eta1
    p['eta2' + sfx] = rs.randn(Dout, M)
    return p


def sgpr_problem(N=1000, M=50, Din=1, Dout=1, seed=0):
    """SGPR regression: x ~ U(-3,3), y = sin(3x) + N(0, 0.1^2) (summed over Din)."""
    rs = np.random.RandomState(seed)
    x = rs.uniform(-3, 3, size=(N, Din))
    y = np.sin(3 * x).sum(1, keepdims=True) + 0.1 * rs.randn(N, 1)
    if Dout > 1:
        y = np.hstack([y + 0.1 * rs.randn(N, 1) for _ in range(Dout)])
    params = layer_params(rs, Din, Dout, M, x=x)
    params['sn'] = np.array([np.log(0.1)])
    return x, y, params


def sgplvm_problem(N=4096, Q=8, D=32, M=128, seed=1):
    """SGPLVM: q(x_n) = N(m, v), m ~ N(0,1), v ~ U(0.01, 0.5); y ~ N(0,1).

    The x factors are the natural-parameter differences to the N(0,1) prior
    (base_models.py:895-911): x1 = m / v, x2 = log(1/v - 1) / 2.
    """
    rs = np.random.RandomState(seed)
    m = rs.randn(N, Q)
    v = rs.uniform(0.01, 0.5, size=(N, Q))
    y = rs.randn(N, D)
    params = layer_params(rs, Q, D, M, x=m)
    params['sn'] = np.array([np.log(0.5)])
    params['x1'] = m / v
    params['x2'] = np.log(1.0 / v - 1.0) / 2.0
    return y, params


def sdgpr_problem(N=20000, widths=(10, 5, 5, 1), M=100, seed=2):
    """Deep GP regression: x ~ N(0, I), y = sum_k sin(x_k) + N(0, 0.05^2)."""
    rs = np.random.RandomState(seed)
    Din = widths[0]
    x = rs.randn(N, Din)
    y = np.sin(x).sum(1, keepdims=True) + 0.05 * rs.randn(N, 1)
    if widths[-1] > 1:
        y = np.hstack([y] * widths[-1])
    params = {}
    Ms = [M] * (len(widths) - 1)
    for i in range(len(widths) - 1):
        if i == 0:
            params.update(layer_params(rs, widths[0], widths[1], M, x=x, sfx='_0'))
        else:
            params.update(layer_params(rs, widths[i], widths[i + 1], M, sfx='_%d' % i,
                                       ls=np.log(1.0), sf=np.log(1.0)))
    params['sn'] = np.array([np.log(0.05)])
    hidden = list(widths[1:-1])
    return x, y, params, Ms, hidden


def sgpssm_problem(T=10000, Din=4, Dout=8, M=200, seed=3):
    """GPSSM: x_t = 0.9 tanh(W x_{t-1}) + 0.05 eps (W seeded orthogonal),
    y_t = C x_t + 0.1 eps, C ~ N(0, 1/4).  q(x_t) starts at the true path with
    variance 0.1, split into three tied factors (base_models.py:1646-1650)."""
    rs = np.random.RandomState(seed)
    W, _ = np.linalg.qr(rs.randn(Din, Din))
    xs = np.zeros((T, Din))
    xprev = rs.randn(Din)
    for t in range(T):
        xprev = 0.9 * np.tanh(W.dot(xprev)) + 0.05 * rs.randn(Din)
        xs[t] = xprev
    C = 0.5 * rs.randn(Dout, Din)
    y = xs.dot(C.T) + 0.1 * rs.randn(T, Dout)
    params = layer_params(rs, Din, Din, M, x=xs[:-1], sfx='_dynamic')
    post_v = 0.1
    params['x_factor_1'] = (xs / post_v) / 3.0
    params['x_factor_2'] = np.log(1.0 / post_v / 3.0) / 2.0 * np.ones((T, Din))
    params['C_emission'] = C + 0.05 * rs.randn(Dout, Din)
    params['R_emission'] = np.log(0.1) * np.ones(Dout)
    params['sn'] = np.array([np.log(0.05)])
    return y, params


def serving_models(n_models=64, seed=4, Ms=(32, 64, 128, 256)):
    """64 resident fitted models: M_i = Ms[i mod 4], Din_i = 1 + (7 i mod 10);
    even i SGPR (Dout=1), odd i SDGPR with one hidden layer of min(Din, 5)."""
    rs = np.random.RandomState(seed)
    models = []
    for i in range(n_models):
        M = Ms[i % len(Ms)]
        Din = 1 + (7 * i) % 10
        Ntr = 4 * M
        x = rs.randn(Ntr, Din)
        if i % 2 == 0:
            p = layer_params(rs, Din, 1, M, x=x)
            p['sn'] = np.array([np.log(0.1)])
            models.append(dict(kind='sgpr', Din=Din, Ms=[M], N=Ntr, params=p, hidden=[]))
        else:
            H = min(Din, 5)
            p = layer_params(rs, Din, H, M, x=x, sfx='_0')
            p.update(layer_params(rs, H, 1, M, sfx='_1', ls=np.log(1.0), sf=np.log(1.0)))
            p['sn'] = np.array([np.log(0.1)])
            models.append(dict(kind='sdgpr', Din=Din, Ms=[M, M], N=Ntr, params=p, hidden=[H]))
    return models
