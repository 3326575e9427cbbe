"""Parameter initialisation (``init_hypers``) of the layers and models.

Host-side numpy/scipy set-up, run once before optimisation, with the reference's
random draws in the same order (the legacy global numpy RNG, and scipy's kmeans2 on
that RNG), so a seeded run starts from the same point:

  SGP layer        geepee/base_models.py:518-597
  Gaussian lik     geepee/lik_layers.py:251-262
  linear emission  geepee/lik_layers.py:511-524
  SGPR / SDGPR     geepee/base_models.py:1036-1049, 1290-1312
  SGPLVM           geepee/base_models.py:839-873
  SGPSSM           geepee/base_models.py:1597-1683

Where the reference initialises q(u) by fitting a VFE regression with L-BFGS (SGPLVM,
SGPSSM) this module sets q(u) to that regression's optimum for the initial
hyper-parameters in closed form (Titsias' optimal q(u)), which is where the
reference's fit heads, and keeps the hyper-parameters at their initial values.  The
SGPSSM latent path for Din != Dout, which the reference takes from a linear dynamical
system fitted by Gibbs sampling and EM, starts from the PCA projection of y instead.
"""
import numpy as np
from scipy.cluster.vq import kmeans2
from scipy.spatial.distance import cdist

from .config import JITTER


def _se(ls, sf, x, z):
    """SE-ARD kernel with log-lengthscales ls and log-scale sf (kernels.py:10-22)."""
    ell2 = np.exp(2.0 * ls)
    d2 = cdist(x / np.sqrt(ell2), z / np.sqrt(ell2), 'sqeuclidean')
    return np.exp(2.0 * sf) * np.exp(-0.5 * d2)


def _pack_R(theta1, M):
    """eta1_R row: upper Cholesky factor of theta1 with log diagonal, packed in
    np.triu_indices(M) order (base_models.py:581-588)."""
    R = np.linalg.cholesky(theta1).T
    d = np.diag_indices(M)
    R[d] = np.log(R[d])
    return R[np.triu_indices(M)]


def sgp_layer(N, M, Din, Dout, nat_param=True, x_train=None, key_suffix=''):
    """Initial ls, sf, zu, eta1_R, eta2 of one sparse-GP layer (base_models.py:518-597):
    zu from kmeans2 on x (random subset when N >= 10000), ls = log(median pairwise
    distance / 2), sf = log 0.5; without x, ls = log(1 + 0.1 U), sf = 0, zu on a line.
    q(u) has mean linspace(-1, 1) and diagonal variances 0.5 U per output dim."""
    if x_train is None:
        ls = np.log(np.ones((Din,)) + 0.1 * np.random.rand(Din))
        sf = np.log(np.array([1.0]))
        zu = np.tile(np.linspace(-1, 1, M).reshape((M, 1)), (1, Din))
    else:
        x_train = np.asarray(x_train, dtype=np.float64)
        if N < 10000:
            zu, _ = kmeans2(x_train, M, minit='points')
        else:
            zu = x_train[np.random.permutation(N)[:M], :]
        if N < 1000:
            X1 = np.copy(x_train)
        else:
            X1 = x_train[np.random.permutation(N)[:1000], :]
        dist = cdist(X1, X1, 'euclidean')
        med = np.median(dist[np.triu_indices(X1.shape[0])])
        ls = np.log(med / 2 + 1e-16) * np.ones((Din,))
        sf = np.log(np.array([0.5]))
    Kuu = _se(ls, sf, zu, zu) + np.diag(JITTER * np.ones((M,)))
    Kuuinv = np.linalg.inv(Kuu)
    K = M * (M + 1) // 2
    eta1_R = np.zeros((Dout, K))
    eta2 = np.zeros((Dout, M))
    for d in range(Dout):
        mu = np.linspace(-1, 1, M).reshape((M, 1))
        alpha = 0.5 * np.random.rand(M)
        if nat_param:
            theta1 = np.diag(1 / alpha)
            theta2 = theta1.dot(mu)
        else:
            Su = np.linalg.inv(np.diag(1 / alpha) + Kuuinv)
            theta1 = Su
            theta2 = Su.dot(mu / alpha.reshape((M, 1)))
        eta1_R[d] = _pack_R(theta1, M)
        eta2[d] = theta2.reshape((M,))
    return {'sf' + key_suffix: sf, 'ls' + key_suffix: ls, 'zu' + key_suffix: zu,
            'eta1_R' + key_suffix: eta1_R, 'eta2' + key_suffix: eta2}


def gauss_lik(key_suffix=''):
    return {'sn' + key_suffix: np.log(0.01) * np.ones(1)}


def gauss_emis(Dout, Din, key_suffix=''):
    return {'C' + key_suffix: np.ones((Dout, Din)) / (Dout * Din),
            'R' + key_suffix: np.log(0.01) * np.ones(Dout)}


def pca_reduce(X, Q):
    """Projection of the centred data on its Q leading principal directions
    (utils.py:21-34)."""
    evals, evecs = np.linalg.eigh(np.cov(X.T))
    order = np.argsort(evals)[::-1]
    return (X - X.mean(0)).dot(evecs[:, order[:Q]])


def optimal_qu(params, sfx, x, y, M, nat_param, sn):
    """Replace eta1_R / eta2 of layer `sfx` by the collapsed-bound optimum of q(u) for a
    Gaussian regression y ~ f(x) + N(0, sn2) with the layer's hyper-parameters: the
    site precision theta1 = Kuu^-1 Kuf Kfu Kuu^-1 / sn2 and theta2 = Kuu^-1 Kuf y / sn2
    (natural form), or Sigma = (Kuu^-1 + theta1)^-1, mu = Sigma theta2."""
    ls, sf, zu = params['ls' + sfx], params['sf' + sfx], params['zu' + sfx]
    sn2 = float(np.exp(2.0 * np.asarray(sn).ravel()[0]))
    Kuu = _se(ls, sf, zu, zu) + np.diag(JITTER * np.ones((M,)))
    Kuuinv = np.linalg.inv(Kuu)
    Kuuinv = 0.5 * (Kuuinv + Kuuinv.T)
    P = Kuuinv.dot(_se(ls, sf, x, zu).T)                  # Kuu^-1 Kuf (M x N)
    site1 = P.dot(P.T) / sn2
    site1 = 0.5 * (site1 + site1.T)
    site2 = P.dot(y) / sn2                                 # (M, Dout)
    Dout = y.shape[1]
    K = M * (M + 1) // 2
    eta1_R = np.zeros((Dout, K))
    eta2 = np.zeros((Dout, M))
    if not nat_param:
        Su = np.linalg.inv(Kuuinv + site1)
        Su = 0.5 * (Su + Su.T)
    for d in range(Dout):
        if nat_param:
            theta1, theta2 = site1, site2[:, d]
        else:
            theta1, theta2 = Su, Su.dot(site2[:, d])
        w, V = np.linalg.eigh(theta1)
        theta1 = (V * np.maximum(w, 1e-10 * max(w.max(), 1e-300))).dot(V.T)
        eta1_R[d] = _pack_R(0.5 * (theta1 + theta1.T), M)
        eta2[d] = theta2
    params['eta1_R' + sfx] = eta1_R
    params['eta2' + sfx] = eta2
    return params


def sgplvm(y, Din, M, nat_param=True):
    """SGPLVM (base_models.py:839-873): q(x) from the standardised PCA projection of y
    with variance 0.1, the layer initialised on it, q(u) at the regression optimum."""
    N = y.shape[0]
    post_m = pca_reduce(y, Din)
    post_m = (post_m - np.mean(post_m, axis=0)) / (np.std(post_m, axis=0) + 1e-5)
    post_v = 0.1 * np.ones_like(post_m)
    if nat_param:
        post_2 = 1.0 / post_v
        x_params = {'x1': post_2 * post_m, 'x2': np.log(post_2 - 1) / 2}
    else:
        x_params = {'x1': post_m, 'x2': np.log(post_v) / 2}
    params = sgp_layer(N, M, Din, y.shape[1], nat_param, x_train=post_m)
    lik = gauss_lik()
    optimal_qu(params, '', post_m, y, M, nat_param, lik['sn'])
    params.update(lik)
    params.update(x_params)
    return params


def sgpssm(y, Din, M, x_control=None, gp_emi=False, control_to_emi=True, nat_param=True):
    """SGPSSM (base_models.py:1597-1683): the latent path from y (Din == Dout) or its PCA
    projection, variance 0.1 split over the three tied factors; the dynamics layer on
    (x_t [, u_t]) -> x_{t+1} and, with gp_emi, the emission layer on (x_t [, u_t]) -> y_t,
    each with q(u) at the regression optimum; linear emission C = I (Din == Dout) or
    the PCA loadings' least-squares map, R = log 0.01."""
    N, Dout = y.shape
    post_m = np.copy(y) if Din == Dout else pca_reduce(y, Din)
    post_v = 0.1 * np.ones_like(post_m)
    out = {'sn': np.log(0.01) * np.ones(1)}
    if nat_param:
        post_2 = 1.0 / post_v
        out['x_factor_1'] = post_2 * post_m / 3
        out['x_factor_2'] = np.log(post_2 / 3) / 2
    else:
        out['x_factor_1'] = np.copy(post_m)
        out['x_factor_2'] = np.log(post_v) / 2
    x = post_m[:N - 1, :]
    if x_control is not None:
        x = np.hstack((x, x_control[:N - 1, :]))
    dyn = sgp_layer(N - 1, M, x.shape[1], Din, nat_param, x_train=x, key_suffix='_dynamic')
    optimal_qu(dyn, '_dynamic', x, post_m[1:, :], M, nat_param, out['sn'])
    out.update(dyn)
    if gp_emi:
        xe = post_m
        if x_control is not None and control_to_emi:
            xe = np.hstack((xe, x_control))
        emi = sgp_layer(N, M, xe.shape[1], Dout, nat_param, x_train=xe, key_suffix='_emission')
        lik = gauss_lik('_emission')
        optimal_qu(emi, '_emission', xe, y, M, nat_param, lik['sn_emission'])
        out.update(emi)
        out.update(lik)
    else:
        Dx = Din + (x_control.shape[1] if (x_control is not None and control_to_emi) else 0)
        emi = gauss_emis(Dout, Dx, '_emission')
        if Din == Dout and Dx == Din:
            emi['C_emission'] = np.eye(Din)
        elif Dx == Din:
            emi['C_emission'] = np.linalg.lstsq(post_m, y - y.mean(0), rcond=None)[0].T
        out.update(emi)
    return out
