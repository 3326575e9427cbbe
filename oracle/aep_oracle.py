"""CPU restatement (numpy float64) of geepee's approximate-EP (AEP) hot path.

TEST INFRASTRUCTURE ONLY.  This module is the parity oracle: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it, and only as the checker or as the timed CPU baseline.  The product path
(``geepee_amd``) never imports it.

Every function restates the reference's arithmetic from
``geepee/{kernels,lik_layers,base_models,aep_models,utils}.py`` (cited per function)
in this module's own structure: layers are plain state objects built by
``layer_update`` and driven by free functions, not the reference's class tree.

Pins (see ``tests/test_oracle_pins.py``): SE kernel by brute force; psi statistics
by Gauss-Hermite quadrature, their vx->0 limits and the independent GPy-form
expression; every analytic gradient by central finite differences; the Gaussian
tilted normaliser and the GPSSM transition normaliser by quadrature; the whole
SGPR / SDGPR / SGPLVM / SGPSSM energy at N=1, alpha=1 against the exact marginal
likelihood; SGPR and SGPLVM energies at alpha->0 with the optimal q(u) against the
collapsed (Titsias / Titsias-Lawrence) bounds.  Parity unpinned: the ordering of
minibatch draws beyond using numpy's legacy global RNG exactly as the reference
does (``aep_models.py:628, 714, 912, 1051``).

``FAITHFUL = True`` switches the einsums to numpy's default unoptimised
contraction (the reference's cost structure, e.g. the O(D*M^4) three-operand
einsums at ``aep_models.py:252, 258, 556`` and ``base_models.py:497``) and the psi
statistics to the scalar C loop in ``oracle/csrc/psi_loop.c``; it is used only
to time the CPU baseline.
"""
import ctypes
import os
from types import SimpleNamespace

import numpy as np
from scipy.spatial.distance import cdist

JITTER = 1e-5          # geepee/config.py:11
GH_DEGREE = 10         # geepee/config.py:12
PROP_MM = 'MM'         # geepee/config.py:13
PROP_MC = 'MC'         # geepee/config.py:15
MC_NO_SAMPLES = 5      # geepee/config.py:16

FAITHFUL = False
# 'lu': np.linalg.inv as the reference calls it.  'chol': a Cholesky-based inverse,
# used only to measure how much the oracle's own outputs move when a different (equally
# valid) inversion algorithm is used -- the conditioning noise floor of a problem.
INV = 'lu'
_HERE = os.path.dirname(os.path.abspath(__file__))


def _inv(A):
    if INV == 'lu':
        return np.linalg.inv(A)
    from scipy.linalg import cho_factor, cho_solve
    A = np.asarray(A)
    if A.ndim == 2:
        return cho_solve(cho_factor(A, lower=True), np.eye(A.shape[0]))
    return np.stack([_inv(a) for a in A])


def ein(spec, *ops):
    """einsum; unoptimised (numpy default, as the reference calls it) when FAITHFUL."""
    return np.einsum(spec, *ops, optimize=not FAITHFUL)


# ----------------------------------------------------------------------------
# kernels.py
# ----------------------------------------------------------------------------

def se_kernel(ls, sf, x, z):
    """ARD squared-exponential K = sf2 exp(-r2/2), r2 = sum_q (x_q - z_q)^2 / ell_q^2.

    Follows ``compute_kernel(2*ls, 2*sf, x, z)`` (kernels.py:10-22) as called at
    base_models.py:277, 462 and aep_models.py:154: V = exp(2 ls) = ell^2 and the
    scale is exp(2 sf), through scipy's seuclidean distance (kernels.py:20).
    """
    ell2 = np.exp(2.0 * ls)
    sf2 = np.exp(2.0 * sf)
    x = np.atleast_2d(x)
    z = np.atleast_2d(z)
    r2 = cdist(x, z, 'seuclidean', V=ell2) ** 2.0
    return sf2 * np.exp(-0.5 * r2)


_psi_lib = None


def _load_psi_lib():
    global _psi_lib
    if _psi_lib is None:
        path = os.path.join(_HERE, 'libpsi_oracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle/libpsi_oracle.so is not built; run oracle/build.sh')
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.psi_loop.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 dp, dp, dp, dp, dp, dp, dp, dp]
        lib.psi_loop.restype = None
        _psi_lib = lib
    return _psi_lib


def psi_stats(ls, sf, mx, vx, z, chunk=256):
    """psi1 [N,M] and psi2 [N,M,M] of the SE kernel under x ~ N(mx, diag(vx)).

    Restates ``compute_psi_weave(2*ls, 2*sf, mx, vx, z)`` (kernels.py:181-240, loop
    at :201-234; call sites aep_models.py:194, base_models.py:299):
      psi1 = sf2 exp(sum_q [-(mu-z)^2/(2(v+l)) + 0.5 log(l/(l+v))])
      psi2 = sf2^2 exp(sum_q [-(z1-z2)^2/(4l) - (mu-(z1+z2)/2)^2/(2v+l)
                              + 0.5 log(l/(l+2v))])
    with l = ell^2 and sf2 = exp(2 sf).  psi2 is exactly symmetric in (m1, m2).
    """
    ell2 = np.exp(2.0 * ls)
    sf2 = float(np.exp(2.0 * sf).ravel()[0])
    mx = np.ascontiguousarray(mx, dtype=np.float64)
    vx = np.ascontiguousarray(vx, dtype=np.float64)
    z = np.ascontiguousarray(z, dtype=np.float64)
    N, Q = mx.shape
    M = z.shape[0]
    ld1 = 0.5 * np.log(ell2 / (ell2 + vx))
    ld2 = 0.5 * np.log(ell2 / (ell2 + 2.0 * vx))
    if FAITHFUL:
        lib = _load_psi_lib()
        psi1 = np.empty((N, M))
        psi2 = np.empty((N, M, M))
        dp = ctypes.POINTER(ctypes.c_double)
        e2 = np.ascontiguousarray(ell2, dtype=np.float64)
        ld1c = np.ascontiguousarray(ld1)
        ld2c = np.ascontiguousarray(ld2)
        lib.psi_loop(N, M, Q, sf2, mx.ctypes.data_as(dp), vx.ctypes.data_as(dp),
                     z.ctypes.data_as(dp), e2.ctypes.data_as(dp), ld1c.ctypes.data_as(dp),
                     ld2c.ctypes.data_as(dp), psi1.ctypes.data_as(dp), psi2.ctypes.data_as(dp))
        return psi1, psi2
    psi1 = np.empty((N, M))
    psi2 = np.empty((N, M, M))
    zhat = 0.5 * (z[:, None, :] + z[None, :, :])
    dz2 = np.sum((z[:, None, :] - z[None, :, :]) ** 2 / (4.0 * ell2), axis=-1)
    for s in range(0, N, chunk):
        e = slice(s, min(N, s + chunk))
        mu = mx[e]
        v = vx[e]
        a1 = -(mu[:, None, :] - z[None, :, :]) ** 2 / 2.0 / (v + ell2)[:, None, :]
        psi1[e] = sf2 * np.exp(np.sum(a1 + ld1[e][:, None, :], axis=-1))
        a2 = (mu[:, None, None, :] - zhat[None]) ** 2 / (2.0 * v + ell2)[:, None, None, :]
        psi2[e] = sf2 * sf2 * np.exp(-dz2[None] - np.sum(a2, axis=-1)
                                      + np.sum(ld2[e], axis=-1)[:, None, None])
    return psi1, psi2


def psi1_grads(dL_dpsi1, psi1, variance, lengthscale, Z, mu, S):
    """Adjoint of psi1 (kernels.py:355-378): returns dvar, dlengthscale, dZ, dmu, dS.

    ``variance`` is sf2 and ``lengthscale`` is ell (not ell^2), as passed at
    aep_models.py:236.
    """
    lengthscale2 = np.square(lengthscale)
    Lpsi1 = dL_dpsi1 * psi1
    Zmu = Z[None, :, :] - mu[:, None, :]
    denom = 1.0 / (S + lengthscale2)
    Zmu2_denom = np.square(Zmu) * denom[:, None, :]
    dvar = Lpsi1.sum() / variance
    dmu = ein('nm,nmq,nq->nq', Lpsi1, Zmu, denom)
    dS = ein('nm,nmq,nq->nq', Lpsi1, (Zmu2_denom - 1.0), denom) / 2.0
    dZ = -ein('nm,nmq,nq->mq', Lpsi1, Zmu, denom)
    dl = ein('nm,nmq,nq->q', Lpsi1, (Zmu2_denom + (S / lengthscale2)[:, None, :]),
             denom * lengthscale)
    return dvar, dl, dZ, dmu, dS


def psi2_grads(dL_dpsi2, psi2, variance, lengthscale, Z, mu, S):
    """Adjoint of psi2 (kernels.py:402-444): dL/dpsi2 is symmetrised first (:415-418)."""
    N, M, Q = mu.shape[0], Z.shape[0], mu.shape[1]
    lengthscale2 = np.square(lengthscale)
    denom = 1.0 / (2 * S + lengthscale2)
    denom2 = np.square(denom)
    dL_dpsi2 = (dL_dpsi2 + np.swapaxes(dL_dpsi2, 1, 2)) / 2
    Lpsi2 = dL_dpsi2 * psi2
    Lpsi2sum = Lpsi2.reshape(N, M * M).sum(1)
    tmp = Lpsi2.reshape(N * M, M).dot(Z).reshape(N, M, Q)
    Lpsi2Z = tmp.sum(1)
    Lpsi2Z2 = Lpsi2.reshape(N * M, M).dot(np.square(Z)).reshape(N, M, Q).sum(1)
    Lpsi2Z2p = (tmp * Z[None, :, :]).sum(1)
    Lpsi2Zhat = Lpsi2Z
    Lpsi2Zhat2 = (Lpsi2Z2 + Lpsi2Z2p) / 2
    dvar = Lpsi2sum.sum() * 2 / variance
    dmu = (-2 * denom) * (mu * Lpsi2sum[:, None] - Lpsi2Zhat)
    dS = (2 * np.square(denom)) * (np.square(mu) * Lpsi2sum[:, None]
                                   - 2 * mu * Lpsi2Zhat + Lpsi2Zhat2) \
        - denom * Lpsi2sum[:, None]
    Lpsi2_N = Lpsi2.sum(0)
    Lpsi2_M = Lpsi2.sum(2)
    dZ = -Lpsi2_N.sum(0)[:, None] * Z / lengthscale2 + Lpsi2_N.dot(Z) / lengthscale2 \
        + 2 * Lpsi2_M.T.dot(mu * denom) - Lpsi2_M.T.dot(denom) * Z \
        - (Lpsi2.reshape(N, M * M).T.dot(denom).reshape(M, M, Q) * Z[None, :, :]).sum(1)
    dl = 2 * lengthscale * ((S / lengthscale2 * denom + np.square(mu * denom)) * Lpsi2sum[:, None]
                            + (Lpsi2Z2 - Lpsi2Z2p) / (2 * np.square(lengthscale2))
                            - (2 * mu * denom2) * Lpsi2Zhat + denom2 * Lpsi2Zhat2).sum(axis=0)
    return dvar, dl, dZ, dmu, dS


def psi_grads(dpsi1, psi1, dpsi2, psi2, ell, sf2, mx, vx, z):
    """Sum of the psi1 and psi2 adjoints (kernels.py:302-309)."""
    a = psi1_grads(dpsi1, psi1, sf2, ell, z, mx, vx)
    b = psi2_grads(dpsi2, psi2, sf2, ell, z, mx, vx)
    return tuple(p + q for p, q in zip(a, b))


def kfu_grads(dL_dkfu, kfu, variance, lengthscale, Z, mu, grad_x=False):
    """Adjoint of K(x, Z) (kernels.py:381-399); lengthscale is ell."""
    lengthscale2 = np.square(lengthscale)
    Lpsi1 = dL_dkfu * kfu
    Zmu = Z[None, :, :] - mu[:, None, :]
    dvar = Lpsi1.sum() / variance
    dZ = -ein('nm,nmq->mq', Lpsi1, Zmu / lengthscale2)
    dl = ein('nm,nmq->q', Lpsi1, np.square(Zmu) / lengthscale ** 3)
    if grad_x:
        dx = ein('nm,nmq->nq', Lpsi1, Zmu / lengthscale2)
        return dvar, dl, dZ, dx
    return dvar, dl, dZ


def dtrace_MKzz(ls, sf, z, Mmat, Kzz):
    """d tr(M Kzz) / d(lsf, lls, z) for Kzz = SE without jitter (kernels.py:447-475).

    Called as d_trace_MKzz_dhypers(2*ls, 2*sf, z, M, Kuu - JITTER*I)
    (aep_models.py:292-294, 498-500); ``ls`` / ``sf`` here are the raw log params.
    """
    ell2 = np.exp(2.0 * ls)
    gr_lsf = np.sum(Mmat * Kzz)
    Ml = 0.5 * Mmat * Kzz
    Xl = z / np.sqrt(ell2)[None, :]
    ones = np.ones(Ml.shape[0])
    gr_lls = ones.dot(Ml.T.dot(Xl ** 2)) + ones.dot(Ml.dot(Xl ** 2)) \
        - 2.0 * ones.dot(Xl * Ml.dot(Xl))
    Xbar = z / ell2[None, :]
    Mbar1 = -Mmat.T * Kzz
    Mbar2 = -Mmat * Kzz
    gr_z = (Xbar * ones.dot(Mbar1)[:, None] - Mbar1.dot(Xbar)) \
        + (Xbar * ones.dot(Mbar2)[:, None] - Mbar2.dot(Xbar))
    return gr_lsf, gr_lls, gr_z


# ----------------------------------------------------------------------------
# Sparse-GP layer state (base_models.py:168-658, aep_models.py:26-586)
# ----------------------------------------------------------------------------

def eta_to_R(eta1_R, M):
    """Upper-triangular R_d from packed row-major triu values, diagonal exponentiated
    (base_models.py:645-653)."""
    D = eta1_R.shape[0]
    triu = np.triu_indices(M)
    diag = np.diag_indices(M)
    R = np.zeros((D, M, M))
    for d in range(D):
        Rd = np.zeros((M, M))
        Rd[triu] = eta1_R[d]
        Rd[diag] = np.exp(Rd[diag])
        R[d] = Rd
    return R


def pack_dR(R, dtheta1):
    """d/d eta1_R from d/d theta1 for theta1 = R^T R (aep_models.py:267-280, 573-584,
    base_models.py:504-514): dR = R (dT + dT^T), diagonal times R_ii, triu packed."""
    D, M, _ = R.shape
    triu = np.triu_indices(M)
    diag = np.diag_indices(M)
    dtheta1_R = ein('dab,dbc->dac', R, dtheta1 + np.transpose(dtheta1, [0, 2, 1]))
    out = np.zeros((D, M * (M + 1) // 2))
    for d in range(D):
        g = dtheta1_R[d]
        g[diag] = g[diag] * R[d][diag]
        out[d] = g[triu]
    return out


def layer_update(params, sfx, M, N, nat_param=True):
    """update_hypers + compute_kuu + update_posterior (base_models.py:630-658,
    454-464, 466-488)."""
    L = SimpleNamespace()
    L.M, L.N, L.nat_param = M, N, nat_param
    L.ls = np.asarray(params['ls' + sfx], dtype=np.float64)
    L.sf = np.asarray(params['sf' + sfx], dtype=np.float64)
    L.zu = np.asarray(params['zu' + sfx], dtype=np.float64)
    L.Din = L.zu.shape[1]
    L.R = eta_to_R(params['eta1_R' + sfx], M)
    L.Dout = L.R.shape[0]
    L.theta1 = ein('dba,dbc->dac', L.R, L.R)
    L.theta2 = np.array(params['eta2' + sfx], dtype=np.float64)
    L.Kuu = se_kernel(L.ls, L.sf, L.zu, L.zu) + np.diag(JITTER * np.ones(M))
    L.Kuuinv = _inv(L.Kuu)
    if nat_param:
        L.Suinv = L.Kuuinv + L.theta1
        L.Su = _inv(L.Suinv)
        L.mu = ein('dab,db->da', L.Su, L.theta2)
    else:
        L.Su = L.theta1
        L.Suinv = _inv(L.Su)
        L.mu = L.theta2
    L.Splusmm = L.Su + ein('da,db->dab', L.mu, L.mu)
    L.A = ein('ab,db->da', L.Kuuinv, L.mu)
    L.B_sto = -L.Kuuinv + ein('ab,dbc->dac', L.Kuuinv, ein('dab,bc->dac', L.Splusmm, L.Kuuinv))
    L.B_det = -L.Kuuinv + ein('ab,dbc->dac', L.Kuuinv, ein('dab,bc->dac', L.Su, L.Kuuinv))
    return L


def layer_cavity(L, alpha):
    """compute_cavity (aep_models.py:513-546): beta = (N - alpha)/N."""
    beta = (L.N - alpha) * 1.0 / L.N
    Kuuinv = L.Kuuinv
    if L.nat_param:
        L.Suhatinv = Kuuinv + beta * L.theta1
        L.Suhat = _inv(L.Suhatinv)
        L.muhat = ein('dab,db->da', L.Suhat, beta * L.theta2)
    else:
        data_f_1 = L.Suinv - Kuuinv
        data_f_2 = ein('dab,db->da', L.Suinv, L.mu)
        L.Suhatinv = Kuuinv + beta * data_f_1
        L.Suhat = _inv(L.Suhatinv)
        L.muhat = ein('dab,db->da', L.Suhat, beta * data_f_2)
    L.Ahat = ein('ab,db->da', Kuuinv, L.muhat)
    L.Splusmmhat = L.Suhat + ein('da,db->dab', L.muhat, L.muhat)
    L.Bhat_sto = -Kuuinv + ein('ab,dbc->dac', Kuuinv, ein('dab,bc->dac', L.Splusmmhat, Kuuinv))
    L.Bhat_det = -Kuuinv + ein('ab,dbc->dac', Kuuinv, ein('dab,bc->dac', L.Suhat, Kuuinv))
    return L


def _solve_vec(A, b):
    return np.linalg.solve(A, b[..., None])[..., 0]


def layer_phi(L, alpha):
    """compute_phi (aep_models.py:62-114): prior + (N/a - 1) post - (N/a) cavity."""
    N = L.N
    scale_post = N * 1.0 / alpha - 1.0
    scale_cav = -N * 1.0 / alpha
    phi_prior = L.Dout * 0.5 * np.linalg.slogdet(L.Kuu)[1]
    phi_post = 0.5 * np.sum(np.linalg.slogdet(L.Su)[1]) \
        + 0.5 * np.sum(L.mu * _solve_vec(L.Su, L.mu))
    phi_cav = 0.5 * np.sum(np.linalg.slogdet(L.Suhat)[1]) \
        + 0.5 * np.sum(L.muhat * _solve_vec(L.Suhat, L.muhat))
    return phi_prior + scale_post * phi_post + scale_cav * phi_cav


def fwd_det_cav(L, x):
    """_forward_prop_deterministic_thru_cav (aep_models.py:142-158)."""
    kff = np.exp(2 * L.sf)
    kfu = se_kernel(L.ls, L.sf, x, L.zu)
    mout = ein('nm,dm->nd', kfu, L.Ahat)
    vout = kff + ein('dab,na,nb->nd', L.Bhat_det, kfu, kfu)
    return mout, vout, kfu


def fwd_mm_cav(L, mx, vx):
    """_forward_prop_random_thru_cav_mm (aep_models.py:182-199)."""
    psi0 = np.exp(2 * L.sf)
    psi1, psi2 = psi_stats(L.ls, L.sf, mx, vx, L.zu)
    mout = ein('nm,dm->nd', psi1, L.Ahat)
    vout = psi0 + ein('dab,nab->nd', L.Bhat_sto, psi2) - mout ** 2
    return mout, vout, psi1, psi2


def fwd_det_post(L, x):
    """_forward_prop_deterministic_thru_post (base_models.py:265-284)."""
    psi0 = np.exp(2 * L.sf)
    kfu = se_kernel(L.ls, L.sf, x, L.zu)
    mout = ein('nm,dm->nd', kfu, L.A)
    vout = psi0 + ein('dab,na,nb->nd', L.B_det, kfu, kfu)
    return mout, vout


def fwd_mm_post(L, mx, vx):
    """_forward_prop_random_thru_post_mm (base_models.py:286-307)."""
    psi0 = np.exp(2.0 * L.sf)
    psi1, psi2 = psi_stats(L.ls, L.sf, mx, vx, L.zu)
    mout = ein('nm,dm->nd', psi1, L.A)
    vout = psi0 + ein('dab,nab->nd', L.B_sto, psi2) - mout ** 2
    return mout, vout


def _cav_grad_u(L, dmu, dSu, alpha):
    """compute_cav_grad_u (aep_models.py:548-586)."""
    beta = (L.N - alpha) * 1.0 / L.N
    if L.nat_param:
        dSu = dSu + ein('da,db->dab', dmu, beta * L.theta2)
        dSuinv = -ein('dab,dbc,dce->dae', L.Suhat, dSu, L.Suhat)
        dKuuinv = np.sum(dSuinv, axis=0)
        dtheta1 = beta * dSuinv
        deta2 = beta * ein('dab,db->da', L.Suhat, dmu)
    else:
        data_f_2 = ein('dab,db->da', L.Suinv, L.mu)
        dSuhat = dSu + ein('da,db->dab', dmu, beta * data_f_2)
        dSuhatinv = -ein('dab,dbc,dce->dae', L.Suhat, dSuhat, L.Suhat)
        dSuinv_1 = beta * dSuhatinv
        Suhatdmu = ein('dab,db->da', L.Suhat, dmu)
        dSuinv = dSuinv_1 + beta * ein('da,db->dab', Suhatdmu, L.mu)
        dtheta1 = -ein('dab,dbc,dce->dae', L.Suinv, dSuinv, L.Suinv)
        deta2 = beta * ein('dab,db->da', L.Suinv, Suhatdmu)
        dKuuinv = (1 - beta) / beta * np.sum(dSuinv_1, axis=0)
    return pack_dR(L.R, dtheta1), deta2, dKuuinv


def _post_grad_u(L, dmu, dSu):
    """compute_posterior_grad_u (base_models.py:490-516)."""
    if L.nat_param:
        dSu = dSu + ein('da,db->dab', dmu, L.theta2)
        dSuinv = -ein('dab,dbc,dce->dae', L.Su, dSu, L.Su)
        dKuuinv = np.sum(dSuinv, axis=0)
        dtheta1 = dSuinv
        deta2 = ein('dab,db->da', L.Su, dmu)
    else:
        deta2 = dmu
        dtheta1 = dSu
        dKuuinv = 0
    return pack_dR(L.R, dtheta1), deta2, dKuuinv


def bwd_det(L, dm, dv, kfu, x, alpha):
    """backprop_grads_reg (aep_models.py:412-511): hyper grads of a layer whose
    inputs are deterministic, including the phi terms of that layer."""
    N = L.N
    ell = np.exp(L.ls)
    sf2 = np.exp(2 * L.sf)
    kfuKuuinv = np.dot(kfu, L.Kuuinv)
    scale_post = N * 1.0 / alpha - 1.0
    scale_cav = -N * 1.0 / alpha
    dkfu = ein('nd,dm->nm', dm, L.Ahat) + 2 * ein('nd,dab,na->nb', dv, L.Bhat_det, kfu)
    dsf2, dls, dzu = kfu_grads(dkfu, kfu, sf2, ell, L.zu, x)
    dls = dls * ell
    dsf2 = dsf2 + np.sum(dv)
    dsf = 2 * sf2 * dsf2
    dmucav = ein('nd,nm->dm', dm, kfuKuuinv)
    dSucav = ein('na,nd,nb->dab', kfuKuuinv, dv, kfuKuuinv)
    Suinvmuhat = ein('dab,db->da', L.Suhatinv, L.muhat)
    dmucav = dmucav + scale_cav * Suinvmuhat
    dSucav = dSucav + scale_cav * (0.5 * L.Suhatinv
                                   - 0.5 * ein('da,db->dab', Suinvmuhat, Suinvmuhat))
    deta1_cav, deta2_cav, dKi_cav = _cav_grad_u(L, dmucav, dSucav, alpha)
    Suinvmu = ein('dab,db->da', L.Suinv, L.mu)
    dmu = scale_post * Suinvmu
    dSu = scale_post * (0.5 * L.Suinv - 0.5 * ein('da,db->dab', Suinvmu, Suinvmu))
    deta1_post, deta2_post, dKi_post = _post_grad_u(L, dmu, dSu)
    dKi_prior = -0.5 * L.Dout * L.Kuu
    dAhat = ein('nd,nm->dm', dm, kfu)
    dKi_m = ein('da,db->ab', dAhat, L.muhat)
    KuuinvSmmd = ein('ab,dbc->dac', L.Kuuinv, L.Suhat)
    dBhat = ein('nd,na,nb->dab', dv, kfu, kfu)
    dKi_v = 2 * ein('dab,dac->bc', KuuinvSmmd, dBhat) - np.sum(dBhat, axis=0)
    dKuuinv = dKi_m + dKi_v + dKi_cav + dKi_post + dKi_prior
    M_inner = -np.dot(L.Kuuinv, np.dot(dKuuinv, L.Kuuinv))
    g_sf, g_ls, g_z = dtrace_MKzz(L.ls, L.sf, L.zu, M_inner,
                                  L.Kuu - np.diag(JITTER * np.ones(L.M)))
    return {'sf': dsf + 2 * g_sf, 'ls': dls + 2 * g_ls, 'zu': dzu + g_z,
            'eta1_R': deta1_cav + deta1_post, 'eta2': deta2_cav + deta2_post}


def bwd_mm(L, m, dm, dv, psi1, psi2, mx, vx, alpha):
    """backprop_grads_lvm_mm (aep_models.py:201-304): hyper grads of a moment-matched
    layer plus the grads w.r.t. its input means / variances."""
    N = L.N
    ell = np.exp(L.ls)
    sf2 = np.exp(2 * L.sf)
    beta = (N - alpha) * 1.0 / N
    scale_post = N * 1.0 / alpha - 1.0
    scale_cav = -N * 1.0 / alpha
    dm_all = dm - 2 * dv * m
    dAhat = ein('nd,nm->dm', dm_all, psi1)
    dBhat = ein('nd,nab->dab', dv, psi2)
    dpsi1 = ein('nd,dm->nm', dm_all, L.Ahat)
    dpsi2 = ein('nd,dab->nab', dv, L.Bhat_sto)
    dsf2, dls, dzu, dmx, dvx = psi_grads(dpsi1, psi1, dpsi2, psi2, ell, sf2, mx, vx, L.zu)
    dls = dls * ell
    dsf2 = dsf2 + np.sum(dv)
    dsf = 2 * sf2 * dsf2
    Kuuinv = L.Kuuinv
    dvcav = ein('ab,dbc,ce->dae', Kuuinv, dBhat, Kuuinv)
    dmcav = 2 * ein('dab,db->da', dvcav, L.muhat) + ein('ab,db->da', Kuuinv, dAhat)
    dvcav = dvcav + beta * ein('da,db->dab', dmcav, L.theta2)
    dvcavinv = -ein('dab,dbc,dce->dae', L.Suhat, dvcav, L.Suhat)
    dtheta1 = beta * dvcavinv
    dtheta2 = beta * ein('dab,db->da', L.Suhat, dmcav)
    dKi_vcav = np.sum(dvcavinv, axis=0)
    dKi_A = ein('da,db->ab', dAhat, L.muhat)
    KuuinvSmmd = ein('ab,dbc->dac', Kuuinv, L.Splusmmhat)
    dKi_B = 2 * ein('dab,dac->bc', KuuinvSmmd, dBhat) - np.sum(dBhat, axis=0)
    dKuuinv = dKi_A + dKi_B + dKi_vcav
    Minner = scale_post * np.sum(L.Splusmm, axis=0) + scale_cav * np.sum(L.Splusmmhat, axis=0) \
        - 2.0 * dKuuinv
    dtheta1 = -0.5 * scale_post * L.Splusmm - 0.5 * scale_cav * beta * L.Splusmmhat + dtheta1
    dtheta2 = scale_post * L.mu + scale_cav * beta * L.muhat + dtheta2
    deta1_R = pack_dR(L.R, dtheta1)
    M_all = 0.5 * (L.Dout * Kuuinv + np.dot(Kuuinv, np.dot(Minner, Kuuinv)))
    g_sf, g_ls, g_z = dtrace_MKzz(L.ls, L.sf, L.zu, M_all, L.Kuu - np.diag(JITTER * np.ones(L.M)))
    grad = {'sf': dsf + 2 * g_sf, 'ls': dls + 2 * g_ls, 'zu': dzu + g_z,
            'eta1_R': deta1_R, 'eta2': dtheta2}
    return grad, {'mx': dmx, 'vx': dvx}


def fwd_mc_cav(L, mx, vx, eps):
    """_forward_prop_random_thru_cav_mc (aep_models.py:160-180): the deterministic
    cavity propagation at x = eps sqrt(vx) + mx for standard-normal draws eps (S, B, Din),
    which the reference takes from np.random.randn(MC_NO_SAMPLES, B, Din).  Returns the
    (S, B, Dout) means and variances and the stacked kfu (S*B, M) and x (S*B, Din)."""
    S, B, Din = eps.shape
    x = eps * np.sqrt(vx) + mx
    x_stk = x.reshape(S * B, Din)
    m_stk, v_stk, kfu_stk = fwd_det_cav(L, x_stk)
    return m_stk.reshape(S, B, -1), v_stk.reshape(S, B, -1), kfu_stk, x_stk


def bwd_mc(L, dm, dv, kfu, x, alpha):
    """backprop_grads_lvm_mc (aep_models.py:307-410): hyper grads of a layer evaluated at
    the stacked Monte Carlo inputs x (S*B rows; dm, dv already scaled, any shape with
    S*B*Dout entries), and the grads w.r.t. those inputs."""
    N, M = L.N, L.M
    ell = np.exp(L.ls)
    sf2 = np.exp(2 * L.sf)
    dm = dm.reshape(kfu.shape[0], L.Dout)
    dv = dv.reshape(kfu.shape[0], L.Dout)
    beta = (N - alpha) * 1.0 / N
    scale_post = N * 1.0 / alpha - 1.0
    scale_cav = -N * 1.0 / alpha
    kfuKuuinv = np.dot(kfu, L.Kuuinv)
    dkfu = ein('nd,dm->nm', dm, L.Ahat) + 2 * ein('nd,dab,na->nb', dv, L.Bhat_det, kfu)
    dsf2, dls, dzu, dx = kfu_grads(dkfu, kfu, sf2, ell, L.zu, x, grad_x=True)
    dls = dls * ell
    dsf2 = dsf2 + np.sum(dv)
    dsf = 2 * sf2 * dsf2
    SKf = ein('dab,nb->nda', L.Suhat, kfuKuuinv)
    dSinv = -ein('nda,nd,ndb->dab', SKf, dv, SKf) - ein('nda,nd,db->dab', SKf, dm, L.muhat)
    dSinvM = ein('nda,nd->da', SKf, dm)
    dtheta1 = -0.5 * scale_post * L.Splusmm - 0.5 * scale_cav * beta * L.Splusmmhat + beta * dSinv
    dtheta2 = scale_post * L.mu + scale_cav * beta * L.muhat + beta * dSinvM
    dAhat = ein('nd,nm->dm', dm, kfu)
    dKi_m = ein('da,db->ab', dAhat, L.muhat)
    KuuinvSmmd = ein('ab,dbc->dac', L.Kuuinv, L.Suhat)
    dBhat = ein('nd,na,nb->dab', dv, kfu, kfu)
    dKi_v1 = 2 * ein('dab,dac->bc', KuuinvSmmd, dBhat) - np.sum(dBhat, axis=0)
    dKuuinv = dKi_m + dKi_v1 + np.sum(dSinv, axis=0)
    Minner = scale_post * np.sum(L.Splusmm, axis=0) + scale_cav * np.sum(L.Splusmmhat, axis=0) \
        - 2.0 * dKuuinv
    M_all = 0.5 * (L.Dout * L.Kuuinv + np.dot(L.Kuuinv, np.dot(Minner, L.Kuuinv)))
    g_sf, g_ls, g_z = dtrace_MKzz(L.ls, L.sf, L.zu, M_all, L.Kuu - np.diag(JITTER * np.ones(M)))
    grad = {'sf': dsf + 2 * g_sf, 'ls': dls + 2 * g_ls, 'zu': dzu + g_z,
            'eta1_R': pack_dR(L.R, dtheta1), 'eta2': dtheta2}
    return grad, dx


def reparam_bwd(dx, vx, eps):
    """backprop_grads_reparam (base_models.py:373-388): grads w.r.t. the input means and
    variances of x = eps sqrt(vx) + mx."""
    dx = dx.reshape(eps.shape)
    # zero-variance columns (control inputs) divide by zero here as in the reference; their
    # gradients are discarded by the callers
    with np.errstate(divide='ignore', invalid='ignore'):
        return np.sum(dx, axis=0), np.sum(dx * eps, axis=0) / (2 * np.sqrt(vx))


def fwd_mc_post(L, mx, vx, eps):
    """_forward_prop_random_thru_post_mc (base_models.py:309-332) for draws eps."""
    S, B, Din = eps.shape
    x = (eps * np.sqrt(vx) + mx).reshape(S * B, Din)
    m, v = fwd_det_post(L, x)
    return m.reshape(S, B, -1), v.reshape(S, B, -1)


# ----------------------------------------------------------------------------
# lik_layers.py
# ----------------------------------------------------------------------------

def gauss_logZ(sn, mout, vout, y, alpha):
    """Gauss_Layer.compute_log_Z, 2-D branch (lik_layers.py:104-138).  Returns the
    tilted variance separately instead of mutating ``vout`` (the reference adds
    sn2/alpha to its argument in place at :122)."""
    sn2 = np.exp(2.0 * sn)
    vt = vout + sn2 / alpha
    D = mout.shape[1]
    logZ = np.sum(-0.5 * (np.log(2 * np.pi * vt) + (y - mout) ** 2 / vt))
    logZ += y.shape[0] * D * (0.5 * np.log(2 * np.pi * sn2 / alpha)
                              - 0.5 * alpha * np.log(2 * np.pi * sn2))
    dm = (y - mout) / vt
    dv = -0.5 / vt + 0.5 * (y - mout) ** 2 / vt ** 2
    return float(np.ravel(logZ)[0]), dm, dv


def _lse_weights(logZ):
    """Per-sample log-sum-exp over axis 0 minus log S, and the softmax weights (the
    3-D branches of lik_layers.py:139-152 and aep_models.py:1356-1367)."""
    mx = np.max(logZ, axis=0)
    e = np.exp(logZ - mx)
    se = np.sum(e, axis=0)
    return mx + np.log(se) - np.log(logZ.shape[0]), e / se


def gauss_logZ_mc(sn, mout, vout, y, alpha):
    """Gauss_Layer.compute_log_Z, 3-D branch (lik_layers.py:139-152): a log-mean-exp over
    the S Monte Carlo samples of the per-sample tilted normaliser."""
    sn2 = np.exp(2.0 * sn)
    vt = vout + sn2 / alpha
    lz = -0.5 * (np.log(2 * np.pi * vt) + (y - mout) ** 2 / vt)
    lz = lz + (0.5 * np.log(2 * np.pi * sn2 / alpha) - 0.5 * alpha * np.log(2 * np.pi * sn2))
    lse, w = _lse_weights(lz)
    dm = w * (y - mout) / vt
    dv = w * (-0.5 / vt + 0.5 * (y - mout) ** 2 / vt ** 2)
    return float(np.sum(lse)), dm, dv


def probit_logZ(mout, vout, y, alpha):
    """Probit_Layer.compute_log_Z (lik_layers.py:303-411), 2-D (B, D) or 3-D (S, B, D)
    inputs: the closed form at alpha = 1, Gauss-Hermite quadrature of degree GH_DEGREE
    otherwise, and the small constants the reference adds (1e-16, or 1e-8 in the 2-D
    quadrature branch, to Z and to the gradients)."""
    from scipy import special
    three = mout.ndim == 3
    if alpha == 1.0:
        t = y * mout / np.sqrt(1 + vout)
        Z = 0.5 * (1 + special.erf(t / np.sqrt(2)))
        eps = 1e-16
        lz = np.log(Z + eps)
        g = 1 / (Z + eps) / np.sqrt(2 * np.pi) * np.exp(-t ** 2.0 / 2)
        dm = g * (y / np.sqrt(1 + vout))
        dv = g * (-0.5 * y * mout / (1 + vout) ** 1.5)
        if not three:
            return float(np.sum(lz)), dm, dv
        lse, w = _lse_weights(lz)
        return float(np.sum(lse)), w * dm, w * dv
    gx, gw = np.polynomial.hermite.hermgauss(GH_DEGREE)
    shape = (-1,) + (1,) * mout.ndim
    gx, gw = gx.reshape(shape), gw.reshape(shape)
    eps = 1e-16 if three else 1e-8
    ts = gx * np.sqrt(2 * vout)[None] + mout[None]
    pdfs = 0.5 * (1 + special.erf(y * ts / np.sqrt(2))) + eps
    Zt = np.sum(pdfs ** alpha * gw, axis=0) / np.sqrt(np.pi)
    a = pdfs ** (alpha - 1.0) * np.exp(-ts ** 2 / 2)
    dZdm = np.sum(gw * a, axis=0) * y * alpha / np.pi / np.sqrt(2)
    dZdv = np.sum(gw * (a * gx), axis=0) * y * alpha / np.pi / np.sqrt(2) / np.sqrt(2 * vout)
    if not three:
        return float(np.sum(np.log(Zt))), dZdm / Zt + eps, dZdv / Zt + eps
    lse, w = _lse_weights(np.log(Zt))
    return float(np.sum(lse)), w * dZdm / Zt + eps, w * dZdv / Zt + eps


def gauss_dsn(sn, dv, B, D, alpha, scale):
    """Gauss_Layer.backprop_grads (lik_layers.py:154-181)."""
    sn2 = np.exp(2.0 * sn)
    return (np.sum(dv) * 2 * sn2 / alpha + B * D * (1 - alpha)) * scale


def emission_tilted(C, Rpar, y, mx, vx, alpha, scale):
    """Gauss_Emis.compute_emission_tilted (lik_layers.py:573-627) with R = exp(2 R_param)
    (:501-511).  Returns (logZ*scale, {'mx','vx'}, {'C','R'})."""
    R = np.exp(2 * Rpar)
    Nb = mx.shape[0]
    Dout = C.shape[0]
    Vy = np.diag(R / alpha) + ein('da,na,ab->ndb', C, vx, C.T)
    Ydiff = y - ein('da,na->nd', C, mx)
    VinvY = _solve_vec(Vy, Ydiff)
    quad_term = -0.5 * np.sum(Ydiff * VinvY)
    CVC = ein('da,na,ab->ndb', C, vx, C.T)
    ICVCR = np.eye(Dout)[None] + alpha * (CVC / R)
    Vlogdet_term = -0.5 * np.sum(np.linalg.slogdet(ICVCR)[1])
    const_term = -Nb * Dout * 0.5 * alpha * np.log(2 * np.pi)
    Rlogdet_term = -0.5 * Nb * alpha * np.sum(np.log(R))
    logZ = const_term + Rlogdet_term + Vlogdet_term + quad_term
    Vyinv = _inv(Vy)
    dR = (-0.5 * np.sum(np.diagonal(Vyinv, axis1=1, axis2=2), axis=0)
          + 0.5 * np.sum(VinvY ** 2, axis=0)) / alpha
    dR += 0.5 * Nb * (1 - alpha) / R
    dR *= 2 * R
    dSigma = -0.5 * Vyinv + 0.5 * ein('na,nb->nab', VinvY, VinvY)
    dC = ein('na,nb->ab', VinvY, mx) + 2 * ein('nc,bc,nab->ac', vx, C, dSigma)
    dmx = ein('na,ab->nb', VinvY, C)
    dvx = ein('nab,da,db->nd', dSigma, C.T, C.T)
    return logZ * scale, {'mx': dmx * scale, 'vx': dvx * scale}, {'C': dC * scale, 'R': dR * scale}


def emission_predict(C, Rpar, mf, vf):
    """Gauss_Emis.output_probabilistic (lik_layers.py:558-571)."""
    R = np.exp(2 * Rpar)
    my = ein('ab,nb->na', C, mf)
    vy_noiseless = ein('ab,nb,bc->nac', C, vf, C.T)
    return my, vy_noiseless, vy_noiseless + np.diag(R)


# ----------------------------------------------------------------------------
# Models (aep_models.py:589-1437)
# ----------------------------------------------------------------------------

def _finish(energy, grads, fixed, divide_by):
    for p in fixed:
        grads[p] = np.zeros_like(grads[p])
    if divide_by is not None:
        energy = energy / divide_by
        grads = {k: v / divide_by for k, v in grads.items()}
    return float(np.ravel(energy)[0]), grads


def _lik(name, params, mout, vout, y, alpha, B, scale):
    """(logZ, dm, dv, lik grads) of the likelihood layer: Gauss_Layer (with its sn
    gradient) or Probit_Layer (no hyper-parameters; lik_layers.py:285-471)."""
    if name.lower() == 'probit':
        logZ, dm, dv = probit_logZ(mout, vout, y, alpha)
        return logZ, dm, dv, {}
    logZ, dm, dv = gauss_logZ(params['sn'], mout, vout, y, alpha)
    return logZ, dm, dv, {'sn': gauss_dsn(params['sn'], dv, B, y.shape[1], alpha, scale)}


def sgpr_objective(x, y, params, M, mb_size, alpha=1.0, nat_param=True, fixed=(), idxs=None,
                   lik='Gaussian'):
    """SGPR.objective_function (aep_models.py:610-667).  Draws the minibatch with
    np.random.choice(N, mb_size, replace=False) like :628 unless ``idxs`` is given."""
    N = y.shape[0]
    if idxs is None and mb_size < N:
        idxs = np.random.choice(N, mb_size, replace=False)
    xb, yb = (x, y) if idxs is None else (x[idxs], y[idxs])
    B = yb.shape[0]
    scale_logZ = -N * 1.0 / B / alpha
    L = layer_update(params, '', M, N, nat_param)
    layer_cavity(L, alpha)
    mout, vout, kfu = fwd_det_cav(L, xb)
    logZ, dm, dv, glik = _lik(lik, params, mout, vout, yb, alpha, B, scale_logZ)
    grads = bwd_det(L, scale_logZ * dm, scale_logZ * dv, kfu, xb, alpha)
    grads.update(glik)
    energy = scale_logZ * logZ + layer_phi(L, alpha)
    return _finish(energy, grads, fixed, N)


def sdgpr_objective(x, y, params, Ms, mb_size, alpha=1.0, fixed=(), idxs=None, lik='Gaussian'):
    """SDGPR.objective_function (aep_models.py:895-988); layer i uses suffix '_i'."""
    N = y.shape[0]
    Lnum = len(Ms)
    if idxs is None and mb_size < N:
        idxs = np.random.choice(N, mb_size, replace=False)
    xb, yb = (x, y) if idxs is None else (x[idxs], y[idxs])
    B = yb.shape[0]
    scale_logZ = -N * 1.0 / B / alpha
    layers = [layer_update(params, '_%d' % i, Ms[i], N, True) for i in range(Lnum)]
    for Li in layers:
        layer_cavity(Li, alpha)
    ms, vs, p1, p2 = [], [], [], []
    for i, Li in enumerate(layers):
        if i == 0:
            m, v, kfu = fwd_det_cav(Li, xb)
            ms.append(m); vs.append(v); p1.append(kfu); p2.append(None)
        else:
            m, v, a, b = fwd_mm_cav(Li, ms[i - 1], vs[i - 1])
            ms.append(m); vs.append(v); p1.append(a); p2.append(b)
    logZ, dm, dv, glik = _lik(lik, params, ms[-1], vs[-1], yb, alpha, B, scale_logZ)
    dmi, dvi = scale_logZ * dm, scale_logZ * dv
    grads = {}
    for i in range(Lnum - 1, -1, -1):
        Li = layers[i]
        if i == 0:
            g = bwd_det(Li, dmi, dvi, p1[0], xb, alpha)
        else:
            g, gin = bwd_mm(Li, ms[i], dmi, dvi, p1[i], p2[i], ms[i - 1], vs[i - 1], alpha)
            dmi, dvi = gin['mx'], gin['vx']
        for k, val in g.items():
            grads[k + '_%d' % i] = val
    grads.update(glik)
    energy = scale_logZ * logZ + sum(layer_phi(Li, alpha) for Li in layers)
    return _finish(energy, grads, fixed, N)


def sgplvm_objective(y, params, M, mb_size, alpha=1.0, prior_mean=0.0, prior_var=1.0,
                     nat_param=True, fixed=(), idxs=None, prop_mode=PROP_MM, eps=None,
                     lik='Gaussian'):
    """SGPLVM.objective_function (aep_models.py:694-815) with x-factor helpers
    (:817-867) and Base_SGPLVM.update_hypers / compute_posterior_grad_x
    (base_models.py:895-929).  The energy is NOT divided by N (:808-815)."""
    N = y.shape[0]
    if idxs is None:
        idxs = np.arange(N) if mb_size == N else np.random.choice(N, mb_size, replace=False)
    yb = y[idxs]
    B = yb.shape[0]
    Din = params['x1'].shape[1]
    scale_logZ = -N * 1.0 / B / alpha
    L = layer_update(params, '', M, N, nat_param)
    prior_x1 = prior_mean / prior_var
    prior_x2 = 1.0 / prior_var
    factor_x1 = params['x1']
    factor_x2 = np.exp(2 * params['x2'])
    if nat_param:
        x_post_1 = prior_x1 + factor_x1
        x_post_2 = prior_x2 + factor_x2
    else:
        x_post_1 = factor_x1 / factor_x2
        x_post_2 = 1.0 / factor_x2
    layer_cavity(L, alpha)
    # get_cavity_x (:840-861)
    if nat_param:
        cav_t1 = prior_x1 + (1.0 - alpha) * factor_x1[idxs]
        cav_t2 = prior_x2 + (1.0 - alpha) * factor_x2[idxs]
    else:
        mp, vp = factor_x1[idxs], factor_x2[idxs]
        cav_t1 = prior_x1 + (mp / vp - prior_x1) * (1 - alpha)
        cav_t2 = prior_x2 + (1 / vp - prior_x2) * (1 - alpha)
    vcav = 1.0 / cav_t2
    mcav = cav_t1 / cav_t2
    vpost = 1.0 / x_post_2[idxs]
    mpost = x_post_1[idxs] / x_post_2[idxs]
    if prop_mode == PROP_MC:
        # :745-760; the draws follow the minibatch draw on the legacy global RNG
        if eps is None:
            eps = np.random.randn(MC_NO_SAMPLES, B, Din)
        m3, v3, kfu_s, x_s = fwd_mc_cav(L, mcav, vcav, eps)
        if lik.lower() == 'probit':
            logZ, dm, dv = probit_logZ(m3, v3, yb, alpha)
        else:
            logZ, dm, dv = gauss_logZ_mc(params['sn'], m3, v3, yb, alpha)
        grads, dx = bwd_mc(L, scale_logZ * dm, scale_logZ * dv, kfu_s, x_s, alpha)
        gmx, gvx = reparam_bwd(dx, vcav, eps)
        gin = {'mx': gmx, 'vx': gvx}
    else:
        mout, vout, psi1, psi2 = fwd_mm_cav(L, mcav, vcav)
        if lik.lower() == 'probit':
            logZ, dm, dv = probit_logZ(mout, vout, yb, alpha)
        else:
            logZ, dm, dv = gauss_logZ(params['sn'], mout, vout, yb, alpha)
        grads, gin = bwd_mm(L, mout, scale_logZ * dm, scale_logZ * dv, psi1, psi2, mcav, vcav, alpha)
    if lik.lower() != 'probit':
        grads['sn'] = gauss_dsn(params['sn'], dv, B, y.shape[1], alpha, scale_logZ)

    def phi_x(m, v):   # compute_phi_x (:863-867)
        return np.sum(0.5 * (m ** 2 / v + np.log(v))), m / v, 0.5 * (-m ** 2 / v ** 2 + 1 / v)
    phi_prior_x = 0.5 * (prior_mean ** 2 / prior_var + np.log(prior_var)) * N * Din
    phi_cav_x, dmc, dvc = phi_x(mcav, vcav)
    phi_post_x, dmp, dvp = phi_x(mpost, vpost)
    sc_cav = -N * 1.0 / B / alpha
    sc_post = -N * 1.0 / B * (1.0 - 1.0 / alpha)
    x_contrib = phi_prior_x + sc_cav * phi_cav_x + sc_post * phi_post_x
    dmc = sc_cav * dmc + gin['mx']
    dvc = sc_cav * dvc + gin['vx']
    dmp = sc_post * dmp
    dvp = sc_post * dvp
    # compute_cav_grad_x (:817-838)
    t1 = mcav / vcav
    t2 = 1 / vcav
    dt1 = (1.0 - alpha) * dmc / t2
    dt2 = (1.0 - alpha) * (-dmc * t1 / t2 ** 2 - dvc / t2 ** 2)
    if nat_param:
        dt2 = dt2 * 2 * factor_x2[idxs]
    else:
        mp, vp = factor_x1[idxs], factor_x2[idxs]
        dmm = dt1 / vp
        dvv = -dt1 * mp / vp ** 2 - dt2 / vp ** 2
        dt1 = dmm
        dt2 = dvv * 2 * factor_x2[idxs]
    g1 = np.zeros_like(factor_x1)
    g2 = np.zeros_like(factor_x2)
    g1[idxs] = dt1
    g2[idxs] = dt2
    # Base_SGPLVM.compute_posterior_grad_x (base_models.py:913-929)
    h1 = np.zeros_like(x_post_1)
    h2 = np.zeros_like(x_post_1)
    if nat_param:
        p1, p2 = x_post_1[idxs], x_post_2[idxs]
        h1[idxs] = dmp / p2
        h2[idxs] = (-dmp * p1 / p2 ** 2 - dvp / p2 ** 2) * (2 * factor_x2[idxs])
    else:
        h1[idxs] = dmp
        h2[idxs] = dvp * (2 * factor_x2[idxs])
    grads['x1'] = g1 + h1
    grads['x2'] = g2 + h2
    energy = scale_logZ * logZ + x_contrib + layer_phi(L, alpha)
    return _finish(energy, grads, fixed, None)


def transition_tilted(sn, m_prop, v_prop, m_t, v_t, alpha, scale, Din):
    """SGPSSM.compute_transition_tilted, 2-D branch (aep_models.py:1317-1343)."""
    sn2 = np.exp(2 * sn)
    v_sum = v_t + v_prop + sn2 / alpha
    m_diff = m_t - m_prop
    logZ = -0.5 * m_diff ** 2 / v_sum - 0.5 * np.log(1 + alpha * (v_t + v_prop) / sn2) \
        - 0.5 * alpha * np.log(2 * np.pi * sn2)
    logZ = scale * np.sum(logZ)
    dvt = scale * (-0.5 / v_sum + 0.5 * m_diff ** 2 / v_sum ** 2)
    dmt = scale * (-m_diff / v_sum)
    dsn = np.sum(dvt) * 2 * sn2 / alpha + scale * m_prop.shape[0] * Din * (1 - alpha)
    return logZ, -dmt, dvt, dmt, dvt, dsn


def transition_tilted_mc(sn, m_prop, v_prop, m_t, v_t, alpha, scale, Din):
    """SGPSSM.compute_transition_tilted, 3-D branch (aep_models.py:1344-1372): m_prop,
    v_prop are (S, B, D) Monte Carlo propagations, m_t, v_t the (B, D) cavities.
    Returns (logZ*scale, dmprop, dvprop, dmt, dvt, dsn), all scaled."""
    sn2 = np.exp(2 * sn)
    v_sum = v_t + v_prop + sn2 / alpha
    m_diff = m_t - m_prop
    lz = -0.5 * m_diff ** 2 / v_sum - 0.5 * np.log(1 + alpha * (v_t + v_prop) / sn2) \
        - 0.5 * alpha * np.log(2 * np.pi * sn2)
    lse, w = _lse_weights(lz)
    dl = scale * w
    dl_dm = dl * m_diff / v_sum
    dl_dv = dl * (-0.5 / v_sum + 0.5 * m_diff ** 2 / v_sum ** 2)
    dsn = np.sum(dl_dv) * 2 * sn2 / alpha + scale * m_prop.shape[1] * Din * (1 - alpha)
    return scale * np.sum(lse), dl_dm, dl_dv, -np.sum(dl_dm, axis=0), np.sum(dl_dv, axis=0), dsn


def sgpssm_objective(y, params, M, Din, mb_size, alpha=1.0, prior_mean=0.0, prior_var=1.0,
                     x_control=None, gp_emi=False, control_to_emi=True, fixed=(),
                     start_idx=None, prop_mode=PROP_MM, eps=None):
    """SGPSSM.objective_function (aep_models.py:1024-1206) with Base_SGPSSM.update_hypers
    (base_models.py:1705-1728) and the x-factor helpers (aep_models.py:1208-1437).
    The dynamics layer is built with N-1 training points (:1018)."""
    N, Dout = y.shape
    Dcon = 0 if x_control is None else x_control.shape[1]
    Dcon_emi = Dcon if control_to_emi else 0
    if mb_size >= N:
        dyn_idxs = np.arange(0, N - 1)
        emi_idxs = np.arange(0, N)
        yb = y
    else:
        if start_idx is None:
            start_idx = np.random.randint(0, N - mb_size)
        end_idx = start_idx + mb_size
        emi_idxs = np.arange(start_idx, end_idx)
        dyn_idxs = np.arange(start_idx, end_idx - 1)
        yb = y[emi_idxs]
    Bd = dyn_idxs.shape[0]
    Be = emi_idxs.shape[0]
    scale_dyn = -(N - 1) * 1.0 / Bd / alpha
    scale_emi = -N * 1.0 / Be / alpha
    dyn = layer_update(params, '_dynamic', M, N - 1, True)
    emi = layer_update(params, '_emission', M, N, True) if gp_emi else None
    sn = params['sn']
    x_prior_1 = prior_mean / prior_var
    x_prior_2 = 1.0 / prior_var
    xf1 = params['x_factor_1']
    xf2 = np.exp(2 * params['x_factor_2'])
    x_post_1 = 3 * xf1
    x_post_2 = 3 * xf2
    x_post_1[[0, -1], :] = 2 * xf1[[0, -1], :]
    x_post_2[[0, -1], :] = 2 * xf2[[0, -1], :]
    x_post_1[0, :] += x_prior_1
    x_post_2[0, :] += x_prior_2
    layer_cavity(dyn, alpha)
    if gp_emi:
        layer_cavity(emi, alpha)
    cav_1 = x_post_1 - alpha * xf1
    cav_2 = x_post_2 - alpha * xf2
    cav_m = cav_1 / (cav_2 + 1e-16)
    cav_v = 1.0 / (cav_2 + 1e-16)
    cav_t_m, cav_t_v = cav_m[dyn_idxs + 1], cav_v[dyn_idxs + 1]
    cav_tm1_m, cav_tm1_v = cav_m[dyn_idxs], cav_v[dyn_idxs]
    if Dcon > 0:
        cav_tm1_mc = np.hstack((cav_tm1_m, x_control[dyn_idxs]))
        cav_tm1_vc = np.hstack((cav_tm1_v, np.zeros((Bd, Dcon))))
    else:
        cav_tm1_mc, cav_tm1_vc = cav_tm1_m, cav_tm1_v
    cav_up_m, cav_up_v = cav_m[emi_idxs], cav_v[emi_idxs]
    if Dcon_emi > 0:
        cav_up_mc = np.hstack((cav_up_m, x_control[emi_idxs]))
        cav_up_vc = np.hstack((cav_up_v, np.zeros((Be, Dcon_emi))))
    else:
        cav_up_mc, cav_up_vc = cav_up_m, cav_up_v
    mc = prop_mode == PROP_MC
    if mc:
        # :1114-1143; dynamics draws first, then the GP emission's
        eps_dyn = np.random.randn(MC_NO_SAMPLES, Bd, Din + Dcon) if eps is None else eps[0]
        m3, v3, kfu_s, x_s = fwd_mc_cav(dyn, cav_tm1_mc, cav_tm1_vc, eps_dyn)
        logZ_dyn, dmprop, dvprop, dmt, dvt, dsn = transition_tilted_mc(
            sn, m3, v3, cav_t_m, cav_t_v, alpha, scale_dyn, Din)
        g_dyn, dx = bwd_mc(dyn, dmprop, dvprop, kfu_s, x_s, alpha)
        gmx, gvx = reparam_bwd(dx, cav_tm1_vc, eps_dyn)
        gin_dyn = {'mx': gmx, 'vx': gvx}
    else:
        mprop, vprop, psi1, psi2 = fwd_mm_cav(dyn, cav_tm1_mc, cav_tm1_vc)
        logZ_dyn, dmprop, dvprop, dmt, dvt, dsn = transition_tilted(
            sn, mprop, vprop, cav_t_m, cav_t_v, alpha, scale_dyn, Din)
        g_dyn, gin_dyn = bwd_mm(dyn, mprop, dmprop, dvprop, psi1, psi2, cav_tm1_mc, cav_tm1_vc,
                                alpha)
    grads = {'sn': dsn}
    for k, v in g_dyn.items():
        grads[k + '_dynamic'] = v
    if gp_emi and mc:
        eps_emi = np.random.randn(MC_NO_SAMPLES, Be, Din + Dcon_emi) if eps is None else eps[1]
        m3, v3, kfu_s, x_s = fwd_mc_cav(emi, cav_up_mc, cav_up_vc, eps_emi)
        logZ_e, dm, dv = gauss_logZ_mc(params['sn_emission'], m3, v3, yb, alpha)
        logZ_emi = scale_emi * logZ_e
        g_emi, dx = bwd_mc(emi, scale_emi * dm, scale_emi * dv, kfu_s, x_s, alpha)
        gmx, gvx = reparam_bwd(dx, cav_up_vc, eps_emi)
        gin_emi = {'mx': gmx, 'vx': gvx}
    elif gp_emi:
        mout, vout, e1, e2 = fwd_mm_cav(emi, cav_up_mc, cav_up_vc)
        logZ_e, dm, dv = gauss_logZ(params['sn_emission'], mout, vout, yb, alpha)
        logZ_emi = scale_emi * logZ_e
        g_emi, gin_emi = bwd_mm(emi, mout, scale_emi * dm, scale_emi * dv, e1, e2,
                                cav_up_mc, cav_up_vc, alpha)
    if gp_emi:
        for k, v in g_emi.items():
            grads[k + '_emission'] = v
        grads['sn_emission'] = gauss_dsn(params['sn_emission'], dv, Be, Dout, alpha, scale_emi)
    else:
        logZ_emi, gin_emi, g_emi = emission_tilted(
            params['C_emission'], params['R_emission'], y[emi_idxs], cav_up_mc, cav_up_vc,
            alpha, scale_emi)
        for k, v in g_emi.items():
            grads[k + '_emission'] = v
    dmcav_up = gin_emi['mx'][:, :Din]
    dvcav_up = gin_emi['vx'][:, :Din]
    dmcav_next = gin_dyn['mx'][:, :Din]
    dvcav_next = gin_dyn['vx'][:, :Din]
    # compute_posterior_grad_x (:1208-1232)
    dpost_1 = x_post_1 / x_post_2
    dpost_2 = -0.5 * x_post_1 ** 2 / x_post_2 ** 2 - 0.5 / x_post_2
    sxp = -(1.0 - 1.0 / alpha) * np.ones((N, 1))
    sxp[0:N - 1] += 1.0 / alpha
    sxp[1:N] += 1.0 / alpha
    gp1 = 3.0 * (sxp * dpost_1)
    gp2 = 6.0 * (sxp * dpost_2) * xf2
    gp1[[0, -1], :] = 2.0 * (sxp * dpost_1)[[0, -1], :]
    gp2[[0, -1], :] = 4.0 * (sxp * dpost_2)[[0, -1], :] * xf2[[0, -1], :]
    # compute_cavity_grad_x (:1287-1315) and compute_logZ_grad_x (:1234-1285) share
    # the cavity chain rule factor (3 - alpha), (2 - alpha) at the ends.
    sxc = (3.0 - alpha) * np.ones((N, 1))
    sxc[0] = 2.0 - alpha
    sxc[-1] = 2.0 - alpha
    sc = -1.0 / alpha * np.ones((N, 1))
    sc[0:N - 1] += -1.0 / alpha
    sc[1:N] += -1.0 / alpha
    dcav_1 = cav_1 / cav_2
    dcav_2 = -0.5 * cav_1 ** 2 / cav_2 ** 2 - 0.5 / cav_2
    gc1 = sc * dcav_1 * sxc
    gc2 = sc * dcav_2 * sxc * 2 * xf2
    gz1 = np.zeros_like(cav_1)
    gz2 = np.zeros_like(cav_1)
    gz1[emi_idxs] = dmcav_up / cav_2[emi_idxs]
    gz2[emi_idxs] = -dmcav_up * cav_1[emi_idxs] / cav_2[emi_idxs] ** 2 \
        - dvcav_up / cav_2[emi_idxs] ** 2
    ip = np.arange(emi_idxs[0] + 1, emi_idxs[-1] + 1)
    gz1[ip] += dmt / cav_2[ip]
    gz2[ip] += -dmt * cav_1[ip] / cav_2[ip] ** 2 - dvt / cav_2[ip] ** 2
    inx = np.arange(emi_idxs[0], emi_idxs[-1])
    gz1[inx] += dmcav_next / cav_2[inx]
    gz2[inx] += -dmcav_next * cav_1[inx] / cav_2[inx] ** 2 - dvcav_next / cav_2[inx] ** 2
    gz1 = gz1 * sxc
    gz2 = gz2 * sxc * 2 * xf2
    grads['x_factor_1'] = gp1 + gc1 + gz1
    grads['x_factor_2'] = gp2 + gc2 + gz2
    # phi terms (:1389-1437)
    m0 = x_prior_1 / x_prior_2
    v0 = 1.0 / x_prior_2
    phi_prior_x = 0.5 * Din * (m0 ** 2 / v0 + np.log(v0))
    phi_post = 0.5 * (x_post_1 ** 2 / x_post_2 - np.log(x_post_2))
    phi_post_x = np.sum(sxp * phi_post)
    phi_cav = 0.5 * (cav_1 ** 2 / cav_2 - np.log(cav_2))
    phi_cav_x = np.sum(sc * phi_cav)
    x_contrib = phi_prior_x + phi_post_x + phi_cav_x
    energy = logZ_dyn + logZ_emi + x_contrib + layer_phi(dyn, alpha) \
        + (layer_phi(emi, alpha) if gp_emi else 0.0)
    return _finish(energy, grads, fixed, N)


def transition_tilted_h(sn, m_prop, v_prop, m_t, v_t, alpha, scale, D):
    """SDGPR_H.compute_transition_tilted (aep_models.py:1702-1735): the hidden-layer
    transition normaliser, all outputs times scale."""
    sn2 = np.exp(2 * sn)
    v_sum = v_t + v_prop + sn2 / alpha
    m_diff = m_t - m_prop
    logZ = np.sum(-0.5 * m_diff ** 2 / v_sum - 0.5 * np.log(2 * np.pi * v_sum)
                  + 0.5 * (1 - alpha) * np.log(2 * np.pi * sn2) - 0.5 * np.log(alpha))
    dv = -0.5 / v_sum + 0.5 * m_diff ** 2 / v_sum ** 2
    dsn = np.sum(dv) * 2 * sn2 / alpha + m_prop.shape[0] * D * (1 - alpha)
    return scale * logZ, scale * m_diff / v_sum, scale * dv, -scale * m_diff / v_sum, scale * dv, scale * dsn


def sdgpr_h_objective(x, y, params, Ms, mb_size, alpha=1.0, fixed=(), lik='Gaussian'):
    """SDGPR_H.objective_function (aep_models.py:1506-1618): a deep GP whose hidden
    layer values carry Gaussian factors (h_factor_1_i, h_factor_2_i), tied twice per value
    (compute_cavity_h :1737-1752, compute_grads_hidden :1620-1660, log-partitions
    :1662-1700).  Full batch only: the reference's hidden-factor gradients broadcast the
    batch rows against all N rows (:1633-1652), which fails for a proper minibatch."""
    N = y.shape[0]
    if mb_size < N:
        raise NotImplementedError('SDGPR_H objective is defined for the full batch only')
    Lnum = len(Ms)
    scale_logZ = -1.0 / alpha
    layers = [layer_update(params, '_%d' % i, Ms[i], N, True) for i in range(Lnum)]
    for Li in layers:
        layer_cavity(Li, alpha)
    h1 = [np.asarray(params['h_factor_1_%d' % i]) for i in range(Lnum - 1)]
    h2 = [np.exp(2 * np.asarray(params['h_factor_2_%d' % i])) for i in range(Lnum - 1)]
    cav_1 = [a * (2.0 - alpha) for a in h1]
    cav_2 = [b * (2.0 - alpha) for b in h2]
    cav_m = [a / b for a, b in zip(cav_1, cav_2)]
    cav_v = [1.0 / b for b in cav_2]
    sn = np.asarray(params['sn_hidden'], dtype=np.float64)
    dsn = np.zeros_like(sn)
    grads, dmc, dvc = {}, [], []
    logZ = 0.0
    for i in range(Lnum - 1):
        Li = layers[i]
        if i == 0:
            mp, vp, kfu = fwd_det_cav(Li, x)
        else:
            mp, vp, psi1, psi2 = fwd_mm_cav(Li, cav_m[i - 1], cav_v[i - 1])
        lz, dmp, dvp, dmi, dvi, dsni = transition_tilted_h(sn[i], mp, vp, cav_m[i], cav_v[i], alpha,
                                                           scale_logZ, Li.Dout)
        if i == 0:
            g = bwd_det(Li, dmp, dvp, kfu, x, alpha)
        else:
            g, gin = bwd_mm(Li, mp, dmp, dvp, psi1, psi2, cav_m[i - 1], cav_v[i - 1], alpha)
            dmc[i - 1] = dmc[i - 1] + gin['mx']
            dvc[i - 1] = dvc[i - 1] + gin['vx']
        logZ += lz
        for k, val in g.items():
            grads[k + '_%d' % i] = val
        dmc.append(dmi)
        dvc.append(dvi)
        dsn[i] = dsni
    i = Lnum - 1
    Li = layers[-1]
    mp, vp, psi1, psi2 = fwd_mm_cav(Li, cav_m[i - 1], cav_v[i - 1])
    lz, dm, dv, glik = _lik(lik, params, mp, vp, y, alpha, N, scale_logZ)
    g, gin = bwd_mm(Li, mp, scale_logZ * dm, scale_logZ * dv, psi1, psi2, cav_m[i - 1], cav_v[i - 1],
                    alpha)
    logZ += scale_logZ * lz
    dmc[i - 1] = dmc[i - 1] + gin['mx']
    dvc[i - 1] = dvc[i - 1] + gin['vx']
    for k, val in g.items():
        grads[k + '_%d' % i] = val
    grads.update(glik)
    sp, sc = -(1.0 - 1.0 / alpha), -1.0 / alpha
    phi_h = 0.0
    for j in range(Lnum - 1):
        p1, p2 = 2 * h1[j], 2 * h2[j]
        d1_post = 2 * (p1 / p2)
        d2_post = -p1 ** 2 / p2 ** 2 - 1 / p2
        d1_cav = (2 - alpha) * (cav_1[j] / cav_2[j])
        d2_cav = (2 - alpha) * (-0.5 * cav_1[j] ** 2 / cav_2[j] ** 2 - 0.5 / cav_2[j])
        d1_z = (2 - alpha) * (dmc[j] / cav_2[j])
        d2_z = (2 - alpha) * (-dmc[j] * cav_1[j] / cav_2[j] ** 2 - dvc[j] / cav_2[j] ** 2)
        grads['h_factor_1_%d' % j] = d1_z + sc * d1_cav + sp * d1_post
        grads['h_factor_2_%d' % j] = 2 * (d2_z + sc * d2_cav + sp * d2_post) * h2[j]
        phi_h += sp * np.sum(0.5 * (p1 ** 2 / p2 - np.log(p2))) \
            + sc * np.sum(0.5 * (cav_1[j] ** 2 / cav_2[j] - np.log(cav_2[j])))
    grads['sn_hidden'] = dsn
    energy = logZ + sum(layer_phi(Lj, alpha) for Lj in layers) + phi_h
    return _finish(energy, grads, fixed, N)


def sdgpr_h_predict(params, Ms, N, xs, noisy=False):
    """SDGPR_H.predict_f / predict_y (aep_models.py:1754-1815): as SDGPR."""
    return sdgpr_predict(params, Ms, N, xs, noisy)


# ----------------------------------------------------------------------------
# Prediction (base_models.py:235-307, 734-761, 985-1034, 1140-1158, 1264-1275,
# 1441-1560)
# ----------------------------------------------------------------------------

def sgpr_predict(params, M, N, xs, nat_param=True, noisy=False):
    L = layer_update(params, '', M, N, nat_param)
    mf, vf = fwd_det_post(L, xs)
    if noisy:
        vf = vf + np.exp(2.0 * params['sn'])
    return mf, vf


def sdgpr_predict(params, Ms, N, xs, noisy=False):
    mf = vf = None
    for i, Mi in enumerate(Ms):
        L = layer_update(params, '_%d' % i, Mi, N, True)
        if i == 0:
            mf, vf = fwd_det_post(L, xs)
        else:
            mf, vf = fwd_mm_post(L, mf, vf)
    if noisy:
        vf = vf + np.exp(2.0 * params['sn'])
    return mf, vf


def sdgpr_predict_mc(params, Ms, N, xs, no_samples):
    """Base_SDGPR.predict_f_mc (base_models.py:1159-1180): samples pushed through the
    layers' posteriors with draws from the legacy global RNG -- (no_samples, N, D_1) for
    the first layer, one per sample and output for each later one.  Returns (samples,
    mf, vf) of the last layer."""
    samples = xs
    mf = vf = None
    for i, Mi in enumerate(Ms):
        L = layer_update(params, '_%d' % i, Mi, N, True)
        mf, vf = fwd_det_post(L, samples)
        if i == 0:
            eps = np.random.randn(no_samples, mf.shape[0], mf.shape[1])
            samples = (np.sqrt(vf) * eps + mf).reshape(no_samples * mf.shape[0], mf.shape[1])
        else:
            eps = np.random.randn(mf.shape[0], mf.shape[1])
            samples = np.sqrt(vf) * eps + mf
    return samples.reshape(no_samples, xs.shape[0], -1), mf, vf


def sgplvm_predict(params, M, N, xs, noisy=False):
    """Base_SGPLVM.predict_f / predict_y (base_models.py:734-761): the layer's posterior
    at deterministic inputs, plus exp(2 sn) for predict_y (lik_layers.py:238-249)."""
    L = layer_update(params, '', M, N, True)
    mf, vf = fwd_det_post(L, xs)
    if noisy:
        vf = vf + np.exp(2.0 * params['sn'])
    return mf, vf


def sgpssm_predict(params, M, T, xs, noisy=False, gp_emi=False):
    """Base_SGPSSM.predict_f (base_models.py:1441-1451): the dynamics layer's posterior
    at deterministic inputs.  predict_y (:1544-1560): through the GP emission layer's
    moment-matched posterior plus exp(2 sn_emission), or through the linear-Gaussian
    emission, keeping the diagonal of C diag(vf) C^T + diag(R) (lik_layers.py:558-571)."""
    dyn = layer_update(params, '_dynamic', M, T - 1, True)
    mf, vf = fwd_det_post(dyn, xs)
    if not noisy:
        return mf, vf
    if gp_emi:
        emi = layer_update(params, '_emission', M, T, True)
        mg, vg = fwd_mm_post(emi, mf, vf)
        return mg, vg + np.exp(2.0 * params['sn_emission'])
    my, _, vy = emission_predict(params['C_emission'], params['R_emission'], mf, vf)
    return my, np.diagonal(vy, axis1=1, axis2=2).copy()


def sgpssm_posterior_x(params, prior_mean=0.0, prior_var=1.0):
    """Base_SGPSSM.get_posterior_x (base_models.py:1562-1576) over the tied factors of
    update_hypers (:1705-1728)."""
    xf1 = np.asarray(params['x_factor_1'])
    xf2 = np.exp(2 * np.asarray(params['x_factor_2']))
    p1, p2 = 3 * xf1, 3 * xf2
    p1[[0, -1], :] = 2 * xf1[[0, -1], :]
    p2[[0, -1], :] = 2 * xf2[[0, -1], :]
    p1[0, :] += prior_mean / prior_var
    p2[0, :] += 1.0 / prior_var
    return p1 / p2, 1.0 / p2


def sgpssm_posterior_y(params, M, T_train, x_control=None, control_to_emi=True, gp_emi=False,
                       prior_mean=0.0, prior_var=1.0):
    """Base_SGPSSM.get_posterior_y (base_models.py:1578-1595): (my, vf, vyn)."""
    mx, vx = sgpssm_posterior_x(params, prior_mean, prior_var)
    if x_control is not None and control_to_emi:
        mx = np.hstack((mx, x_control))
        vx = np.hstack((vx, np.zeros_like(x_control)))
    if gp_emi:
        emi = layer_update(params, '_emission', M, T_train, True)
        mf, vf = fwd_mm_post(emi, mx, vx)
        return mf, vf, vf + np.exp(2.0 * params['sn_emission'])
    my, vy, vyn = emission_predict(params['C_emission'], params['R_emission'], mx, vx)
    return my, np.diagonal(vy, axis1=1, axis2=2).copy(), np.diagonal(vyn, axis1=1, axis2=2).copy()


def sgpssm_predict_forward(params, M, T_train, Din, T, x_control=None, control_to_emi=True,
                           gp_emi=False, prior_mean=0.0, prior_var=1.0):
    """Base_SGPSSM.predict_forward_mm (base_models.py:1463-1502): T moment-matched steps of
    the dynamics posterior from the last latent state, each emitted; returns (mx, vx, my,
    vy_noiseless, vy)."""
    dyn = layer_update(params, '_dynamic', M, T_train - 1, True)
    emi = layer_update(params, '_emission', M, T_train, True) if gp_emi else None
    Dcon = 0 if x_control is None else x_control.shape[1]
    Dcon_emi = Dcon if control_to_emi else 0
    Dout = params['C_emission'].shape[0] if not gp_emi else params['eta2_emission'].shape[0]
    out = [np.zeros((T, Din)), np.zeros((T, Din)), np.zeros((T, Dout)), np.zeros((T, Dout)),
           np.zeros((T, Dout))]
    pm, pv = sgpssm_posterior_x(params, prior_mean, prior_var)
    m, v = pm[[-1], :], pv[[-1], :]
    for t in range(T):
        if Dcon > 0:
            m = np.hstack((m, x_control[[t], :]))
            v = np.hstack((v, np.zeros((1, Dcon))))
        mt, vt = fwd_mm_post(dyn, m, v)
        if Dcon_emi > 0:
            mtc, vtc = np.hstack((mt, x_control[[t], :])), np.hstack((vt, np.zeros((1, Dcon_emi))))
        else:
            mtc, vtc = mt, vt
        if gp_emi:
            mf, vf = fwd_mm_post(emi, mtc, vtc)
            my, vyn = mf, vf + np.exp(2.0 * params['sn_emission'])
        else:
            my, vy0, vyn = emission_predict(params['C_emission'], params['R_emission'], mt, vt)
            vf = np.diagonal(vy0, axis1=1, axis2=2)
            vyn = np.diagonal(vyn, axis1=1, axis2=2)
        out[0][t], out[1][t], out[2][t], out[3][t], out[4][t] = mt, vt, my, vf, vyn
        m, v = mt, vt
    return tuple(out)


def sgpssm_predict_forward_mc(params, M, T_train, Din, T, no_samples, x_control=None,
                              control_to_emi=True, gp_emi=False, prior_mean=0.0, prior_var=1.0):
    """Base_SGPSSM.predict_forward_mc (base_models.py:1504-1542): samples of the latent
    path with draws from the legacy global RNG (no_samples x Din per step, one set before
    the first step); returns (x, my, vy) with a leading time axis."""
    dyn = layer_update(params, '_dynamic', M, T_train - 1, True)
    emi = layer_update(params, '_emission', M, T_train, True) if gp_emi else None
    Dcon = 0 if x_control is None else x_control.shape[1]
    if Dcon > 0 and control_to_emi:
        # the reference's control-to-emission branch does not run (:1531, a misplaced
        # parenthesis in np.tile)
        raise NotImplementedError('Monte Carlo rollout with control inputs to the emission')
    Dout = params['C_emission'].shape[0] if not gp_emi else params['eta2_emission'].shape[0]
    x = np.zeros((T, no_samples, Din))
    my = np.zeros((T, no_samples, Dout))
    vy = np.zeros((T, no_samples, Dout))
    pm, pv = sgpssm_posterior_x(params, prior_mean, prior_var)
    eps = np.random.randn(no_samples, Din)
    xs = eps * np.sqrt(pv[[-1], :]) + pm[[-1], :]
    for t in range(T):
        xc = np.hstack((xs, np.tile(x_control[[t], :], [no_samples, 1]))) if Dcon > 0 else xs
        mt, vt = fwd_det_post(dyn, xc)
        eps = np.random.randn(no_samples, Din)
        xs = eps * np.sqrt(vt) + mt
        if gp_emi:
            mf, vf = fwd_det_post(emi, xs)
            myt, vyt = mf, vf + np.exp(2.0 * params['sn_emission'])
        else:
            myt, _, vyt = emission_predict(params['C_emission'], params['R_emission'], xs,
                                           np.zeros_like(xs))
            vyt = np.diagonal(vyt, axis1=1, axis2=2)
        x[t], my[t], vy[t] = xs, myt, vyt
    return x, my, vy


# ----------------------------------------------------------------------------
# utils.py
# ----------------------------------------------------------------------------

def flatten_dict(params):
    """utils.py:68-81: sorted keys, ravelled values, split indices."""
    keys = list(params.keys())
    shapes = {}
    vals = []
    sizes = []
    for key in sorted(keys):
        val = np.asarray(params[key])
        shapes[key] = val.shape
        vals.append(val.ravel())
        sizes.append(val.size)
    vec = np.concatenate(vals) if vals else np.array([])
    return vec, (keys, np.cumsum(sizes)[:-1], shapes)


def unflatten_dict(vec, args):
    """utils.py:84-90."""
    keys, indices, shapes = args
    vals = np.split(vec, indices)
    return {k: np.reshape(vals[i], shapes[k]) for i, k in enumerate(sorted(keys))}


def adam(func, x0, maxiter, step_size=1e-3, b1=0.9, b2=0.999, eps=1e-8, args=()):
    """utils.py:93-117 without printing: bias-corrected Adam."""
    x = x0
    m = np.zeros_like(x)
    v = np.zeros_like(x)
    fs = []
    for i in range(maxiter):
        f, g = func(x, *args)
        m = (1 - b1) * g + b1 * m
        v = (1 - b2) * (g ** 2) + b2 * v
        mhat = m / (1 - b1 ** (i + 1))
        vhat = v / (1 - b2 ** (i + 1))
        x = x - step_size * mhat / (np.sqrt(vhat) + eps)
        fs.append(f)
    return x, np.array(fs)
