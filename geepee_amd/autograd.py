"""torch.autograd.Function wrappers over the hand-written adjoint kernels.

``PsiStatistics``  (mx, vx, z, ls, sf) -> (psi1, psi2)
    forward gpx_psi_fwd (geepee/kernels.py:181-240), backward gpx_psi_bwd
    (kernels.py:302-309, 355-444).  psi2's incoming gradient is symmetrised first
    (gpx_sym), as the reference does at kernels.py:415-418.

``MomentMatch``  (mx, vx, z, ls, sf, Ahat, Bhat) -> (m, v)
    the uncertain-input projection of geepee/aep_models.py:182-199,
      m = psi1 Ahat^T,  v = sf2 + <Bhat[d], psi2_n> - m^2,
    with the adjoint of aep_models.py:201-260 (dAhat, dBhat, and the psi adjoint).
    With ``recompute=True`` psi2 is never stored: the forward keeps psi1 and the
    symmetrised Bhat only, and the backward recomputes psi2 inside the fused kernels.

Both run entirely on the device; every tensor is float64 and contiguous.  Gradients are
returned for the log-parameters ls and sf (ell = exp(ls), sf2 = exp(2 sf)), as the
reference's parameter dicts hold them.
"""
import torch

from . import dops


def _c(t):
    return t.detach().contiguous()


def _zeros(n, like):
    return torch.zeros(n, dtype=like.dtype, device=like.device)


def _loghyper_grads(ls, sf, dell, dsf2, sumdv):
    """ell = exp(ls), sf2 = exp(2 sf): chain dL/dell, dL/dsf2 (+ an additive sf2 term
    of weight sumdv) to dL/dls, dL/dsf."""
    return dops.hyper(ls, sf, dell, _zeros(ls.numel(), ls), dsf2, sumdv, _zeros(1, sf))


class PsiStatistics(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mx, vx, z, ls, sf):
        mx, vx, z, ls, sf = [_c(t) for t in (mx, vx, z, ls, sf)]
        psi1, psi2 = dops.psi_fwd(mx, vx, z, ls, sf)
        ctx.save_for_backward(mx, vx, z, ls, sf, psi1, psi2)
        return psi1, psi2

    @staticmethod
    def backward(ctx, dpsi1, dpsi2):
        mx, vx, z, ls, sf, psi1, psi2 = ctx.saved_tensors
        dpsi1 = _zeros(psi1.shape, psi1) if dpsi1 is None else _c(dpsi1)
        # gpx_psi_bwd takes a symmetric dL/dpsi2; psi2 is symmetric, so the symmetrised
        # incoming gradient gives the same adjoint (kernels.py:415-418)
        dpsi2 = _zeros(psi2.shape, psi2) if dpsi2 is None else dops.sym(_c(dpsi2), 0.5)
        dsf2, dell, dz, dmx, dvx = dops.psi_bwd(mx, vx, z, ls, sf, dpsi1, psi1, dpsi2, psi2)
        dls, dsf = _loghyper_grads(ls, sf, dell, dsf2, _zeros(1, sf))
        return dmx, dvx, dz, dls, dsf.reshape(sf.shape)


class MomentMatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mx, vx, z, ls, sf, Ahat, Bhat, recompute):
        mx, vx, z, ls, sf, Ahat, Bhat = [_c(t) for t in (mx, vx, z, ls, sf, Ahat, Bhat)]
        D, M = Ahat.shape
        if recompute:
            if not dops.fused_fits(M, mx.shape[1], D):
                raise ValueError('recompute needs M <= 256, Q <= 15, D <= 32 (got M=%d Q=%d D=%d)'
                                 % (M, mx.shape[1], D))
            Bp = dops.sym_pairmajor(Bhat)
            psi1, m, v = dops.fused_mm_fwd(mx, vx, z, ls, sf, Ahat, Bp)
            ctx.save_for_backward(mx, vx, z, ls, sf, Ahat, Bhat, Bp, psi1, m)
        else:
            N = mx.shape[0]
            psi1, psi2 = dops.psi_fwd(mx, vx, z, ls, sf)
            m = dops.mm(psi1, Ahat, tB=True)
            Bv = dops.mm(psi2.reshape(N, M * M), Bhat.reshape(D, M * M), tB=True)
            v = dops.mm_var(m, Bv, sf)
            ctx.save_for_backward(mx, vx, z, ls, sf, Ahat, Bhat, psi2, psi1, m)
        ctx.recompute = bool(recompute)
        return m, v

    @staticmethod
    def backward(ctx, dm, dv):
        mx, vx, z, ls, sf, Ahat, Bhat, X, psi1, m = ctx.saved_tensors
        N, D = m.shape
        M = Ahat.shape[1]
        dm = _zeros(m.shape, m) if dm is None else _c(dm)
        dv = _zeros(m.shape, m) if dv is None else _c(dv)
        dm_all = dops.dmall(dm, dv, m)
        dAhat = dops.mm(dm_all, psi1, tA=True)
        if ctx.recompute:
            dBhat, dsf2, dell, dz, dmx, dvx = dops.fused_mm_bwd(mx, vx, z, ls, sf, Ahat, X, psi1,
                                                                dm_all, dv)
        else:
            psi2 = X
            dBhat = dops.mm(dv, psi2.reshape(N, M * M), tA=True).reshape(D, M, M)
            dpsi1 = dops.mm(dm_all, Ahat)
            Bsym = dops.sym(Bhat, 0.5)
            dpsi2 = dops.mm(dv, Bsym.reshape(D, M * M)).reshape(N, M, M)
            dsf2, dell, dz, dmx, dvx = dops.psi_bwd(mx, vx, z, ls, sf, dpsi1, psi1, dpsi2, psi2)
        sumdv = dops.colsum(dv, N * D, 1)
        dls, dsf = _loghyper_grads(ls, sf, dell, dsf2, sumdv)
        return dmx, dvx, dz, dls, dsf.reshape(sf.shape), dAhat, dBhat, None


def psi_statistics(mx, vx, z, ls, sf):
    return PsiStatistics.apply(mx, vx, z, ls, sf)


def moment_match(mx, vx, z, ls, sf, Ahat, Bhat, recompute=None):
    """recompute=None picks the fused (psi2-free) path whenever the shape fits it."""
    if recompute is None:
        recompute = dops.fused_fits(Ahat.shape[1], mx.shape[1], Ahat.shape[0])
    return MomentMatch.apply(mx, vx, z, ls, sf, Ahat, Bhat, recompute)
