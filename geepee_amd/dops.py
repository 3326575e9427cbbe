"""Device operations: typed wrappers that allocate outputs (torch, as plain device
memory) and launch the libgpx kernels on the current stream.  Nothing here computes
on the host or falls back to torch arithmetic."""
import torch

from . import _lib
from ._lib import P, S, call, query, Probe, set_probe  # noqa: F401

S_ = S

F64 = torch.float64


def empty(*shape, like=None, device='cuda'):
    return torch.empty(*shape, dtype=F64, device=like.device if like is not None else device)


def mm(A, B, C=None, tA=False, tB=False, alpha=1.0, beta=0.0, splits=None, E=None, gamma=0.0):
    """C = alpha op(A) op(B) + beta C + gamma E (2-D or batched 3-D; 2-D operands
    broadcast; E is added in the product's own epilogue).
    Picks a split-K factor when the output has few tiles and K is long (a tuned plan
    for the shape, if the autotuner stored one, wins)."""
    if C is None:
        batch = max(t.shape[0] if t.dim() == 3 else 1 for t in (A, B))
        m = A.shape[-1] if tA else A.shape[-2]
        n = B.shape[-2] if tB else B.shape[-1]
        C = empty(batch, m, n, like=A) if (A.dim() == 3 or B.dim() == 3) else empty(m, n, like=A)
        beta = 0.0
    tile = 0
    if splits is None:
        m, n = C.shape[-2:]
        k = A.shape[-2] if tA else A.shape[-1]
        batch = C.shape[0] if C.dim() == 3 else 1
        from . import autotune
        hit = autotune.lookup('gemm', (m, n, k, batch, int(tA), int(tB)))
        if hit:
            splits, tile = hit['splits'], hit['tile']
        else:
            tiles = -(-m // 64) * -(-n // 64) * batch
            splits = 1
            if tiles < 256 and k >= 256:
                # long, skinny contractions: split K so that enough workgroups run
                splits = max(1, min(-(-256 // tiles), k // 128, 64))
    return _lib.mm(A, B, C, tA=tA, tB=tB, alpha=alpha, beta=beta, splits=splits, tile=tile,
                   E=E, gamma=gamma)


def spd_inv(A, out=None):
    """(inverse, logdet[b], info[b]) of a batch (b, M, M) or one (M, M) SPD matrix."""
    batch = A.shape[0] if A.dim() == 3 else 1
    if out is None:
        out = torch.empty_like(A)
    ld = empty(batch, like=A)
    info = torch.empty(batch, dtype=torch.int32, device=A.device)
    _lib.spd_inv(A, out, ld, info)
    return out, ld, info


def status_fold(infos, status):
    """status[0] = max of the int32 info tensors (0: every SPD launch succeeded),
    status[1] = max(status[1], status[0]); one launch, no host synchronisation."""
    import ctypes
    infos = [t for t in infos if t is not None and t.numel() > 0]
    if len(infos) > 16:
        raise ValueError('at most 16 info arrays per fold')
    ptrs = (ctypes.c_void_p * max(1, len(infos)))(*[t.data_ptr() for t in infos])
    counts = (ctypes.c_int * max(1, len(infos)))(*[t.numel() for t in infos])
    call('gpx_status_fold', len(infos), ctypes.cast(ptrs, ctypes.c_void_p),
         ctypes.cast(counts, ctypes.c_void_p), P(status), S())
    return status


def lincomb(out, a, x, b=0.0, y=None, c=0.0, w=None):
    call('gpx_lincomb', out.numel(), float(a), P(x), x.numel(), float(b), P(y),
         y.numel() if y is not None else 1, float(c), P(w), w.numel() if w is not None else 1,
         P(out), S())
    return out


def colsum(Pm, rows, cols, out=None, alpha=1.0, beta=0.0):
    """out = alpha * column sums of the (rows, cols) matrix Pm + beta * out."""
    if out is None:
        out = empty(cols, like=Pm)
        beta = 0.0
    if rows >= 4096 and cols <= 65535:
        work = empty(query('gpx_colsum_workspace', rows, cols), like=Pm)
        call('gpx_colsum_tall', rows, cols, P(Pm), float(alpha), float(beta), P(out), P(work), S())
        return out
    call('gpx_colsum', rows, cols, P(Pm), float(alpha), float(beta), P(out), S())
    return out


def se(x, z, ls, sf, jitter=0.0, out=None):
    N, Q = x.shape
    M = z.shape[0]
    if out is None:
        out = empty(N, M, like=x)
    call('gpx_se_kernel', N, M, Q, P(x), x.stride(0), P(z), P(ls), P(sf), float(jitter), P(out), S())
    return out


def eta_to_R(eta, M):
    D = eta.shape[0]
    R = empty(D, M, M, like=eta)
    call('gpx_eta_to_R', D, M, P(eta), P(R), S())
    return R


def kuu_batched(zs, lss, sfs, jitter, out):
    """out[l] = se(zs[l], zs[l], lss[l], sfs[l]) + jitter I for layers of one M, one launch."""
    import ctypes
    nl = len(zs)
    vp = lambda ts: (ctypes.c_void_p * nl)(*[t.data_ptr() for t in ts])
    zp, lp, sp = vp(zs), vp(lss), vp(sfs)
    qs = (ctypes.c_int * nl)(*[int(z.shape[1]) for z in zs])
    for z in zs:
        if not z.is_contiguous():
            raise ValueError('kuu_batched needs contiguous inducing inputs')
    call('gpx_kuu_batched', nl, zs[0].shape[0], ctypes.cast(zp, ctypes.c_void_p), ctypes.cast(qs, ctypes.c_void_p),
         ctypes.cast(lp, ctypes.c_void_p), ctypes.cast(sp, ctypes.c_void_p), float(jitter), P(out), S())
    return out


def prec_form(Ds, X, Ki, beta, out):
    """Per layer l (D = Ds[l] rows of X, stacked): out holds [X_l + Ki[l]; beta X_l + Ki[l]]
    (2 D matrices), the layers one after another; X (sum Ds, M, M), Ki (len(Ds), M, M)."""
    import ctypes
    arr = (ctypes.c_int * len(Ds))(*[int(d) for d in Ds])
    call('gpx_prec_form', len(Ds), ctypes.cast(arr, ctypes.c_void_p), X.shape[-1], P(X), P(Ki), float(beta),
         P(out), S())
    return out


def pack_dR(R, G, out=None):
    D, M, _ = R.shape
    if out is None:
        out = empty(D, M * (M + 1) // 2, like=R)
    call('gpx_pack_dR', D, M, P(R), P(G), P(out), S())
    return out


def sym(X, a=0.5):
    out = torch.empty_like(X)
    batch = X.shape[0] if X.dim() == 3 else 1
    call('gpx_sym', batch, X.shape[-1], float(a), P(X), P(out), S())
    return out


def rowscale(w, X):
    N, D = w.shape
    M = X.shape[1]
    out = empty(D, N, M, like=X)
    call('gpx_rowscale', N, D, M, P(w), P(X), P(out), S())
    return out


def wsum(w, T, out, alpha=1.0, beta=0.0):
    N, D = w.shape
    M = T.shape[2]
    call('gpx_wsum', N, D, M, float(alpha), P(w), P(T), float(beta), P(out), S())
    return out


def dmall(dm, dv, m):
    out = torch.empty_like(dm)
    call('gpx_dmall', dm.numel(), P(dm), P(dv), P(m), P(out), S())
    return out


def hyper(ls, sf, dell, gls, dsf2, sumdv, gsf):
    dls = torch.empty_like(ls)
    dsf = torch.empty_like(sf)
    call('gpx_hyper_grads', ls.numel(), P(ls), P(sf), P(dell), P(gls), P(dsf2), P(sumdv), P(gsf),
         P(dls), P(dsf), S())
    return dls, dsf


def phi(D, ldK, ldP, sgnP, dotP, ldC, sgnC, dotC, sp, sc):
    out = empty(1, like=ldK)
    call('gpx_phi', D, P(ldK), P(ldP), float(sgnP), P(dotP), P(ldC), float(sgnC), P(dotC),
         float(sp), float(sc), P(out), S())
    return out


def phi_nat(D, M, ldK, ldP, ldC, mu, muhat, theta2, beta, sp, sc):
    out = empty(1, like=ldK)
    call('gpx_phi_nat', D, M, P(ldK), P(ldP), P(ldC), P(mu), P(muhat), P(theta2), float(beta),
         float(sp), float(sc), P(out), S())
    return out


def batched_dot(x, y):
    b = x.shape[0]
    out = empty(b, like=x)
    call('gpx_batched_dot', b, x[0].numel(), P(x), P(y), P(out), S())
    return out


def det_var(T, kfu, sf):
    D, N, M = T.shape
    v = empty(N, D, like=T)
    call('gpx_det_var', N, D, M, P(T), P(kfu), P(sf), P(v), S())
    return v


def mm_var(m, Bp, sf):
    v = torch.empty_like(m)
    call('gpx_mm_var', m.numel(), P(m), P(Bp), P(sf), P(v), S())
    return v


def psi_fwd(mx, vx, z, ls, sf, with_psi2=True):
    N, Q = mx.shape
    M = z.shape[0]
    psi1 = empty(N, M, like=mx)
    psi2 = empty(N, M, M, like=mx) if with_psi2 else None
    call('gpx_psi_fwd', N, M, Q, P(mx), P(vx), P(z), P(ls), P(sf), P(psi1), P(psi2), S())
    return psi1, psi2


def psi_bwd(mx, vx, z, ls, sf, dpsi1, psi1, dpsi2, psi2):
    N, Q = mx.shape
    M = z.shape[0]
    outs = [empty(1, like=mx), empty(Q, like=mx), empty(M, Q, like=mx), empty(N, Q, like=mx),
            empty(N, Q, like=mx)]
    work = empty(query('gpx_psi_bwd_workspace', N, M, Q), like=mx)
    call('gpx_psi_bwd', N, M, Q, P(mx), P(vx), P(z), P(ls), P(sf), P(dpsi1), P(psi1), P(dpsi2),
         P(psi2), *[P(o) for o in outs], P(work), S())
    return outs


def kfu_bwd(dK, K, x, z, ls, sf, want_dx=False):
    N, Q = x.shape
    M = z.shape[0]
    outs = [empty(1, like=x), empty(Q, like=x), empty(M, Q, like=x),
            empty(N, Q, like=x) if want_dx else None]
    work = empty(query('gpx_kfu_bwd_workspace', N, M, Q), like=x)
    call('gpx_kfu_bwd', N, M, Q, P(dK), P(K), P(x), P(z), P(ls), P(sf), *[P(o) for o in outs],
         P(work), S())
    return outs


def dtrace(Mm, z, ls, sf):
    M, Q = z.shape
    outs = [empty(1, like=z), empty(Q, like=z), empty(M, Q, like=z)]
    work = empty(query('gpx_dtrace_workspace', M, Q), like=z)
    call('gpx_dtrace_kzz', M, Q, P(Mm), P(z), P(ls), P(sf), *[P(o) for o in outs], P(work), S())
    return outs


def gauss_tilted(m, v, y, sn, alpha, scale, want_dmall=False):
    B, D = m.shape
    dm, dv = torch.empty_like(m), torch.empty_like(m)
    dma = torch.empty_like(m) if want_dmall else None
    scal = empty(2 + 2 * 64, like=m)
    call('gpx_gauss_tilted', B, D, P(m), P(v), P(y), P(sn), float(alpha), float(scale), P(dm), P(dv),
         P(dma), P(scal), S())
    return dm, dv, dma, scal


def transition(mp, vp, mt, vt, sn, alpha, scale):
    B, D = mp.shape
    outs = [torch.empty_like(mp) for _ in range(4)]
    scal = empty(2 + 2 * 64, like=mp)
    call('gpx_transition_tilted', B, D, P(mp), P(vp), P(mt), P(vt), P(sn), float(alpha),
         float(scale), *[P(o) for o in outs], P(scal), S())
    return outs + [scal]


def gauss_tilted_mc(m3, v3, y, sn, alpha, scale):
    """Gaussian tilted normaliser over the samples of (S, B, D) propagations."""
    S, B, D = m3.shape
    dm, dv = torch.empty_like(m3), torch.empty_like(m3)
    scal = empty(2 + 2 * 64, like=m3)
    call('gpx_gauss_tilted_mc', S, B, D, P(m3), P(v3), P(y), P(sn), float(alpha), float(scale),
         P(dm), P(dv), P(scal), S_())
    return dm, dv, scal


def probit_tilted(m, v, y, alpha, scale, want_dmall=False):
    """Probit tilted normaliser of (B, D) or (S, B, D) propagations; scal[0] = logZ."""
    three = m.dim() == 3
    Sn = m.shape[0] if three else 1
    B, D = m.shape[-2:]
    dm, dv = torch.empty_like(m), torch.empty_like(m)
    dma = torch.empty_like(m) if want_dmall else None
    scal = empty(2 + 2 * 64, like=m)
    call('gpx_probit_tilted', Sn, int(three), B, D, P(m), P(v), P(y), float(alpha), float(scale),
         P(dm), P(dv), P(dma), P(scal), S_())
    return dm, dv, dma, scal


def transition_mc(mp3, vp3, mt, vt, sn, alpha, scale):
    Sn, B, D = mp3.shape
    dmp, dvp = torch.empty_like(mp3), torch.empty_like(mp3)
    dmt, dvt = torch.empty_like(mt), torch.empty_like(mt)
    scal = empty(2 + 2 * 64, like=mp3)
    call('gpx_transition_tilted_mc', Sn, B, D, P(mp3), P(vp3), P(mt), P(vt), P(sn), float(alpha),
         float(scale), P(dmp), P(dvp), P(dmt), P(dvt), P(scal), S_())
    return dmp, dvp, dmt, dvt, scal


def mc_inputs(eps, mx, vx):
    """x[s] = eps[s] sqrt(vx) + mx for eps (S, N, Q)."""
    Sn, N, Q = eps.shape
    x = torch.empty_like(eps)
    call('gpx_mc_inputs', Sn, N, Q, P(eps), P(mx), P(vx), P(x), S_())
    return x


def mc_reparam_bwd(dx, eps, vx):
    Sn, N, Q = eps.shape
    dmx, dvx = torch.empty_like(vx), torch.empty_like(vx)
    call('gpx_mc_reparam_bwd', Sn, N, Q, P(dx), P(eps), P(vx), P(dmx), P(dvx), S_())
    return dmx, dvx


def ssm_posterior(xf1, x2, prior_mean, prior_var):
    """Posterior means / variances of the GPSSM latent path from its tied factors."""
    N, Q = xf1.shape
    pm, pv = torch.empty_like(xf1), torch.empty_like(xf1)
    call('gpx_ssm_x_posterior', N, Q, P(xf1), P(x2), float(prior_mean), float(prior_var), P(pm), P(pv), S_())
    return pm, pv


def hidden_cavity(h1, h2, alpha):
    N, D = h1.shape
    cm, cv = torch.empty_like(h1), torch.empty_like(h1)
    call('gpx_hidden_cavity', N, D, P(h1), P(h2), float(alpha), P(cm), P(cv), S_())
    return cm, cv


def hidden_grads(h1, h2, alpha, dmc, dvc):
    N, D = h1.shape
    g1, g2 = torch.empty_like(h1), torch.empty_like(h1)
    scal = empty(2 + 64, like=h1)
    call('gpx_hidden_grads', N, D, P(h1), P(h2), float(alpha), P(dmc.contiguous()), P(dvc.contiguous()),
         P(g1), P(g2), P(scal), S_())
    return g1, g2, scal


def emission(C, Rp, y, mx, vx, alpha, scale):
    Nb, Din = mx.shape
    Dout = C.shape[0]
    outs = [empty(Nb, Din, like=mx), empty(Nb, Din, like=mx), empty(Dout, Din, like=mx),
            empty(Dout, like=mx), empty(1, like=mx)]
    work = empty(query('gpx_emission_workspace', Nb, Dout, Din), like=mx)
    call('gpx_emission_tilted', Nb, Dout, Din, P(C), P(Rp), P(y), y.stride(0), P(mx), mx.stride(0),
         P(vx), vx.stride(0), float(alpha), float(scale), *[P(o) for o in outs], P(work), S())
    return outs


def emission_predict(C, Rp, mf, vf):
    Nb, Din = mf.shape
    Dout = C.shape[0]
    my = empty(Nb, Dout, like=mf)
    vy = empty(Nb, Dout, like=mf)
    call('gpx_emission_predict', Nb, Dout, Din, P(C), P(Rp), P(mf), P(vf), P(my), P(vy), S())
    return my, vy


def fused_fits(M, Q, D):
    """Shape limits of the fused moment-matched kernels (larger layers use the
    materialising kernels)."""
    return M <= 256 and Q <= 15 and D <= 32


def fused_plan(N, M, Q=None, D=None, tuned=True):
    """Launch shape of the fused moment-matched kernels: points per block (16*nsub), pair-tile
    splits of the forward / per-point pass, point ranges of the per-pair pass.  A plan
    stored by the autotuner for (N, M, Q, D) wins over the heuristic."""
    if tuned and Q is not None:
        from . import autotune
        hit = autotune.lookup('fused', (N, M, Q, D))
        if hit:
            return hit
    T = (M + 15) // 16
    ntiles = T * (T + 1) // 2
    nsub = 4 if N >= 4096 else (2 if N >= 512 else 1)
    fwd_blocks = -(-N // (16 * nsub))
    splits_f = max(1, min(ntiles, -(-1024 // fwd_blocks)))
    pt_blocks = -(-N // 16)
    splits_b = max(1, min(T, -(-1024 // pt_blocks)))
    nsplit = max(1, min(-(-N // 16), -(-1024 // (ntiles * 4))))
    if D is not None and (D > 16 or N >= 1000):
        # the merged adjoint kernel: one workgroup per CU, one round of 256 over the tiles
        nsplit = max(1, min(-(-N // 16), 256 // ntiles))
    return {'nsub': nsub, 'splits_f': splits_f, 'splits_b': splits_b, 'nsplit': nsplit}


def sym_pairmajor(B):
    D, M, _ = B.shape
    out = empty(M, M, D, like=B)
    call('gpx_sym_pairmajor', D, M, P(B), P(out), S())
    return out


def fused_mm_fwd(mx, vx, z, ls, sf, Ahat, Bp, plan=None):
    N, Q = mx.shape
    M = z.shape[0]
    D = Ahat.shape[0]
    plan = plan or fused_plan(N, M, Q, D)
    psi1 = empty(N, M, like=mx)
    mout = empty(N, D, like=mx)
    vout = empty(N, D, like=mx)
    work = empty(query('gpx_fused_fwd_workspace', N, M, D, plan['splits_f']), like=mx)
    call('gpx_fused_mm_fwd', N, M, Q, D, P(mx), P(vx), P(z), P(ls), P(sf), P(Ahat), P(Bp),
         plan['splits_f'], plan['nsub'], P(psi1), P(mout), P(vout), P(work), S())
    return psi1, mout, vout


def fused_mm_fwd_noalloc(mx, vx, z, ls, sf, Ahat, Bp, psi1, mout, vout, plan=None):
    """fused_mm_fwd into caller-provided outputs through the allocation-free entry point
    (its workspace is the library's fixed per-GPU scratch; overlapping calls on one GPU
    race)."""
    N, Q = mx.shape
    M = z.shape[0]
    D = Ahat.shape[0]
    plan = plan or fused_plan(N, M, Q, D)
    call('gpx_fused_mm_fwd_noalloc', N, M, Q, D, P(mx), P(vx), P(z), P(ls), P(sf), P(Ahat), P(Bp),
         plan['splits_f'], plan['nsub'], P(psi1), P(mout), P(vout), S())
    return psi1, mout, vout


def fused_mm_bwd(mx, vx, z, ls, sf, Ahat, Bp, psi1, dm_all, dv, plan=None):
    N, Q = mx.shape
    M = z.shape[0]
    D = Ahat.shape[0]
    plan = plan or fused_plan(N, M, Q, D)
    outs = [empty(D, M, M, like=mx), empty(1, like=mx), empty(Q, like=mx), empty(M, Q, like=mx),
            empty(N, Q, like=mx), empty(N, Q, like=mx)]
    work = empty(query('gpx_fused_bwd_workspace', N, M, Q, D, plan['splits_b'], plan['nsplit']), like=mx)
    call('gpx_fused_mm_bwd', N, M, Q, D, P(mx), P(vx), P(z), P(ls), P(sf), P(Ahat), P(Bp), P(psi1),
         P(dm_all), P(dv), plan['splits_b'], plan['nsplit'], *[P(o) for o in outs], P(work), S())
    return outs


class FusedBwdRest(object):
    """The rest of a split fused adjoint (the per-pair pass and the tail): call() on the
    stream it should run on, after that stream has waited for the first part, to get
    (dBhat, dsf2, dell, dZ).  The tensors in `keep` must stay alive until then."""

    def __init__(self, N, M, Q, D, plan, z, ls, sf, Bp, dv, work):
        self.args = (N, M, Q, D, plan, z, ls, sf, Bp, dv)
        self.work = work
        self.keep = [z, ls, sf, Bp, dv, work]

    def __call__(self):
        N, M, Q, D, plan, z, ls, sf, Bp, dv = self.args
        dBhat = empty(D, M, M, like=z)
        call('gpx_fused_mm_bwd_pairs', N, M, Q, D, P(z), P(ls), P(sf), P(Bp), P(dv), plan['splits_b'],
             plan['nsplit'], P(dBhat), P(self.work), S())
        outs = [empty(1, like=z), empty(Q, like=z), empty(M, Q, like=z)]
        call('gpx_fused_mm_bwd_tail', N, M, Q, D, P(z), P(ls), P(sf), plan['splits_b'], plan['nsplit'],
             *[P(o) for o in outs], P(self.work), S())
        return [dBhat] + outs


def fused_mm_bwd_split(mx, vx, z, ls, sf, Ahat, Bp, psi1, dm_all, dv, plan=None):
    """fused_mm_bwd's input gradients now: (dmx, dvx, rest), rest a FusedBwdRest giving
    (dBhat, dsf2, dell, dZ) on the stream it is called on."""
    N, Q = mx.shape
    M = z.shape[0]
    D = Ahat.shape[0]
    plan = plan or fused_plan(N, M, Q, D)
    dmx, dvx = empty(N, Q, like=mx), empty(N, Q, like=mx)
    work = empty(query('gpx_fused_bwd_workspace', N, M, Q, D, plan['splits_b'], plan['nsplit']), like=mx)
    call('gpx_fused_mm_bwd_points', N, M, Q, D, P(mx), P(vx), P(z), P(ls), P(sf), P(Ahat), P(Bp), P(psi1),
         P(dm_all), P(dv), plan['splits_b'], plan['nsplit'], P(dmx), P(dvx), P(work), S())
    return dmx, dvx, FusedBwdRest(N, M, Q, D, plan, z, ls, sf, Bp, dv, work)
