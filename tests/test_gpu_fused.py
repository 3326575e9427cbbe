"""Fused moment-matched layer (psi2 recomputed, never stored) against the materialising
path and the oracle (GPU only)."""
import sys

import numpy as np
import pytest
import torch

from oracle import aep_oracle as O
from geepee_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def _spread(M, Q):
    """Inducing inputs on a grid (Q <= 2), so Kuu stays well conditioned."""
    if Q == 1:
        return np.linspace(-3, 3, M)[:, None]
    k = int(np.ceil(np.sqrt(M)))
    g = np.stack(np.meshgrid(np.linspace(-2.5, 2.5, k), np.linspace(-2.5, 2.5, k)), -1).reshape(-1, 2)
    return g[np.random.RandomState(1).permutation(len(g))[:M]]


# D covers every slice count of the D contraction (1, 2, 4, 8), with D = 4 x slices and
# not; M and N not multiples of 16 exercise the padded columns and points
@pytest.mark.parametrize('N,M,Q,D', [(37, 20, 3, 2), (130, 40, 2, 5), (300, 100, 5, 5),
                                     (256, 128, 8, 32), (64, 33, 1, 17), (100, 50, 4, 4),
                                     (70, 24, 2, 8), (50, 16, 2, 16),
                                     # the corners fused_fits admits: M = 256 with Q = 15 and
                                     # D = 32 (the largest LDS request), M in (200, 256]
                                     (70, 256, 15, 32), (45, 210, 9, 3),
                                     # point ranges of >= 256 points per split (the two-pass
                                     # form's four-waves-per-SIMD per-pair build at N = 900; the
                                     # merged kernel at N = 1000) and, at D > 16, 8 point blocks
                                     # a range with a ragged last one
                                     (900, 200, 4, 4), (1000, 200, 4, 4), (999, 120, 2, 20),
                                     # N >= 1000 with D <= 16: the merged adjoint kernel
                                     (4100, 50, 3, 6)])
def test_fused_layer_matches_plain(N, M, Q, D):
    from geepee_amd import dops
    from geepee_amd.aep_models import DeviceParams, _shapes
    from geepee_amd.layers import SGP_Layer
    say('fused layer', N, M, Q, D)
    rs = np.random.RandomState(N + M)
    p = W.layer_params(rs, Q, D, M, ls=np.log(0.3 if Q > 2 else 0.2), sf=np.log(0.8))
    if Q <= 2:
        p['zu'] = _spread(M, Q)
    mx = rs.randn(N, Q)
    vx = rs.uniform(0.05, 0.5, (N, Q))
    dp = DeviceParams(_shapes(p))
    dp.upload(p)
    L = SGP_Layer(4 * N, Q, D, M)
    L.update_hypers(dp.views)
    L.compute_cavity(0.5)
    mxd = torch.as_tensor(mx, device='cuda')
    vxd = torch.as_tensor(vx, device='cuda')
    m0, v0, p10, p20 = L.forward_prop_thru_cav(mxd, vxd)
    L.fused = True
    m1, v1, p11, p21 = L.forward_prop_thru_cav(mxd, vxd)
    assert p21 is None
    for a, b, name in [(m1, m0, 'm'), (v1, v0, 'v'), (p11, p10, 'psi1')]:
        assert nrel(a.cpu().numpy(), b.cpu().numpy()) < 1e-12, name
    dm = torch.as_tensor(rs.randn(N, D), device='cuda')
    dv = torch.as_tensor(rs.randn(N, D), device='cuda')
    # kernel level: the fused adjoint against the materialising pipeline (psi2 and
    # dL/dpsi2 stored), both fed the same dm_all, dv; these outputs are well conditioned
    dm_all = dops.dmall(dm, dv, m0)
    fused = dops.fused_mm_bwd(mxd, vxd, L.zu, L.ls, L.sf, L.Ahat, L._Bp, p11, dm_all, dv)
    Bsym = dops.sym(L.Bhat_sto, 0.5)
    dBhat = dops.mm(dv, p20.reshape(N, M * M), tA=True).reshape(D, M, M)
    dpsi2 = dops.mm(dv, Bsym.reshape(D, M * M)).reshape(N, M, M)
    plain = [dBhat] + list(dops.psi_bwd(mxd, vxd, L.zu, L.ls, L.sf, dops.mm(dm_all, L.Ahat), p10, dpsi2, p20))
    for a, b, name in zip(fused, plain, ('dBhat', 'dsf2', 'dell', 'dZ', 'dmx', 'dvx')):
        assert nrel(a.cpu().numpy(), b.cpu().numpy()) < 1e-12, name
    # the split entry (per-point part here, per-pair part and tail on another stream)
    # gives the same bits
    side = torch.cuda.Stream()
    dmx2, dvx2, rest = dops.fused_mm_bwd_split(mxd, vxd, L.zu, L.ls, L.sf, L.Ahat, L._Bp, p11, dm_all, dv)
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        t4 = rest()
    torch.cuda.current_stream().wait_stream(side)
    for a, b, name in zip(list(t4) + [dmx2, dvx2], fused, ('dBhat', 'dsf2', 'dell', 'dZ', 'dmx', 'dvx')):
        assert torch.equal(a, b), name
    # so does _main (per-point then per-pair part) followed by _tail, through the C-ABI
    from geepee_amd import _lib
    plan = dops.fused_plan(N, M, Q, D)
    sb, ns = plan['splits_b'], plan['nsplit']
    work = torch.empty(_lib.load().gpx_fused_bwd_workspace(N, M, Q, D, sb, ns), dtype=torch.float64, device='cuda')
    outs = [torch.empty_like(t) for t in fused]
    P = _lib.P
    _lib.call('gpx_fused_mm_bwd_main', N, M, Q, D, P(mxd), P(vxd), P(L.zu), P(L.ls), P(L.sf), P(L.Ahat), P(L._Bp),
              P(p11), P(dm_all), P(dv), sb, ns, P(outs[0]), P(outs[4]), P(outs[5]), P(work), _lib.S())
    _lib.call('gpx_fused_mm_bwd_tail', N, M, Q, D, P(L.zu), P(L.ls), P(L.sf), sb, ns, P(outs[1]), P(outs[2]),
              P(outs[3]), P(work), _lib.S())
    for a, b, name in zip(outs, fused, ('dBhat', 'dsf2', 'dell', 'dZ', 'dmx', 'dvx')):
        assert torch.equal(a, b), name
    # layer level: the rest of the chain is shared code through Kuu^-1; the problems are
    # chosen so that a 1e-15 relative change of psi2 moves no layer gradient by more than
    # ~1e-12 (inducing points spread out against the lengthscale)
    L.fused = False
    g0, i0 = L.backprop_grads_lvm_mm(m0, v0, dm, dv, p10, p20, mxd, vxd, 0.5)
    L.fused = True
    g1, i1 = L.backprop_grads_lvm_mm(m1, v1, dm, dv, p11, None, mxd, vxd, 0.5)
    for k in g0:
        assert nrel(g1[k].cpu().numpy(), g0[k].cpu().numpy()) < 1e-10, k
    for k in i0:
        assert nrel(i1[k].cpu().numpy(), i0[k].cpu().numpy()) < 1e-11, k


def test_fused_models_match_oracle():
    from geepee_amd import aep_models as G
    from tests.test_gpu_models import compare, oracle_ref
    say('fused models')
    y, p = W.sgplvm_problem(N=200, Q=3, D=6, M=24, seed=5)
    model = G.SGPLVM(y, 3, 24).set_fused()
    got = model.objective_function(p, 200, alpha=0.5)
    want, floor = oracle_ref(lambda: O.sgplvm_objective(y, p, 24, 200, alpha=0.5))
    compare(got, want, floor=floor)
    x, y, p, Ms, hidden = W.sdgpr_problem(N=400, widths=(4, 3, 2, 1), M=20, seed=6)
    model = G.SDGPR(x, y, Ms, hidden).set_fused()
    np.random.seed(2)
    got = model.objective_function(p, 150, alpha=0.5)
    want, floor = oracle_ref(lambda: (np.random.seed(2), O.sdgpr_objective(x, y, p, Ms, 150, alpha=0.5))[1])
    compare(got, want, floor=floor)
