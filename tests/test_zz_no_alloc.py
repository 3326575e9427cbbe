"""The allocation-free fused forward (gpx_fused_mm_fwd_noalloc) against gpx_fused_mm_fwd
with a caller workspace, through the C-ABI (GPU only).

Both entry points run the same kernels in the same order, so the outputs must be equal
bit for bit.  Two calls in a row, the second on a smaller problem, check that the
shared scratch carries nothing over; a shape whose workspace exceeds the scratch is
rejected before any launch.
"""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(rs, N, M, Q, D):
    from geepee_amd import dops
    dev = 'cuda'
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    mx = t(rs.randn(N, Q))
    vx = t(rs.uniform(0.05, 0.5, (N, Q)))
    z = t(rs.randn(M, Q))
    ls = t(np.log(rs.uniform(0.6, 1.4, Q)))
    sf = t(np.array([np.log(0.8)]))
    Ahat = t(rs.randn(D, M))
    B = rs.randn(D, M, M) * 0.1
    Bp = dops.sym_pairmajor(t(B))
    return mx, vx, z, ls, sf, Ahat, Bp


def _both(N, M, Q, D, seed):
    from geepee_amd import _lib, dops
    P, S = _lib.P, _lib.S
    rs = np.random.RandomState(seed)
    mx, vx, z, ls, sf, Ahat, Bp = _inputs(rs, N, M, Q, D)
    plan = dops.fused_plan(N, M, Q, D)
    sp, ns = plan['splits_f'], plan['nsub']
    work = torch.empty(_lib.query('gpx_fused_fwd_workspace', N, M, D, sp), dtype=torch.float64,
                       device='cuda')
    ref = [torch.empty(N, M, dtype=torch.float64, device='cuda'),
           torch.empty(N, D, dtype=torch.float64, device='cuda'),
           torch.empty(N, D, dtype=torch.float64, device='cuda')]
    got = [torch.full_like(r, float('nan')) for r in ref]
    _lib.call('gpx_fused_mm_fwd', N, M, Q, D, P(mx), P(vx), P(z), P(ls), P(sf), P(Ahat), P(Bp), sp, ns,
              P(ref[0]), P(ref[1]), P(ref[2]), P(work), S())
    _lib.call('gpx_fused_mm_fwd_noalloc', N, M, Q, D, P(mx), P(vx), P(z), P(ls), P(sf), P(Ahat), P(Bp), sp,
              ns, P(got[0]), P(got[1]), P(got[2]), S())
    torch.cuda.synchronize()
    for a, b, name in zip(got, ref, ('psi1', 'mout', 'vout')):
        assert torch.equal(a, b), name
    assert torch.isfinite(ref[2]).all()
    return ref


def test_noalloc_forward_matches_workspace_forward():
    from geepee_amd import _lib
    say('no-alloc fused forward: N=300 M=100 Q=5 D=5, then N=77 M=40 Q=3 D=2')
    big = _both(300, 100, 5, 5, seed=1)
    small = _both(77, 40, 3, 2, seed=2)
    assert big[0].shape != small[0].shape
    # the benchmarked SGPLVM layer fits the scratch
    say('no-alloc fused forward: N=4096 M=128 Q=8 D=32')
    _both(4096, 128, 8, 32, seed=3)
    # a workspace beyond the scratch is refused before any launch
    cap = _lib.load().gpx_fused_fwd_noalloc_max()
    assert cap == 4 << 20
    N, M, Q, D, sp = 40000, 128, 8, 32, 16
    assert _lib.query('gpx_fused_fwd_workspace', N, M, D, sp) > cap
    rc = _lib.load().gpx_fused_mm_fwd_noalloc(N, M, Q, D, 1, 1, 1, 1, 1, 1, 1, sp, 1, 1, 1, 1, None)
    assert rc == _lib.GPX_ETOOBIG
    with pytest.raises(_lib.GpxError, match='fixed scratch'):
        _lib.call('gpx_fused_mm_fwd_noalloc', N, M, Q, D, 1, 1, 1, 1, 1, 1, 1, sp, 1, 1, 1, 1, None)
