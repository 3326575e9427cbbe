"""Kernel-level parity of libgpx against the oracle / numpy float64 (GPU only).

Tolerances are normwise relative errors ||gpu - ref|| / ||ref|| (FP64 throughout).
"""
import sys

import numpy as np
import pytest
import torch

from oracle import aep_oracle as O

pytestmark = pytest.mark.gpu


def _lib():
    from geepee_amd import _lib
    _lib.require_device()
    return _lib


_KEEP = []


def dev(a):
    """Device copy of a host array.  Every copy is kept alive until the test ends: a
    temporary whose data_ptr() is handed to a C call must not be freed (and its block
    reused by the caching allocator) before the kernel has run."""
    t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    _KEEP.append(t)
    return t


@pytest.fixture(autouse=True)
def _release_temporaries():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def nrel(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('m,n,k,tA,tB,batch,splits', [
    (16, 16, 4, 0, 0, 1, 1), (37, 53, 29, 0, 0, 1, 1), (64, 64, 64, 1, 0, 2, 1),
    (100, 100, 100, 0, 1, 5, 1), (130, 70, 300, 1, 1, 3, 1), (200, 33, 1000, 0, 1, 1, 4),
    (5, 1, 7, 0, 0, 4, 1), (1, 300, 1, 0, 0, 1, 1),
    # matrix-vector (n = 1) and outer-product (k = 1) kernels, every transpose
    (128, 1, 128, 1, 0, 3, 1), (100, 1, 4096, 0, 1, 1, 1), (70, 1, 65, 1, 1, 2, 1),
    (128, 128, 1, 0, 1, 4, 1), (33, 70, 1, 1, 0, 2, 1), (9, 5, 1, 1, 1, 3, 1),
    # split-K reductions: splits not a multiple of the reduction's 8 groups, more splits
    # than groups, batched, and a ragged last output workgroup
    (32, 128, 4096, 1, 0, 1, 32), (45, 7, 2000, 0, 1, 3, 13), (128, 128, 8192, 1, 0, 2, 64),
    (3, 5, 700, 1, 1, 2, 5)])
def test_gemm(m, n, k, tA, tB, batch, splits):
    L = _lib()
    say('gemm', m, n, k, tA, tB, batch, splits)
    rs = np.random.RandomState(m + n + k)
    A = rs.randn(batch, *((k, m) if tA else (m, k)))
    B = rs.randn(batch, *((n, k) if tB else (k, n)))
    C0 = rs.randn(batch, m, n)
    ref = 0.7 * np.einsum('bij,bjk->bik', A.transpose(0, 2, 1) if tA else A,
                          B.transpose(0, 2, 1) if tB else B) - 0.3 * C0
    C = dev(C0)
    L.mm(dev(A), dev(B), C, tA=bool(tA), tB=bool(tB), alpha=0.7, beta=-0.3, splits=splits)
    assert nrel(host(C), ref) < 1e-14


def test_gemm_broadcast_operand():
    L = _lib()
    say('gemm broadcast')
    rs = np.random.RandomState(1)
    K = rs.randn(20, 20)
    X = rs.randn(3, 20, 20)
    C = dev(np.zeros((3, 20, 20)))
    L.mm(dev(K), dev(X), C)
    assert nrel(host(C), np.einsum('ab,dbc->dac', K, X)) < 1e-14


@pytest.mark.parametrize('m,n,k,batch,splits,bcast', [
    (100, 100, 100, 5, 1, True), (100, 100, 100, 5, 1, False), (100, 1, 100, 5, 1, False),
    (64, 64, 300, 2, 1, False), (100, 100, 1000, 1, 4, True), (33, 17, 9, 3, 1, False)])
def test_gemm_bias_epilogue(m, n, k, batch, splits, bcast):
    """C = alpha A B + beta C + gamma E on every kernel path (full-K, K-looped, split-K),
    with E per batch member or broadcast."""
    L = _lib()
    say('gemm_ex', m, n, k, batch, splits, bcast)
    rs = np.random.RandomState(m * 7 + k)
    A, B = rs.randn(batch, m, k), rs.randn(batch, k, n)
    C0 = rs.randn(batch, m, n)
    E = rs.randn(m, n) if bcast else rs.randn(batch, m, n)
    ref = 0.7 * np.einsum('bij,bjk->bik', A, B) - 0.3 * C0 + 1.9 * E
    C = dev(C0)
    L.mm(dev(A), dev(B), C, alpha=0.7, beta=-0.3, splits=splits, E=dev(E), gamma=1.9)
    assert nrel(host(C), ref) < 1e-14
    C = dev(np.full((batch, m, n), np.nan))   # beta = 0 never reads C
    L.mm(dev(A), dev(B), C, alpha=0.7, splits=splits, E=dev(E), gamma=-1.0)
    assert nrel(host(C), 0.7 * np.einsum('bij,bjk->bik', A, B) - E) < 1e-14


def test_gemm_outer_product_matches_mfma_path():
    """k = 1 products through the outer-product kernel give the bits of the matrix-core
    kernel (forced with tile = 32), epilogue included."""
    L = _lib()
    say('gemm k=1 bits')
    rs = np.random.RandomState(5)
    u, v, X, E = rs.randn(6, 100, 1), rs.randn(6, 1, 100), rs.randn(6, 100, 100), rs.randn(100, 100)
    outs = []
    for tile in (0, 32):
        C = dev(X)
        L.mm(dev(u), dev(v), C, alpha=0.37, beta=1.0, E=dev(E), gamma=-0.5, tile=tile)
        outs.append(host(C))
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize('D,M', [(2, 100), (1, 200), (3, 64)])
def test_pack_dR_sizes(D, M):
    """pack_dR over several k slices, against the oracle."""
    L = _lib()
    say('pack_dR', D, M)
    rs = np.random.RandomState(M)
    eta = rs.randn(D, M * (M + 1) // 2) * 0.1
    Rref = O.eta_to_R(eta, M)
    G = rs.randn(D, M, M)
    out = dev(np.zeros((D, M * (M + 1) // 2)))
    L.call('gpx_pack_dR', D, M, L.P(dev(Rref)), L.P(dev(G)), L.P(out), L.S())
    assert nrel(host(out), O.pack_dR(Rref, G)) < 1e-14


def test_phi_nat_matches_solve_form():
    """gpx_phi_nat (quadratic terms from theta2) against the log-partitions written the
    way aep_models.py:62-114 computes them: slogdet and solve with Sigma, Sigma-hat."""
    L = _lib()
    say('phi_nat')
    rs = np.random.RandomState(3)
    D, M, N, alpha = 3, 12, 50, 0.5
    beta = (N - alpha) / N
    G = rs.randn(M, M)
    Kuu = G.dot(G.T) / M + np.eye(M)
    Ki = np.linalg.inv(Kuu)
    th1 = np.stack([(lambda R: R.T.dot(R))(np.triu(rs.randn(M, M)) * 0.3) for _ in range(D)])
    th2 = rs.randn(D, M)
    Su = np.linalg.inv(Ki + th1)
    Suhat = np.linalg.inv(Ki + beta * th1)
    mu = np.einsum('dab,db->da', Su, th2)
    muhat = beta * np.einsum('dab,db->da', Suhat, th2)
    sp, sc = N / alpha - 1.0, -N / alpha
    want = D * 0.5 * np.linalg.slogdet(Kuu)[1]
    want += sp * (0.5 * np.sum(np.linalg.slogdet(Su)[1]) +
                  0.5 * np.sum(mu * np.linalg.solve(Su, mu[..., None])[..., 0]))
    want += sc * (0.5 * np.sum(np.linalg.slogdet(Suhat)[1]) +
                  0.5 * np.sum(muhat * np.linalg.solve(Suhat, muhat[..., None])[..., 0]))
    ldK = dev([np.linalg.slogdet(Kuu)[1]])
    ldP = dev([np.linalg.slogdet(Ki + t)[1] for t in th1])
    ldC = dev([np.linalg.slogdet(Ki + beta * t)[1] for t in th1])
    out = dev([0.0])
    L.call('gpx_phi_nat', D, M, L.P(ldK), L.P(ldP), L.P(ldC), L.P(dev(mu)), L.P(dev(muhat)),
           L.P(dev(th2)), beta, sp, sc, L.P(out), L.S())
    assert abs(host(out)[0] - want) < 1e-10 * max(abs(want), 1.0)


@pytest.mark.parametrize('M,batch', [(5, 3), (33, 2), (100, 4), (128, 2), (200, 2), (256, 1)])
def test_spd_inverse_logdet(M, batch):
    L = _lib()
    say('spd', M, batch)
    rs = np.random.RandomState(M)
    A = np.stack([(lambda G: G.dot(G.T) / M + np.eye(M))(rs.randn(M, M)) for _ in range(batch)])
    out = dev(np.zeros_like(A))
    ld = dev(np.zeros(batch))
    info = torch.zeros(batch, dtype=torch.int32, device='cuda')
    L.spd_inv(dev(A), out, ld, info)
    assert host(info).tolist() == [0] * batch
    assert nrel(host(out), np.linalg.inv(A)) < 1e-12
    assert nrel(host(ld), np.linalg.slogdet(A)[1]) < 1e-13


def test_spd_reports_non_spd_member():
    L = _lib()
    say('spd non-SPD')
    A = np.stack([np.eye(6), np.diag([1.0, 2.0, -1.0, 4.0, 5.0, 6.0])])
    out = dev(np.zeros_like(A))
    info = torch.zeros(2, dtype=torch.int32, device='cuda')
    L.spd_inv(dev(A), out, None, info)
    assert host(info).tolist() == [0, 3]
    assert np.allclose(host(out)[0], np.eye(6))
    # a failing pivot inside a later pivot block, and padding of M to the block size
    d = np.ones(13)
    d[9] = -1.0
    A = np.stack([np.diag(d), np.diag(np.arange(1.0, 14.0))])
    out = dev(np.zeros_like(A))
    info = torch.zeros(2, dtype=torch.int32, device='cuda')
    L.spd_inv(dev(A), out, None, info)
    assert host(info).tolist() == [10, 0]
    assert np.allclose(host(out)[1], np.diag(1.0 / np.arange(1.0, 14.0)))


def test_se_kernel():
    L = _lib()
    say('se kernel')
    rs = np.random.RandomState(2)
    x, z = rs.randn(70, 3), rs.randn(40, 3)
    ls, sf = np.log(rs.uniform(0.5, 2, 3)), np.array([0.2])
    K = dev(np.zeros((70, 40)))
    L.call('gpx_se_kernel', 70, 40, 3, L.P(dev(x)), 3, L.P(dev(z)), L.P(dev(ls)), L.P(dev(sf)), 0.0,
           L.P(K), L.S())
    assert nrel(host(K), O.se_kernel(ls, sf, x, z)) < 1e-14


def _psi_inputs(seed, N, M, Q):
    rs = np.random.RandomState(seed)
    return (np.log(rs.uniform(0.7, 1.6, Q)), np.array([np.log(rs.uniform(0.5, 1.5))]),
            rs.randn(N, Q), rs.uniform(0.05, 0.6, (N, Q)), rs.randn(M, Q), rs)


@pytest.mark.parametrize('N,M,Q', [(7, 5, 1), (33, 40, 3), (64, 128, 8), (10, 100, 5)])
def test_psi_fwd(N, M, Q):
    L = _lib()
    say('psi fwd', N, M, Q)
    ls, sf, mx, vx, z, _ = _psi_inputs(N + M, N, M, Q)
    p1 = dev(np.zeros((N, M)))
    p2 = dev(np.zeros((N, M, M)))
    L.call('gpx_psi_fwd', N, M, Q, L.P(dev(mx)), L.P(dev(vx)), L.P(dev(z)), L.P(dev(ls)),
           L.P(dev(sf)), L.P(p1), L.P(p2), L.S())
    r1, r2 = O.psi_stats(ls, sf, mx, vx, z)
    g2 = host(p2)
    assert nrel(host(p1), r1) < 1e-14
    assert nrel(g2, r2) < 1e-14
    assert np.array_equal(g2, np.swapaxes(g2, 1, 2))


@pytest.mark.parametrize('N,M,Q', [(9, 6, 2), (50, 33, 3), (40, 128, 8)])
def test_psi_bwd(N, M, Q):
    L = _lib()
    say('psi bwd', N, M, Q)
    ls, sf, mx, vx, z, rs = _psi_inputs(N * M, N, M, Q)
    r1, r2 = O.psi_stats(ls, sf, mx, vx, z)
    d1 = rs.randn(N, M)
    d2 = rs.randn(N, M, M)
    d2 = 0.5 * (d2 + d2.transpose(0, 2, 1))
    ell, sf2 = np.exp(ls), np.exp(2 * sf)
    ref = O.psi_grads(d1, r1, d2, r2, ell, sf2, mx, vx, z)
    outs = [dev(np.zeros(s)) for s in [(1,), (Q,), (M, Q), (N, Q), (N, Q)]]
    work = torch.empty(L.query('gpx_psi_bwd_workspace', N, M, Q), dtype=torch.float64, device='cuda')
    L.call('gpx_psi_bwd', N, M, Q, L.P(dev(mx)), L.P(dev(vx)), L.P(dev(z)), L.P(dev(ls)),
           L.P(dev(sf)), L.P(dev(d1)), L.P(dev(r1)), L.P(dev(d2)), L.P(dev(r2)),
           *[L.P(o) for o in outs], L.P(work), L.S())
    for name, got, want in zip(['dsf2', 'dell', 'dZ', 'dmx', 'dvx'], outs, ref):
        assert nrel(host(got), want) < 1e-12, name


@pytest.mark.parametrize('N,M,Q', [(150, 30, 3), (1000, 100, 10), (4100, 256, 16), (5000, 64, 5)])
def test_kfu_bwd(N, M, Q):
    L = _lib()
    say('kfu bwd', N, M, Q)
    rs = np.random.RandomState(5)
    x, z = rs.randn(N, Q), rs.randn(M, Q)
    ls, sf = np.log(rs.uniform(0.7, 1.5, Q)), np.array([0.1])
    K = O.se_kernel(ls, sf, x, z)
    G = rs.randn(N, M)
    ref = O.kfu_grads(G, K, np.exp(2 * sf), np.exp(ls), z, x, grad_x=True)
    outs = [dev(np.zeros(s)) for s in [(1,), (Q,), (M, Q), (N, Q)]]
    work = torch.empty(L.query('gpx_kfu_bwd_workspace', N, M, Q), dtype=torch.float64, device='cuda')
    L.call('gpx_kfu_bwd', N, M, Q, L.P(dev(G)), L.P(dev(K)), L.P(dev(x)), L.P(dev(z)), L.P(dev(ls)),
           L.P(dev(sf)), *[L.P(o) for o in outs], L.P(work), L.S())
    for name, got, want in zip(['dsf2', 'dell', 'dZ', 'dx'], outs, ref):
        assert nrel(host(got), want) < 1e-12, name


@pytest.mark.parametrize('M,Q', [(45, 4), (7, 1), (100, 5), (100, 10), (256, 16)])
def test_dtrace(M, Q):
    L = _lib()
    say('dtrace', M, Q)
    rs = np.random.RandomState(6)
    z = rs.randn(M, Q)
    ls, sf = np.log(rs.uniform(0.7, 1.5, Q)), np.array([0.3])
    Mm = rs.randn(M, M)
    Kzz = O.se_kernel(ls, sf, z, z)
    ref = O.dtrace_MKzz(ls, sf, z, Mm, Kzz)
    outs = [dev(np.zeros(s)) for s in [(1,), (Q,), (M, Q)]]
    work = torch.empty(L.query('gpx_dtrace_workspace', M, Q), dtype=torch.float64, device='cuda')
    L.call('gpx_dtrace_kzz', M, Q, L.P(dev(Mm)), L.P(dev(z)), L.P(dev(ls)), L.P(dev(sf)),
           *[L.P(o) for o in outs], L.P(work), L.S())
    for name, got, want in zip(['gsf', 'gls', 'gz'], outs, ref):
        assert nrel(host(got), want) < 1e-12, name


def test_eta_R_and_pack():
    L = _lib()
    say('eta/pack')
    rs = np.random.RandomState(7)
    D, M = 3, 17
    eta = rs.randn(D, M * (M + 1) // 2)
    R = dev(np.zeros((D, M, M)))
    L.call('gpx_eta_to_R', D, M, L.P(dev(eta)), L.P(R), L.S())
    Rref = O.eta_to_R(eta, M)
    assert nrel(host(R), Rref) < 1e-15
    G = rs.randn(D, M, M)
    out = dev(np.zeros((D, M * (M + 1) // 2)))
    L.call('gpx_pack_dR', D, M, L.P(R), L.P(dev(G)), L.P(out), L.S())
    assert nrel(host(out), O.pack_dR(Rref, G)) < 1e-14


def test_gauss_tilted():
    L = _lib()
    say('gauss tilted')
    rs = np.random.RandomState(8)
    B, D, alpha, scale = 300, 3, 0.5, -7.0
    m, v, y = rs.randn(B, D), rs.uniform(0.1, 1, (B, D)), rs.randn(B, D)
    sn = np.array([np.log(0.3)])
    logZ, dm, dv = O.gauss_logZ(sn, m, v, y, alpha)
    dsn = O.gauss_dsn(sn, dv, B, D, alpha, scale)
    o = [dev(np.zeros((B, D))) for _ in range(3)] + [dev(np.zeros(130))]
    L.call('gpx_gauss_tilted', B, D, L.P(dev(m)), L.P(dev(v)), L.P(dev(y)), L.P(dev(sn)), alpha, scale,
           *[L.P(t) for t in o], L.S())
    assert nrel(host(o[0]), scale * dm) < 1e-14
    assert nrel(host(o[1]), scale * dv) < 1e-14
    assert nrel(host(o[2]), scale * dm - 2 * scale * dv * m) < 1e-14
    sc = host(o[3])[:2]
    assert abs(sc[0] - logZ) < 1e-12 * abs(logZ)
    assert abs(sc[1] - dsn[0]) < 1e-12 * abs(dsn[0])


@pytest.mark.parametrize('B', [99, 5000])
def test_transition_and_emission(B):
    L = _lib()
    say('transition/emission', B)
    rs = np.random.RandomState(9)
    D, alpha, scale = 4, 0.5, -3.0
    mp, vp, mt, vt = rs.randn(B, D), rs.uniform(0.1, 1, (B, D)), rs.randn(B, D), rs.uniform(0.1, 1, (B, D))
    sn = np.array([np.log(0.2)])
    lz, dmp, dvp, dmt, dvt, dsn = O.transition_tilted(sn, mp, vp, mt, vt, alpha, scale, D)
    # scal: [logZ, dsn] and up to 64 blocks' partials
    o = [dev(np.zeros((B, D))) for _ in range(4)] + [dev(np.zeros(2 + 2 * 64))]
    L.call('gpx_transition_tilted', B, D, L.P(dev(mp)), L.P(dev(vp)), L.P(dev(mt)), L.P(dev(vt)),
           L.P(dev(sn)), alpha, scale, *[L.P(t) for t in o], L.S())
    assert nrel(host(o[0]), dmp - 2 * dvp * mp) < 1e-14
    assert nrel(host(o[1]), dvp) < 1e-14
    assert nrel(host(o[2]), dmt) < 1e-14
    assert nrel(host(o[3]), dvt) < 1e-14
    sc = host(o[4])
    assert abs(sc[0] - lz) < 1e-12 * abs(lz) and abs(sc[1] - dsn[0]) < 1e-12 * abs(dsn[0])
    Nb, Dout, Din = 150, 8, 4
    C, Rp = rs.randn(Dout, Din), np.log(rs.uniform(0.2, 0.6, Dout))
    y, mx, vx = rs.randn(Nb, Dout), rs.randn(Nb, Din), rs.uniform(0.1, 0.5, (Nb, Din))
    lz, gin, gh = O.emission_tilted(C, Rp, y, mx, vx, alpha, scale)
    o = [dev(np.zeros(s)) for s in [(Nb, Din), (Nb, Din), (Dout, Din), (Dout,), (1,)]]
    work = torch.empty(L.query('gpx_emission_workspace', Nb, Dout, Din), dtype=torch.float64, device='cuda')
    L.call('gpx_emission_tilted', Nb, Dout, Din, L.P(dev(C)), L.P(dev(Rp)), L.P(dev(y)), Dout,
           L.P(dev(mx)), Din, L.P(dev(vx)), Din, alpha, scale, *[L.P(t) for t in o], L.P(work), L.S())
    assert nrel(host(o[0]), gin['mx']) < 1e-12
    assert nrel(host(o[1]), gin['vx']) < 1e-12
    assert nrel(host(o[2]), gh['C']) < 1e-12
    assert nrel(host(o[3]), gh['R']) < 1e-12
    assert abs(host(o[4])[0] - lz) < 1e-12 * abs(lz)


def test_prec_form_matches_numpy():
    """Both precisions of several layers (Kuu^-1 + R^T R, Kuu^-1 + beta R^T R) from the
    stacked R^T R in one launch, the layers of ragged output counts."""
    from geepee_amd import dops
    say('prec_form')
    rs = np.random.RandomState(7)
    Ds, M, beta = [3, 1, 2], 37, 0.75
    X = rs.randn(sum(Ds), M, M)
    Ki = rs.randn(len(Ds), M, M)
    out = dev(np.full((2 * sum(Ds), M, M), np.nan))
    dops.prec_form(Ds, dev(X), dev(Ki), beta, out)
    got = host(out)
    o = x = 0
    for l, d in enumerate(Ds):
        assert np.array_equal(got[o:o + d], X[x:x + d] + Ki[l])
        # beta x + k may be one fused multiply-add on the device: within an ulp of the
        # rounded product plus k
        want = beta * X[x:x + d] + Ki[l]
        assert np.all(np.abs(got[o + d:o + 2 * d] - want) <= 2.3e-16 * (np.abs(beta * X[x:x + d]) + np.abs(want)))
        o += 2 * d
        x += d


def test_kuu_batched_matches_se():
    """Kuu of several layers (different input dims) in one launch: the bits of the
    single-layer squared-exponential kernel plus jitter."""
    from geepee_amd import dops
    say('kuu_batched')
    rs = np.random.RandomState(8)
    M, Qs = 45, [10, 5, 1]
    zs = [dev(rs.randn(M, q)) for q in Qs]
    lss = [dev(0.3 * rs.randn(q)) for q in Qs]
    sfs = [dev(0.2 * rs.randn(1)) for _ in Qs]
    K = dev(np.full((len(Qs), M, M), np.nan))
    dops.kuu_batched(zs, lss, sfs, 1e-5, K)
    for l in range(len(Qs)):
        want = dops.se(zs[l], zs[l], lss[l], sfs[l], jitter=1e-5)
        assert torch.equal(K[l], want), l
