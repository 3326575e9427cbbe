"""Host-side logic that needs no GPU: launch plans, the autotuner cache, request
generation."""
import json

import numpy as np
import pytest

from geepee_amd import autotune, dops
from geepee_amd.serving import make_requests


def _ntiles(M):
    T = (M + 15) // 16
    return T * (T + 1) // 2


@pytest.mark.parametrize('N', [1, 15, 16, 17, 1000, 4096, 20000])
@pytest.mark.parametrize('M', [1, 16, 33, 100, 128, 200, 256])
def test_fused_plan_within_kernel_limits(N, M, monkeypatch, tmp_path):
    monkeypatch.setenv('GEEPEE_AMD_TUNE_CACHE', str(tmp_path / 'none.json'))
    autotune.clear()
    p = dops.fused_plan(N, M, 4, 8)
    assert p['nsub'] in (1, 2, 4)
    assert 1 <= p['splits_f'] <= _ntiles(M)
    assert 1 <= p['splits_b'] <= (M + 15) // 16
    assert 1 <= p['nsplit'] <= -(-N // 16)


@pytest.mark.parametrize('N,M', [(16, 16), (1000, 100), (4096, 128), (10000, 200)])
def test_fused_candidates_within_kernel_limits(N, M):
    fwd, bpt, bpr = autotune.fused_candidates(N, M)
    assert fwd and bpt and bpr
    for c in fwd:
        assert c['nsub'] in (1, 2, 4) and 1 <= c['splits_f'] <= _ntiles(M)
    assert all(1 <= s <= (M + 15) // 16 for s in bpt)
    assert all(1 <= s <= -(-N // 16) for s in bpr)


def test_gemm_candidates():
    c = autotune.gemm_candidates(128, 128, 4096, 1)
    assert {'tile': 64, 'splits': 1} in c and {'tile': 32, 'splits': 1} in c
    assert all(4096 // x['splits'] >= 64 for x in c)


def test_autotune_cache_round_trip(monkeypatch, tmp_path):
    path = tmp_path / 'tune.json'
    monkeypatch.setenv('GEEPEE_AMD_TUNE_CACHE', str(path))
    autotune.clear()
    assert autotune.lookup('fused', (100, 20, 3, 4)) is None
    plan = {'nsub': 2, 'splits_f': 3, 'splits_b': 5, 'nsplit': 7}
    autotune.store('fused', (100, 20, 3, 4), plan)
    assert json.loads(path.read_text())
    autotune.clear()
    assert autotune.lookup('fused', (100, 20, 3, 4)) == plan
    assert dops.fused_plan(100, 20, 3, 4) == plan
    assert dops.fused_plan(100, 20, 3, 4, tuned=False) != plan
    autotune.clear()


def test_make_requests_shape():
    reqs = make_requests(8, 500, 1000.0, [1 + i for i in range(8)], seed=3)
    t = np.array([r[0] for r in reqs])
    assert np.all(np.diff(t) >= 0)
    assert abs(len(reqs) / t[-1] - 1000.0) < 200.0
    for _, mid, X in reqs:
        assert 0 <= mid < 8 and X.shape[1] == 1 + mid and 1 <= X.shape[0] <= 4096


def _unpack_R(eta, M):
    R = np.zeros((M, M))
    R[np.triu_indices(M)] = eta
    R[np.diag_indices(M)] = np.exp(np.diag(R))
    return R


def test_layer_init_matches_reference_recipe():
    """base_models.py:518-597: kmeans2 on x, median-distance lengthscales, sf = log 0.5,
    theta1 = diag(1 / (0.5 U)) packed as an upper Cholesky factor with log diagonal,
    theta2 = theta1 linspace(-1, 1); the draws come from the global legacy RNG."""
    from geepee_amd import init
    from scipy.cluster.vq import kmeans2
    from scipy.spatial.distance import cdist
    rs = np.random.RandomState(0)
    x = rs.randn(40, 3)
    np.random.seed(5)
    p = init.sgp_layer(40, 7, 3, 2, True, x_train=x, key_suffix='_0')
    np.random.seed(5)
    zu, _ = kmeans2(x, 7, minit='points')
    alphas = [0.5 * np.random.rand(7) for _ in range(2)]
    assert np.array_equal(p['zu_0'], zu)
    d = cdist(x, x)
    assert np.allclose(p['ls_0'], np.log(np.median(d[np.triu_indices(40)]) / 2))
    assert np.allclose(p['sf_0'], np.log(0.5))
    for dd in range(2):
        R = _unpack_R(p['eta1_R_0'][dd], 7)
        assert np.allclose(R.T.dot(R), np.diag(1 / alphas[dd]))
        assert np.allclose(p['eta2_0'][dd], np.linspace(-1, 1, 7) / alphas[dd])
    np.random.seed(5)
    q = init.sgp_layer(40, 7, 3, 2, True, x_train=x, key_suffix='_0')
    assert all(np.array_equal(p[k], q[k]) for k in p)


def test_optimal_qu_is_the_collapsed_optimum():
    """Sigma = Kuu (Kuu + Kuf Kfu / sn2)^-1 Kuu and mu = Sigma Kuu^-1 Kuf y / sn2 (the
    optimum of the collapsed bound), in both parameterisations."""
    from geepee_amd import init
    rs = np.random.RandomState(1)
    x, y = rs.randn(30, 2), rs.randn(30, 2)
    np.random.seed(0)
    base = init.sgp_layer(30, 6, 2, 2, True, x_train=x)
    sn = np.log(0.3)
    Kuu = init._se(base['ls'], base['sf'], base['zu'], base['zu']) + 1e-5 * np.eye(6)
    Kfu = init._se(base['ls'], base['sf'], x, base['zu'])
    S = Kuu.dot(np.linalg.solve(Kuu + Kfu.T.dot(Kfu) / np.exp(2 * sn), Kuu))
    mu = S.dot(np.linalg.solve(Kuu, Kfu.T.dot(y))) / np.exp(2 * sn)
    for nat in (True, False):
        p = init.optimal_qu(dict(base), '', x, y, 6, nat, sn)
        for d in range(2):
            R = _unpack_R(p['eta1_R'][d], 6)
            t1 = R.T.dot(R)
            if nat:
                Sig = np.linalg.inv(np.linalg.inv(Kuu) + t1)
                m = Sig.dot(p['eta2'][d])
            else:
                Sig, m = t1, p['eta2'][d]
            assert np.allclose(Sig, S, rtol=1e-6, atol=1e-9)
            assert np.allclose(m, mu[:, d], rtol=1e-6, atol=1e-9)


def test_model_init_keys_and_shapes():
    from geepee_amd import init, workloads as W
    y, want = W.sgplvm_problem(N=60, Q=3, D=5, M=8, seed=2)
    np.random.seed(0)
    got = init.sgplvm(y, 3, 8)
    assert sorted(got) == sorted(want)
    assert all(np.shape(got[k]) == np.shape(want[k]) for k in want)
    y, want = W.sgpssm_problem(T=80, Din=2, Dout=3, M=6, seed=3)
    np.random.seed(0)
    got = init.sgpssm(y, 2, 6)
    assert sorted(got) == sorted(want)
    assert all(np.shape(got[k]) == np.shape(want[k]) for k in want)
    # the PCA projection is the reference's: leading eigenvectors of cov(y)
    P = init.pca_reduce(y, 2)
    w, V = np.linalg.eigh(np.cov(y.T))
    assert np.allclose(np.abs(P), np.abs((y - y.mean(0)).dot(V[:, ::-1][:, :2])))


def test_admission_respects_the_budget():
    from geepee_amd.serving import admit
    # fits: three requests of 100 bytes under a 350-byte budget
    assert admit([(100, 1), (100, 1), (100, 1), (100, 1)], 0, 350, 1 << 16, idle=True) == 3
    # the point cap stops a launch as the byte budget does
    assert admit([(10, 40), (10, 40)], 0, 1000, 64, idle=True) == 1
    # in-flight bytes count: the head waits while another launch holds the memory
    assert admit([(100, 1)], 300, 350, 1 << 16, idle=False) == 0
    # a request larger than the whole budget waits while anything is in flight ...
    assert admit([(500, 1), (10, 1)], 100, 350, 1 << 16, idle=False) == 0
    assert admit([(500, 1)], 0, 350, 1 << 16, idle=False) == 0
    # ... and then runs alone
    assert admit([(500, 1), (10, 1)], 0, 350, 1 << 16, idle=True) == 1
    # the queue is read only as far as the launch reaches
    seen = []

    def sizes():
        for i in range(10):
            seen.append(i)
            yield (100, 1)
    assert admit(sizes(), 0, 250, 1 << 16, idle=True) == 2
    assert seen == [0, 1, 2]


def test_spd_status_raises_linalg_error():
    """check_status reads the device status pair written by gpx_status_fold: entry 0 is
    the last evaluation's largest info (k+1 for a failed pivot k), entry 1 the sticky
    flag of the optimiser loop."""
    import types
    import torch
    from geepee_amd.aep_models import Base_Model
    m = types.SimpleNamespace(_status=torch.tensor([0, 0], dtype=torch.int32))
    Base_Model.check_status(m)
    Base_Model.check_status(m, sticky=True)
    m._status = torch.tensor([4, 4], dtype=torch.int32)
    with pytest.raises(np.linalg.LinAlgError, match='pivot 3'):
        Base_Model.check_status(m)
    m._status = torch.tensor([0, 2], dtype=torch.int32)
    Base_Model.check_status(m)
    with pytest.raises(np.linalg.LinAlgError, match='during optimisation'):
        Base_Model.check_status(m, sticky=True)
