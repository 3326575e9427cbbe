"""The fused HIP path against the oracle at exactly the benchmarked shapes (GPU only):

* SGPLVM N=4096, Q=8, D=32, M=128, full batch, alpha=0.5 (the timed evaluation);
* SDGPR 10->5->5->1, N=20000, M=100, minibatch 1000 drawn with the legacy global RNG;
* SGPSSM T=10000, latent 4, Dout=8, M=200, full batch.

Bar per tensor: normwise relative error <= max(1e-9, 10 x the oracle's own spread when
its LU inverses are replaced by Cholesky inverses).  The oracle runs on the host beside
the GPU (the GPSSM case takes about a minute of host time and ~13 GB of host memory:
its psi2 adjoint holds several T x M x M arrays).
"""
import sys
import time

import numpy as np
import pytest

from oracle import aep_oracle as O
from geepee_amd import workloads as W
from tests.test_gpu_models import compare, oracle_ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(model, oracle_fn, mb, seed):
    t = time.time()
    np.random.seed(seed)
    got = model.objective_function(params(model), mb, alpha=0.5)
    want, floor = oracle_ref(lambda: (np.random.seed(seed), oracle_fn())[1])
    worst = compare(got, want, floor=floor)
    say('  worst normwise error %.2e (floor max %.2e), %.0f s' % (worst, max(floor.values()), time.time() - t))


def params(model):
    return model._bench_params


@pytest.mark.timeout(900)
def test_sgplvm_c2_fused_matches_oracle():
    from geepee_amd import aep_models as G
    say('SGPLVM N=4096 Q=8 D=32 M=128 vs oracle')
    y, p = W.sgplvm_problem(N=4096, Q=8, D=32, M=128, seed=1)
    model = G.SGPLVM(y, 8, 128).set_fused()
    model._bench_params = p
    _check(model, lambda: O.sgplvm_objective(y, p, 128, 4096, alpha=0.5), 4096, 0)


@pytest.mark.timeout(900)
def test_sdgpr_c3_fused_matches_oracle():
    from geepee_amd import aep_models as G
    say('SDGPR 10-5-5-1 N=20000 M=100 B=1000 vs oracle')
    x, y, p, Ms, hidden = W.sdgpr_problem(N=20000, widths=(10, 5, 5, 1), M=100, seed=2)
    model = G.SDGPR(x, y, Ms, hidden).set_fused()
    model._bench_params = p
    _check(model, lambda: O.sdgpr_objective(x, y, p, Ms, 1000, alpha=0.5), 1000, 11)


@pytest.mark.timeout(900)
def test_sgpssm_c4_fused_matches_oracle():
    from geepee_amd import aep_models as G
    say('SGPSSM T=10000 Din=4 Dout=8 M=200 vs oracle')
    y, p = W.sgpssm_problem(T=10000, Din=4, Dout=8, M=200, seed=3)
    model = G.SGPSSM(y, 4, 200).set_fused()
    model._bench_params = p
    _check(model, lambda: O.sgpssm_objective(y, p, 200, 4, 10000, alpha=0.5), 10000, 0)
