"""Writes the committed golden fixtures tests/golden/*.npz.

Each fixture holds the seeded inputs of a small problem and the outputs of the CPU
oracle (oracle/aep_oracle.py) -- energy and every gradient -- so the GPU tests have a
frozen expected answer and the CPU tests catch drift in the oracle itself.  The
reference cannot be imported (Python 2 with weave), so these vectors come from the
pinned restatement, not from the reference.  Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import aep_oracle as O  # noqa: E402
from geepee_amd import workloads as W  # noqa: E402


def _save(name, inputs, params, energy, grads, meta):
    d = {}
    for k, v in inputs.items():
        d['in_' + k] = np.asarray(v)
    for k, v in params.items():
        d['p_' + k] = np.asarray(v)
    for k, v in grads.items():
        d['g_' + k] = np.asarray(v)
    d['energy'] = np.array(energy)
    for k, v in meta.items():
        d['meta_' + k] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, name + '.npz'), **d)


def cases():
    out = {}
    x, y, p = W.sgpr_problem(N=40, M=8, Din=2, Dout=2, seed=10)
    idx = np.random.RandomState(1).choice(40, 16, replace=False)
    out['sgpr'] = (dict(x=x, y=y, idx=idx), p, dict(M=8, mb=16, alpha=0.5),
                   lambda x=x, y=y, p=p, idx=idx: O.sgpr_objective(x, y, p, 8, 16, alpha=0.5, idxs=idx))
    x, y, p, Ms, hidden = W.sdgpr_problem(N=50, widths=(3, 2, 2, 1), M=6, seed=11)
    out['sdgpr'] = (dict(x=x, y=y), p, dict(Ms=Ms, hidden=hidden, mb=50, alpha=0.5),
                    lambda x=x, y=y, p=p, Ms=Ms: O.sdgpr_objective(x, y, p, Ms, 50, alpha=0.5))
    y, p = W.sgplvm_problem(N=30, Q=2, D=3, M=6, seed=12)
    out['sgplvm'] = (dict(y=y), p, dict(M=6, Q=2, mb=30, alpha=0.5),
                     lambda y=y, p=p: O.sgplvm_objective(y, p, 6, 30, alpha=0.5))
    y, p = W.sgpssm_problem(T=40, Din=2, Dout=3, M=6, seed=13)
    out['sgpssm'] = (dict(y=y), p, dict(M=6, Din=2, mb=40, alpha=0.5),
                     lambda y=y, p=p: O.sgpssm_objective(y, p, 6, 2, 40, alpha=0.5))
    return out


def main():
    for name, (inputs, params, meta, fn) in cases().items():
        e, g = fn()
        _save(name, inputs, params, e, g, meta)
        print('wrote', name, e)


if __name__ == '__main__':
    main()
