"""Loader for the committed golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
NAMES = ('sgpr', 'sdgpr', 'sgplvm', 'sgpssm')


def load(name):
    z = np.load(os.path.join(HERE, name + '.npz'))
    out = {'in': {}, 'p': {}, 'g': {}, 'meta': {}}
    for k in z.files:
        if k == 'energy':
            out['energy'] = float(z[k])
            continue
        pre, key = k.split('_', 1)
        out[pre][key] = z[k]
    return out
