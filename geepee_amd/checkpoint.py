"""Fitted-model files: the params dict that save_model writes and load_model reads
(geepee/base_models.py:149-165 pickles get_hypers() with cPickle).

Writing keeps that format (a pickle of {key: float64 ndarray}), so files move both ways
between the reference and this package.  Reading never runs code from the file: the
unpickler resolves only the globals numpy needs to rebuild arrays and refuses every
other one, so a checkpoint is loaded as weights only.  Python 2 pickles of the reference
are read with latin-1 string decoding, the form numpy documents for them.
"""
import pickle

import numpy as np

# (module, name) pairs an array pickle may reference: numpy's array reconstructor and
# dtype, in the module paths numpy 1.x and 2.x write, and the codec Python 2 pickles use
# for raw byte strings.
_ALLOWED = {
    ('numpy.core.multiarray', '_reconstruct'), ('numpy._core.multiarray', '_reconstruct'),
    ('numpy.core.multiarray', 'scalar'), ('numpy._core.multiarray', 'scalar'),
    ('numpy.core.numeric', '_frombuffer'), ('numpy._core.numeric', '_frombuffer'),
    ('numpy', 'ndarray'), ('numpy', 'dtype'),
    ('_codecs', 'encode'),
}


class _ArraysOnly(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _ALLOWED:
            return super(_ArraysOnly, self).find_class(module, name)
        raise pickle.UnpicklingError('checkpoint refers to %s.%s: only numpy arrays are loaded'
                                     % (module, name))


def save_params(params, fname):
    """Write a params dict as the reference's save_model does."""
    with open(fname, 'wb') as f:
        pickle.dump({k: np.asarray(v, dtype=np.float64) for k, v in params.items()}, f, protocol=2)


def load_params(fname):
    """Read a params dict written by save_model here or in the reference; raises
    pickle.UnpicklingError for any file that is not a dict of numpy arrays."""
    with open(fname, 'rb') as f:
        obj = _ArraysOnly(f, encoding='latin1').load()
    if not isinstance(obj, dict):
        raise pickle.UnpicklingError('checkpoint is not a params dict')
    out = {}
    for k, v in obj.items():
        k = k.decode('latin1') if isinstance(k, bytes) else k
        if not isinstance(k, str):
            raise pickle.UnpicklingError('checkpoint key %r is not a string' % (k,))
        a = np.asarray(v)
        if a.dtype.kind not in 'fiu':
            raise pickle.UnpicklingError('checkpoint entry %s is not numeric' % k)
        out[k] = a.astype(np.float64)
    return out
