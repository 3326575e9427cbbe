"""Fitted-model files (geepee/base_models.py:149-165): the params-dict pickle is written
in the reference's format and read back as arrays only (CPU)."""
import builtins
import pickle

import numpy as np
import pytest

from geepee_amd import checkpoint
from geepee_amd import workloads as W


def test_round_trip_keeps_every_array(tmp_path):
    x, y, p = W.sgpr_problem(N=50, M=7, Din=2, Dout=2, seed=3)
    f = str(tmp_path / 'model.pickle')
    checkpoint.save_params(p, f)
    back = checkpoint.load_params(f)
    assert sorted(back) == sorted(p)
    for k in p:
        assert back[k].dtype == np.float64 and np.array_equal(back[k], p[k]), k
    # a plain pickle.dump of the dict (what the reference writes), any protocol
    for proto in (0, 2, pickle.HIGHEST_PROTOCOL):
        with open(f, 'wb') as fh:
            pickle.dump(p, fh, protocol=proto)
        back = checkpoint.load_params(f)
        for k in p:
            assert np.array_equal(back[k], p[k]), (proto, k)


class _Payload(object):
    def __reduce__(self):
        return (builtins.eval, ("__import__('os').getpid()",))


def test_code_in_a_checkpoint_is_refused(tmp_path):
    f = str(tmp_path / 'bad.pickle')
    with open(f, 'wb') as fh:
        pickle.dump({'ls': np.zeros(2), 'sf': _Payload()}, fh)
    with pytest.raises(pickle.UnpicklingError, match='builtins.eval'):
        checkpoint.load_params(f)
    with open(f, 'wb') as fh:
        pickle.dump([np.zeros(2)], fh)
    with pytest.raises(pickle.UnpicklingError, match='not a params dict'):
        checkpoint.load_params(f)
    with open(f, 'wb') as fh:
        pickle.dump({'ls': np.array(['a', 'b'])}, fh)
    with pytest.raises(pickle.UnpicklingError, match='not numeric'):
        checkpoint.load_params(f)
