"""Host-side optimiser plumbing with the reference's contracts (geepee/utils.py:37-117).

These run on the host in numpy: they move one flat parameter vector per iteration,
while every objective evaluation runs on the GPU.
"""
import numpy as np


class ObjectiveWrapper(object):
    """utils.py:37-53: unflatten, evaluate, flatten; non-finite gradients become 0 and
    the last finite point is remembered for KeyboardInterrupt recovery."""

    def __init__(self):
        self.previous_x = None

    def __call__(self, params, params_args, obj, idxs, alpha, prop_mode):
        params_dict = unflatten_dict(params, params_args)
        f, grad_dict = obj.objective_function(params_dict, idxs, alpha=alpha, prop_mode=prop_mode)
        g, _ = flatten_dict(grad_dict)
        g_is_fin = np.isfinite(g)
        if np.all(g_is_fin):
            self.previous_x = params
            return f, g
        print("Warning: inf or nan in gradient: replacing with zeros")
        return f, np.where(g_is_fin, g, 0.)


def flatten_dict(params):
    """utils.py:68-81: values ravelled in sorted-key order."""
    keys = list(params.keys())
    shapes = {}
    sizes = []
    vals = []
    for key in sorted(keys):
        val = np.asarray(params[key], dtype=np.float64)
        shapes[key] = val.shape
        vals.append(val.ravel())
        sizes.append(val.size)
    vec = np.concatenate(vals) if vals else np.zeros(0)
    indices = np.cumsum(sizes)[:-1]
    return vec, (keys, indices, shapes)


def unflatten_dict(params, params_args):
    """utils.py:84-90."""
    keys, indices, shapes = params_args[0], params_args[1], params_args[2]
    vals = np.split(params, indices)
    return {key: np.reshape(vals[i], shapes[key]) for i, key in enumerate(sorted(keys))}


def adam(func, init_params, callback=None, maxiter=1000, step_size=0.001, b1=0.9, b2=0.999,
         eps=1e-8, args=None, disp=True, return_cost=False):
    """utils.py:93-117: Adam (Kingma & Ba) with bias correction; prints every 10 iterations."""
    x = init_params
    m = np.zeros_like(x)
    v = np.zeros_like(x)
    fs = []
    for i in range(maxiter):
        f, g = func(x, *args)
        if disp and i % 10 == 0:
            print('iter %d \t obj %.3f' % (i, f))
        if callback:
            callback(x, i, args)
        m = (1 - b1) * g + b1 * m
        v = (1 - b2) * (g ** 2) + b2 * v
        mhat = m / (1 - b1 ** (i + 1))
        vhat = v / (1 - b2 ** (i + 1))
        x = x - step_size * mhat / (np.sqrt(vhat) + eps)
        fs.append(f)
    if return_cost:
        return x, np.array(fs)
    return x
