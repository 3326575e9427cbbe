"""HIP-graph replay of a whole energy+gradient evaluation.

The objective of a model is a fixed sequence of libgpx launches once shapes are fixed,
so it is captured once into a hipGraph (through torch.cuda.CUDAGraph, which records the
launches our ctypes calls make on the capture stream) and replayed per evaluation.
Parameters and minibatch indices enter through static device buffers that are
refilled before each replay; the flat output [energy, grads in sorted-key order]
lives in the graph's memory pool.  No launch inside the graph uses hipMemsetAsync.
"""
import numpy as np
import torch

from .aep_models import DeviceParams, _shapes
from .config import PROP_MM


class GraphedObjective(object):
    """Captured ``objective_function`` of ``model`` for a fixed minibatch size and alpha.

    ``__call__(params)`` accepts a numpy params dict (uploaded with one copy) or None to
    reuse the resident parameters, draws the minibatch with numpy's legacy global RNG
    exactly like the eager path, replays the graph and returns the flat device output.
    ``objective_function`` keeps the reference's (energy, grad dict) contract.
    """

    def __init__(self, model, params, mb_size, alpha=1.0, prop_mode=PROP_MM, warmup=2):
        if prop_mode != PROP_MM:
            raise NotImplementedError('only moment-matching propagation is captured')
        self.model = model
        self.alpha = alpha
        self.mb = min(mb_size, model.N)
        self.shapes = _shapes(params)
        self.dp = DeviceParams(self.shapes)
        self.dp.upload(params)
        model._params_host = params
        self.keys = sorted(self.shapes)
        self.minibatch = self.mb < model.N and not hasattr(model, 'dyn_layer')
        if hasattr(model, 'dyn_layer') and mb_size < model.N:
            raise NotImplementedError('SGPSSM minibatch windows are not captured')
        self.idx = None
        if self.minibatch:
            self.idx = torch.arange(self.mb, dtype=torch.int64, device='cuda')
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self._run()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = self._run()
        torch.cuda.synchronize()

    def _run(self):
        m = self.model
        energy, grads, div = m._objective(self.dp.views, self.mb, self.alpha, idx=self.idx)
        return m._flat_device(energy, grads, div)

    def __call__(self, params=None):
        if params is not None:
            self.dp.upload(params)
            self.model._params_host = params
        if self.minibatch:
            idx = np.random.choice(self.model.N, self.mb, replace=False)
            self.idx.copy_(torch.from_numpy(idx))
        self.graph.replay()
        return self.out

    def objective_function(self, params, mb_size=None, alpha=None, prop_mode=PROP_MM):
        if mb_size is not None and min(mb_size, self.model.N) != self.mb:
            raise ValueError('graph captured for mb_size=%d' % self.mb)
        if alpha is not None and alpha != self.alpha:
            raise ValueError('graph captured for alpha=%g' % self.alpha)
        out = self(params).cpu().numpy()
        self.model.check_status()
        energy = float(out[0])
        grads = {}
        o = 1
        for k in self.keys:
            n = int(np.prod(self.shapes[k]))
            grads[k] = out[o:o + n].reshape(self.shapes[k])
            o += n
        return energy, grads
