"""Approximate-EP sparse GP models on the GPU: SGPR, SDGPR, SGPLVM, SGPSSM.

Drop-in for geepee/aep_models.py:589-1437 (with the model bases of
geepee/base_models.py:15-165, 661-1752): same constructors, ``objective_function(params,
mb_size, alpha, prop_mode) -> (energy, grad_dict)``, ``predict_f`` / ``predict_y``,
``optimise``, ``get_hypers`` / ``update_hypers``.  Parameters may be numpy arrays (the
reference's contract: uploaded in one copy, gradients returned as numpy in one copy)
or device tensors (``objective_function_device`` keeps everything resident).
Minibatches are drawn on the host with numpy's legacy global RNG exactly as the
reference does, so the same seed selects the same points.
"""
import numpy as np
import torch

from . import _lib, checkpoint, config, dops
from .config import PROP_MM, PROP_MC, JITTER
from .dops import mm, lincomb, colsum
from . import init
from .layers import SGP_Layer
from .utils import ObjectiveWrapper, flatten_dict, unflatten_dict, adam

F64 = torch.float64


class DeviceParams(object):
    """A params dict laid out as one flat device buffer in sorted-key order (the order of
    geepee/utils.py:68-81), with one view per key."""

    def __init__(self, shapes, device='cuda'):
        self.keys = sorted(shapes)
        self.shapes = {k: tuple(shapes[k]) for k in self.keys}
        self.sizes = [int(np.prod(self.shapes[k])) for k in self.keys]
        self.total = int(sum(self.sizes))
        self.buf = torch.empty(self.total, dtype=F64, device=device)
        self.views = {}
        o = 0
        for k, n in zip(self.keys, self.sizes):
            self.views[k] = self.buf[o:o + n].view(self.shapes[k])
            o += n

    def upload(self, params):
        vec = np.concatenate([np.asarray(params[k], dtype=np.float64).ravel() for k in self.keys])
        self.buf.copy_(torch.from_numpy(vec))
        return self.views

    def __getitem__(self, k):
        return self.views[k]


def _shapes(params):
    return {k: np.shape(v) for k, v in params.items()}


class Gauss_Layer(object):
    """Gaussian likelihood (geepee/lik_layers.py:87-282) on the GPU."""

    def __init__(self, N, D):
        self.N = N
        self.D = D

    def compute_log_Z(self, mout, vout, y, alpha=1.0, scale=1.0, want_dmall=False):
        """Returns (dm, dv, dm_all, scal) with dm, dv scaled by ``scale`` and
        scal = [logZ, dsn * scale].  Unlike lik_layers.py:122 the caller's vout is not
        modified in place (nothing on the path reads the mutated value)."""
        return dops.gauss_tilted(mout, vout, y, self.sn, alpha, scale, want_dmall)

    def output_probabilistic(self, mf, vf, alpha=1.0):
        """mf, vf + exp(2 sn) / alpha (lik_layers.py:238-249)."""
        sn2 = float(np.exp(2.0 * self.sn.detach().cpu().numpy().ravel()[0]))
        out = torch.empty_like(vf)
        ones = torch.ones(1, dtype=F64, device=vf.device)
        return mf, lincomb(out, 1.0, vf, sn2 / alpha, ones)

    def compute_log_Z_mc(self, m3, v3, y, alpha=1.0, scale=1.0):
        """3-D branch (lik_layers.py:139-152): (S, B, D) Monte Carlo propagations;
        returns (dm, dv, scal) scaled like compute_log_Z."""
        return dops.gauss_tilted_mc(m3, v3, y, self.sn, alpha, scale)

    def hyper_grads(self, scal, key_suffix=''):
        return {'sn' + key_suffix: scal[1:2]}

    def update_hypers(self, params, key_suffix=''):
        self.sn = params['sn' + key_suffix]


class Probit_Layer(object):
    """Probit likelihood for y in {-1, +1} (geepee/lik_layers.py:285-471) on the GPU:
    closed form at alpha = 1, 10-point Gauss-Hermite quadrature otherwise.  It has no
    hyper-parameters."""

    def __init__(self, N, D):
        self.N = N
        self.D = D

    def compute_log_Z(self, mout, vout, y, alpha=1.0, scale=1.0, want_dmall=False):
        return dops.probit_tilted(mout, vout, y, alpha, scale, want_dmall)

    def compute_log_Z_mc(self, m3, v3, y, alpha=1.0, scale=1.0):
        dm, dv, _, scal = dops.probit_tilted(m3, v3, y, alpha, scale)
        return dm, dv, scal

    def hyper_grads(self, scal, key_suffix=''):
        return {}

    def output_probabilistic(self, mf, vf, alpha=1.0):
        raise NotImplementedError('TODO: return probablity of y=1')   # lik_layers.py:457-471

    def update_hypers(self, params, key_suffix=''):
        pass


def _make_lik(lik, N, D):
    if lik.lower() == 'gaussian':
        return Gauss_Layer(N, D)
    if lik.lower() == 'probit':
        return Probit_Layer(N, D)
    raise NotImplementedError('likelihood not implemented')


class Base_Model(object):
    """Common driver (geepee/base_models.py:15-165)."""

    def __init__(self, y_train):
        self.y_train = np.asarray(y_train, dtype=np.float64)
        self.N = self.y_train.shape[0]
        self.fixed_params = []
        self.updated = False
        self._dp = None
        _lib.require_device()
        self.y_dev = torch.as_tensor(self.y_train, device='cuda')
        # SPD status of the last evaluation ([0]) and a flag that stays set until
        # optimise_device clears it ([1]); written on the device by gpx_status_fold
        self._status = torch.zeros(2, dtype=torch.int32, device='cuda')

    # parameter transfer -------------------------------------------------
    def _device_params(self, params):
        if isinstance(next(iter(params.values())), torch.Tensor):
            return params
        shapes = _shapes(params)
        if self._dp is None or self._dp.shapes != {k: tuple(v) for k, v in shapes.items()}:
            self._dp = DeviceParams(shapes)
        self._dp.upload(params)
        self._params_host = params
        return self._dp.views

    def objective_function(self, params, mb_size, alpha=1.0, prop_mode=PROP_MM):
        """(energy, grad dict) as numpy, the reference contract (aep_models.py:610, 694,
        895, 1024)."""
        energy, grads, div = self._objective(params, mb_size, alpha, prop_mode=self._mode(prop_mode))
        return self._finish(energy, grads, params, div)

    def objective_function_device(self, params, mb_size, alpha=1.0, prop_mode=PROP_MM):
        """Device-resident variant: params is a dict of device tensors; returns one flat
        device vector [energy, grads in sorted-key order] without any host copy."""
        energy, grads, div = self._objective(params, mb_size, alpha, prop_mode=self._mode(prop_mode))
        return self._flat_device(energy, grads, div)

    # the models whose objective branches on prop_mode (SGPLVM, SGPSSM) accept MM and MC;
    # SGPR and SDGPR ignore it, as the reference's do (aep_models.py:611, 895)
    _prop_modes = None

    def _mode(self, prop_mode):
        if self._prop_modes is None:
            return PROP_MM
        if prop_mode not in self._prop_modes:
            raise NotImplementedError('propagation mode %s not implemented' % prop_mode)
        return prop_mode

    def _flat_device(self, energy_dev, grads_dev, divide_by):
        keys = sorted(grads_dev)
        for p in self.fixed_params:
            grads_dev[p] = torch.zeros_like(grads_dev[p])
        flat = torch.cat([energy_dev.reshape(-1)] + [grads_dev[k].reshape(-1) for k in keys])
        if divide_by is not None:
            lincomb(flat, 1.0 / divide_by, flat)
        dops.status_fold(self._infos(), self._status)
        return flat

    def _infos(self):
        out = []
        for L in self._layers():
            for name in ('info_Kuu', 'info_Su', 'info_Suhat'):
                out.append(getattr(L, name, None))
        return out

    def check_status(self, sticky=False):
        """Raise numpy.linalg.LinAlgError (what np.linalg.inv raises at
        base_models.py:464, 471 and aep_models.py:525) if an SPD inverse of the last
        evaluation -- or, with sticky=True, of any evaluation since the flag was
        cleared -- met a non-positive pivot.  One device-to-host copy."""
        st = self._status.cpu().numpy()
        if st[1 if sticky else 0] > 0:
            raise np.linalg.LinAlgError(
                'matrix not positive definite (pivot %d of an SPD inverse)%s'
                % (int(st[1 if sticky else 0]) - 1, ' during optimisation' if sticky else ''))

    def _finish(self, energy_dev, grads_dev, params, divide_by):
        """Zero fixed params, divide by N, return (float, numpy dict) with one D2H copy."""
        keys = sorted(grads_dev)
        flat = self._flat_device(energy_dev, grads_dev, divide_by)
        out = flat.cpu().numpy()
        self.check_status()
        energy = float(out[0])
        grads = {}
        o = 1
        for k in keys:
            shape = np.shape(params[k])
            n = int(np.prod(shape))
            grads[k] = out[o:o + n].reshape(shape)
            o += n
        return energy, grads

    def set_fused(self, fused=True):
        """Moment-matched layers recompute psi2 inside fused kernels instead of writing it."""
        for L in self._layers():
            L.fused = fused
        return self

    def _layers(self):
        out = []
        for name in ('sgp_layer', 'dyn_layer', 'emi_layer'):
            if isinstance(getattr(self, name, None), SGP_Layer):
                out.append(getattr(self, name))
        return out + list(getattr(self, 'sgp_layers', []))

    # optimisation (base_models.py:59-134) --------------------------------
    def optimise(self, method='L-BFGS-B', tol=None, reinit_hypers=True, callback=None,
                 maxfun=100000, maxiter=1000, alpha=0.5, mb_size=None, adam_lr=0.001,
                 prop_mode=PROP_MM, disp=True, return_cost=False, **kargs):
        self.updated = False
        init_params_dict = self.init_hypers(self.y_train) if reinit_hypers else self.get_hypers()
        init_params_vec, params_args = flatten_dict(init_params_dict)
        objective_wrapper = ObjectiveWrapper()
        if mb_size is None:
            mb_size = self.N
        costs = []
        try:
            if method.lower() == 'adam':
                results = adam(objective_wrapper, init_params_vec, step_size=adam_lr,
                               maxiter=maxiter, args=(params_args, self, mb_size, alpha, prop_mode),
                               disp=disp, callback=callback, return_cost=return_cost)
                if return_cost:
                    final_params, costs = results
                else:
                    final_params = results
            else:
                from scipy.optimize import minimize
                options = {'maxfun': maxfun, 'maxiter': maxiter, 'disp': disp, 'gtol': 1e-6,
                           'ftol': 1e-6}
                results = minimize(fun=objective_wrapper, x0=init_params_vec,
                                   args=(params_args, self, self.N, alpha, prop_mode),
                                   method=method, jac=True, tol=tol, callback=callback,
                                   options=options)
                final_params = results.x
        except KeyboardInterrupt:
            print('Caught KeyboardInterrupt ...')
            final_params = objective_wrapper.previous_x
            costs = []
        final_params = unflatten_dict(final_params, params_args)
        self.update_hypers(final_params)
        if return_cost and method.lower() == 'adam':
            return final_params, costs
        return final_params

    def optimise_device(self, params=None, maxiter=1000, adam_lr=0.001, alpha=0.5, mb_size=None,
                        graph=True, b1=0.9, b2=0.999, eps=1e-8):
        """Adam (utils.py:93-117) with the parameter vector, both moments and the gradient
        resident on the device: per iteration one objective evaluation (a hipGraph
        replay when graph=True) and one gpx_adam_step launch; minibatches are drawn
        with numpy's legacy RNG in the reference's order.  Starts from `params` (default:
        the current hyper-parameters), updates the model, and returns the final params
        dict and the objective value of every iteration (one device-to-host copy)."""
        from ._lib import P, S, call
        from .graphs import GraphedObjective
        params = self.get_hypers() if params is None else params
        mb = self.N if mb_size is None else mb_size
        if graph:
            g = GraphedObjective(self, params, mb, alpha)
            dp, step = g.dp, g
        else:
            dp = DeviceParams(_shapes(params))
            dp.upload(params)
            step = lambda: self.objective_function_device(dp.views, mb, alpha=alpha)
        n = dp.total
        mom = torch.zeros(2, n, dtype=F64, device=dp.buf.device)
        costs = torch.empty(max(maxiter, 1), dtype=F64, device=dp.buf.device)
        self._status[1:].zero_()
        for i in range(maxiter):
            out = step()
            call('gpx_adam_step', n, P(dp.buf), P(out[1:]), P(mom[0]), P(mom[1]), float(adam_lr),
                 float(b1), float(b2), float(eps), i + 1, P(out), P(costs[i:]), S())
        # the reference's np.linalg.inv raises inside the loop; here the device flag is read
        # once, after the last step
        self.check_status(sticky=True)
        vec = dp.buf.cpu().numpy()
        final, o = {}, 0
        for k, sz in zip(dp.keys, dp.sizes):
            final[k] = vec[o:o + sz].reshape(dp.shapes[k]).copy()
            o += sz
        self.update_hypers(final)
        return final, costs[:maxiter].cpu().numpy()

    def set_fixed_params(self, params):
        if isinstance(params, (list, tuple)):
            for p in params:
                if p not in self.fixed_params:
                    self.fixed_params.append(p)
        else:
            self.fixed_params.append(params)

    def save_model(self, fname='/tmp/model.pickle'):
        """Pickle get_hypers() as base_models.py:149-156 does."""
        checkpoint.save_params(self.get_hypers(), fname)

    def load_model(self, fname='/tmp/model.pickle'):
        """base_models.py:158-165, reading the file as arrays only (no code from the
        file runs; see checkpoint.load_params)."""
        self.update_hypers(checkpoint.load_params(fname))

    def get_hypers(self):
        return {k: np.array(v, copy=True) for k, v in self._params_host.items()}


def _check_mode(prop_mode):
    if prop_mode not in (PROP_MM, PROP_MC):
        raise NotImplementedError('prop_mode %s unknown' % prop_mode)


def _mc_draws(*shape):
    """Standard-normal draws in the reference's call (np.random.randn on the legacy global
    RNG, aep_models.py:171), uploaded in one copy."""
    return torch.as_tensor(np.random.randn(*shape), device='cuda')


class SGPR(Base_Model):
    """aep_models.py:589-667 + base_models.py:932-1070."""

    def __init__(self, x_train, y_train, no_pseudo, lik='Gaussian', nat_param=True):
        super(SGPR, self).__init__(y_train)
        self.x_train = np.asarray(x_train, dtype=np.float64)
        self.Din = self.x_train.shape[1]
        self.Dout = self.y_train.shape[1]
        self.M = no_pseudo
        self.nat_param = nat_param
        self.x_dev = torch.as_tensor(self.x_train, device='cuda')
        self.lik_layer = _make_lik(lik, self.N, self.Dout)
        self.sgp_layer = SGP_Layer(self.N, self.Din, self.Dout, self.M, nat_param)

    def update_hypers(self, params):
        dp = self._device_params(params)
        self.sgp_layer.update_hypers(dp)
        self.lik_layer.update_hypers(dp)
        return dp

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        N = self.N
        if idx is None and mb_size < N:
            idx = torch.as_tensor(np.random.choice(N, mb_size, replace=False), device='cuda')
        if idx is None:
            xb, yb = self.x_dev, self.y_dev
        else:
            xb, yb = self.x_dev.index_select(0, idx), self.y_dev.index_select(0, idx)
        B = yb.shape[0]
        scale_logZ = -N * 1.0 / B / alpha
        dp = self._device_params(params)
        layer = self.sgp_layer
        layer.prepare(dp, alpha)
        self.lik_layer.update_hypers(dp)
        mout, vout, kfu = layer.forward_prop_thru_cav(xb)
        dm, dv, _, scal = self.lik_layer.compute_log_Z(mout, vout, yb, alpha, scale_logZ)
        grads = layer.backprop_grads_reg(mout, vout, dm, dv, kfu, xb, alpha)
        grads.update(self.lik_layer.hyper_grads(scal))
        phi = layer.compute_phi(alpha)
        energy = lincomb(torch.empty(1, dtype=F64, device='cuda'), scale_logZ, scal[0:1], 1.0, phi)
        return energy, grads, N

    def predict_f(self, inputs):
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self.sgp_layer.forward_prop_thru_post(x)
        return mf.cpu().numpy(), vf.cpu().numpy()

    def predict_y(self, inputs):
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self.sgp_layer.forward_prop_thru_post(x)
        my, vy = self.lik_layer.output_probabilistic(mf, vf)
        return my.cpu().numpy(), vy.cpu().numpy()

    def init_hypers(self, y_train=None):
        """base_models.py:1036-1049: the layer on the training inputs, sn = log 0.01."""
        p = init.sgp_layer(self.N, self.M, self.Din, self.Dout, self.nat_param, x_train=self.x_train)
        p.update(init.gauss_lik())
        return p


class SDGPR(Base_Model):
    """aep_models.py:870-988 + base_models.py:1073-1336; layer i keys end in '_i'."""

    def __init__(self, x_train, y_train, no_pseudos, hidden_sizes, lik='Gaussian'):
        super(SDGPR, self).__init__(y_train)
        self.x_train = np.asarray(x_train, dtype=np.float64)
        self.Din = self.x_train.shape[1]
        self.Dout = self.y_train.shape[1]
        self.size = [self.Din] + list(hidden_sizes) + [self.Dout]
        self.L = len(self.size) - 1
        self.Ms = list(no_pseudos) if isinstance(no_pseudos, (list, tuple)) else [no_pseudos] * self.L
        self.x_dev = torch.as_tensor(self.x_train, device='cuda')
        self.lik_layer = _make_lik(lik, self.N, self.Dout)
        self.sgp_layers = [SGP_Layer(self.N, self.size[i], self.size[i + 1], self.Ms[i])
                           for i in range(self.L)]

    def init_hypers(self, y_train=None):
        """base_models.py:1290-1312: layer 0 on the training inputs, the hidden layers
        without data, sn = log 0.01."""
        p = {}
        for i, layer in enumerate(self.sgp_layers):
            p.update(init.sgp_layer(self.N, layer.M, layer.Din, layer.Dout, True,
                                    x_train=self.x_train if i == 0 else None, key_suffix='_%d' % i))
        p.update(init.gauss_lik())
        return p

    def update_hypers(self, params):
        dp = self._device_params(params)
        for i, layer in enumerate(self.sgp_layers):
            layer.update_hypers(dp, key_suffix='_%d' % i)
        self.lik_layer.update_hypers(dp)
        return dp

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        N = self.N
        if idx is None and mb_size < N:
            idx = torch.as_tensor(np.random.choice(N, mb_size, replace=False), device='cuda')
        if idx is None:
            xb, yb = self.x_dev, self.y_dev
        else:
            xb, yb = self.x_dev.index_select(0, idx), self.y_dev.index_select(0, idx)
        B = yb.shape[0]
        scale_logZ = -N * 1.0 / B / alpha
        phis, ready = self._prepare_layers(params, alpha)
        ms, vs, p1, p2 = [], [], [], []
        for i, layer in enumerate(self.sgp_layers):
            if ready is not None:
                # layer i's forward needs only its own q(u) set-up
                torch.cuda.current_stream().wait_event(ready[i])
            if i == 0:
                m, v, kfu = layer.forward_prop_thru_cav(xb)
                ms.append(m); vs.append(v); p1.append(kfu); p2.append(None)
            else:
                m, v, a, b = layer.forward_prop_thru_cav(ms[i - 1], vs[i - 1])
                ms.append(m); vs.append(v); p1.append(a); p2.append(b)
        last_mm = self.L > 1
        if ready is not None:
            self._issue_deferred_adjoint(alpha)
        dmi, dvi, dma, scal = self.lik_layer.compute_log_Z(ms[-1], vs[-1], yb, alpha, scale_logZ,
                                                           want_dmall=last_mm)
        grads = {}
        # the hyper-parameter chain of each moment-matched layer forks onto that layer's
        # stream while the input gradients carry on down the layers on this one
        streams = getattr(self, '_streams', None) if self.L > 1 and config.MULTI_STREAM else None
        keep = []
        for i in range(self.L - 1, -1, -1):
            layer = self.sgp_layers[i]
            if i == 0:
                if streams:
                    torch.cuda.current_stream().wait_stream(streams[0])
                # (forking this layer's adjoint over more streams measured slower when
                # replayed as a graph: the extra branches add dependencies, not overlap)
                g = layer.backprop_grads_reg(ms[0], vs[0], dmi, dvi, p1[0], xb, alpha)
            else:
                g, gin = layer.backprop_grads_lvm_mm(ms[i], vs[i], dmi, dvi, p1[i], p2[i],
                                                     ms[i - 1], vs[i - 1], alpha, dm_all=dma,
                                                     side=streams[i] if streams else None, keep=keep)
                dmi, dvi, dma = gin['mx'], gin['vx'], None
                p2[i] = None
            for k, val in g.items():
                grads[k + '_%d' % i] = val
        if streams:
            cur = torch.cuda.current_stream()
            for st in streams:
                cur.wait_stream(st)
        del keep
        grads.update(self.lik_layer.hyper_grads(scal))
        energy = lincomb(torch.empty(1, dtype=F64, device='cuda'), scale_logZ, scal[0:1])
        for phi in phis:
            lincomb(energy, 1.0, energy, 1.0, phi)
        return energy, grads, N

    def _prepare_layers(self, params, alpha):
        """update_hypers, compute_cavity and compute_phi of every layer.  The layers'
        chains are independent, so each runs on a stream of its own forked from the
        current one (parallel branches when captured in a hipGraph).  Returns the phi
        terms and, on the multi-stream path, one event per layer marking its q(u) set-up;
        the caller waits for event i before layer i's forward and joins the streams at
        the end."""
        dp = self._device_params(params)
        self.lik_layer.update_hypers(dp)
        cur = torch.cuda.current_stream()
        if self.L == 1 or not config.MULTI_STREAM:
            phis = []
            for i, layer in enumerate(self.sgp_layers):
                layer.prepare(dp, alpha, key_suffix='_%d' % i)
                phis.append(layer.compute_phi(alpha))
            return phis, None
        if getattr(self, '_streams', None) is None:
            self._streams = [torch.cuda.Stream() for _ in range(self.L)]
        # The SPD inverses of the layers that share an M run batched: one launch for their
        # Kuu, one for their [posterior; cavity] precisions (each launch takes about as
        # long for one matrix as for many, one workgroup per matrix).
        precs = self._batched_inverses(dp, alpha)
        phis, ready = [], []
        for i, (layer, st) in enumerate(zip(self.sgp_layers, self._streams)):
            st.wait_stream(cur)
            j = self._setup_after(i)
            if j is not None:
                st.wait_event(ready[j])
            with torch.cuda.stream(st):
                layer.finish_prepare(*precs[i], alpha=alpha)
                phis.append(layer.compute_phi(alpha))
                ev = torch.cuda.Event()
                ev.record(st)
                ready.append(ev)
                # the adjoint's q(u)-only terms follow on the same stream; the forward
                # waits only for the q(u) set-up above, the layer's backward for these
                if not self._defer_adjoint(i):
                    layer.adjoint_constants('reg' if i == 0 else 'mm', alpha)
        return phis, ready

    def _setup_after(self, i):
        """Index of the layer whose q(u) set-up layer i's set-up waits for (None: it starts
        right after the batched inverses).  The first layer's set-up runs alone and its
        forward starts sooner; the others follow beside that forward (10-5-5-1 deep GP at
        M = 100, same-box replays: 0.759-0.764 ms against 0.776-0.798 ms with all three set-ups
        together, 0.83 ms with each waiting for the previous one)."""
        return 0 if i > 0 else None

    def _issue_deferred_adjoint(self, alpha):
        """The deferred layers' adjoint constants, each on its layer's stream after the work
        issued so far on the current one."""
        for i in range(self.L - 1, -1, -1):
            if self._defer_adjoint(i):
                st = self._streams[i]
                st.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(st):
                    self.sgp_layers[i].adjoint_constants('reg' if i == 0 else 'mm', alpha)

    def _defer_adjoint(self, i):
        """Whether layer i's adjoint constants are issued on its stream after the forwards
        (in _objective) instead of right after its q(u) set-up.  The first layer's
        backward comes last, so its terms wait: run beside the forwards they slowed them
        (a replay of the 10-5-5-1 deep GP at M = 100 took 0.725 against 0.647 ms)."""
        return i == 0

    def _batched_inverses(self, dp, alpha):
        """prepare's Kuu and precision inverses for every layer, batched over the layers
        of each M; returns per layer (prec, cov, logdet, info) views for finish_prepare."""
        layers = self.sgp_layers
        groups = {}
        for i, layer in enumerate(layers):
            groups.setdefault(layer.M, []).append(i)
        # theta1's factors of a group in one launch when their packed parameters lie one after
        # another in the parameter buffer (sorted keys: eta1_R_0, eta1_R_1, ...)
        Rv, R_all = {}, {}
        for M, idx in groups.items():
            etas = [dp['eta1_R_%d' % i] for i in idx]
            if len(idx) > 1 and all(etas[j + 1].data_ptr() == etas[j].data_ptr() + etas[j].numel() * 8
                                    for j in range(len(idx) - 1)):
                P = etas[0].shape[1]
                Ds = [layers[i].Dout for i in idx]
                R_all[M] = dops.eta_to_R(torch.as_strided(etas[0], (sum(Ds), P), (P, 1)), M)
                o = 0
                for i, d in zip(idx, Ds):
                    Rv[i] = R_all[M][o:o + d]
                    o += d
        for i, layer in enumerate(layers):
            layer.prepare_params(dp, key_suffix='_%d' % i, R=Rv.get(i))
        precs = {}
        for M, idx in groups.items():
            K = dops.empty(len(idx), M, M, like=layers[idx[0]].zu)
            if len(idx) > 1:
                dops.kuu_batched([layers[i].zu for i in idx], [layers[i].ls for i in idx],
                                 [layers[i].sf for i in idx], JITTER, K)
                for g, i in enumerate(idx):
                    layers[i].Kuu = K[g]
            else:
                layers[idx[0]].kuu_into(K[0])
            Ki, ld, info = dops.spd_inv(K)
            for g, i in enumerate(idx):
                layers[i].set_kuu_inverse(Ki[g], ld[g:g + 1], info[g:g + 1])
            sizes = [2 * layers[i].Dout for i in idx]
            prec = dops.empty(sum(sizes), M, M, like=K)
            o, spans = 0, []
            for i, n in zip(idx, sizes):
                spans.append((i, o, n))
                o += n
            # R^T R of every layer's output dims in one product, then both precisions of
            # every layer (Kuu^-1 + theta1, Kuu^-1 + beta theta1) in one launch
            if len(idx) == 1:
                Rs = layers[idx[0]].theta_1_R
            elif M in R_all:
                Rs = R_all[M]
            else:
                Rs = torch.cat([layers[i].theta_1_R for i in idx])
            X = mm(Rs, Rs, tA=True)
            beta = (layers[idx[0]].N - alpha) * 1.0 / layers[idx[0]].N
            dops.prec_form([layers[i].Dout for i in idx], X, Ki, beta, prec)
            cov, ldp, infop = dops.spd_inv(prec)
            for i, o, n in spans:
                precs[i] = (prec[o:o + n], cov[o:o + n], ldp[o:o + n], infop[o:o + n])
        return precs

    def predict_f(self, inputs, prop_mode=PROP_MM, no_samples=200):
        """Base_SDGPR.predict_f (base_models.py:1132-1138): moment-matched propagation
        through the layers' posteriors (predict_f_mm, :1140-1158), or predict_f_mc."""
        _check_mode(prop_mode)
        if prop_mode == PROP_MC:
            return self.predict_f_mc(inputs, no_samples)
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self._predict_dev(x)
        return mf.cpu().numpy(), vf.cpu().numpy()

    def predict_f_mc(self, inputs, no_samples):
        """base_models.py:1159-1180: samples through the layers; the first layer draws
        (no_samples, N, D_1) normals, each later layer one normal per sample and output,
        on the legacy global RNG.  Returns (samples (no_samples, N, Dout), mf, vf of the
        last layer at the propagated samples)."""
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        N = x.shape[0]
        samples = x
        mf = vf = None
        for i, layer in enumerate(self.sgp_layers):
            mf, vf = layer.forward_prop_thru_post(samples)
            if i == 0:
                eps = _mc_draws(no_samples, mf.shape[0], mf.shape[1])
            else:
                eps = _mc_draws(1, mf.shape[0], mf.shape[1])
            samples = dops.mc_inputs(eps, mf, vf).reshape(-1, mf.shape[1])
        out = samples.reshape(no_samples, N, self.sgp_layers[-1].Dout)
        return out.cpu().numpy(), mf.cpu().numpy(), vf.cpu().numpy()

    def _predict_dev(self, x):
        mf = vf = None
        for i, layer in enumerate(self.sgp_layers):
            if i == 0:
                mf, vf = layer.forward_prop_thru_post(x)
            else:
                mf, vf = layer.forward_prop_thru_post(mf, vf)
        return mf, vf

    def predict_y(self, inputs):
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self._predict_dev(x)
        my, vy = self.lik_layer.output_probabilistic(mf, vf)
        return my.cpu().numpy(), vy.cpu().numpy()


class SGPLVM(Base_Model):
    """aep_models.py:670-867 + base_models.py:661-929.  The energy is not divided by N
    (aep_models.py:808-815)."""

    def __init__(self, y_train, hidden_size, no_pseudo, lik='Gaussian', prior_mean=0, prior_var=1,
                 nat_param=True):
        super(SGPLVM, self).__init__(y_train)
        self.Dout = self.y_train.shape[1]
        self.Din = hidden_size
        self.M = no_pseudo
        self.nat_param = nat_param
        self.prior_mean = prior_mean
        self.prior_var = prior_var
        self.lik_layer = _make_lik(lik, self.N, self.Dout)
        self.sgp_layer = SGP_Layer(self.N, self.Din, self.Dout, self.M, nat_param)
        from . import xfactors
        self._xf = xfactors

    def init_hypers(self, y_train=None):
        """base_models.py:839-873 (q(u) at the regression optimum instead of an L-BFGS fit)."""
        y = self.y_train if y_train is None else np.asarray(y_train, dtype=np.float64)
        return init.sgplvm(y, self.Din, self.M, self.nat_param)

    def update_hypers(self, params):
        dp = self._device_params(params)
        self.sgp_layer.update_hypers(dp)
        self.lik_layer.update_hypers(dp)
        return dp

    _prop_modes = (PROP_MM, PROP_MC)

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        N = self.N
        if idx is None and mb_size != N:
            idx = torch.as_tensor(np.random.choice(N, mb_size, replace=False), dtype=torch.int64,
                                  device='cuda')
        yb = self.y_dev if idx is None else self.y_dev.index_select(0, idx)
        B = yb.shape[0]
        scale_logZ = -N * 1.0 / B / alpha
        dp = self._device_params(params)
        layer = self.sgp_layer
        layer.prepare(dp, alpha)
        # (a second stream for phi and the adjoint's q(u)-only terms beside the forward
        # measured slower here: the fused forward already fills the GPU)
        phi = layer.compute_phi(alpha)
        self.lik_layer.update_hypers(dp)
        xs = self._xf.lvm_x_state(dp['x1'], dp['x2'], idx, alpha, self.prior_mean,
                                  self.prior_var, self.nat_param)
        if prop_mode == PROP_MC:
            # aep_models.py:745-760: the draws follow the minibatch draw
            eps = _mc_draws(config.MC_NO_SAMPLES, B, self.Din)
            m3, v3, kfu, x = layer.forward_prop_thru_cav(xs['mcav'], xs['vcav'], PROP_MC, eps=eps)
            dm, dv, scal = self.lik_layer.compute_log_Z_mc(m3, v3, yb, alpha, scale_logZ)
            grads, dx = layer.backprop_grads_lvm_mc(m3, v3, dm, dv, kfu, x, alpha)
            gin = layer.backprop_grads_reparam(dx, xs['mcav'], xs['vcav'], eps)
        else:
            mout, vout, psi1, psi2 = layer.forward_prop_thru_cav(xs['mcav'], xs['vcav'])
            dm, dv, dma, scal = self.lik_layer.compute_log_Z(mout, vout, yb, alpha, scale_logZ,
                                                              want_dmall=True)
            grads, gin = layer.backprop_grads_lvm_mm(mout, vout, dm, dv, psi1, psi2, xs['mcav'],
                                                     xs['vcav'], alpha, dm_all=dma)
            del psi2
        grads.update(self.lik_layer.hyper_grads(scal))
        gx1, gx2, xcontrib = self._xf.lvm_x_grads(xs, gin['mx'], gin['vx'], N, B, alpha,
                                                 self.prior_mean, self.prior_var, self.Din,
                                                 self.nat_param)
        grads['x1'] = gx1
        grads['x2'] = gx2
        energy = lincomb(torch.empty(1, dtype=F64, device='cuda'), scale_logZ, scal[0:1], 1.0,
                         xcontrib, 1.0, phi)
        return energy, grads, None

    def predict_f(self, inputs):
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self.sgp_layer.forward_prop_thru_post(x)
        return mf.cpu().numpy(), vf.cpu().numpy()

    def predict_y(self, inputs):
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self.sgp_layer.forward_prop_thru_post(x)
        my, vy = self.lik_layer.output_probabilistic(mf, vf)
        return my.cpu().numpy(), vy.cpu().numpy()


class Gauss_Emis(object):
    """Linear-Gaussian emission y_t = C x_t + N(0, diag R), R = exp(2 R_param)
    (geepee/lik_layers.py:474-676) on the GPU."""

    def __init__(self, y, Dout, Din):
        self.y = y
        self.N = y.shape[0]
        self.Dout = Dout
        self.Din = Din

    def update_hypers(self, params, key_suffix=''):
        self.C = params['C' + key_suffix]
        self.Rp = params['R' + key_suffix]

    def compute_emission_tilted(self, mx, vx, alpha, scale, rows):
        """(dmx, dvx, dC, dR, logZ*scale) for the emission rows [rows)."""
        return dops.emission(self.C, self.Rp, self.y[rows], mx, vx, alpha, scale)

    def output_probabilistic(self, mf, vf):
        """my = C mf, diagonal of C diag(vf) C^T + diag(R) (lik_layers.py:558-571)."""
        return dops.emission_predict(self.C, self.Rp, mf, vf)


class SGPSSM(Base_Model):
    """aep_models.py:991-1437 + base_models.py:1339-1752.  The dynamics layer sees N-1
    training points (aep_models.py:1018); energy and gradients are divided by N."""

    def __init__(self, y_train, hidden_size, no_pseudo, lik='Gaussian', prior_mean=0, prior_var=1,
                 x_control=None, gp_emi=False, control_to_emi=True):
        super(SGPSSM, self).__init__(y_train)
        if lik.lower() not in ('gaussian', 'probit') or (lik.lower() == 'probit' and not gp_emi):
            # base_models.py:1404-1414: a probit emission needs the GP emission layer (the
            # reference's linear probit emission class does not exist)
            raise NotImplementedError('likelihood %s not implemented' % lik)
        self.Dout = self.y_train.shape[1]
        self.Din = hidden_size
        self.M = no_pseudo
        self.gp_emi = gp_emi
        self.prior_mean = prior_mean
        self.prior_var = prior_var
        if x_control is not None:
            self.x_control = np.asarray(x_control, dtype=np.float64)
            self.Dcon_dyn = self.x_control.shape[1]
            self.Dcon_emi = self.Dcon_dyn if control_to_emi else 0
            self.u_dev = torch.as_tensor(self.x_control, device='cuda')
        else:
            self.x_control = None
            self.Dcon_dyn = self.Dcon_emi = 0
        self.dyn_layer = SGP_Layer(self.N - 1, self.Din + self.Dcon_dyn, self.Din, self.M)
        if gp_emi:
            self.emi_layer = SGP_Layer(self.N, self.Din + self.Dcon_emi, self.Dout, self.M)
            self.lik_layer = _make_lik(lik, self.N, self.Dout)
        else:
            self.emi_layer = Gauss_Emis(self.y_dev, self.Dout, self.Din + self.Dcon_emi)

    def init_hypers(self, y_train=None):
        """base_models.py:1597-1683 (latent path from y or its PCA projection; q(u) at the
        regression optimum instead of L-BFGS fits)."""
        y = self.y_train if y_train is None else np.asarray(y_train, dtype=np.float64)
        return init.sgpssm(y, self.Din, self.M, self.x_control, self.gp_emi, self.Dcon_emi > 0)

    def update_hypers(self, params):
        dp = self._device_params(params)
        self.dyn_layer.update_hypers(dp, key_suffix='_dynamic')
        self.emi_layer.update_hypers(dp, key_suffix='_emission')
        if self.gp_emi:
            self.lik_layer.update_hypers(dp, key_suffix='_emission')
        self.sn = dp['sn']
        return dp

    def _with_control(self, m, v, s, e, D):
        if D == 0:
            return m, v
        u = self.u_dev[s:e]
        return (torch.cat([m, u], dim=1), torch.cat([v, torch.zeros_like(u)], dim=1))

    _prop_modes = (PROP_MM, PROP_MC)

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        if idx is not None:
            raise NotImplementedError('SGPSSM draws contiguous windows, not index sets')
        N, Din = self.N, self.Din
        if mb_size >= N:
            s, e = 0, N
        else:
            s = int(np.random.randint(0, N - mb_size))
            e = s + mb_size
        Bd, Be = e - 1 - s, e - s
        scale_dyn = -(N - 1) * 1.0 / Bd / alpha
        scale_emi = -N * 1.0 / Be / alpha
        dp = self._device_params(params)
        dyn = self.dyn_layer
        dyn.prepare(dp, alpha, key_suffix='_dynamic')
        if self.gp_emi:
            self.emi_layer.prepare(dp, alpha, key_suffix='_emission')
            self.lik_layer.update_hypers(dp, key_suffix='_emission')
        else:
            self.emi_layer.update_hypers(dp, key_suffix='_emission')
        self.sn = dp['sn']
        cm, cv = self._xf().ssm_cavity(dp['x_factor_1'], dp['x_factor_2'], alpha, self.prior_mean,
                                       self.prior_var)
        tm1_m, tm1_v = self._with_control(cm[s:e - 1], cv[s:e - 1], s, e - 1, self.Dcon_dyn)
        up_m, up_v = self._with_control(cm[s:e], cv[s:e], s, e, self.Dcon_emi)
        if self.Dcon_dyn == 0:
            tm1_m, tm1_v = tm1_m.contiguous(), tm1_v.contiguous()
        if self.Dcon_emi == 0:
            up_m, up_v = up_m.contiguous(), up_v.contiguous()
        mc = prop_mode == PROP_MC
        if mc:
            # aep_models.py:1114-1126: the dynamics draws come first
            eps = _mc_draws(config.MC_NO_SAMPLES, Bd, Din + self.Dcon_dyn)
            m3, v3, kfu, x = dyn.forward_prop_thru_cav(tm1_m, tm1_v, PROP_MC, eps=eps)
            dmp, dvp, dmt, dvt, scal_t = dops.transition_mc(m3, v3, cm[s + 1:e].contiguous(),
                                                            cv[s + 1:e].contiguous(), self.sn, alpha,
                                                            scale_dyn)
            g_dyn, dx = dyn.backprop_grads_lvm_mc(m3, v3, dmp, dvp, kfu, x, alpha)
            gin_dyn = dyn.backprop_grads_reparam(dx, tm1_m, tm1_v, eps)
        else:
            mprop, vprop, psi1, psi2 = dyn.forward_prop_thru_cav(tm1_m, tm1_v)
            dmall, dvprop, dmt, dvt, scal_t = dops.transition(mprop, vprop, cm[s + 1:e].contiguous(),
                                                              cv[s + 1:e].contiguous(), self.sn, alpha,
                                                              scale_dyn)
            g_dyn, gin_dyn = dyn.backprop_grads_lvm_mm(mprop, vprop, None, dvprop, psi1, psi2, tm1_m,
                                                       tm1_v, alpha, dm_all=dmall)
            del psi2
        grads = {'sn': scal_t[1:2]}
        for k, v in g_dyn.items():
            grads[k + '_dynamic'] = v
        energy = lincomb(torch.empty(1, dtype=F64, device='cuda'), 1.0, scal_t[0:1], 1.0,
                         dyn.compute_phi(alpha))
        if self.gp_emi:
            emi = self.emi_layer
            ye = self.y_dev[s:e].contiguous()
            if mc:
                eps = _mc_draws(config.MC_NO_SAMPLES, Be, Din + self.Dcon_emi)
                m3, v3, kfu, x = emi.forward_prop_thru_cav(up_m, up_v, PROP_MC, eps=eps)
                dm, dv, scal_e = self.lik_layer.compute_log_Z_mc(m3, v3, ye, alpha, scale_emi)
                g_emi, dx = emi.backprop_grads_lvm_mc(m3, v3, dm, dv, kfu, x, alpha)
                gin_emi = emi.backprop_grads_reparam(dx, up_m, up_v, eps)
            else:
                mout, vout, e1, e2 = emi.forward_prop_thru_cav(up_m, up_v)
                dm, dv, dma, scal_e = self.lik_layer.compute_log_Z(mout, vout, ye, alpha, scale_emi,
                                                                   want_dmall=True)
                g_emi, gin_emi = emi.backprop_grads_lvm_mm(mout, vout, dm, dv, e1, e2, up_m, up_v, alpha,
                                                           dm_all=dma)
                del e2
            for k, v in g_emi.items():
                grads[k + '_emission'] = v
            grads.update(self.lik_layer.hyper_grads(scal_e, '_emission'))
            lincomb(energy, 1.0, energy, scale_emi, scal_e[0:1], 1.0, emi.compute_phi(alpha))
            dup, vup = gin_emi['mx'], gin_emi['vx']
        else:
            dup, vup, dC, dR, lz = dops.emission(self.emi_layer.C, self.emi_layer.Rp,
                                                 self.y_dev[s:e], up_m, up_v, alpha, scale_emi)
            grads['C_emission'] = dC
            grads['R_emission'] = dR
            lincomb(energy, 1.0, energy, 1.0, lz)
        g1, g2, xc = self._xf().ssm_grads(dp['x_factor_1'], dp['x_factor_2'], alpha, self.prior_mean,
                                          self.prior_var, s, e, dup, vup, dmt, dvt,
                                          gin_dyn['mx'], gin_dyn['vx'])
        grads['x_factor_1'] = g1
        grads['x_factor_2'] = g2
        lincomb(energy, 1.0, energy, 1.0, xc)
        return energy, grads, N

    @staticmethod
    def _xf():
        from . import xfactors
        return xfactors

    # posterior latent path and rollouts (base_models.py:1453-1542, 1562-1600) -------
    def _posterior_x_dev(self):
        dp = self._dp.views if self._dp is not None else self._device_params(self.get_hypers())
        return dops.ssm_posterior(dp['x_factor_1'], dp['x_factor_2'], self.prior_mean, self.prior_var)

    def get_posterior_x(self, idxs=None):
        """Base_SGPSSM.get_posterior_x (base_models.py:1562-1576)."""
        pm, pv = self._posterior_x_dev()
        pm, pv = pm.cpu().numpy(), pv.cpu().numpy()
        if idxs is not None:
            pm, pv = pm[idxs], pv[idxs]
        return pm, pv

    def get_posterior_y(self):
        """Base_SGPSSM.get_posterior_y (base_models.py:1578-1595): (my, vf, vyn)."""
        pm, pv = self._posterior_x_dev()
        if self.Dcon_emi > 0:
            pm = torch.cat([pm, self.u_dev], dim=1)
            pv = torch.cat([pv, torch.zeros_like(self.u_dev)], dim=1)
        if self.gp_emi:
            mf, vf = self.emi_layer.forward_prop_thru_post(pm, pv)
            my, vyn = self.lik_layer.output_probabilistic(mf, vf)
        else:
            my, vyn = self.emi_layer.output_probabilistic(pm, pv)
            _, vf = dops.emission_predict(self.emi_layer.C, self._no_noise(), pm, pv)
        return my.cpu().numpy(), vf.cpu().numpy(), vyn.cpu().numpy()

    def _no_noise(self):
        """Emission log-noise of -inf: gpx_emission_predict then returns the noiseless
        variance diag(C diag(v) C^T) exactly."""
        D = self.Dout
        if getattr(self, '_ninf', None) is None or self._ninf.numel() != D:
            self._ninf = torch.full((D,), float('-inf'), dtype=F64, device='cuda')
        return self._ninf

    def predict_forward(self, T, x_control=None, prop_mode=PROP_MM, no_samples=200):
        """Base_SGPSSM.predict_forward (base_models.py:1453-1461)."""
        if prop_mode == PROP_MM:
            return self.predict_forward_mm(T, x_control)
        if prop_mode == PROP_MC:
            return self.predict_forward_mc(T, x_control, no_samples)
        if prop_mode == 'LIN':
            raise NotImplementedError('TODO')
        raise NotImplementedError('unknown prop mode %s' % prop_mode)

    def _emit(self, m, v, sn2):
        """(my, vf, vyn) of latent inputs m, v (control columns already appended)."""
        if self.gp_emi:
            mf, vf = self.emi_layer.forward_prop_thru_post(m, v)
            ones = torch.ones(1, dtype=F64, device='cuda')
            return mf, vf, lincomb(torch.empty_like(vf), 1.0, vf, sn2, ones)
        my, vyn = dops.emission_predict(self.emi_layer.C, self.emi_layer.Rp, m, v)
        _, vf = dops.emission_predict(self.emi_layer.C, self._no_noise(), m, v)
        return my, vf, vyn

    def predict_forward_mm(self, T, x_control=None):
        """base_models.py:1463-1502: T moment-matched steps of the dynamics posterior from
        the last latent state, each emitted; every step stays on the device and the whole
        path comes back in one copy.  Returns (mx, vx, my, vy_noiseless, vy)."""
        pm, pv = self._posterior_x_dev()
        u = None if x_control is None else torch.as_tensor(np.asarray(x_control, dtype=np.float64),
                                                           device='cuda')
        sn2 = float(np.exp(2.0 * np.asarray(self.get_hypers()['sn_emission']).ravel()[0])) \
            if self.gp_emi else 0.0
        mx = torch.empty(T, self.Din, dtype=F64, device='cuda')
        vx = torch.empty_like(mx)
        my = torch.empty(T, self.Dout, dtype=F64, device='cuda')
        vf = torch.empty_like(my)
        vy = torch.empty_like(my)
        m, v = pm[-1:], pv[-1:]
        for t in range(T):
            if self.Dcon_dyn > 0:
                m = torch.cat([m, u[t:t + 1]], dim=1)
                v = torch.cat([v, torch.zeros_like(u[t:t + 1])], dim=1)
            mt, vt = self.dyn_layer.forward_prop_thru_post(m, v)
            if self.gp_emi:
                if self.Dcon_emi > 0:
                    mtc = torch.cat([mt, u[t:t + 1]], dim=1)
                    vtc = torch.cat([vt, torch.zeros_like(u[t:t + 1])], dim=1)
                else:
                    mtc, vtc = mt, vt
                a, b, c = self._emit(mtc, vtc, sn2)
            else:
                # the linear emission reads the latent state alone (base_models.py:1490)
                a, b, c = self._emit(mt, vt, sn2)
            mx[t:t + 1].copy_(mt)
            vx[t:t + 1].copy_(vt)
            my[t:t + 1].copy_(a)
            vf[t:t + 1].copy_(b)
            vy[t:t + 1].copy_(c)
            m, v = mt, vt
        return tuple(o.cpu().numpy() for o in (mx, vx, my, vf, vy))

    def predict_forward_mc(self, T, x_control=None, no_samples=200):
        """base_models.py:1504-1542: sampled latent paths; draws from numpy's legacy global
        RNG as the reference makes them.  Returns (x, my, vy) with a leading time axis."""
        if self.Dcon_emi > 0:
            raise NotImplementedError('Monte Carlo rollout with control inputs to the emission')
        pm, pv = self._posterior_x_dev()
        u = None if x_control is None else torch.as_tensor(np.asarray(x_control, dtype=np.float64),
                                                           device='cuda')
        sn2 = float(np.exp(2.0 * np.asarray(self.get_hypers()['sn_emission']).ravel()[0])) \
            if self.gp_emi else 0.0
        S, Din = no_samples, self.Din
        x = torch.empty(T, S, Din, dtype=F64, device='cuda')
        my = torch.empty(T, S, self.Dout, dtype=F64, device='cuda')
        vy = torch.empty_like(my)
        xs = dops.mc_inputs(_mc_draws(S, 1, Din), pm[-1:], pv[-1:]).reshape(S, Din)
        zeros = torch.zeros(S, Din, dtype=F64, device='cuda')
        for t in range(T):
            xc = torch.cat([xs, u[t:t + 1].expand(S, -1)], dim=1) if self.Dcon_dyn > 0 else xs
            mt, vt = self.dyn_layer.forward_prop_thru_post(xc)
            xs = dops.mc_inputs(_mc_draws(1, S, Din), mt, vt).reshape(S, Din)
            if self.gp_emi:
                mf, vf = self.emi_layer.forward_prop_thru_post(xs)
                ones = torch.ones(1, dtype=F64, device='cuda')
                a, c = mf, lincomb(torch.empty_like(vf), 1.0, vf, sn2, ones)
            else:
                a, c = dops.emission_predict(self.emi_layer.C, self.emi_layer.Rp, xs, zeros)
            x[t].copy_(xs)
            my[t].copy_(a)
            vy[t].copy_(c)
        return x.cpu().numpy(), my.cpu().numpy(), vy.cpu().numpy()

    def predict_f(self, inputs):
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self.dyn_layer.forward_prop_thru_post(x)
        return mf.cpu().numpy(), vf.cpu().numpy()

    def predict_y(self, inputs):
        x = torch.as_tensor(np.asarray(inputs, dtype=np.float64), device='cuda')
        mf, vf = self.dyn_layer.forward_prop_thru_post(x)
        if self.gp_emi:
            mg, vg = self.emi_layer.forward_prop_thru_post(mf, vf)
            my, vy = self.lik_layer.output_probabilistic(mg, vg)
        else:
            my, vy = self.emi_layer.output_probabilistic(mf, vf)
        return my.cpu().numpy(), vy.cpu().numpy()


class SDGPR_H(SDGPR):
    """Deep GP regression with Gaussian factors on the hidden layer values
    (aep_models.py:1440-1863): each hidden value h_i[n] carries two tied factors with
    natural parameters (h_factor_1_i, exp(2 h_factor_2_i)); the layers are linked by
    Gaussian transitions of noise exp(2 sn_hidden[i]).  Layer i's output keys end in '_i'.
    The objective is defined for the full batch (the reference's hidden-factor gradients
    broadcast minibatch rows against all N rows, :1633-1652)."""

    def __init__(self, x_train, y_train, no_pseudos, hidden_sizes, lik='Gaussian'):
        super(SDGPR_H, self).__init__(x_train, y_train, no_pseudos, hidden_sizes, lik)

    def init_hypers(self, y_train=None):
        """aep_models.py:1817-1840: the layers as in SDGPR, sn_hidden = log 0.001, hidden
        factors with h_factor_1 = 0 and h_factor_2 = log 0.01."""
        p = super(SDGPR_H, self).init_hypers(y_train)
        p['sn_hidden'] = np.log(0.001) * np.ones(self.L - 1)
        for i in range(self.L - 1):
            p['h_factor_1_%d' % i] = np.zeros((self.N, self.size[i + 1]))
            p['h_factor_2_%d' % i] = np.log(0.01) * np.ones((self.N, self.size[i + 1]))
        return p

    def _objective(self, params, mb_size, alpha, idx=None, prop_mode=PROP_MM):
        N, L = self.N, self.L
        if idx is not None or mb_size < N:
            raise NotImplementedError('SDGPR_H objective is defined for the full batch only')
        if L < 2:
            raise ValueError('SDGPR_H needs at least one hidden layer')
        scale = -1.0 / alpha
        dp = self._device_params(params)
        self.lik_layer.update_hypers(dp)
        energy = torch.zeros(1, dtype=F64, device='cuda')
        for i, layer in enumerate(self.sgp_layers):
            layer.prepare(dp, alpha, key_suffix='_%d' % i)
            lincomb(energy, 1.0, energy, 1.0, layer.compute_phi(alpha))
        h = [(dp['h_factor_1_%d' % i], dp['h_factor_2_%d' % i]) for i in range(L - 1)]
        cav = [dops.hidden_cavity(h1, h2, alpha) for h1, h2 in h]
        sn = dp['sn_hidden']
        grads, dmc, dvc, dsn = {}, [None] * (L - 1), [None] * (L - 1), []
        for i in range(L - 1):
            layer = self.sgp_layers[i]
            if i == 0:
                mp, vp, kfu = layer.forward_prop_thru_cav(self.x_dev)
            else:
                mp, vp, p1, p2 = layer.forward_prop_thru_cav(cav[i - 1][0], cav[i - 1][1])
            dmall, dvp, dmt, dvt, scal = dops.transition(mp, vp, cav[i][0], cav[i][1], sn[i:i + 1],
                                                         alpha, scale)
            if i == 0:
                dmp = lincomb(torch.empty_like(dmt), -1.0, dmt)
                g = layer.backprop_grads_reg(mp, vp, dmp, dvp, kfu, self.x_dev, alpha)
            else:
                g, gin = layer.backprop_grads_lvm_mm(mp, vp, None, dvp, p1, p2, cav[i - 1][0],
                                                     cav[i - 1][1], alpha, dm_all=dmall)
                del p2
                lincomb(dmc[i - 1], 1.0, dmc[i - 1], 1.0, gin['mx'])
                lincomb(dvc[i - 1], 1.0, dvc[i - 1], 1.0, gin['vx'])
            dmc[i], dvc[i] = dmt, dvt
            lincomb(energy, 1.0, energy, 1.0, scal[0:1])
            dsn.append(scal[1:2])
            for k, val in g.items():
                grads[k + '_%d' % i] = val
        i = L - 1
        layer = self.sgp_layers[i]
        mp, vp, p1, p2 = layer.forward_prop_thru_cav(cav[i - 1][0], cav[i - 1][1])
        dm, dv, dma, scal = self.lik_layer.compute_log_Z(mp, vp, self.y_dev, alpha, scale, want_dmall=True)
        g, gin = layer.backprop_grads_lvm_mm(mp, vp, dm, dv, p1, p2, cav[i - 1][0], cav[i - 1][1], alpha,
                                             dm_all=dma)
        del p2
        lincomb(dmc[i - 1], 1.0, dmc[i - 1], 1.0, gin['mx'])
        lincomb(dvc[i - 1], 1.0, dvc[i - 1], 1.0, gin['vx'])
        lincomb(energy, 1.0, energy, scale, scal[0:1])
        for k, val in g.items():
            grads[k + '_%d' % i] = val
        grads.update(self.lik_layer.hyper_grads(scal))
        for j, (h1, h2) in enumerate(h):
            g1, g2, hs = dops.hidden_grads(h1, h2, alpha, dmc[j], dvc[j])
            grads['h_factor_1_%d' % j] = g1
            grads['h_factor_2_%d' % j] = g2
            lincomb(energy, 1.0, energy, 1.0, hs[0:1])
        grads['sn_hidden'] = torch.cat(dsn)
        return energy, grads, N
