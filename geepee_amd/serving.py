"""Serving resident fitted models: grouped predict launches with continuous batching.

A fitted model is its params dict (the format Base_Model.get_hypers / save_model produce,
geepee/base_models.py:149-165, 599-628).  ``ModelStore`` keeps, per model and layer, the
inducing inputs, hyper-parameters and the posterior predictive state A, B_det / B_sto of
Base_SGP_Layer.update_posterior (base_models.py:466-488) resident in HBM, plus one
device descriptor table.  ``ModelStore.predict_many`` answers any set of requests with
a single launch of gpx_predict_grouped.

``PredictServer`` is the request loop: requests arrive over time; whenever a launch
slot (one of two streams) is free it admits every arrived request that fits the HBM
budget for request I/O, copies their inputs with one host-to-device transfer, launches
one grouped predict and copies the results back, while the other slot's launch is still
running.  Latency = time from a request's arrival to its result being host-visible.
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib, checkpoint
from .aep_models import DeviceParams, _shapes
from .layers import SGP_Layer

_c_int, _c_long, _c_p = ctypes.c_int, ctypes.c_long, ctypes.c_void_p


class _Layer(ctypes.Structure):
    _fields_ = [('type', _c_int), ('M', _c_int), ('Din', _c_int), ('Dout', _c_int),
                ('zu', _c_p), ('ls', _c_p), ('sf', _c_p), ('A', _c_p), ('B', _c_p)]


class _Model(ctypes.Structure):
    _fields_ = [('nlayers', _c_int), ('Din', _c_int), ('Dout', _c_int), ('pad_', _c_int),
                ('sn', _c_p), ('layer', _Layer * 2)]


class _Item(ctypes.Structure):
    _fields_ = [('model', _c_int), ('rows', _c_int), ('in_off', _c_long), ('out_off', _c_long)]


def _check_layout():
    L = _lib.load()
    mb, ib = ctypes.c_int(), ctypes.c_int()
    L.gpx_serve_desc_sizes(ctypes.byref(mb), ctypes.byref(ib))
    assert mb.value == ctypes.sizeof(_Model) and ib.value == ctypes.sizeof(_Item), \
        (mb.value, ctypes.sizeof(_Model), ib.value, ctypes.sizeof(_Item))


class ModelStore(object):
    """Resident SGPR / two-layer SDGPR models in posterior (predictive) form."""

    def __init__(self, specs):
        """specs: list of dicts with kind ('sgpr' | 'sdgpr'), Din, Ms, N, hidden and either
        params (a params dict) or checkpoint (the path of a file save_model wrote, read as
        arrays only)."""
        _lib.require_device()
        _check_layout()
        self.tile = _lib.query('gpx_serve_tile')
        specs = [sp if 'params' in sp else dict(sp, params=checkpoint.load_params(sp['checkpoint']))
                 for sp in specs]
        self.specs = specs
        self.layers = []
        self._keep = []
        descs = (_Model * len(specs))()
        for i, sp in enumerate(specs):
            dp = DeviceParams(_shapes(sp['params']))
            dp.upload(sp['params'])
            self._keep.append(dp)
            if sp['kind'] == 'sgpr':
                sizes = [sp['Din'], sp.get('Dout', 1)]
            else:
                sizes = [sp['Din']] + list(sp['hidden']) + [sp.get('Dout', 1)]
            nl = len(sizes) - 1
            if nl > 2:
                raise ValueError('serving supports at most two layers')
            lay = []
            for l in range(nl):
                sfx = '' if sp['kind'] == 'sgpr' else '_%d' % l
                L = SGP_Layer(sp['N'], sizes[l], sizes[l + 1], sp['Ms'][l])
                L.update_hypers(dp.views, key_suffix=sfx)
                L._posterior_predictive()
                lay.append(L)
                d = descs[i].layer[l]
                d.type, d.M, d.Din, d.Dout = (0 if l == 0 else 1), L.M, L.Din, L.Dout
                d.zu, d.ls, d.sf = L.zu.data_ptr(), L.ls.data_ptr(), L.sf.data_ptr()
                d.A = L.A.data_ptr()
                d.B = (L.B_det if l == 0 else L.B_sto).data_ptr()
            descs[i].nlayers, descs[i].Din, descs[i].Dout = nl, sizes[0], sizes[-1]
            descs[i].sn = dp.views['sn'].data_ptr() if 'sn' in dp.views else None
            if _lib.load().gpx_serve_validate(ctypes.byref(descs[i])) != 0:
                raise ValueError('model %d exceeds the serving kernel limits' % i)
            self.layers.append(lay)
        raw = np.frombuffer(bytes(descs), dtype=np.uint8)
        self.desc_dev = torch.from_numpy(raw.copy()).cuda()
        torch.cuda.synchronize()
        self.Din = [int(sp['Din']) for sp in specs]
        self.Dout = [int(descs[i].Dout) for i in range(len(specs))]
        self._din = np.asarray(self.Din, dtype=np.int64)
        self._dout = np.asarray(self.Dout, dtype=np.int64)
        self._bufs = {}
        assert self.ITEM_DTYPE.itemsize == ctypes.sizeof(_Item)
        self.state_bytes = sum(t.numel() * 8 for L in self.layers for l in L
                               for t in (l.A, l.B_det, l.B_sto, l.zu))

    # -- one grouped launch ------------------------------------------------------
    ITEM_DTYPE = np.dtype([('model', '<i4'), ('rows', '<i4'), ('in_off', '<i8'), ('out_off', '<i8')])

    def plan(self, reqs):
        """Work-item table (structured array in the GpxServeItem layout) and I/O offsets
        for requests [(model_id, X)], built with vectorised numpy."""
        mids = np.fromiter((m for m, _ in reqs), dtype=np.int64, count=len(reqs))
        ns = np.fromiter((X.shape[0] for _, X in reqs), dtype=np.int64, count=len(reqs))
        din = self._din[mids]
        dout = self._dout[mids]
        in_sz, out_sz = ns * din, ns * dout
        in_off = np.concatenate(([0], np.cumsum(in_sz)[:-1]))
        out_off = np.concatenate(([0], np.cumsum(out_sz)[:-1]))
        ntile = -(-ns // self.tile)
        ridx = np.repeat(np.arange(len(reqs)), ntile)
        first = np.concatenate(([0], np.cumsum(ntile)[:-1]))
        j = np.arange(int(ntile.sum())) - np.repeat(first, ntile)
        items = np.empty(len(ridx), dtype=self.ITEM_DTYPE)
        items['model'] = mids[ridx]
        items['rows'] = np.minimum(self.tile, ns[ridx] - j * self.tile)
        items['in_off'] = in_off[ridx] + j * self.tile * din[ridx]
        items['out_off'] = out_off[ridx] + j * self.tile * dout[ridx]
        offs = list(zip(in_off.tolist(), out_off.tolist(), ns.tolist()))
        return items, int(in_sz.sum()), int(out_sz.sum()), offs

    def _slot_buffers(self, key, nin, nout, nitems):
        """Per-stream pinned host and device buffers, grown on demand and reused."""
        b = self._bufs.get(key)
        need = (max(nin, 1), max(nout, 1), max(nitems, 1))
        if b is None or any(x > y for x, y in zip(need, b['cap'])):
            cap = tuple(max(2 * x, 1024) for x in need)
            b = {'cap': cap,
                 'xh': torch.empty(cap[0], dtype=torch.float64, pin_memory=True),
                 'ih': torch.empty(cap[2] * self.ITEM_DTYPE.itemsize, dtype=torch.uint8, pin_memory=True),
                 'mh': torch.empty(cap[1], dtype=torch.float64, pin_memory=True),
                 'vh': torch.empty(cap[1], dtype=torch.float64, pin_memory=True),
                 'xd': torch.empty(cap[0], dtype=torch.float64, device='cuda'),
                 'idev': torch.empty(cap[2] * self.ITEM_DTYPE.itemsize, dtype=torch.uint8, device='cuda'),
                 'md': torch.empty(cap[1], dtype=torch.float64, device='cuda'),
                 'vd': torch.empty(cap[1], dtype=torch.float64, device='cuda')}
            self._bufs[key] = b
        return b

    def launch(self, reqs, noisy=False, stream=None, bufs=None):
        """Enqueue one grouped predict for reqs on `stream`; returns a handle whose
        results are valid after handle['done'] (an event) completes.  The stream's
        buffers are reused, so a stream holds one launch in flight at a time."""
        items, nin, nout, offs = self.plan(reqs)
        stream = stream or torch.cuda.current_stream()
        b = self._slot_buffers(stream.cuda_stream, nin, nout, len(items))
        xv = b['xh'].numpy()
        o = 0
        for _, X in reqs:
            xv[o:o + X.size] = np.asarray(X, dtype=np.float64).ravel()
            o += X.size
        ib = items.view(np.uint8)
        b['ih'].numpy()[:ib.size] = ib
        with torch.cuda.stream(stream):
            b['xd'][:max(nin, 1)].copy_(b['xh'][:max(nin, 1)], non_blocking=True)
            b['idev'][:ib.size].copy_(b['ih'][:ib.size], non_blocking=True)
            _lib.call('gpx_predict_grouped', len(items), self.desc_dev.data_ptr(), b['idev'].data_ptr(),
                      b['xd'].data_ptr(), b['md'].data_ptr(), b['vd'].data_ptr(), int(bool(noisy)),
                      stream.cuda_stream)
            b['mh'][:max(nout, 1)].copy_(b['md'][:max(nout, 1)], non_blocking=True)
            b['vh'][:max(nout, 1)].copy_(b['vd'][:max(nout, 1)], non_blocking=True)
            done = torch.cuda.Event()
            done.record(stream)
        return {'done': done, 'mh': b['mh'], 'vh': b['vh'], 'offs': offs, 'reqs': reqs,
                'bytes': 8 * (nin + 2 * nout)}

    def collect(self, h):
        h['done'].synchronize()
        out = []
        for (io, oo, n), (mid, X) in zip(h['offs'], h['reqs']):
            D = self.Dout[mid]
            out.append((h['mh'].numpy()[oo:oo + n * D].reshape(n, D).copy(),
                        h['vh'].numpy()[oo:oo + n * D].reshape(n, D).copy()))
        return out

    def predict_many(self, reqs, noisy=False):
        return self.collect(self.launch(reqs, noisy))

    # -- per-request plain path (one op per reference call, psi2 materialised) --
    def predict_plain(self, mid, X, noisy=False):
        x = torch.as_tensor(np.asarray(X, dtype=np.float64), device='cuda')
        mf = vf = None
        for l, L in enumerate(self.layers[mid]):
            mf, vf = L.forward_prop_thru_post(x) if l == 0 else L.forward_prop_thru_post(mf, vf)
        if noisy:
            sn = self.specs[mid]['params']['sn']
            sn2 = float(np.exp(2.0 * np.asarray(sn).ravel()[0]))
            ones = torch.ones(1, dtype=torch.float64, device='cuda')
            vf = _lib_lincomb(vf, sn2, ones)
        return mf.cpu().numpy(), vf.cpu().numpy()


def _lib_lincomb(v, c, ones):
    from .dops import lincomb
    return lincomb(torch.empty_like(v), 1.0, v, c, ones)


def make_requests(n_models, n_req, rate, dins, seed=0, max_batch=4096):
    """Poisson arrivals at `rate` per second, uniform over models, batch sizes
    floor(exp(U(0, ln max_batch))), points N(0, I)."""
    rs = np.random.RandomState(seed)
    gaps = rs.exponential(1.0 / rate, size=n_req)
    t = np.cumsum(gaps)
    reqs = []
    for i in range(n_req):
        mid = int(rs.randint(n_models))
        b = max(1, int(np.floor(np.exp(rs.uniform(0, np.log(max_batch))))))
        reqs.append((float(t[i]), mid, rs.randn(b, dins[mid])))
    return reqs


def admit(sizes, used, budget, max_points, idle):
    """How many queued requests (in arrival order) one launch takes.  sizes: an iterable
    of (bytes, points) over a non-empty queue, read only as far as the launch reaches;
    used: request I/O bytes of the launches in flight.  Requests
    are taken while the in-flight bytes stay within `budget` and the points within
    `max_points`.  A request larger than the whole budget runs alone, and only when
    nothing else is in flight (`idle`); otherwise the queue head waits for memory."""
    n, need, pts = 0, 0, 0
    for b, p in sizes:
        if used + need + b > budget or pts + p > max_points:
            break
        n += 1
        need += b
        pts += p
    if n == 0 and idle and used == 0:
        return 1
    return n


class PredictServer(object):
    """Continuous batching over two launch slots with an HBM budget for request I/O."""

    def __init__(self, store, budget_bytes=256 << 20, max_points=1 << 16, slots=2):
        self.store = store
        self.budget = budget_bytes
        self.max_points = max_points
        self.streams = [torch.cuda.Stream() for _ in range(slots)]

    def run(self, requests, noisy=False):
        """requests: [(arrival_s, model_id, X)] sorted by arrival.  Replays the arrival
        process against the wall clock; returns latencies (s), outputs, wall time."""
        n = len(requests)
        lat = np.zeros(n)
        outs = [None] * n
        nxt = 0
        queue = []
        inflight = [None] * len(self.streams)
        used = 0
        t0 = time.perf_counter()
        done_count = 0
        while done_count < n:
            now = time.perf_counter() - t0
            while nxt < n and requests[nxt][0] <= now:
                queue.append(nxt)
                nxt += 1
            for s, h in enumerate(inflight):
                if h is not None and h['done'].query():
                    res = self.store.collect(h)
                    tdone = time.perf_counter() - t0
                    for j, r in zip(h['ids'], res):
                        outs[j] = r
                        lat[j] = tdone - requests[j][0]
                    used -= h['bytes']
                    done_count += len(h['ids'])
                    inflight[s] = None
            free = [s for s, h in enumerate(inflight) if h is None]
            ntake = 0
            if queue and free:
                sizes = ((8 * requests[j][2].shape[0] * (self.store.Din[requests[j][1]]
                                                         + 2 * self.store.Dout[requests[j][1]]),
                          requests[j][2].shape[0]) for j in queue)
                ntake = admit(sizes, used, self.budget, self.max_points,
                              idle=all(h is None for h in inflight))
            if ntake:
                take = queue[:ntake]
                queue = queue[ntake:]
                s = free[0]
                h = self.store.launch([(requests[j][1], requests[j][2]) for j in take], noisy,
                                      stream=self.streams[s])
                h['ids'] = take
                used += h['bytes']
                inflight[s] = h
            elif nxt < n and not queue and all(h is None for h in inflight):
                wait = requests[nxt][0] - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
        wall = time.perf_counter() - t0
        return lat, outs, wall

    def run_plain(self, requests, noisy=False):
        """Each request alone in arrival order through the per-op path."""
        n = len(requests)
        lat = np.zeros(n)
        outs = [None] * n
        t0 = time.perf_counter()
        for j, (ta, mid, X) in enumerate(requests):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            outs[j] = self.store.predict_plain(mid, X, noisy)
            lat[j] = time.perf_counter() - t0 - ta
        return lat, outs, time.perf_counter() - t0
