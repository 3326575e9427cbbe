"""Launch-shape autotuner for the fused moment-matched kernels and the batched GEMM.

Each candidate launch shape is timed with HIP events on the stream it runs on, over
synthetic inputs of the requested shape; the fastest is kept per (kernel, shape,
device) in a JSON cache.  Every candidate computes the same quantity; they differ only
in how the work is split over workgroups (and so in the fixed order of the partial
sums), so any of them meets the same tolerance.

Knobs
  fused forward   nsub (16*nsub points per workgroup), splits_f (pair-tile splits)
  fused backward  splits_b (column-tile splits of the per-point pass),
                  nsplit (point ranges of the per-pair pass)
  gemm            tile (32 or 64 rows/cols per workgroup), split-K factor

Cache: ``$GEEPEE_AMD_TUNE_CACHE`` if set, else ``~/.cache/geepee_amd/tune.json``, over
the plans shipped in ``tuned_plans.json`` (the benchmark shapes, tuned on MI355X).
``lookup`` is consulted by dops.fused_plan and dops.mm before their heuristics.
"""
import itertools
import json
import os

import numpy as np
import torch

from . import _lib, dops

_cache = None


def cache_path():
    return os.environ.get('GEEPEE_AMD_TUNE_CACHE') or os.path.join(
        os.path.expanduser('~'), '.cache', 'geepee_amd', 'tune.json')


_tag = None


def _device_tag():
    global _tag
    if _tag is None:
        try:
            _tag = torch.cuda.get_device_properties(0).gcnArchName.split(':')[0]
        except (RuntimeError, AssertionError, AttributeError):
            _tag = 'gfx950'
    return _tag


SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tuned_plans.json')


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _load():
    """Plans shipped with the package (tuned on MI355X for the benchmark shapes),
    overridden by the user's cache."""
    global _cache
    if _cache is None:
        _cache = dict(_read(SHIPPED)) if os.path.exists(SHIPPED) else {}
        p = cache_path()
        if os.path.exists(p):
            _cache.update(_read(p))
    return _cache


def _key(kind, shape):
    return '%s|%s|%s' % (_device_tag(), kind, ','.join(str(int(s)) for s in shape))


def lookup(kind, shape):
    c = _load()
    return c.get(_key(kind, shape)) if c else None


def store(kind, shape, plan, save=True):
    c = _load()
    c[_key(kind, shape)] = plan
    if save:
        p = cache_path()
        os.makedirs(os.path.dirname(p), exist_ok=True)
        user = _read(p) if os.path.exists(p) else {}
        user[_key(kind, shape)] = plan
        tmp = p + '.tmp%d' % os.getpid()
        with open(tmp, 'w') as f:
            json.dump(user, f, indent=1, sort_keys=True)
        os.replace(tmp, p)


def clear(save=False):
    """Forget the in-memory table (the next lookup re-reads the cache file); with
    save=True also delete the file."""
    global _cache
    _cache = None
    if save and os.path.exists(cache_path()):
        os.remove(cache_path())


def _time(fn, reps):
    st = torch.cuda.current_stream()
    fn()
    st.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(st)
    for _ in range(reps):
        fn()
    t1.record(st)
    t1.synchronize()
    return t0.elapsed_time(t1) / reps


def _pow2_upto(lo, hi):
    out, v = [], lo
    while v <= hi:
        out.append(v)
        v *= 2
    if out[-1] != hi:
        out.append(hi)
    return out


def fused_candidates(N, M):
    T = (M + 15) // 16
    ntiles = T * (T + 1) // 2
    fwd = [{'nsub': ns, 'splits_f': s} for ns in (1, 2, 4) for s in _pow2_upto(1, min(ntiles, 64))
           if -(-N // (16 * ns)) * s >= 64]
    bwd_points = [s for s in _pow2_upto(1, min(T, 64)) if -(-N // 16) * s >= 64]
    bwd_pairs = [s for s in _pow2_upto(1, min(-(-N // 16), 64)) if ntiles * 4 * s >= 64]
    # point-range counts that fill whole rounds of 256 workgroups over the tiles (the
    # merged adjoint kernel of D > 16 runs one workgroup per CU)
    bwd_pairs = sorted(set(bwd_pairs) | {k * 256 // ntiles for k in (1, 2, 3, 4)
                                         if 1 <= k * 256 // ntiles <= min(-(-N // 16), 64)})
    return fwd or [{'nsub': 1, 'splits_f': min(ntiles, 64)}], bwd_points or [min(T, 64)], \
        bwd_pairs or [min(-(-N // 16), 64)]


def tune_fused(N, M, Q, D, reps=5, seed=0, save=True, verbose=False):
    """Time the fused forward and backward launch shapes at (N, M, Q, D); store and
    return the best plan and the table of timings (ms)."""
    _lib.require_device()
    if not dops.fused_fits(M, Q, D):
        raise ValueError('shape outside the fused kernels (M <= 256, Q <= 15, D <= 32)')
    rs = np.random.RandomState(seed)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    mx, vx, z = dev(rs.randn(N, Q)), dev(rs.uniform(0.01, 0.5, (N, Q))), dev(rs.randn(M, Q))
    ls, sf = dev(np.zeros(Q)), dev(np.zeros(1))
    Ahat = dev(rs.randn(D, M) * 0.1)
    Bp = dops.sym_pairmajor(dev(rs.randn(D, M, M) * 0.01))
    dm_all, dv = dev(rs.randn(N, D)), dev(rs.randn(N, D))
    fwd, bpt, bpr = fused_candidates(N, M)
    base = dops.fused_plan(N, M, Q, D, tuned=False)
    table = {'fwd': {}, 'bwd': {}}
    best_f, tf = None, float('inf')
    for c in fwd:
        t = _time(lambda: dops.fused_mm_fwd(mx, vx, z, ls, sf, Ahat, Bp, plan=dict(base, **c)), reps)
        table['fwd']['%d/%d' % (c['nsub'], c['splits_f'])] = t
        if t < tf:
            best_f, tf = c, t
    psi1, _, _ = dops.fused_mm_fwd(mx, vx, z, ls, sf, Ahat, Bp, plan=dict(base, **best_f))
    best_b, tb = None, float('inf')
    for sb, ns in itertools.product(bpt, bpr):
        p = dict(base, splits_b=sb, nsplit=ns)
        t = _time(lambda: dops.fused_mm_bwd(mx, vx, z, ls, sf, Ahat, Bp, psi1, dm_all, dv, plan=p), reps)
        table['bwd']['%d/%d' % (sb, ns)] = t
        if t < tb:
            best_b, tb = {'splits_b': sb, 'nsplit': ns}, t
    plan = dict(best_f, **best_b)
    plan['ms_fwd'], plan['ms_bwd'] = tf, tb
    store('fused', (N, M, Q, D), plan, save=save)
    if verbose:
        print('fused N=%d M=%d Q=%d D=%d -> %s' % (N, M, Q, D, plan), flush=True)
    return plan, table


def gemm_candidates(m, n, k, batch):
    tiles = [32, 64]
    out = []
    for tile in tiles:
        nt = -(-m // tile) * -(-n // tile) * batch
        for s in _pow2_upto(1, 64):
            if s > 1 and (k // s < 64 or nt * s > 4096):
                continue
            out.append({'tile': tile, 'splits': s})
    return out


def tune_gemm(m, n, k, batch=1, tA=False, tB=False, reps=10, seed=0, save=True):
    """Time the GEMM tile / split-K candidates for one shape; store and return the best."""
    _lib.require_device()
    rs = np.random.RandomState(seed)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device='cuda')
    A = dev(rs.randn(batch, k, m) if tA else rs.randn(batch, m, k))
    B = dev(rs.randn(batch, n, k) if tB else rs.randn(batch, k, n))
    C = dops.empty(batch, m, n, like=A)
    best, bt, table = None, float('inf'), {}
    for c in gemm_candidates(m, n, k, batch):
        t = _time(lambda: _lib.mm(A, B, C, tA=tA, tB=tB, splits=c['splits'], tile=c['tile']), reps)
        table['%d/%d' % (c['tile'], c['splits'])] = t
        if t < bt:
            best, bt = c, t
    plan = dict(best, ms=bt)
    store('gemm', (m, n, k, batch, int(tA), int(tB)), plan, save=save)
    return plan, table
