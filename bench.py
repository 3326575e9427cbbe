"""Benchmark: AEP energy + analytic-gradient evaluations per second on one MI355X.

Default workload: the SGPLVM uncertain-input objective (N=4096 latent Gaussians, Q=8,
D=32 outputs, M=128 inducing points, alpha=0.5, full batch, FP64): one step = one
``objective_function`` evaluation (psi0/psi1/psi2 forward, likelihood, hand-written
backward through the psi statistics and the q(u) / Kuu chain) with the parameters
already resident in HBM and the gradient left on the device.  Default path: the fused
moment-matched kernels (psi2 recomputed, never stored); --path plain materialises psi2,
--path graph replays the fused evaluation as one hipGraph.  Other workloads: sdgpr,
sgpssm, sgpr, serving, copy (counter calibration).

Prints one JSON line.  ``roofline`` reports the dominant kernel timed with HIP events
on the launch stream over the timed region.  The default line also carries
``c3_fused_vs_plain`` (the deep-GP evaluation on the plain path against the fused path
replayed as a hipGraph) and ``serving`` (served predictions at three Poisson arrival
rates, continuous batching against one request at a time), and ``cpu_baseline``: the CPU
oracle on the same SGPLVM evaluation, faithful cost on one thread, with the all-thread and
optimised-contraction variants beside it (rank 0, N=1 only; about a minute of host time).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 (vector and matrix)

# HBM bytes per launch of the roofline entry points, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (KiB per dispatch, summed over the entry point's kernels) on the benchmark
# shapes.  On MI355X a plain copy of 512 MiB reads as FETCH_SIZE 262,200 KiB and
# WRITE_SIZE 524,288 KiB, so FETCH_SIZE is scaled by 2.0 and WRITE_SIZE by 1.0.
TRAFFIC = {
    ('sgplvm', 'gpx_psi_fwd'): 2.0 * 2185 * 1024 + 1.0 * 528384 * 1024,
    ('sgplvm', 'gpx_fused_mm_bwd'): 2.0 * 59085 * 1024 + 1.0 * 85058 * 1024,
}

WORKLOADS = {
    'sgplvm': dict(desc='SGPLVM uncertain-input AEP objective, N=4096 latent Gaussians '
                        'N(0,1)/U(0.01,0.5), Din=8, Dout=32, M=128, alpha=0.5, full batch',
                   N=4096, Q=8, D=32, M=128, mb=4096),
    'sdgpr': dict(desc='Deep GP SDGPR 10->5->5->1, N=20000 inputs N(0,I), M=100 per layer, '
                       'minibatch 1000, alpha=0.5', N=20000, M=100, mb=1000),
    'sgpssm': dict(desc='GPSSM latent dim 4, T=10000, Dout=8 linear-Gaussian emission, M=200, '
                        'alpha=0.5, full batch', T=10000, Q=4, D=8, M=200, mb=10000),
    'sgpr': dict(desc='SGPR N=1000 x~U(-3,3), y=sin(3x)+N(0,0.01), M=50, alpha=0.5, full batch',
                 N=1000, M=50, mb=1000),
}


def build_model(name, w):
    """(model, params) of one evaluation workload, seeded."""
    from geepee_amd import aep_models as G
    from geepee_amd import workloads as W
    if name == 'sgplvm':
        y, p = W.sgplvm_problem(N=w['N'], Q=w['Q'], D=w['D'], M=w['M'], seed=1)
        return G.SGPLVM(y, w['Q'], w['M']), p
    if name == 'sdgpr':
        x, y, p, Ms, hidden = W.sdgpr_problem(N=w['N'], widths=(10, 5, 5, 1), M=w['M'], seed=2)
        return G.SDGPR(x, y, Ms, hidden), p
    if name == 'sgpssm':
        y, p = W.sgpssm_problem(T=w['T'], Din=w['Q'], Dout=w['D'], M=w['M'], seed=3)
        return G.SGPSSM(y, w['Q'], w['M']), p
    if name == 'sgpr':
        x, y, p = W.sgpr_problem(N=w['N'], M=w['M'], seed=0)
        return G.SGPR(x, y, w['M']), p
    raise ValueError(name)


def _cpu_info():
    model = platform.processor() or ''
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def _blas_info():
    from threadpoolctl import threadpool_info
    for p in threadpool_info():
        if p.get('user_api') == 'blas':
            return '%s %s (%s, %s)' % (p.get('internal_api'), p.get('version'), p.get('threading_layer'),
                                       p.get('architecture'))
    return 'unknown'


def _host_threads():
    """Threads this process may use on the host: the CPU affinity set, capped by
    OMP_NUM_THREADS when the environment sets it (the GPU box gives each GPU 16)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get('OMP_NUM_THREADS')
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline_sgplvm(w, evals=3, variant_evals=2):
    """The oracle on the benchmarked SGPLVM evaluation itself (N=4096, Q=8, D=32, M=128)
    in four settings: faithful cost (the reference's naive einsums, scalar C psi loop, LU
    inverses) and optimised contraction order (einsum optimize=True, vectorised psi),
    each at one BLAS thread and at every host thread the process may use.  The faithful
    single-thread median over `evals` evaluations is the reported value; the other three
    settings take the median of `variant_evals`."""
    from threadpoolctl import threadpool_limits
    from oracle import aep_oracle as O
    from geepee_amd import workloads as W
    y, p = W.sgplvm_problem(N=w['N'], Q=w['Q'], D=w['D'], M=w['M'], seed=1)
    nthreads = _host_threads()
    runs = {}
    for mode in ('faithful', 'optimised'):
        for threads in (1, nthreads):
            key = '%s_%dt' % (mode, threads)
            if key in runs:
                continue
            O.FAITHFUL = mode == 'faithful'
            ts = []
            n_ev = evals if key == 'faithful_1t' else variant_evals
            try:
                with threadpool_limits(limits=threads):
                    for _ in range(n_ev):
                        t0 = time.perf_counter()
                        O.sgplvm_objective(y, p, w['M'], w['N'], alpha=0.5)
                        ts.append(time.perf_counter() - t0)
            finally:
                O.FAITHFUL = False
            runs[key] = {'median_s': float(np.median(ts)), 'evals': n_ev, 'threads': threads,
                         'evals_per_s': 1.0 / float(np.median(ts))}
            print('cpu baseline %s: %s' % (key, json.dumps(runs[key])), file=sys.stderr, flush=True)
    main = runs['faithful_1t']
    return {'value': main['evals_per_s'], 'unit': 'evals/s', 'cores': 1, 'kind': 'port',
            'sample': ('the oracle on the full SGPLVM evaluation (N=%d, Q=%d, D=%d, M=%d, alpha=0.5), '
                       'median of %d evaluations, faithful cost (numpy einsums without path '
                       'optimisation, LU inverses, the scalar C psi loop) on one thread'
                       % (w['N'], w['Q'], w['D'], w['M'], evals)),
            'host': {'cpu': _cpu_info(), 'nproc': os.cpu_count(), 'threads_available': nthreads,
                     'blas': _blas_info(), 'numpy': np.__version__},
            'variants': runs}


def mm_layers(name, w):
    """(N_rows, M, Q, D) of every moment-matched (uncertain-input) layer evaluated per step."""
    if name == 'sgplvm':
        return [(w['N'], w['M'], w['Q'], w['D'])]
    if name == 'sdgpr':
        return [(w['mb'], w['M'], 5, 5), (w['mb'], w['M'], 5, 1)]
    if name == 'sgpssm':
        return [(w['T'] - 1, w['M'], w['Q'], w['Q'])]
    return []


def roofline_model(entry, name, w):
    """Algorithmic work of one launch of `entry` summed over the workload's MM layers:
    ('hbm', bytes) for the materialising psi forward, ('mfma', flops) for the fused ones."""
    tot, launches = 0.0, 0
    for (N, M, Q, D) in mm_layers(name, w):
        launches += 1
        if entry == 'gpx_psi_fwd':
            # psi1 (N x M) and psi2 (N x M x M) written once; mx, vx (N x Q) and z (M x Q) read once
            tot += 8.0 * (N * M * M + N * M + 2 * N * Q + M * Q)
        elif entry == 'gpx_fused_mm_fwd':
            # per point and unordered pair: the K=2Q+2 exponent product and the D-long
            # contraction with Bhat (2 flops per multiply-add); exp counted as one flop
            pairs = M * (M + 1) / 2.0
            tot += N * pairs * (2.0 * (2 * Q + 2) + 1.0 + 2.0 * D)
        elif entry == 'gpx_fused_mm_bwd':
            # per point and unordered pair, each product formed once: the exponent (2K) and
            # its exp (1), G = <dv, Bhat> (2D), the dBhat and LW contractions (2D + 2(Q+1)),
            # and the row sums and V terms of the per-point statistics (2 + 2Q)
            pairs = M * (M + 1) / 2.0
            K = 2 * Q + 2
            tot += N * pairs * (2.0 * K + 1.0 + 4.0 * D + 2.0 * (Q + 1) + 2.0 + 2.0 * Q)
    bound = 'hbm' if entry == 'gpx_psi_fwd' else 'mfma'
    return bound, (tot / launches if launches else 0.0)


def run_evals(args, name, w):
    import torch
    from geepee_amd import dops
    from geepee_amd.graphs import GraphedObjective
    model, p = build_model(name, w)
    if args.path in ('fused', 'graph'):
        model.set_fused()
    mb, alpha = w['mb'], 0.5
    np.random.seed(0)
    if args.path == 'graph':
        g = GraphedObjective(model, p, mb, alpha)
        step = g
    else:
        dp = model._device_params(p)
        step = lambda: model.objective_function_device(dp, mb, alpha=alpha)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    probe = dops.Probe(args.probe) if (args.probe and args.path != 'graph') else None
    dops.set_probe(probe)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dops.set_probe(None)
    dt = (t1 - t0) / args.steps
    res = {'value': 1.0 / dt, 'ms_per_step': dt * 1e3}
    if probe is not None and probe.count():
        ms = probe.mean_ms()
        bound, work = roofline_model(args.probe, name, w)
        if bound == 'hbm':
            achieved, peak, unit = work / (ms * 1e-3) / 1e9, HBM_PEAK_GBS, 'GB/s'
        else:
            achieved, peak, unit = work / (ms * 1e-3) / 1e12, FP64_PEAK_TFLOPS, 'TFLOP/s'
        res['roofline'] = {'bound': bound, 'kernel': args.probe, 'achieved': achieved, 'peak': peak,
                           'unit': unit, 'frac': achieved / peak, 'traffic': TRAFFIC.get((name, args.probe)),
                           'avg_launch_ms': ms, 'launches': probe.count(),
                           ('bytes_per_launch' if bound == 'hbm' else 'flops_per_launch'): work}
    _ = out
    return res


def run_serving(args):
    """C5: 64 resident models, Poisson arrivals, log-uniform batch sizes 1..4096.
    Continuous batching (two launch slots, HBM budget) against each request alone on
    the per-op plain path, both replaying the same arrival times."""
    import torch
    from geepee_amd import workloads as W
    from geepee_amd.serving import ModelStore, PredictServer, make_requests
    specs = W.serving_models(n_models=64, seed=4)
    store = ModelStore(specs)
    dins = [s['Din'] for s in specs]
    warm = make_requests(64, 64, rate=1e9, dins=dins, seed=99)
    server = PredictServer(store)
    server.run(warm)
    server.run_plain(warm[:16])
    torch.cuda.synchronize()
    out = {}
    for rate in args.rates:
        reqs = make_requests(64, args.requests, rate=rate, dins=dins, seed=int(rate))
        lat, _, wall = server.run(reqs)
        pts = sum(r[2].shape[0] for r in reqs)
        out['%g' % rate] = {'batched': {'qps': len(reqs) / wall, 'points_per_s': pts / wall,
                                        'p50_ms': 1e3 * float(np.percentile(lat, 50)),
                                        'p99_ms': 1e3 * float(np.percentile(lat, 99))}}
        nplain = min(len(reqs), args.plain_requests)
        latp, _, wallp = server.run_plain(reqs[:nplain])
        out['%g' % rate]['plain'] = {'qps': nplain / wallp, 'requests': nplain,
                                     'p50_ms': 1e3 * float(np.percentile(latp, 50)),
                                     'p99_ms': 1e3 * float(np.percentile(latp, 99))}
        print('rate %g: %s' % (rate, json.dumps(out['%g' % rate])), file=sys.stderr, flush=True)
    top = out['%g' % args.rates[-1]]
    return {'metric': 'served predict queries/s (p99 latency beside)', 'value': top['batched']['qps'],
            'unit': 'queries/s', 'ms_per_step': top['batched']['p99_ms'], 'rates': out,
            'workload': '64 resident SGPR/2-layer SDGPR models, M in {32,64,128,256}, Din 1-10; '
                        'Poisson arrivals, %d requests per rate, batch log-uniform 1-4096, N(0,I) points'
                        % args.requests}


def run_c3_ratio(steps=20, warmup=3):
    """The deep-GP evaluation (SDGPR 10-5-5-1, N=20000, M=100, minibatch 1000) on the plain
    path (one kernel per reference operation, psi2 materialised, eager) and on the fused
    path replayed as a hipGraph, same parameters and minibatch draws."""
    import argparse as _ap
    w = WORKLOADS['sdgpr']
    out = {'workload': w['desc']}
    for path, n in (('plain', max(5, steps // 2)), ('graph', steps)):
        a = _ap.Namespace(path=path, steps=n, warmup=warmup, probe=None)
        r = run_evals(a, 'sdgpr', w)
        out[path + '_ms_per_eval'] = r['ms_per_step']
        out[path + '_steps'] = n
        print('c3 %s: %.3f ms/eval' % (path, r['ms_per_step']), file=sys.stderr, flush=True)
    out['fused_graph_evals_per_s'] = 1e3 / out['graph_ms_per_eval']
    out['plain_evals_per_s'] = 1e3 / out['plain_ms_per_eval']
    out['ratio'] = out['plain_ms_per_eval'] / out['graph_ms_per_eval']
    out['target_ratio'] = 5.0
    return out


def run_copy(args):
    """Calibration for HBM counters: gpx_lincomb as a plain copy (out = 1 * x) of
    64 Mi doubles, i.e. 512 MiB read and 512 MiB written per launch."""
    import torch
    from geepee_amd import dops
    n = 64 << 20
    x = torch.ones(n, dtype=torch.float64, device='cuda')
    out = torch.empty_like(x)
    for _ in range(args.warmup):
        dops.lincomb(out, 1.0, x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dops.lincomb(out, 1.0, x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {'metric': 'copy GB/s (counter calibration)', 'value': 16.0 * n / dt / 1e9, 'unit': 'GB/s',
            'bytes_read_per_launch': 8 * n, 'bytes_written_per_launch': 8 * n, 'ms_per_step': dt * 1e3}


def run_fill(args):
    """HBM write ceiling: gpx_lincomb broadcasting one value into 64 Mi doubles (512 MiB
    written per launch, one double read), the reference rate for write-bound kernels."""
    import torch
    from geepee_amd import dops
    n = 64 << 20
    one = torch.ones(1, dtype=torch.float64, device='cuda')
    out = torch.empty(n, dtype=torch.float64, device='cuda')
    for _ in range(args.warmup):
        dops.lincomb(out, 1.0, one)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dops.lincomb(out, 1.0, one)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    return {'metric': 'fill GB/s (HBM write ceiling)', 'value': 8.0 * n / dt / 1e9, 'unit': 'GB/s',
            'bytes_written_per_launch': 8 * n, 'ms_per_step': dt * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--workload', default='sgplvm', choices=sorted(WORKLOADS) + ['serving', 'copy', 'fill'])
    ap.add_argument('--path', default='fused', choices=['plain', 'fused', 'graph'])
    ap.add_argument('--probe', default=None,
                    help='C-ABI entry point timed for the roofline (default: the path\'s dominant one)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-extras', action='store_true',
                    help='skip the deep-GP fused/plain ratio and the serving run of the default line')
    ap.add_argument('--rates', type=float, nargs='+', default=[1e2, 1e3, 1e4])
    ap.add_argument('--requests', type=int, default=600)
    ap.add_argument('--plain-requests', type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build runs the hot path on a single MI355X '
              '(the multi-GPU sharding is a design note); rerun with --gpus 1', file=sys.stderr)
        sys.exit(2)
    if args.workload in ('copy', 'fill'):
        print(json.dumps(run_copy(args) if args.workload == 'copy' else run_fill(args)), flush=True)
        return
    if args.workload == 'serving':
        r = run_serving(args)
        print(json.dumps({'metric': r['metric'], 'value': r['value'], 'unit': r['unit'], 'n_gpus': 1,
                          'steps': args.requests, 'warmup': 64, 'ms_per_step': r['ms_per_step'],
                          'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
                          'dtype': 'f64', 'data': 'synthetic (seeded)',
                          'config': {'workload': r['workload'], 'path': 'continuous batching'},
                          'rates': r['rates']}), flush=True)
        return
    w = WORKLOADS[args.workload]
    if args.probe is None:
        args.probe = 'gpx_psi_fwd' if args.path == 'plain' else 'gpx_fused_mm_bwd'
    res = run_evals(args, args.workload, w)
    line = {
        'metric': 'AEP energy+grad evals/s',
        'value': res['value'],
        'unit': 'evals/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': res['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded)',
        'config': {'workload': w['desc'], 'path': args.path, 'params': 'device-resident'},
    }
    if 'roofline' in res:
        line['roofline'] = res['roofline']
    if not args.no_extras and args.workload == 'sgplvm':
        # the other two halves of the headline metric, measured in the same run: the
        # deep-GP fused+graph against plain ratio, and served predictions against one
        # request at a time
        line['c3_fused_vs_plain'] = run_c3_ratio()
        r = run_serving(args)
        line['serving'] = {'metric': r['metric'], 'workload': r['workload'], 'rates': r['rates']}
    if not args.no_cpu and args.workload == 'sgplvm':
        line['cpu_baseline'] = cpu_baseline_sgplvm(w)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
