"""Benchmark: AEP energy + analytic-gradient evaluations per second on one MI355X.

Default workload: the SGPLVM uncertain-input objective (N=4096 latent Gaussians, Q=8,
D=32 outputs, M=128 inducing points, alpha=0.5, full batch, FP64): one step = one
``objective_function`` evaluation (psi0/psi1/psi2 forward, likelihood, hand-written
backward through the psi statistics and the q(u) / Kuu chain) with the parameters
already resident in HBM and the gradient left on the device.

Prints one JSON line.  ``roofline`` reports the dominant kernel timed with HIP events
on the launch stream over the timed region; ``cpu_baseline`` times the CPU oracle in
its faithful-cost mode on a bounded sample (rank 0, N=1 only).
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

WORKLOADS = {
    'sgplvm': dict(desc='SGPLVM uncertain-input AEP objective, N=4096 latent Gaussians '
                        'N(0,1)/U(0.01,0.5), Din=8, Dout=32, M=128, alpha=0.5, full batch',
                   N=4096, Q=8, D=32, M=128),
    'sdgpr': dict(desc='Deep GP SDGPR 10->5->5->1, N=20000, M=100/layer, minibatch 1000, alpha=0.5',
                  N=20000, M=100, B=1000),
}


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


def cpu_baseline_sgplvm(w, budget_s=40.0):
    """Faithful-cost oracle (naive einsums, scalar C psi loop, LU inverses) on the same
    synthetic SGPLVM problem at N=512 and N=1024 rows, extrapolated linearly in N to
    the full N (the per-point work is linear in N; the D*M^4 einsums are not)."""
    from oracle import aep_oracle as O
    from geepee_amd import workloads as W
    O.FAITHFUL = True
    try:
        times = {}
        for n in (512, 1024):
            y, p = W.sgplvm_problem(N=n, Q=w['Q'], D=w['D'], M=w['M'], seed=1)
            t0 = time.perf_counter()
            O.sgplvm_objective(y, p, w['M'], n, alpha=0.5)
            times[n] = time.perf_counter() - t0
    finally:
        O.FAITHFUL = False
    slope = (times[1024] - times[512]) / 512.0
    t_full = times[1024] + slope * (w['N'] - 1024)
    threads = os.environ.get('OMP_NUM_THREADS', str(os.cpu_count()))
    return {'value': 1.0 / t_full, 'unit': 'evals/s', 'cores': int(threads) if threads.isdigit() else None,
            'kind': 'port',
            'sample': ('oracle faithful-cost SGPLVM eval timed at N=512 (%.2fs) and N=1024 (%.2fs), '
                       'linear in N to N=%d (%.1fs/eval); %s; numpy %s'
                       % (times[512], times[1024], w['N'], t_full, _cpu_info(), np.__version__))}


def run_sgplvm(args, w):
    import torch
    from geepee_amd import aep_models as G
    from geepee_amd import dops, workloads as W
    N, Q, D, M = w['N'], w['Q'], w['D'], w['M']
    y, p = W.sgplvm_problem(N=N, Q=Q, D=D, M=M, seed=1)
    model = G.SGPLVM(y, Q, M)
    dp = model._device_params(p)
    for _ in range(args.warmup):
        model.objective_function_device(dp, N, alpha=0.5)
    torch.cuda.synchronize()
    probe = dops.Probe(args.probe) if args.probe else None
    dops.set_probe(probe)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = model.objective_function_device(dp, N, alpha=0.5)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dops.set_probe(None)
    dt = (t1 - t0) / args.steps
    res = {'value': 1.0 / dt, 'ms_per_step': dt * 1e3}
    if probe is not None:
        ms = probe.mean_ms()
        # gpx_psi_fwd writes psi1 (N x M) and psi2 (N x M x M) and reads the N x Q inputs
        bytes_alg = 8.0 * (N * M * M + N * M + 2 * N * Q + M * Q)
        achieved = bytes_alg / (ms * 1e-3) / 1e9
        res['roofline'] = {'bound': 'hbm', 'kernel': 'psi2_kernel (gpx_psi_fwd)',
                           'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                           'frac': achieved / HBM_PEAK_GBS, 'traffic': None,
                           'avg_launch_ms': ms, 'launches': probe.count(),
                           'bytes_per_launch': bytes_alg}
    _ = out
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--workload', default='sgplvm', choices=sorted(WORKLOADS))
    ap.add_argument('--probe', default='gpx_psi_fwd')
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build runs the hot path on a single MI355X '
              '(the multi-GPU sharding is a design note); rerun with --gpus 1', file=sys.stderr)
        sys.exit(2)
    w = WORKLOADS[args.workload]
    res = run_sgplvm(args, w)
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
        'config': {'workload': w['desc'], 'path': 'plain', 'params': 'device-resident'},
    }
    if 'roofline' in res:
        line['roofline'] = res['roofline']
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline_sgplvm(w)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
