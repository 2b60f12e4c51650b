"""Benchmark: powder-averaged time-domain propagation on one MI355X.

Workload: static 8-spin solid (5x1H, 2x13C, 1x15N, 14.1 T, random CSA and
dipolar couplings, IK-0 level 3 basis of dimension 1789), 4000 Haar-random
orientations, 4096 steps of 5 us, rho0 = Lx(1H), coil = L+(1H).

A step is one full pass of the hot path over the whole grid: sp_execute on a
plan whose inputs (reduced generator components, grid, states) are already
resident in HBM. The value is orientation-steps/s = 4000*4096/(time per pass).

Also reported: the roofline of the propagation kernel (FP64, compute bound)
from HIP-event kernel durations, the relative FID error of the GPU result
against the CPU oracle on a sample of orientations, and the CPU baseline (the
oracle's faithful restatement of the reference algorithm, OpenMP over
orientations) timed on this host on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fp64|fp32]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ('orientation-steps/s on 1 MI355X at ≤1e-6 rel FID error vs FP64; surrogate speedup')
PEAK_FP64 = 78.6e12   # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_FP32_VALU = 78.6e12  # non-packed FP32 vector rate (packed FP32 ops are disabled in this build)


def algorithmic_flops_per_orientation(n, nsteps, taylor_k, ps_q, sexp, block):
    """Real flops of the dense propagation per orientation on the reduced
    dimension n (padding excluded): complex GEMM = 8 n^3, complex matvec =
    8 n^2, complex dot = 8 n. Matrix products: q-1 powers and K//q Horner
    steps of the Paterson-Stockmeyer Taylor evaluation, s squarings,
    log2(B) squarings to Q = P^B; vector-matrix work: B-1 rows of the U block
    and M-1 steps of the V chain; detection: (nsteps+1) dots."""
    npts = nsteps + 1
    nblk = math.ceil(npts / block)
    gemms = (ps_q - 1) + taylor_k // ps_q + sexp + int(round(math.log2(block)))
    return gemms * 8.0 * n ** 3 + (block - 1 + nblk - 1) * 8.0 * n ** 2 + npts * 8.0 * n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--mode', default='fp64', choices=['fp64', 'fp32'])
    ap.add_argument('--orientations', type=int, default=4000)
    ap.add_argument('--cpu-sample', type=int, default=48, help='orientations in the CPU baseline sample')
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: every box in this setup has a single MI355X, and the '
              'multi-GPU path (orientation sharding + one FID all-reduce) is not built. Run with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from spinach_amd import _native, grids, workloads

    t0 = time.time()
    w = workloads.static_8spin(n_orient=args.orientations)
    print(f'[bench] workload built in {time.time() - t0:.1f}s: dim {w["I"].shape[0]}', file=sys.stderr, flush=True)
    n_or = len(w['grid']['weights'])
    nsteps = w['nsteps']
    plan = _native.plan_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], nsteps, mode=args.mode)
    for _ in range(args.warmup):
        plan.execute()
    times, kms = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        st = plan.execute()
        times.append(time.perf_counter() - t1)
        kms.append(st['kernel_ms'])
    total = time.perf_counter() - t_start
    ms_per_step = 1e3 * total / args.steps
    value = n_or * nsteps / (total / args.steps)
    fid, per = plan.result(per_orientation=True)
    print(f'[bench] {args.steps} steps, {ms_per_step:.2f} ms/step, kernel {np.mean(kms):.2f} ms', file=sys.stderr,
          flush=True)

    # roofline of the propagation kernel
    n = st['reduced_dim']
    mean_s = st['sum_squarings_lo'] / n_or
    flops = algorithmic_flops_per_orientation(n, nsteps, st['taylor_terms'], st['ps_block'], mean_s, st['block'])
    k_avg_s = float(np.mean(kms)) * 1e-3
    achieved = flops * n_or / k_avg_s
    roof = {'bound': 'mfma', 'achieved': round(achieved / 1e12, 3), 'peak': round(PEAK_FP64 / 1e12, 1),
            'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP64, 4), 'traffic': None,
            'kernel': 'spk::k_powder', 'kernel_ms_avg': round(float(np.mean(kms)), 3),
            'flops_per_orientation': flops, 'orientations_per_launch': n_or}

    # parity on a sample against the oracle, and the CPU baseline
    cpu = None
    err = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import pyoracle as O
        ncores = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else os.cpu_count()
        ncores = min(ncores, 64)
        m = min(args.cpu_sample, n_or)
        sub = grids.subset(w['grid'], np.arange(m))
        t1 = time.perf_counter()
        ref, rper = O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], nsteps, variant=0,
                                    nthreads=ncores, per_orientation=True)
        tc = time.perf_counter() - t1
        cpu = {'value': round(m * nsteps / tc, 1), 'unit': 'orientation-steps/s', 'cores': ncores, 'kind': 'port',
               'sample': f'{m} of the {n_or} orientations x {nsteps} steps, faithful reference algorithm '
                         f'(per-orientation ZTE reduce, expm, sequential detection loop), FP64, {tc:.1f}s'}
        wsum = np.asarray(sub['weights'])
        gsum = (wsum[:, None] * per[:m]).sum(axis=0)
        rsum = (wsum[:, None] * rper).sum(axis=0)
        err = float(np.linalg.norm(gsum - rsum) / np.linalg.norm(rsum))
        print(f'[bench] cpu baseline {cpu["value"]:.3g} orientation-steps/s on {ncores} cores; '
              f'GPU rel FID error on the sample {err:.2e}', file=sys.stderr, flush=True)

    line = {
        'metric': METRIC,
        'value': round(value, 1),
        'unit': 'orientation-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ms_per_step, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'c128' if args.mode == 'fp64' else 'c64',
        'data': 'synthetic (seeded spin system and Haar grid)',
        'config': {'workload': 'static 8-spin solid (5x1H,2x13C,1x15N), IK-0 level 3 (dim 1789 -> 185 '
                               'coil-connected), 4000 orientations x 4096 steps at 5 us, FP64',
                   'orientations': n_or, 'nsteps': nsteps, 'full_dim': st['full_dim'],
                   'reduced_dim': n, 'mode': args.mode},
        'fid_rel_err_vs_oracle': err,
        'roofline': roof,
        'cpu_baseline': cpu,
    }
    print(json.dumps(line))


if __name__ == '__main__':
    main()
