"""Benchmark: powder-averaged time-domain propagation on one MI355X.

Default workload (the headline configuration): static 8-spin solid (5x1H,
2x13C, 1x15N, 14.1 T, random CSA and dipolar couplings, IK-0 level 3 basis of
dimension 1789 reduced to its 185-state coil-connected subspace), 4000
Haar-random orientations, 4096 steps of 5 us, rho0 = Lx(1H), coil = L+(1H),
complex FP64.

A step is one full pass of the hot path over the whole grid: sp_execute on a
plan whose inputs (reduced generator components, grid, states) are already
resident in HBM. value = orientations * nsteps / (time per pass).

Reported beside it: the roofline of the propagation kernel (algorithmic
flops / HIP-event kernel time against the FP64 peak), the relative FID error
of the GPU result against the CPU oracle on a sample of orientations, and the
CPU baseline (the oracle's faithful restatement of the reference algorithm,
OpenMP over orientations) timed on this host on a bounded sample.

Other workloads (--workload): static2 (two-spin 1H-13C, 256 x 2047), batch
(512 systems of 2-10 spins, 100-1000 orientations each, 1024 steps, one
persistent launch), mas (6-spin Hilbert-space magic-angle spinning, 3000
orientations x 8192 steps, complex FP32 by default).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode fp64|fp32]
                       [--workload static8|static2|batch|mas]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'orientation-steps/s on 1 MI355X at ≤1e-6 rel FID error vs FP64; surrogate speedup'
PEAK = {'fp64': 78.6e12, 'fp32': 157.3e12}  # MI355X dense MFMA FP64 / FP32, FLOP/s (FP64 vector = matrix)
# LDS bandwidth of MI355X: 128 bytes per clock per CU, 256 CUs, 2.4 GHz
LDS_PEAK = 128 * 256 * 2.4e9
# what bounds each kernel: FP64/FP32 vector ALU, matrix cores, LDS
BOUND = {'spk::k_tridiag': 'valu', 'spk::k_dceig': 'valu', 'spk::k_synth': 'mfma', 'spk::k_mas': 'mfma', 'spk::k_mas_tiled': 'mfma',
         'sgk::k_gemm_bf16': 'mfma'}
# HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs of this script (counters scaled by a device copy of known size:
# FETCH x2.000, WRITE x1.000); see profiles/k_spectral_static8spin_fp64.md and
# profiles/k_powder_static8spin_fp64.md and profiles/k_tridiag_w_batch512_fp64.md (the batch's
# tridiagonalisation stage: k_tridiag_w and k_tridiag together) and profiles/k_mas_tiled_6spin_fp64.md
TRAFFIC = {
    ('batch', 'fp64', 'spk::k_tridiag'): (178670707 + 159255) * 1024 * 2.0 + (344737580 + 2384462) * 1024 * 1.0,
    ('static8', 'fp64', 'spk::k_tridiag'): 6203 * 1024 * 2.0 + 58669 * 1024 * 1.0,
    ('static8', 'fp64', 'spk::k_powder<double,false>'): 165530392 * 1024 * 2.0 + 53150413 * 1024 * 1.0,
    ('static8', 'fp64', 'spk::k_powder<double,true>'): 139595089 * 1024 * 2.0 + 49487957 * 1024 * 1.0,
    ('mas', 'fp64', 'spk::k_mas_tiled'): 4100499 * 1024 * 2.0 + 3624603 * 1024 * 1.0,
}


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def powder_flops_per_orientation(n, nsteps, taylor_k, ps_q, sexp, block, real=False):
    """Real flops of the dense propagation per orientation on the reduced
    dimension n (padding excluded): complex product = 8 n^3, complex
    vector-matrix product = 8 n^2, complex dot = 8 n. Matrix products: q-1
    powers and K//q Horner steps of the Paterson-Stockmeyer Taylor evaluation,
    s squarings, log2(B) squarings to Q = P^B; B-1 rows of the U block;
    M-1 steps of the V chain; (nsteps+1) detection dots. real: the
    real-generator kernel, whose powers are real products (2 n^3), Horner
    steps real x complex (4 n^3) and squarings two real products (4 n^3)."""
    npts = nsteps + 1
    nblk = math.ceil(npts / block)
    sq = sexp + int(round(math.log2(block)))
    if real:
        mat = (ps_q - 1) * 2.0 * n ** 3 + (taylor_k // ps_q) * 4.0 * n ** 3 + sq * 4.0 * n ** 3
    else:
        mat = ((ps_q - 1) + taylor_k // ps_q + sq) * 8.0 * n ** 3
    return mat + (block - 1 + nblk - 1) * 8.0 * n ** 2 + npts * 8.0 * n


def mas_flops_per_orientation(n, nsteps, ticks, taylor_k, ps_q, sexp):
    """Per orientation: every tick an expm (q-1 + K//q products + s squarings),
    the U update and C_j = U'*coil*U (3 products); two products per rotor
    period for the rho chain; detection conj(C) . rho over n^2 elements."""
    npts = nsteps + 1
    n_rho = math.ceil(npts / ticks)
    per_tick = (ps_q - 1) + taylor_k // ps_q + sexp + 3
    return ticks * per_tick * 8.0 * n ** 3 + 2 * (n_rho - 1) * 8.0 * n ** 3 + ticks * n_rho * 8.0 * n * n


def mas_tiled_flops(products, ndet, n_or, nsteps):
    """The same algorithm on 16 x 16 tiles: the real tile products the kernel counted (per
    tick the expm of every tile at its own Taylor degree, the U update and C_j = U_s' coil_st
    U_t per tile pair, W = coil_st U_t in two real products when the coil is real; per rotor
    period two complex products per pair for the rho chain; a complex product is four real
    ones, 2 * 16^3 FLOP each), plus the detection of every time point over the ndet elements
    where both C_j and rho_k can be nonzero."""
    return products * 2.0 * 16 ** 3 + n_or * (nsteps + 1) * 8.0 * ndet


def spectral_stage_flops(n, npts):
    """Algorithmic flops per orientation of the spectral path on the reduced
    dimension n: Householder tridiagonalisation 4/3 n^3 (symmetric
    matrix-vector product 2 m^2 and rank-2 update 2 m^2 per column of a
    trailing block of size m) plus 4 n^2 per transformed vector (two complex
    vectors = four real ones); the divide-and-conquer tridiagonal eigensolver
    is O(n^2) scalar work per merge level, counted as 0 (bound by its
    secular-equation iterations, not by throughput); synthesis 8 n flops per
    time point (complex multiply-add)."""
    return {'spk::k_tridiag': 4.0 / 3.0 * n ** 3 + 16.0 * n ** 2, 'spk::k_dceig': 0.0,
            'spk::k_synth': 8.0 * n * npts}


def prologue_bytes(np_):
    """HBM bytes per orientation of k_tridiag_w's global-memory prologue (padded dimension
    np_ > 256, 32 x 32 FP64 blocks of the lower triangle): the workspace cleared and read once
    for the first product, one read and one write of the live blocks (block rows and columns
    >= (j + 1) / 32) for each of the first k0 reflections, and the trailing block read into
    the register layout."""
    k0 = -(-(np_ - 256) // 32) * 32
    nbk = -(-np_ // 32)
    blk = 32 * 32 * 8
    total = nbk * (nbk + 1) // 2 * blk * 2
    for j in range(k0):
        m = nbk - ((j + 1) >> 5)
        total += m * (m + 1) // 2 * blk * 2
    m2 = nbk - k0 // 32
    return total + m2 * (m2 + 1) // 2 * blk


def step_product_work(n, nnz, real=True):
    """Algorithmic work of one sparse product of the stepping path (k_step) on
    the reduced dimension n with nnz pattern entries: flops 4 per entry for a
    real generator times a complex vector (8 for a complex generator) plus 6
    per row (the term scaling and rho += next); LDS bytes per entry: the value
    (8 or 16), its column (4) and the gathered vector element (16); per row
    the new term written (16) and rho read and written (32)."""
    flops = (4.0 if real else 8.0) * nnz + 6.0 * n
    lds = ((8 if real else 16) + 4 + 16) * nnz + 48.0 * n
    return flops, lds


def ncores():
    c = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else os.cpu_count()
    return max(1, min(c, 64))


def time_plan(plan, steps, warmup, autotune=0):
    """Untimed: autotune passes per candidate (sp_plan_autotune, if > 0) and the
    warm-up; timed: `steps` passes."""
    import numpy as np
    plan.tuning_chosen = plan.autotune(passes=autotune) if autotune > 0 else None
    if plan.tuning_chosen:
        log(f'autotune: {plan.tuning_chosen}')
    for _ in range(warmup):
        plan.execute()
    kms, stages = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = plan.execute()
        kms.append(st['kernel_ms'])
        stages.append(st['stage_ms'])
    total = time.perf_counter() - t0
    st['stage_ms'] = [float(x) for x in np.mean(np.asarray(stages), axis=0)]
    return total / steps, float(np.mean(kms)), st


def stage_table(st, prop_flops, spec_flops, prop_kernel=None, mode='fp64'):
    """{kernel: (ms, flops)} for the kernels that ran in a pass."""
    ms = st['stage_ms']
    out = {}
    if prop_flops:
        if prop_kernel is None:
            prop_kernel = 'spk::k_powder<%s,%s>' % ('float' if mode == 'fp32' else 'double',
                                                    'true' if st['real_systems'] else 'false')
        out[prop_kernel] = (ms[0], prop_flops)
    if spec_flops:
        for k, name in enumerate(('spk::k_tridiag', 'spk::k_dceig', 'spk::k_synth')):
            out[name] = (ms[1 + k], spec_flops[name])
    return out


def bench_static(args, two_spin=False):
    import numpy as np
    from spinach_amd import _native, grids, workloads
    if two_spin:
        w = workloads.static_2spin()
        desc = ('static 2-spin 1H-13C, complete basis (dim 16 -> 2 coil-connected), zcwn 256 orientations '
                'x 2047 steps at 10 us')
    else:
        w = workloads.static_8spin(n_orient=args.orientations)
        desc = ('static 8-spin solid (5x1H,2x13C,1x15N), IK-0 level 3 (dim 1789 -> 185 coil-connected), '
                f'{len(w["grid"]["weights"])} orientations x 4096 steps at 5 us')
    n_or = len(w['grid']['weights'])
    nsteps = w['nsteps']
    plan = _native.plan_powder(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], nsteps,
                               mode=args.mode + args.path_flag)
    sec, kms, st = time_plan(plan, args.steps, args.warmup, args.autotune)
    fid, per = plan.result(per_orientation=True)
    log(f'{args.steps} steps, {1e3 * sec:.2f} ms/step, kernels {kms:.2f} ms, stages {st["stage_ms"]}')
    n = st['reduced_dim']
    krylov = st['krylov_systems'] > 0
    if krylov:
        f1, b1 = step_product_work(n, st['reduced_nnz'], real=st['real_systems'] > 0)
        name = 'spk::k_step<%s,true>' % ('true' if st['real_systems'] else 'false')
        stages = {name: (st['stage_ms'][0], f1 * st['step_products'], b1 * st['step_products'])}
    elif st['spectral_systems']:
        stages = stage_table(st, 0.0, {k: v * n_or for k, v in spectral_stage_flops(n, nsteps + 1).items()})
    else:
        stages = stage_table(st, n_or * powder_flops_per_orientation(
            n, nsteps, st['taylor_terms'], st['ps_block'], st['sum_squarings_lo'] / n_or, st['block'],
            real=st['real_systems'] > 0), None, mode=args.mode)
    cpu = err = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import pyoracle as O
        m = min(args.cpu_sample if not two_spin else n_or, n_or)
        if krylov:
            m = min(m, 100)
        sub = grids.subset(w['grid'], np.arange(m))
        tols = O.make_tols(krylov_tol=2) if krylov else O.DEFAULT_TOLS
        # the reference algorithm as written (per-orientation ZTE and path tracing), then the
        # optimised CPU variant (one union-pattern reduction, the same propagation)
        t1 = time.perf_counter()
        ref, rper = O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], nsteps, variant=0,
                                    nthreads=ncores(), per_orientation=True, tols=tols)
        tc0 = time.perf_counter() - t1
        t1 = time.perf_counter()
        O.refcpu_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], nsteps, variant=2, nthreads=ncores(),
                        tols=tols)
        tc2 = time.perf_counter() - t1
        algo = ('step() per time step (krylov.m), sparse FP64' if krylov else
                'expm and the sequential detection loop, FP64')
        cpu = {'value': round(m * nsteps / tc2, 1), 'unit': 'orientation-steps/s', 'cores': ncores(),
               'kind': 'port',
               'sample': f'{m} of the {n_or} orientations x {nsteps} steps: {algo}; value = one union-pattern '
                         f'reduction ({tc2:.1f}s), faithful_value = per-orientation ZTE and path tracing as the '
                         f'reference runs it ({tc0:.1f}s)',
               'faithful_value': round(m * nsteps / tc0, 1)}
        wsum = np.asarray(sub['weights'])
        err = O.rel_err((wsum[:, None] * per[:m]).sum(axis=0), (wsum[:, None] * rper).sum(axis=0))
    cfg = {'workload': desc, 'orientations': n_or, 'nsteps': nsteps, 'full_dim': st['full_dim'],
           'reduced_dim': n, 'mode': args.mode, 'plan_ms': round(st['plan_ms'], 2),
           'path': 'krylov' if krylov else ('spectral' if st['spectral_systems'] else 'propagator')}
    if krylov:
        cfg['step_products_per_orientation'] = round(st['step_products'] / n_or, 1)
    if not st['spectral_systems'] and not krylov:
        cfg['time_block'] = st['block']
    if plan.tuning_chosen:
        cfg['tuning'] = {k: plan.tuning_chosen[k] for k in ('log2_block', 'prop_wg_per_cu', 'tri_wg_per_cu',
                                                            'syn_wg_per_cu', 'step_wg_per_cu', 'candidates')}
    return n_or * nsteps / sec, sec, kms, stages, cfg, cpu, err


def bench_batch(args):
    import numpy as np
    from spinach_amd import _native, grids, workloads
    t0 = time.time()
    b = workloads.batch_uneven(n_systems=args.systems)
    log(f'{len(b["systems"])} systems built in {time.time() - t0:.1f}s')
    n_or_tot = sum(len(y['grid']['weights']) for y in b['systems'])
    plan = _native.plan_powder_batch(b['systems'], b['dt'], b['nsteps'], mode=args.mode + args.path_flag)
    sec, kms, st = time_plan(plan, args.steps, args.warmup, args.autotune)
    fids = plan.result_batch()
    log(f'{args.steps} steps, {1e3 * sec:.2f} ms/step, kernel {kms:.2f} ms, {n_or_tot} orientations')
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import pyoracle as O
    prop_flops = 0.0
    spec_flops = {k: 0.0 for k in ('spk::k_tridiag', 'spk::k_dceig', 'spk::k_synth')}
    dims = []
    spectral_ok = st['spectral_systems'] > 0
    tri_bytes = 0.0  # the prologue's HBM traffic (systems above 256 states)
    for y in b['systems']:
        n = len(O.union_states(y['I'], y['Q'], y['coil']))
        dims.append(n)
        n_or = len(y['grid']['weights'])
        np_ = -(-n // 16) * 16
        # every member's generator is real: in FP64 the spectral path takes padded dimensions <= 448
        if spectral_ok and np_ <= 448:
            for k, v in spectral_stage_flops(n, b['nsteps'] + 1).items():
                spec_flops[k] += n_or * v
            if np_ > 256:
                tri_bytes += n_or * prologue_bytes(np_)
        else:
            prop_flops += n_or * powder_flops_per_orientation(
                n, b['nsteps'], st['taylor_terms'], st['ps_block'], st['sum_squarings_lo'] / n_or_tot, st['block'],
                real=st['real_systems'] > 0)
    stages = stage_table(st, prop_flops, spec_flops if spectral_ok else None, mode=args.mode)
    if spectral_ok and 'spk::k_tridiag' in stages:
        stages['spk::k_tridiag'] = stages['spk::k_tridiag'] + (tri_bytes,)
    cpu = err = None
    if not args.no_cpu:
        # sample: 2 orientations of every 8th system, summed in the same way
        t1 = time.perf_counter()
        work = 0
        errs = []
        for y, f in list(zip(b['systems'], fids))[::8]:
            sub = grids.subset(y['grid'], np.arange(2))
            ref = O.refcpu_powder(y['I'], y['Q'], sub, y['rho0'], y['coil'], b['dt'], b['nsteps'], variant=0,
                                  nthreads=ncores())
            work += 2 * b['nsteps']
        tc = time.perf_counter() - t1
        cpu = {'value': round(work / tc, 1), 'unit': 'orientation-steps/s', 'cores': ncores(), 'kind': 'port',
               'sample': f'2 orientations of every 8th system ({len(b["systems"][::8])} systems) x '
                         f'{b["nsteps"]} steps, faithful reference algorithm, FP64, {tc:.1f}s'}
        for y, f in list(zip(b['systems'], fids))[::64]:
            ref = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], b['dt'], b['nsteps'],
                                  variant=0, nthreads=ncores())
            errs.append(O.rel_err(f, ref))
        err = max(errs)
    cfg = {'workload': f'{len(b["systems"])} seeded spin systems of 2-10 spins (1H/13C/15N), IK-0 level <=3, '
                       f'100-1000 orientations each ({n_or_tot} total) x {b["nsteps"]} steps at 5 us, '
                       'one persistent launch',
           'systems': len(b['systems']), 'orientations': n_or_tot, 'nsteps': b['nsteps'],
           'reduced_dim_min': int(min(dims)), 'reduced_dim_max': int(max(dims)),
           'mode': args.mode, 'spectral_systems': st['spectral_systems'], 'plan_ms': round(st['plan_ms'], 1)}
    if prop_flops:
        cfg['time_block'] = st['block']
    if plan.tuning_chosen:
        cfg['tuning'] = {k: plan.tuning_chosen[k] for k in ('log2_block', 'prop_wg_per_cu', 'tri_wg_per_cu',
                                                            'syn_wg_per_cu', 'step_wg_per_cu', 'candidates')}
    return n_or_tot * b['nsteps'] / sec, sec, kms, stages, cfg, cpu, err


def bench_mas(args):
    import numpy as np
    from spinach_amd import _native, grids, workloads
    w = workloads.mas_6spin(n_orient=args.orientations if args.orientations != 4000 else 3000)
    n_or = len(w['grid']['weights'])
    plan = _native.plan_mas(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['axis'], w['rate_hz'], w['ticks'],
                            w['nsteps'], mode=args.mode)
    sec, kms, st = time_plan(plan, args.steps, args.warmup)
    fid, per = plan.result(per_orientation=True)
    log(f'{args.steps} steps, {1e3 * sec:.2f} ms/step, kernel {kms:.2f} ms')
    if st['mas_tiles']:
        kname = 'spk::k_mas_tiled'
        flops = mas_tiled_flops(st['step_products'], st['mas_detected'], n_or, w['nsteps'])
    else:
        kname = 'spk::k_mas'
        flops = n_or * mas_flops_per_orientation(64, w['nsteps'], w['ticks'], st['taylor_terms'], st['ps_block'],
                                                 st['sum_squarings_lo'] / n_or)
    cpu = err = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import pyoracle as O
        m = 4
        sub = grids.subset(w['grid'], np.arange(m))
        t1 = time.perf_counter()
        ref, rper = O.mas_powder(w['I'], w['Q'], sub, w['rho0'], w['coil'], w['axis'], w['rate_hz'], w['ticks'],
                                 w['nsteps'], per_orientation=True)
        tc = time.perf_counter() - t1
        cpu = {'value': round(m * w['nsteps'] / tc, 1), 'unit': 'orientation-steps/s', 'cores': ncores(),
               'kind': 'port',
               'sample': f'{m} orientations x {w["nsteps"]} steps, numpy direct stepping (tick propagators '
                         f'by scipy expm, rho <- P rho P\'), FP64, BLAS threads, {tc:.1f}s'}
        wsum = np.asarray(sub['weights'])
        err = O.rel_err((wsum[:, None] * per[:m]).sum(axis=0), (wsum[:, None] * rper).sum(axis=0))
    cfg = {'workload': f'MAS 12.5 kHz, 6 spins (4x13C, 2x1H) in Hilbert space (dim 64), 128 ticks per period '
                       f'x 64 periods, {n_or} orientations', 'orientations': n_or, 'nsteps': w['nsteps'],
           'hilbert_dim': 64, 'mode': args.mode}
    return n_or * w['nsteps'] / sec, sec, kms, stage_table(st, flops, None, kname), cfg, cpu, err


def bench_surrogate(args):
    """Learned powder interpolation on the 8-spin system: 300 training and 500
    held-out orientations from the plain method, a 4 x 256 tanh MLP with
    Fourier features trained in BF16, inference on a 50000-orientation grid.
    The acceptance check (held-out relative error <= 1e-4) picks the network or the plain
    method; value is the fine-grid rate of the path it picked."""
    import numpy as np
    from spinach_amd import grids, surrogate, workloads
    w = workloads.static_8spin(n_orient=4000)
    rng = np.random.default_rng(11)
    tr, ho, fine = grids.haar(300, rng), grids.haar(500, rng), grids.haar(args.fine, rng)
    fid, info = surrogate.learned_powder(w, tr, ho, fine, tol=1e-4, steps=args.train_steps, batch=8192, lr=1e-3)
    log(f'surrogate: {info}')
    n_or = len(fine['weights'])
    nsteps = w['nsteps']
    sec = (info['surrogate_fine_ms'] if info['decision'] == 'surrogate' else info['plain_fine_ms']) * 1e-3
    kms = sec * 1e3
    width, depth, F = 256, 4, 64
    flops_pt = 2.0 * (F * width + (depth - 1) * width * width + 2 * width)
    flops = flops_pt * n_or * (nsteps + 1) if info['decision'] == 'surrogate' else 0.0
    cfg = {'workload': f'learned powder interpolation on the static 8-spin system: 300 training / 500 held-out '
                       f'orientations (plain FP64 method), 4x256 tanh MLP, BF16 MFMA, {n_or}-orientation fine grid '
                       f'x {nsteps} steps', 'orientations': n_or, 'nsteps': nsteps, 'decision': info['decision'],
           'heldout_rel_err': info['heldout_rel_err'],
           'surrogate_fine_rel_err_vs_plain': info['surrogate_fine_rel_err_vs_plain'],
           'surrogate_ms': info['surrogate_fine_ms'], 'plain_ms': info['plain_fine_ms'],
           'surrogate_speedup': info['speedup_surrogate_vs_plain'], 'train_seconds': info['train_seconds'],
           'mode': 'bf16 network / fp64 plain'}
    if info['decision'] == 'surrogate':
        stages = {'sgk::k_gemm_bf16': (kms, flops)}
    else:
        # the plain FP64 method produced the result: its kernels, with the spectral path's work
        st = info['plain_stats']
        spec = {k: n_or * v for k, v in spectral_stage_flops(st['reduced_dim'], nsteps + 1).items()}
        stages = stage_table(st, 0.0, spec)
    # with the network the error is its measured fine-grid error against the plain method; on the
    # fallback the plain method itself ran, whose own error is the spectral path's (not re-measured here)
    err = info['surrogate_fine_rel_err_vs_plain'] if info['decision'] == 'surrogate' else None
    return n_or * nsteps / sec, sec, kms, stages, cfg, None, err


def copy_probe(mib):
    """Device-to-device copy of a known size through the HIP runtime, for
    calibrating the FETCH_SIZE / WRITE_SIZE counters: three copies of `mib`
    MiB (the runtime's copy kernel reads and writes mib MiB per dispatch)."""
    import ctypes as C
    hip = C.CDLL('libamdhip64.so')
    n = int(mib) << 20
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(n)) or hip.hipMalloc(C.byref(b), C.c_size_t(n)):
        raise RuntimeError('hipMalloc failed')
    hip.hipMemset(a, 1, C.c_size_t(n))
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError('hipMemcpy failed')
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({'copy_probe_bytes_per_dispatch': n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--mode', default=None, choices=['fp64', 'fp32'])
    ap.add_argument('--workload', default='static8', choices=['static8', 'static2', 'batch', 'mas', 'surrogate'])
    ap.add_argument('--fine', type=int, default=50000, help='fine-grid orientations (surrogate workload)')
    ap.add_argument('--train-steps', type=int, default=3000, help='Adam steps (surrogate workload)')
    ap.add_argument('--orientations', type=int, default=4000)
    ap.add_argument('--systems', type=int, default=512)
    ap.add_argument('--cpu-sample', type=int, default=1000, help='orientations in the CPU baseline sample')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--path', default='auto', choices=['auto', 'propagator', 'complex', 'krylov'],
                    help='auto: spectral path for real generators (FP64, padded dim <= 448); propagator: the power '
                         'scheme with real products; complex: the general complex power scheme; krylov: '
                         'step() per time step (sys.enable krylov)')
    ap.add_argument('--autotune', type=int, default=2, metavar='PASSES',
                    help='timed passes per candidate of sp_plan_autotune before the warm-up (0: the planner\'s '
                         'launch configuration)')
    ap.add_argument('--copy-probe', type=int, default=0, metavar='MIB',
                    help='only run a device copy of MIB MiB (counter calibration) and exit')
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print('bench.py measures one GPU: the multi-GPU path '
              '(orientation sharding + one FID all-reduce) is a design, not a build. Run with --gpus 1.',
              file=sys.stderr)
        sys.exit(2)
    if args.mode is None:
        # FP64 on every workload: the metric holds results to 1e-6 of FP64, which the FP32 paths
        # miss (the MAS workload's FP32 FID is 6.8e-6 from the oracle)
        args.mode = 'fp64'
    args.path_flag = '' if args.path == 'auto' else '+' + args.path
    if args.workload == 'surrogate':
        args.steps, args.warmup = 1, 0
    if args.workload in ('static8', 'static2'):
        value, sec, kms, stages, cfg, cpu, err = bench_static(args, two_spin=args.workload == 'static2')
    elif args.workload == 'batch':
        value, sec, kms, stages, cfg, cpu, err = bench_batch(args)
    elif args.workload == 'surrogate':
        value, sec, kms, stages, cfg, cpu, err = bench_surrogate(args)
    else:
        value, sec, kms, stages, cfg, cpu, err = bench_mas(args)
    peak = 2.5e15 if args.workload == 'surrogate' and cfg['decision'] == 'surrogate' else PEAK[args.mode]
    # roofline of the dominant kernel (largest share of the pass), HIP-event timed
    kernel = max(stages, key=lambda k: stages[k][0])
    k_ms, k_flops = stages[kernel][:2]
    achieved = k_flops / (k_ms * 1e-3) if k_ms > 0 else 0.0
    if kernel.startswith('spk::k_step'):
        # the stepping path is bound by LDS traffic and the barrier of every Taylor term
        lds = stages[kernel][2] / (k_ms * 1e-3) if k_ms > 0 else 0.0
        roof = {'bound': 'lds', 'achieved': round(lds / 1e9, 1), 'peak': round(LDS_PEAK / 1e9, 1), 'unit': 'GB/s',
                'frac': round(lds / LDS_PEAK, 4), 'traffic': TRAFFIC.get((args.workload, args.mode, kernel)),
                'kernel': kernel, 'kernel_ms_avg': round(k_ms, 3), 'lds_bytes_per_launch': stages[kernel][2],
                'flops_per_launch': k_flops, 'fp64_frac': round(achieved / peak, 5)}
    elif args.workload == 'batch' and kernel == 'spk::k_tridiag' and len(stages[kernel]) > 2:
        # the batch's tridiagonalisation stage (k_tridiag_w and k_tridiag): bound by the HBM
        # traffic of k_tridiag_w's global-memory prologue (systems above 256 states)
        hbm = stages[kernel][2] / (k_ms * 1e-3) if k_ms > 0 else 0.0
        roof = {'bound': 'hbm', 'achieved': round(hbm / 1e9, 1), 'peak': 8000.0, 'unit': 'GB/s',
                'frac': round(hbm / 8e12, 4), 'traffic': TRAFFIC.get((args.workload, args.mode, kernel)),
                'kernel': 'spk::k_tridiag_w + spk::k_tridiag', 'kernel_ms_avg': round(k_ms, 3),
                'algorithmic_bytes_per_launch': stages[kernel][2], 'flops_per_launch': k_flops,
                'fp64_frac': round(achieved / peak, 5)}
    else:
        bound = BOUND.get(kernel, 'mfma' if kernel.startswith('spk::k_powder') else 'valu')
        roof = {'bound': bound, 'achieved': round(achieved / 1e12, 3), 'peak': round(peak / 1e12, 1),
                'unit': 'TFLOP/s', 'frac': round(achieved / peak, 4),
                'traffic': TRAFFIC.get((args.workload, args.mode, kernel)), 'kernel': kernel,
                'kernel_ms_avg': round(k_ms, 3), 'flops_per_launch': k_flops}
        if bound == 'valu':
            roof['peak_note'] = 'FP64 vector peak of MI355X (equal to its FP64 matrix peak)'
    stage_json = {k: {'ms': round(v[0], 3), 'tflops': round(v[1] / (v[0] * 1e-3) / 1e12, 3) if v[0] > 0 else None}
                  for k, v in stages.items()}
    line = {
        'metric': METRIC,
        'value': round(value, 1),
        'unit': 'orientation-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(1e3 * sec, 3),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'c128' if args.mode == 'fp64' else 'c64',
        'data': 'synthetic (seeded spin systems and Haar/zcwn grids)',
        'config': cfg,
        'fid_rel_err_vs_oracle': err,
        'roofline': roof,
        'kernels': stage_json,
        'cpu_baseline': cpu,
    }
    if cpu:
        line['speedup_vs_cpu_baseline'] = round(value / cpu['value'], 1)
    print(json.dumps(line))


if __name__ == '__main__':
    main()
