"""GPU parity of the uneven batch at its real sizes: members with reduced
dimensions 195-441 (the propagator path with np up to 448) against the
oracle, and the full 512-system batch checked by size-independent properties.

Members are rebuilt from the batch's own seeds (seed 2, member s), so they are
exactly the systems of the 512-system workload. Below a reduced dimension of
small_matrix = 200 the reference takes expm (kernel/propagator.m:40-42); at
200 and above it takes chop-Taylor with clean_up's 1e-10 quantisation
(kernel/propagator.m:64-201). The GPU computes the exact exponential, so it is
held to the exact-exponential oracle variants at 1e-10 and the deviation from
the chop variant is printed beside it.
"""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu

# member index -> reduced dimension (seed 2, 2-10 spins, 100-1000 orientations)
LARGE = {363: 441, 304: 360, 74: 308, 0: 208, 86: 195}
NSTEPS = 1024
DT = 5e-6


def _members(n_or=3):
    out = []
    for s in LARGE:
        y = workloads.batch_member(2, s)
        y['grid'] = grids.subset(y['grid'], np.arange(n_or))
        out.append(y)
    return out


@pytest.mark.parametrize('mode', ['fp64', 'fp64+complex'])
def test_large_members_vs_exact_oracle(mode):
    ys = _members()
    plan = _native.plan_powder_batch(ys, DT, NSTEPS, mode=mode)
    st = plan.execute()
    fids = plan.result_batch()
    print(f'[{mode}] stats', st, flush=True)
    assert st['reduced_dim'] == 441 and st['padded_dim'] == 448
    # real generators: the spectral path takes every member in FP64 (k_tridiag_w, with its
    # global-memory prologue above 256 states); '+complex' keeps them on the propagator path
    assert st['spectral_systems'] == (len(LARGE) if mode == 'fp64' else 0)
    worst = 0.0
    for (s, n), y, f in zip(LARGE.items(), ys, fids):
        assert len(O.union_states(y['I'], y['Q'], y['coil'])) == n
        exact = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], DT, NSTEPS, variant=2)
        chop = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], DT, NSTEPS, variant=0)
        e, ec = O.rel_err(f, exact), O.rel_err(f, chop)
        print(f'[{mode}] member {s} (reduced dim {n}): vs exact-expm oracle {e:.2e}, vs chop-Taylor oracle {ec:.2e}',
              flush=True)
        worst = max(worst, e)
    assert worst < 1e-10


def test_largest_member_vs_faithful_reduction():
    """The faithful per-orientation reduction (ZTE + path tracing on the full
    3676-state generator) with the exact exponential: the union-pattern
    reduction keeps every state it keeps."""
    y = workloads.batch_member(2, 363)
    y['grid'] = grids.subset(y['grid'], np.arange(2))
    f, _, st = _native.powder_fid(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], DT, NSTEPS)
    ref = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], DT, NSTEPS, variant=1)
    e = O.rel_err(f, ref)
    print(f'member 363 (3676 -> {st["reduced_dim"]} states): vs faithful-reduction exact oracle {e:.2e}', flush=True)
    assert e < 1e-10


def test_full_batch_properties():
    """All 512 systems x 1024 steps in one plan: every FID finite, f(0) equal
    to sum_n w_n coil'*rho0 (the detection before the first step), and the
    batch equal to single-system calls on a sample."""
    b = workloads.batch_uneven(seed=2, nsteps=NSTEPS)
    plan = _native.plan_powder_batch(b['systems'], b['dt'], b['nsteps'])
    st = plan.execute()
    fids = plan.result_batch()
    total_or = sum(len(y['grid']['weights']) for y in b['systems'])
    print(f'full batch: {len(fids)} systems, {total_or} orientations, {st["kernel_ms"]:.1f} ms, '
          f'{st["spectral_systems"]} spectral', flush=True)
    assert len(fids) == 512
    worst0 = 0.0
    for y, f in zip(b['systems'], fids):
        assert np.all(np.isfinite(f))
        f0 = np.sum(y['grid']['weights']) * np.vdot(y['coil'], y['rho0'])
        worst0 = max(worst0, abs(f[0] - f0) / abs(f0))
    print(f'full batch: worst |f(0) - w.coil\'rho0| / |.| = {worst0:.2e}', flush=True)
    assert worst0 < 1e-12
    for s in (363, 74, 5):
        y = b['systems'][s]
        g, _, _ = _native.powder_fid(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], b['dt'], b['nsteps'])
        e = O.rel_err(fids[s], g)
        print(f'full batch member {s}: batch vs single call {e:.2e}', flush=True)
        assert e < 1e-12


# members of reduced dimension 193-256 (padded 208-256): real generators, so the spectral path
# takes them with the 1024-thread tridiagonalisation; members 5 and 7 (small) share the plan,
# so the 256-thread kernels run on the wide per-item layout beside them
WIDE = {86: 195, 79: 204, 0: 208, 104: 234, 248: 244, 123: 246, 264: 252}


@pytest.mark.parametrize('graph', [False, True])
def test_wide_spectral_members_vs_exact_oracle_and_propagator(graph):
    ys = []
    for s in list(WIDE) + [5, 7]:
        y = workloads.batch_member(2, s)
        y['grid'] = grids.subset(y['grid'], np.arange(3))
        ys.append(y)
    mode = 'fp64+graph' if graph else 'fp64'
    plan = _native.plan_powder_batch(ys, DT, NSTEPS, mode=mode)
    st = plan.execute()
    if graph:
        st = plan.execute()  # a replay
    fids = plan.result_batch()
    print(f'[{mode}] stats', st, flush=True)
    assert st['spectral_systems'] == len(ys) and st['reduced_dim'] == 252
    prop = _native.plan_powder_batch(ys, DT, NSTEPS, mode='fp64+propagator')
    prop.execute()
    pf = prop.result_batch()
    worst = worst_p = 0.0
    for s, y, f, g in zip(list(WIDE) + [5, 7], ys, fids, pf):
        n = len(O.union_states(y['I'], y['Q'], y['coil']))
        if s in WIDE:
            assert n == WIDE[s]
        exact = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], DT, NSTEPS, variant=2)
        e, ep = O.rel_err(f, exact), O.rel_err(f, g)
        print(f'[{mode}] member {s} (reduced dim {n}): spectral vs exact-expm oracle {e:.2e}, '
              f'vs propagator path {ep:.2e}', flush=True)
        worst, worst_p = max(worst, e), max(worst_p, ep)
    assert worst < 1e-10 and worst_p < 1e-10


# one plan with every spectral size class: above 256 states (k_tridiag_w's global-memory
# prologue, eigensolver for 448), 193-256 (register layout only) and up to 192 (256-thread kernel)
GLOBAL = {363: 441, 304: 360, 482: 320, 74: 308, 3: 276}


@pytest.mark.parametrize('graph', [False, True])
def test_global_prologue_members_vs_exact_oracle_and_propagator(graph):
    members = list(GLOBAL) + [264, 5]
    ys = []
    for s in members:
        y = workloads.batch_member(2, s)
        y['grid'] = grids.subset(y['grid'], np.arange(2))
        ys.append(y)
    mode = 'fp64+graph' if graph else 'fp64'
    plan = _native.plan_powder_batch(ys, DT, NSTEPS, mode=mode)
    st = plan.execute()
    if graph:
        st = plan.execute()  # a replay
    fids = plan.result_batch()
    print(f'[{mode}] stats', st, flush=True)
    assert st['spectral_systems'] == len(ys) and st['padded_dim'] == 448
    prop = _native.plan_powder_batch(ys, DT, NSTEPS, mode='fp64+propagator')
    prop.execute()
    pf = prop.result_batch()
    worst = worst_p = 0.0
    for s, y, f, g in zip(members, ys, fids, pf):
        n = len(O.union_states(y['I'], y['Q'], y['coil']))
        if s in GLOBAL:
            assert n == GLOBAL[s]
        exact = O.refcpu_powder(y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], DT, NSTEPS, variant=2)
        e, ep = O.rel_err(f, exact), O.rel_err(f, g)
        print(f'[{mode}] member {s} (reduced dim {n}): spectral vs exact-expm oracle {e:.2e}, '
              f'vs propagator path {ep:.2e}', flush=True)
        worst, worst_p = max(worst, e), max(worst_p, ep)
    assert worst < 1e-10 and worst_p < 1e-10
