"""SP_MODE_GRAPH: a plan whose pass is recorded once as a HIP graph and
replayed must give exactly the eager plan's results, replay after replay, with
other eager work and host synchronisations between the replays (the spectral
path, the propagator path, a batch plan and a MAS plan)."""
import numpy as np
import pytest

from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu


def _replays_match(eager, graph, other, per_orientation=True, replays=3):
    eager.execute()
    ref = eager.result(per_orientation=per_orientation)
    for r in range(replays):
        st = graph.execute()
        got = graph.result(per_orientation=per_orientation)
        print(f'replay {r}: kernel_ms {st["kernel_ms"]:.3f} stage_ms {st["stage_ms"]}', flush=True)
        assert st['stage_ms'] == [0.0, 0.0, 0.0, 0.0]
        if per_orientation:
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        else:
            assert np.array_equal(got, ref)
        # eager work of the program between replays
        other.execute()
        other.result()


def test_graph_replays_spectral_path():
    w = workloads.static_8spin(n_orient=4000)
    sub = grids.subset(w['grid'], np.arange(16))
    args = (w['I'], w['Q'], sub, w['rho0'], w['coil'], w['dt'], w['nsteps'])
    eager = _native.plan_powder(*args, mode='fp64')
    graph = _native.plan_powder(*args, mode='fp64+graph')
    other = _native.plan_powder(*args, mode='fp64+propagator')
    _replays_match(eager, graph, other)
    assert graph.stats['spectral_systems'] == 1


def test_graph_replays_propagator_path():
    w = workloads.static_2spin(npoints=512, n_orient=64)
    args = (w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    for mode in ('fp64+propagator', 'fp32'):
        eager = _native.plan_powder(*args, mode=mode)
        graph = _native.plan_powder(*args, mode=mode + '+graph')
        other = _native.plan_powder(*args, mode='fp64')
        _replays_match(eager, graph, other)


def test_graph_replays_batch_plan():
    rng = np.random.default_rng(5)
    systems = []
    for nsp in (2, 3, 4):
        y = workloads.random_system(rng, nsp)
        systems.append(dict(I=y['I'], Q=y['Q'], grid=grids.haar(8 * nsp, rng), rho0=y['rho0'], coil=y['coil']))
    eager = _native.plan_powder_batch(systems, 1e-5, 256, mode='fp64')
    graph = _native.plan_powder_batch(systems, 1e-5, 256, mode='fp64+graph')
    eager.execute()
    ref = eager.result_batch()
    for r in range(3):
        graph.execute()
        got = graph.result_batch()
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        eager.execute()
        eager.result_batch()


def test_graph_replays_mas_plan():
    w = workloads.mas_6spin(n_orient=3000, periods=2, ticks=16)
    sub = grids.subset(w['grid'], np.arange(8))
    args = (w['I'], w['Q'], sub, w['rho0'], w['coil'], w['axis'], w['rate_hz'], w['ticks'], w['nsteps'])
    eager = _native.plan_mas(*args, mode='fp32')
    graph = _native.plan_mas(*args, mode='fp32+graph')
    other = _native.plan_mas(*args, mode='fp64')
    _replays_match(eager, graph, other)
