"""GPU parity of the magic-angle-spinning Hilbert-space path (sp_mas_fid)
against the direct-stepping oracle."""
import numpy as np
import pytest

import pyoracle as O
from spinach_amd import _native, grids, spinsys as S, workloads

pytestmark = pytest.mark.gpu


def _small(seed=7):
    rng = np.random.default_rng(seed)
    iso = ['13C', '13C', '1H']
    inter = {'zeeman': {'eigs': [list(rng.uniform(-100, 100, 3)) for _ in iso],
                        'euler': [list(rng.uniform(0, 6, 3)) for _ in iso]},
             'coordinates': [np.zeros(3), np.array([1.5, 0.2, 0.1]), np.array([0.3, 1.1, 0.4])]}
    ss = S.basis(S.create({'magnet': 14.1, 'isotopes': iso}, inter), {'formalism': 'zeeman-hilb'})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    return ss, I, S.q_list(Q), S.state(ss, 'Lx', '13C'), S.state(ss, 'L+', '13C'), grids.haar(3, rng)


def test_mas_small_fp64_vs_oracle():
    ss, I, q, rho0, coil, g = _small()
    args = (I, q, g, rho0, coil, (1.0, 1.0, 1.0), 10000.0, 16, 3 * 16)
    fid, per, st = _native.mas_fid(*args, mode='fp64', per_orientation=True)
    ref, rper = O.mas_powder(*args, per_orientation=True)
    e = O.rel_err(fid, ref)
    print(f'MAS 3-spin fp64 rel err {e:.2e}', st, flush=True)
    assert e < 1e-10
    for k in range(len(per)):
        assert O.rel_err(per[k], rper[k]) < 1e-10


def test_mas_axis_along_field_matches_static():
    ss, I, q, rho0, coil, g = _small(3)
    rate, ticks, nsteps = 10000.0, 16, 40
    fid, _, _ = _native.mas_fid(I, q, g, rho0, coil, (0.0, 0.0, 1.0), rate, ticks, nsteps, mode='fp64')
    ref = O.static_hilbert_powder(I, q, g, rho0, coil, 1.0 / (rate * ticks), nsteps)
    assert O.rel_err(fid, ref) < 1e-10


def test_mas_6spin_two_periods_fp64_and_fp32():
    w = workloads.mas_6spin(n_orient=3000)
    g = grids.subset(w['grid'], np.arange(2))
    nsteps = 2 * w['ticks']
    args = (w['I'], w['Q'], g, w['rho0'], w['coil'], w['axis'], w['rate_hz'], w['ticks'], nsteps)
    f64, _, st = _native.mas_fid(*args, mode='fp64')
    ref = O.mas_powder(*args)
    e64 = O.rel_err(f64, ref)
    f32, _, _ = _native.mas_fid(*args, mode='fp32')
    e32 = O.rel_err(f32, ref)
    print(f'MAS 6-spin 2 orientations x {nsteps} steps: fp64 {e64:.2e}, fp32 {e32:.2e}', st, flush=True)
    assert e64 < 1e-10
    assert e32 < 1e-3


def test_mas_tiled_matches_dense_and_oracle(monkeypatch):
    """The 6-spin system's invariant subspaces (at most 12 states) are propagated as 16 x 16
    tiles; the dense 64 x 64 kernel (SPINACH_MAS_DENSE=1) and the oracle must agree with it,
    per orientation, over a record that ends inside a rotor period."""
    w = workloads.mas_6spin(n_orient=3000)
    g = grids.subset(w['grid'], np.arange(3))
    nsteps = 2 * w['ticks'] + 37
    args = (w['I'], w['Q'], g, w['rho0'], w['coil'], w['axis'], w['rate_hz'], w['ticks'], nsteps)
    ft, pt, st = _native.mas_fid(*args, mode='fp64', per_orientation=True)
    monkeypatch.setenv('SPINACH_MAS_DENSE', '1')
    fd, pd, sd = _native.mas_fid(*args, mode='fp64', per_orientation=True)
    ref, rper = O.mas_powder(*args, per_orientation=True)
    print(f"MAS 6-spin tiled {O.rel_err(ft, ref):.2e}, dense {O.rel_err(fd, ref):.2e}, "
          f"tiled vs dense {O.rel_err(ft, fd):.2e}; {st['mas_tiles']} tiles, {st['mas_pairs']} pairs, {st['mas_detected']} detected", flush=True)
    assert st['padded_dim'] == 16 and sd['padded_dim'] == 64
    # 15 subspaces of 12, 8, 8, 6, 6, 4 x 4, 2, 2, 1 x 4 states fill four tiles exactly
    assert st['mas_tiles'] == 4 and 1 <= st['mas_pairs'] <= 16 and sd['mas_tiles'] == 0
    # L+(13C) and Lx(13C) meet on the blocks (m_C, m_H) -> (m_C + 1, m_H): (1*4 + 4*6 + 6*4 + 4*1) *
    # (1 + 4 + 1) = 336 elements at most (fewer where the union pattern splits a block)
    assert 16 <= st['mas_detected'] <= 336
    assert O.rel_err(ft, ref) < 1e-10 and O.rel_err(ft, fd) < 1e-10
    for k in range(len(pt)):
        assert O.rel_err(pt[k], rper[k]) < 1e-10
        assert O.rel_err(pt[k], pd[k]) < 1e-10


@pytest.mark.parametrize('nsteps', [5, 15, 16, 33])
def test_mas_tiled_record_lengths(nsteps):
    """Records shorter than a rotor period, ending on a period boundary and one or two steps past
    it: the whole periods go through the detection product, the rest through dot products."""
    ss, I, q, rho0, coil, g = _small(11)
    args = (I, q, g, rho0, coil, (1.0, 1.0, 1.0), 10000.0, 16, nsteps)
    fid, per, st = _native.mas_fid(*args, mode='fp64', per_orientation=True)
    ref, rper = O.mas_powder(*args, per_orientation=True)
    assert st['mas_tiles'] == 1
    assert O.rel_err(fid, ref) < 1e-10
    for k in range(len(per)):
        assert O.rel_err(per[k], rper[k]) < 1e-10


def _mas_system(iso, seed):
    rng = np.random.default_rng(seed)
    inter = {'zeeman': {'eigs': [list(rng.uniform(-60, 60, 3)) for _ in iso],
                        'euler': [list(rng.uniform(0, 6, 3)) for _ in iso]},
             'coordinates': [np.array(x) for x in workloads._coords(rng, len(iso), box=3.5, min_sep=1.0)]}
    ss = S.basis(S.create({'magnet': 14.1, 'isotopes': iso}, inter), {'formalism': 'zeeman-hilb'})
    I, Q = S.hamiltonian(S.assume(ss, 'nmr'))
    return ss, I, S.q_list(Q)


def test_mas_tiled_eight_tiles_two_per_wave():
    """4 13C, 2 1H and one 15N (Hilbert dimension 128): 30 invariant subspaces of at most 12
    states fill eight tiles, so every wave propagates two tiles and several pairs."""
    ss, I, q = _mas_system(['13C'] * 4 + ['1H'] * 2 + ['15N'], 21)
    g = grids.haar(2, np.random.default_rng(22))
    args = (I, q, g, S.state(ss, 'Lx', '13C'), S.state(ss, 'L+', '13C'), (1.0, 1.0, 1.0), 12500.0, 32, 70)
    fid, per, st = _native.mas_fid(*args, mode='fp64', per_orientation=True)
    ref, rper = O.mas_powder(*args, per_orientation=True)
    e = O.rel_err(fid, ref)
    print(f"MAS 7-spin (dim 128): {st['mas_tiles']} tiles, {st['mas_pairs']} pairs, {st['mas_detected']} detected; "
          f"vs oracle {e:.2e}", flush=True)
    assert st['mas_tiles'] == 8 and e < 1e-10
    for k in range(len(per)):
        assert O.rel_err(per[k], rper[k]) < 1e-10


def test_mas_component_above_a_tile_takes_the_dense_kernel():
    """Six protons: the m = 0 subspace holds 20 states, more than a tile, so the dense kernel
    runs (mas_tiles = 0), still against the oracle."""
    ss, I, q = _mas_system(['1H'] * 6, 23)
    g = grids.haar(2, np.random.default_rng(24))
    args = (I, q, g, S.state(ss, 'Lx', '1H'), S.state(ss, 'L+', '1H'), (1.0, 1.0, 1.0), 12500.0, 32, 40)
    fid, _, st = _native.mas_fid(*args, mode='fp64')
    ref = O.mas_powder(*args)
    e = O.rel_err(fid, ref)
    print(f"MAS 6 protons: mas_tiles {st['mas_tiles']}, padded {st['padded_dim']}, vs oracle {e:.2e}", flush=True)
    assert st['mas_tiles'] == 0 and st['padded_dim'] == 64 and e < 1e-10


def test_mas_tiled_nothing_detected():
    """coil Lz against rho0 Lx: the secular Hamiltonian keeps the transverse state transverse, so
    no component block holds both and the detection sums over no element: the FID is zero (the
    oracle's too), with one padded group of 16 columns in the detection product."""
    ss, I, q, rho0, coil, g = _small(5)
    lz = S.state(ss, 'Lz', '13C')
    args = (I, q, g, rho0, lz, (1.0, 1.0, 1.0), 10000.0, 16, 40)
    fid, _, st = _native.mas_fid(*args, mode='fp64')
    ref = O.mas_powder(*args)
    print(f"MAS nothing detected: max |f| {np.abs(fid).max():.1e}, oracle {np.abs(ref).max():.1e}, "
          f"detected {st['mas_detected']}", flush=True)
    assert st['mas_tiles'] >= 1 and st['mas_detected'] == 16
    assert np.abs(fid).max() == 0.0 and np.abs(ref).max() < 1e-12
