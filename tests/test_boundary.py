"""The drop-in boundary on the CPU, under AddressSanitizer and
UndefinedBehaviorSanitizer:

* the MEX gateway (spinach_amd/mex/powder_hip.cpp) compiled against the
  mex.h stand-in and a recording stub of the C-ABI: argument marshalling
  (sparse and full states, the Q cell layout, spin_system.tols, dead_time,
  sum_up, krylov, precision, the relaxation planes and R0), output layout,
  and every error path with its
  "Spinach:powder_hip:<id>" identifier (etc/mex/prune_cpu.cpp:199-223
  convention); a leak or invalid access fails the sanitizers;
* the host planner (spinach_amd/csrc/planner.hpp) on workload systems and on
  corrupted inputs it must reject;
* the oracle (oracle/refcpu.cpp) against its normal build;
* the native Hamiltonian builder (spinach_amd/csrc/builder.hpp) on
  superoperators whose structure the algebra fixes (Zeeman diagonal,
  flip-flop symmetry and projection conservation, rank-2 slot occupancy)
  and on inputs it must reject.
The host mirror of powder() (spinach_amd/powder.py) is checked for the
arguments it must refuse before it reaches the library.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sps

import boundary_io as B
import pyoracle as O
from spinach_amd import _native, grids, workloads
from spinach_amd import powder as P

SP_MODE_KRYLOV, SP_MODE_FP32 = 128, 1


@pytest.fixture(scope='module')
def native():
    B.build_native('all')
    return B.BUILD


def _sanitizer_clean(r):
    assert 'AddressSanitizer' not in r.stderr and 'runtime error' not in r.stderr, r.stderr[-3000:]
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-3000:])


# ------------------------------------------------------------- MEX gateway

def _stub(native, rhs, nlhs, tmp, fail=None):
    log = os.path.join(tmp, 'stub.json')
    if os.path.exists(log):
        os.remove(log)
    env = {'SP_STUB_LOG': log}
    if fail:
        env['SP_STUB_FAIL'] = fail
    r, outs = B.run_driver(os.path.join(native, 'mex_driver_stub'), rhs, nlhs, tmp, env)
    _sanitizer_clean(r)
    rec = json.load(open(log)) if os.path.exists(log) else None
    return r, outs, rec


def _marker_system():
    """The 2-spin workload with every Q slot replaced by a matrix that names its
    slot (one entry, value slot+1 at row slot), so that the test sees where
    the gateway puts each cell."""
    w = workloads.static_2spin(npoints=32, n_orient=5)
    dim = w['I'].shape[0]
    qn = [[[None] * 3 for _ in range(3)], [[None] * 5 for _ in range(5)]]
    for r, sz in ((0, 3), (1, 5)):
        for k in range(sz):
            for m in range(sz):
                slot = k * 3 + m if r == 0 else 9 + k * 5 + m
                if slot % 7 == 3:
                    continue  # some empty slots
                qn[r][k][m] = sps.csc_matrix(([complex(slot + 1, -slot)], ([slot % dim], [(slot * 5) % dim])),
                                             shape=(dim, dim))
    return w, qn


def test_gateway_marshalling(native, tmp_path):
    w, qn = _marker_system()
    rho0 = np.asarray(w['rho0'], complex) + 0.25j * np.asarray(w['coil'], complex)
    tols = {'liouv_zero': 2e-10, 'zte_tol': 1e-20, 'prop_chop': 3e-10, 'small_matrix': 150, 'krylov_tol': 123}
    opts = {'dead_time': B.mx_dense(np.array([[4.5e-6]])), 'krylov': B.mx_logical(True),
            'precision': B.mx_char('double')}
    rhs = B.powder_hip_rhs(w['I'], qn, w['grid'], rho0, w['coil'], w['dt'], w['nsteps'], tols, opts)
    r, outs, rec = _stub(native, rhs, 2, str(tmp_path))
    assert r.stdout.startswith('OK'), r.stdout
    fid, per = outs
    npts, n_or = w['nsteps'] + 1, len(w['grid']['weights'])
    t = np.arange(npts)
    assert fid.shape == (npts, 1) and np.array_equal(fid[:, 0], (t + 1) + 0.5j * t)
    assert per.shape == (npts, n_or)
    for o in range(n_or):
        assert np.array_equal(per[:, o], o + 1e-3 * t - 1j * o)
    # inputs as the library received them
    Ic = sps.csc_matrix(w['I'])
    Ic.sort_indices()
    got = np.array(rec['I']['entries'])
    cols = np.repeat(np.arange(Ic.shape[1]), np.diff(Ic.indptr))
    assert np.array_equal(got[:, 0], Ic.indices) and np.array_equal(got[:, 1], cols)
    assert np.array_equal(got[:, 2] + 1j * got[:, 3], Ic.data)
    for slot, q in enumerate(rec['Q']):
        if slot % 7 == 3:
            assert q['n'] == 0 and q['entries'] == []
        else:
            (row, col, re, im), = q['entries']
            assert (row, re, im) == (slot % 16, slot + 1, -slot) and col == (slot * 5) % 16
    assert np.array_equal(np.array(rec['rho0']).view(complex), rho0)
    assert np.array_equal(np.array(rec['coil']).view(complex), np.asarray(w['coil'], complex))
    for k in ('alphas', 'betas', 'gammas', 'weights'):
        assert np.array_equal(rec[k], w['grid'][k])
    assert rec['tols'] == [2e-10, 1e-20, 3e-10, 150, 123]
    assert rec['mode'] == SP_MODE_KRYLOV and rec['dead_time'] == 4.5e-6
    assert rec['dt'] == w['dt'] and rec['nsteps'] == w['nsteps']


T1 = np.array([[1.0, 0.3, -0.2], [0.3, 3.0, 0.5], [-0.2, 0.5, 5.0]]) * 40.0
RELAX2 = {'relaxation': ['t1_t2', 'damp'], 'r1_rates': [4.0, T1], 'r2_rates': [T1 * 1.5, 9.0], 'damp_rate': 7.0,
          'rlx_keep': 'secular'}


def test_gateway_relaxation_marshalling(native, tmp_path):
    """opts.rlx_diag (dim x 7, column p = g_p) and opts.R0 reach
    sp_plan_powder_relax as the library's 7*dim planes and CSC matrix; without
    them the gateway calls sp_plan_powder."""
    from spinach_amd import spinsys as S
    w = workloads.static_2spin(npoints=16, n_orient=3, relax=RELAX2)
    g = S.relaxation_terms(w['ss'])['diag']
    dim = w['I'].shape[0]
    R0 = sps.csc_matrix(([0.5 - 0.25j, 3.0], ([1, 2], [2, 1])), shape=(dim, dim))
    for opts, want_r0 in (({'rlx_diag': B.mx_dense(g.T), 'R0': B.mx_sparse(R0)}, True),
                          ({'rlx_diag': B.mx_dense(g.T)}, False)):
        rhs = B.powder_hip_rhs(w['I'], w['Qnested'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'],
                               opts=opts)
        r, outs, rec = _stub(native, rhs, 1, str(tmp_path))
        assert r.stdout.startswith('OK'), r.stdout
        assert np.array_equal(np.array(rec['relax']['diag']), g.ravel())
        if want_r0:
            ent = np.array(rec['relax']['R0']['entries'])
            assert rec['relax']['R0']['n'] == dim
            assert np.array_equal(ent[:, 0], [2, 1]) and np.array_equal(ent[:, 1], [1, 2])
            assert np.array_equal(ent[:, 2] + 1j * ent[:, 3], [3.0, 0.5 - 0.25j])
        else:
            assert rec['relax']['R0'] is None
    rhs = B.powder_hip_rhs(w['I'], w['Qnested'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'])
    r, outs, rec = _stub(native, rhs, 1, str(tmp_path))
    assert r.stdout.startswith('OK') and 'relax' not in rec


def test_gateway_defaults_full_states_single_and_sum_up_false(native, tmp_path):
    w = workloads.static_2spin(npoints=16, n_orient=3)
    rhs = B.powder_hip_rhs(w['I'], w['Qnested'], w['grid'], w['rho0'], w['coil'], w['dt'], w['nsteps'],
                           tols=None, opts={'sum_up': B.mx_logical(False), 'precision': B.mx_char('single')},
                           sparse_states=False)
    r, outs, rec = _stub(native, rhs, 1, str(tmp_path))
    assert r.stdout.startswith('OK'), r.stdout
    per, = outs
    assert per.shape == (16, 3)  # sum_up = false: the per-orientation FIDs are the answer
    assert rec['tols'] == [1e-10, 1e-24, 1e-10, 200, 10000] and rec['mode'] == SP_MODE_FP32
    assert 'dead_time' not in rec and rec['fid_out'] == 1 and rec['per_out'] == 1
    assert np.array_equal(np.array(rec['rho0']).view(complex), np.asarray(w['rho0'], complex))


@pytest.mark.parametrize('case,ident', [
    ('nrhs', 'Spinach:powder_hip:nrhs'),
    ('dense_I', 'Spinach:powder_hip:input'),
    ('q_shape', 'Spinach:powder_hip:input'),
    ('rho_len', 'Spinach:powder_hip:input'),
    ('grid_len', 'Spinach:powder_hip:input'),
    ('precision', 'Spinach:powder_hip:input'),
    ('nsteps', 'Spinach:powder_hip:input'),
    ('tols', 'Spinach:powder_hip:input'),
    ('fail_create', 'Spinach:powder_hip:device'),
    ('fail_plan', 'Spinach:powder_hip:plan'),
    ('fail_dead_time', 'Spinach:powder_hip:input'),
    ('fail_execute', 'Spinach:powder_hip:run'),
    ('fail_result', 'Spinach:powder_hip:run'),
    ('rlx_diag_shape', 'Spinach:powder_hip:input'),
    ('rlx_diag_complex', 'Spinach:powder_hip:input'),
    ('r0_dim', 'Spinach:powder_hip:input'),
])
def test_gateway_errors(native, tmp_path, case, ident):
    w = workloads.static_2spin(npoints=8, n_orient=2)
    g = dict(w['grid'])
    I, rho0, nsteps, tols = w['I'], w['rho0'], w['nsteps'], None
    opts = {'dead_time': B.mx_dense(np.array([[1e-6]]))}
    fail = None
    if case == 'grid_len':
        g['betas'] = g['betas'][:1]
    if case == 'rho_len':
        rho0 = np.asarray(rho0)[:-1]
    if case == 'precision':
        opts['precision'] = B.mx_char('half')
    if case == 'nsteps':
        nsteps = 2.5
    if case == 'tols':
        tols = {'krylov_tol': np.nan}
    rhs = B.powder_hip_rhs(I, w['Qnested'], g, rho0, w['coil'], w['dt'], nsteps, tols=None, opts=opts)
    if case == 'tols':
        rhs[7] = B.mx_struct({'krylov_tol': B.mx_dense(np.array([[1.0, 2.0]]))})
    if case == 'nrhs':
        rhs = rhs[:8]
    if case == 'dense_I':
        rhs[0] = B.mx_dense(I.toarray())
    if case == 'q_shape':
        rhs[1] = B.mx_cell([rhs[1][1][0]], (1, 1))
    dim = I.shape[0]
    if case == 'rlx_diag_shape':
        rhs[8][1]['rlx_diag'] = B.mx_dense(np.zeros((dim, 6)))
    if case == 'rlx_diag_complex':
        rhs[8][1]['rlx_diag'] = B.mx_dense(np.zeros((dim, 7)) + 1j)
    if case == 'r0_dim':
        rhs[8][1]['R0'] = B.mx_sparse(sps.eye(dim + 1, format='csc'))
    if case.startswith('fail_'):
        fail = case[5:]
    r, outs, rec = _stub(native, rhs, 1, str(tmp_path), fail=fail)
    assert r.stdout.startswith('ERR ' + ident + ' '), r.stdout
    if fail:
        assert f'stub failure in {fail}' in r.stdout


# --------------------------------------------------------- planner, oracle

@pytest.mark.parametrize('which', ['static_2spin', 'static_8spin', 'batch_363'])
def test_planner_under_sanitizers(native, tmp_path, which):
    if which == 'static_2spin':
        w = workloads.static_2spin(npoints=8, n_orient=4)
    elif which == 'static_8spin':
        w = workloads.static_8spin(n_orient=4)
    else:
        w = workloads.batch_member(2, 363)
        w['grid'] = grids.subset(w['grid'], np.arange(3))
    path = str(tmp_path / 'p.bin')
    B.write_problem(path, w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], 5e-6, 8)
    r = __import__('subprocess').run([os.path.join(native, 'planner_san'), path], capture_output=True, text=True,
                                     timeout=600)
    _sanitizer_clean(r)
    lines = r.stdout.strip().splitlines()
    print('\n'.join(lines))
    assert len(lines) == 6
    head = dict(kv.split('=') for kv in lines[0].split()[1:])
    n = len(O.union_states(w['I'], w['Q'], w['coil']))
    assert int(head['n']) == n and int(head['np']) == max(16, (n + 15) // 16 * 16)
    assert int(head['real']) == 1 and int(head['sell_nnz']) >= int(head['nnz'])
    G = int(head['lanes'])
    assert G in (1, 2, 4, 8) and int(head['nslices']) == -(-n // (64 // G))
    for line in lines[1:]:
        assert line.startswith('REJECT 1 '), line


def test_oracle_under_sanitizers(native, tmp_path):
    rng = np.random.default_rng(3)
    y = workloads.random_system(rng, 4)
    y['grid'] = grids.haar(3, rng)
    path, out = str(tmp_path / 'p.bin'), str(tmp_path / 'o.bin')
    B.write_problem(path, y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], 5e-6, 40)
    r = __import__('subprocess').run([os.path.join(native, 'refcpu_san'), path, out], capture_output=True,
                                     text=True, timeout=600)
    _sanitizer_clean(r)
    f = np.fromfile(out, '<f8').view(complex).reshape(6, 41)
    args = (y['I'], y['Q'], y['grid'], y['rho0'], y['coil'], 5e-6, 40)
    refs = [O.refcpu_powder(*args, variant=0), O.refcpu_powder(*args, variant=1),
            O.refcpu_powder(*args, variant=2), O.refcpu_powder(*args, variant=1, tols=O.make_tols(krylov_tol=2)),
            O.refcpu_powder(*args, variant=0, dead_time=3.3e-6),
            O.refcpu_powder(*args, variant=1, tols=O.make_tols(merge_dim=2))]
    for k, ref in enumerate(refs):
        assert O.rel_err(f[k], ref) < 1e-12, k


def test_builder_under_sanitizers(native):
    r = __import__('subprocess').run([os.path.join(native, 'builder_san')], capture_output=True, text=True,
                                     timeout=600)
    _sanitizer_clean(r)
    lines = r.stdout.strip().splitlines()
    print('\n'.join(lines))
    assert [l.split()[0] for l in lines] == ['ZEEMAN', 'FLIPFLOP', 'PRUNED', 'ANISO'] + ['REJECT'] * 5 + ['DONE']
    assert all(l.endswith('ok=1') for l in lines[:4]) and lines[-1] == 'DONE failures=0'


# ------------------------------------------------------ host mirror checks

@pytest.mark.parametrize('bad,match', [
    ({'sweep': -1.0}, 'sweep'),
    ({'npoints': 0}, 'npoints'),
    ({'rframes': [1]}, 'rotating frames'),
    ({'decouple': ['1H']}, 'decoupling'),
    ({'homodec_oper': 1}, 'homodecoupling'),
    ({'rho0': lambda x: x}, 'rho0'),
    ({'dead_time': -1e-6}, 'dead_time'),
])
def test_powder_host_mirror_grumble(bad, match):
    w = workloads.static_2spin(npoints=8, n_orient=2)
    params = {'rho0': w['rho0'], 'coil': w['coil'], 'sweep': 1e5, 'npoints': 8, 'grid': w['grid']}
    params.update(bad)
    with pytest.raises((ValueError, NotImplementedError), match=match):
        P.powder(w['ss'], P.acquire, params, 'nmr')


def test_powder_host_mirror_rejects_other_sequences():
    w = workloads.static_2spin(npoints=8, n_orient=2)
    params = {'rho0': w['rho0'], 'coil': w['coil'], 'sweep': 1e5, 'npoints': 8, 'grid': w['grid']}
    with pytest.raises(NotImplementedError):
        P.powder(w['ss'], lambda *a: None, params, 'nmr')


def test_exports_include_new_entry_points():
    lib = _native.load()
    for name in ('sp_plan_set_dead_time', 'sp_powder_fid_noalloc'):
        assert hasattr(lib, name)
