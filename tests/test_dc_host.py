"""The divide-and-conquer tridiagonal eigensolver of the spectral path, built
for the host from the same phase functions the k_dceig kernel runs, against
dense eigendecompositions (numpy.linalg.eigh): eigenvalues, and the detected
trajectory f(t) = sum_k g_k exp(-i lambda_k t) with g_k = (coil' v_k)(v_k' rho)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg as sla

import pyoracle as O
from spinach_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NP = 192    # the solver sized for the 256-thread tridiagonalisation
NPW = 256   # the one sized for the 1024-thread tridiagonalisation (dimensions 193-256)
NPG = 448   # and for its global-memory prologue (dimensions 257-448, 512 threads)


@pytest.fixture(scope='module')
def harness(tmp_path_factory):
    out = tmp_path_factory.mktemp('dc') / 'libdchost.so'
    src = os.path.join(ROOT, 'tests', 'native', 'dc_host_harness.hip')
    subprocess.run(['/opt/rocm/bin/hipcc', '-O2', '-std=c++17', '-fPIC', '-shared', '--offload-arch=gfx950',
                    '-I', os.path.join(ROOT, 'include'), src, '-o', str(out)], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    lib = ctypes.CDLL(str(out))
    lib.dc_host_solve.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_int)]
    return lib


def tridiagonalise(A, vecs):
    """T = Q^T A Q (LAPACK Hessenberg reduction of a symmetric matrix) and Q^T v."""
    H, Q = sla.hessenberg(A, calc_q=True)
    return np.diag(H).copy(), np.diag(H, 1).copy(), [Q.T @ v for v in vecs]


def solve(lib, A, coil, rho, nsteps=4096, nm=None):
    n = len(A)
    nm = nm or (NP if n <= NP else (NPW if n <= NPW else NPG))
    cc = np.conj(coil)
    d, e, (cr, ci, rr, ri) = tridiagonalise(A, [cc.real, cc.imag, rho.real, rho.imag])
    buf = np.zeros(6 * nm)
    buf[:n] = d
    buf[nm:nm + n - 1] = e
    for q, v in enumerate((cr, ci, rr, ri)):
        buf[(2 + q) * nm:(2 + q) * nm + n] = v
    out = np.zeros(6 * nm)
    st, lev = ctypes.c_int(), ctypes.c_int()
    dp = ctypes.POINTER(ctypes.c_double)
    assert lib.dc_host_solve(buf.ctypes.data_as(dp), n, nm, out.ctypes.data_as(dp), ctypes.byref(st),
                             ctypes.byref(lev)) == 0
    lam = out[:n]
    g = out[2 * nm:2 * nm + n] + 1j * out[3 * nm:3 * nm + n]
    w, V = np.linalg.eigh(A)
    gref = (V.T @ cc) * (V.T @ rho)
    t = np.arange(nsteps + 1)
    f = np.exp(-1j * np.outer(t, lam)) @ g
    fref = np.exp(-1j * np.outer(t, w)) @ gref
    return (np.abs(np.sort(lam) - w).max(), np.abs(f - fref).max() / np.abs(fref).max(), st.value, lev.value)


@pytest.mark.parametrize('n', [1, 2, 3, 12, 13, 25, 97, 150, 192, 193, 231, 256, 257, 300, 385, 441, 448])
def test_random_symmetric(harness, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = 0.3 * (A + A.T)
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = rng.standard_normal(n) + 0j
    le, fe, st, lev = solve(harness, A, c, r)
    assert st >= 0 and st <= 30, st
    assert le < 1e-13 and fe < 1e-10, (le, fe)


@pytest.mark.parametrize('n', [30, 96, 192, 252, 444])
def test_degenerate_spectra(harness, n):
    # six-fold degenerate eigenvalues in a random basis: deflation by rotation
    rng = np.random.default_rng(100 + n)
    B = np.diag(np.repeat(rng.standard_normal(n // 6), 6)) * 0.5
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ B @ Q.T
    v = rng.standard_normal(n) + 0j
    le, fe, st, lev = solve(harness, A, v, rng.standard_normal(n) + 0j)
    assert st >= 0 and le < 1e-13 and fe < 1e-10, (le, fe, st)
    # already diagonal: every merge deflates completely
    le, fe, st, lev = solve(harness, B, v, v)
    assert st >= 0 and le == 0.0 and fe < 1e-12


def test_8spin_orientations(harness):
    w = workloads.static_8spin(nsteps=8, n_orient=4000)
    st = O.union_states(w['I'], w['Q'], w['coil'])
    g = w['grid']
    for o in range(2):
        H = O.assemble_dense(w['I'], w['Q'], (g['alphas'][o], g['betas'][o], g['gammas'][o]))[np.ix_(st, st)]
        le, fe, stat, lev = solve(harness, 5e-6 * H.real, w['coil'][st], w['rho0'][st])
        assert lev >= 1 and stat > 0
        assert le < 1e-13 and fe < 1e-10, (le, fe)


def test_wigner_slot_matches_whole_set(harness):
    """k_tridiag computes the 34 Wigner coefficients one slot per thread; the
    per-slot function must reproduce the whole-set one exactly."""
    harness.wigner_slot_mismatches.argtypes = [ctypes.c_double] * 3
    harness.wigner_slot_mismatches.restype = ctypes.c_int
    rng = np.random.default_rng(11)
    for a, b, g in rng.uniform([0, 0, 0], [2 * np.pi, np.pi, 2 * np.pi], (50, 3)):
        assert harness.wigner_slot_mismatches(a, b, g) == 0


def test_select_bit(harness):
    """The second secular pass finds its root as the r-th set bit of a mask."""
    harness.select_bit_mismatches.argtypes = [ctypes.c_int]
    harness.select_bit_mismatches.restype = ctypes.c_int
    assert harness.select_bit_mismatches(20000) == 0
