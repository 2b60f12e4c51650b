"""How close an FP32 path can come to the FP64 FID on the 8-spin workload.

The most favourable FP32 propagation keeps only the storage in FP32: the
exact propagator P = expm(-i dt H) of one orientation, rounded once to
complex64, applied step by step to a complex64 state (the reference's own
loop, evolution.m:562-566). Its error against the exact FID is the floor of
every FP32-storage method; the test measures it on two orientations of the
8-spin system and checks that it lies far above 1e-6, and that it is the
eigenphase drift it is said to be (it grows with t)."""
import numpy as np
import scipy.linalg as sl

import pyoracle as O
from spinach_amd import workloads


def _orientation(w, st, o):
    g = w['grid']
    Ir = w['I'][st][:, st]
    qr = [None if q is None else q[st][:, st] for q in w['Q']]
    return O.assemble_dense(Ir, qr, (g['alphas'][o], g['betas'][o], g['gammas'][o]))


def test_fp32_storage_floor_on_the_8spin_workload():
    w = workloads.static_8spin(n_orient=8)
    st = np.array(O.union_states(w['I'], w['Q'], w['coil']))
    c, r = w['coil'][st], w['rho0'][st]
    nst = w['nsteps']
    for o in (0, 1):
        H = _orientation(w, st, o)
        assert np.abs(H.imag).max() <= 1e-15 * np.abs(H).max()   # real to rounding
        lam, V = np.linalg.eigh(w['dt'] * H.real)
        gk = (V.T @ c.conj()) * (V.T @ r)
        t = np.arange(nst + 1)
        exact = np.exp(-1j * np.outer(t, lam)) @ gk
        P = sl.expm(-1j * w['dt'] * H).astype(np.complex64)
        x = r.astype(np.complex64)
        cc = c.conj().astype(np.complex64)
        f32 = np.empty(nst + 1, complex)
        for k in range(nst + 1):
            f32[k] = cc @ x
            x = P @ x
        e = O.rel_err(f32, exact)
        err = np.abs(f32 - exact)
        early, late = err[:512].mean(), err[-512:].mean()
        print(f'orientation {o}: FP32 storage floor {e:.2e} (mean |error| first/last 512 steps {early:.1e}/{late:.1e})')
        assert 5e-6 < e < 1e-4
        assert late > 3 * early
