"""The planner's real-generator test (sp_generator_is_real, host only).

A system takes the real-product path when H(Omega) = I + sum D^r(k,m) Q{r}{k,m}
is real at every orientation. The library decides it from the slot symmetry
Q{r}{k,m} = (-1)^(k-m) conj(Q{r}{-k,-m}); these tests check that decision
against brute force: the imaginary part of the assembled H at random
orientations (pyoracle.assemble_dense, built from scipy's expm of the angular
momentum matrices, not from the library's Wigner formulas)."""
import numpy as np
import pytest
import scipy.sparse as sps

import pyoracle as O
from spinach_amd import _native, workloads


def _max_imag(w, n=4, seed=0):
    rng = np.random.default_rng(seed)
    worst, scale = 0.0, 0.0
    for _ in range(n):
        e = (rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi))
        h = O.assemble_dense(w['I'], w['Q'], e)
        worst = max(worst, np.abs(h.imag).max())
        scale = max(scale, np.abs(h).max())
    return worst / scale


def _systems():
    rng = np.random.default_rng(11)
    out = [('2spin', workloads.static_2spin(npoints=8, n_orient=4)),
           ('8spin', workloads.static_8spin(nsteps=8, n_orient=4))]
    for n in (3, 5):
        out.append((f'random{n}', workloads.random_system(rng, n)))
    return out


@pytest.mark.parametrize('name,w', _systems(), ids=lambda x: x if isinstance(x, str) else '')
def test_secular_nmr_generators_are_real(name, w):
    assert _max_imag(w) < 1e-13
    assert _native.generator_is_real(w['I'], w['Q'])


def test_complex_hermitian_generator_is_complex():
    # I + i*eps*K with K real antisymmetric stays Hermitian but is complex at every orientation
    w = workloads.static_2spin(npoints=8, n_orient=4)
    d = w['I'].shape[0]
    rng = np.random.default_rng(3)
    K = np.triu(rng.standard_normal((d, d)), 1)
    K = K - K.T
    I = sps.csc_matrix(w['I'].toarray() + 1e3j * K)
    w2 = dict(w, I=I)
    assert _max_imag(w2) > 1e-6
    assert not _native.generator_is_real(I, w['Q'])


# The library's test is a sufficient condition on the unsymmetrised parts: the
# cases below are rejected even where (H+H')/2 happens to stay real, which only
# sends them down the general complex path.
def test_phase_rotated_slot_is_rejected():
    w = workloads.static_2spin(npoints=8, n_orient=4)
    q = list(w['Q'])
    t = next(t for t in range(9, 34) if q[t] is not None and t != 21)  # a rank-2 slot with k or m != 0
    q[t] = q[t] * np.exp(0.3j)
    assert not _native.generator_is_real(w['I'], q)


def test_imaginary_isotropic_diagonal_is_rejected():
    w = workloads.static_2spin(npoints=8, n_orient=4)
    d = w['I'].shape[0]
    I = w['I'].astype(complex) + 1e-3j * sps.eye(d, format='csc')
    assert not _native.generator_is_real(I, w['Q'])


def test_self_partner_slot_must_be_real():
    # slot (k,m) = (0,0) of rank 2 is its own partner: it must be real
    w = workloads.static_2spin(npoints=8, n_orient=4)
    q = list(w['Q'])
    d = w['I'].shape[0]
    q[21] = (q[21] if q[21] is not None else sps.csc_matrix((d, d))) + 1e-2j * sps.eye(d, format='csc')
    assert not _native.generator_is_real(w['I'], q)


def test_empty_partner_is_rejected():
    w = workloads.static_2spin(npoints=8, n_orient=4)
    q = list(w['Q'])
    t = next(t for t in range(9, 34) if q[t] is not None and t != 21)
    partner = 9 + (4 - (t - 9) // 5) * 5 + (4 - (t - 9) % 5)
    assert q[partner] is not None
    q[partner] = None
    assert not _native.generator_is_real(w['I'], q)


def test_invalid_input_reports_an_error():
    w = workloads.static_2spin(npoints=8, n_orient=4)
    q = list(w['Q'])
    q[3] = sps.csc_matrix(([1.0], ([0], [0])), shape=(3, 3))
    with pytest.raises(_native.SpinachPowderError, match='dimension'):
        _native.generator_is_real(w['I'], q)
