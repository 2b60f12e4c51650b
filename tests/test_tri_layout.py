"""The register layout of the tridiagonalisation (k_tridiag), emulated on the CPU:
thread (r, c) of a 16 x 16 grid holds A(r + 16a, c + 16b) for b <= a, the diagonal
blocks weighted by h = 1 (r > c), 1/2 (r == c) or 0 (r < c). Checked here: every
entry of the symmetric matrix is represented exactly once, the product formed from
per-slot row sums (A v at the slot's column) and column sums (A v at its row) is A v,
the per-thread parts of v'A v add up to v'A v, and the h-weighted rank-2 update keeps
representing A - v w' - w v'. (The same index arithmetic as the kernel; the GPU tests
check the kernel itself against the oracle.)"""
import numpy as np

B, NB = 16, 12
N = B * NB


def _h(r, c):
    return 1.0 if r > c else (0.5 if r == c else 0.0)


def _slots():
    for r in range(B):
        for c in range(B):
            for a in range(NB):
                for b in range(a + 1):
                    yield r, c, a, b


def _load(A):
    R = {}
    for r, c, a, b in _slots():
        x = A[r + B * a, c + B * b]
        R[(r, c, a, b)] = _h(r, c) * x if a == b else x
    return R


def _matrix(R):
    """The symmetric matrix the slots represent."""
    M = np.zeros((N, N))
    for (r, c, a, b), x in R.items():
        i, j = r + B * a, c + B * b
        M[i, j] += x
        M[j, i] += x
    return M


def _product(R, v):
    y = np.zeros(N)
    vav = 0.0
    for (r, c, a, b), x in R.items():
        i, j = r + B * a, c + B * b
        y[i] += x * v[j]  # row sum of the slot
        y[j] += x * v[i]  # column sum of the slot
        vav += 2.0 * x * v[i] * v[j]
    return y, vav


def _update(R, v, w):
    for key, x in R.items():
        r, c, a, b = key
        i, j = r + B * a, c + B * b
        h = _h(r, c) if a == b else 1.0
        R[key] = x - h * (v[i] * w[j] + w[i] * v[j])


def test_slots_represent_the_matrix_once():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((N, N))
    A = A + A.T
    R = _load(A)
    assert len(R) == B * B * NB * (NB + 1) // 2  # 78 slots per thread
    np.testing.assert_allclose(_matrix(R), A, rtol=0, atol=1e-12)


def test_product_and_quadratic_form():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((N, N))
    A = A + A.T
    v = rng.standard_normal(N)
    y, vav = _product(_load(A), v)
    np.testing.assert_allclose(y, A @ v, rtol=1e-12, atol=1e-10)
    assert abs(vav - v @ A @ v) <= 1e-10 * abs(v) @ abs(A) @ abs(v)


def test_rank2_update_keeps_the_representation():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((N, N))
    A = A + A.T
    v, w = rng.standard_normal(N), rng.standard_normal(N)
    R = _load(A)
    _update(R, v, w)
    np.testing.assert_allclose(_matrix(R), A - np.outer(v, w) - np.outer(w, v), rtol=0, atol=1e-11)
