"""The register layouts of the tridiagonalisation kernels, emulated on the CPU with the
kernels' index arithmetic (the GPU tests check the kernels themselves against the oracle):

* k_tridiag: thread (r, c) of a 16 x 16 grid holds A(r + 16a, c + 16b), b <= a, 16 x 16
  blocks (78 slots);
* k_tridiag_w: thread (r, c) of a 16 x 32 grid holds A(r + 16a, c + 32b), a >= 2b, blocks of
  16 rows and 32 columns (72 slots).

In both, the blocks that make up a diagonal square (a == b; a = 2b, 2b + 1) hold the whole
square, above the diagonal too. Checked: every entry of the symmetric matrix is represented
exactly once by the rule "a diagonal-square slot stands for its own entry, any other slot for
its entry and the transpose"; the product formed from row sums of every slot and column sums
of the slots below the squares is A v, and the per-thread parts of v'A v add up to v'A v;
the unweighted rank-2 update keeps representing A - v w' - w v'; a column copied from the
slots of one column block is the matrix column. Taking column sums of the diagonal squares
as well (a planted mistake) must fail."""
import numpy as np
import pytest

LAYOUTS = {
    # name: (row block height, column block width, column blocks, row blocks, diagonal-square test)
    'k_tridiag': (16, 16, 12, 12, lambda a, b: a == b),
    'k_tridiag_w': (16, 32, 8, 16, lambda a, b: a // 2 == b),
}


def _slots(name):
    bh, bw, ncb, nrb, _ = LAYOUTS[name]
    for r in range(bh):
        for c in range(bw):
            for b in range(ncb):
                for a in range(nrb):
                    if bh * a + bh - 1 >= bw * b:  # the block reaches the lower triangle
                        yield r, c, a, b


def _n(name):
    bh, _, _, nrb, _ = LAYOUTS[name]
    return bh * nrb


def _ij(name, r, c, a, b):
    bh, bw = LAYOUTS[name][:2]
    return r + bh * a, c + bw * b


def _load(name, A):
    return {key: A[_ij(name, *key)] for key in _slots(name)}


def _matrix(name, R):
    """The symmetric matrix the slots represent."""
    n = _n(name)
    diag = LAYOUTS[name][4]
    M = np.zeros((n, n))
    for key, x in R.items():
        i, j = _ij(name, *key)
        M[i, j] += x
        if not diag(key[2], key[3]):
            M[j, i] += x
    return M


def _product(name, R, v, colsum_diag=False):
    n = _n(name)
    diag = LAYOUTS[name][4]
    y = np.zeros(n)
    vav = 0.0
    for key, x in R.items():
        i, j = _ij(name, *key)
        y[i] += x * v[j]  # row sum of the slot
        vav += x * v[i] * v[j]
        if colsum_diag or not diag(key[2], key[3]):
            y[j] += x * v[i]  # column sum of the slot
            vav += x * v[i] * v[j]
    return y, vav


def _update(name, R, v, w):
    for key, x in R.items():
        i, j = _ij(name, *key)
        R[key] = x - (v[i] * w[j] + w[i] * v[j])


def _column(name, R, jj):
    """The owners of column jj (c == jj mod column-block width) copy their slots of the
    column block of jj to srow[r + bh a]."""
    bh, bw = LAYOUTS[name][:2]
    srow = np.full(_n(name), np.nan)
    for (r, c, a, b), x in R.items():
        if c == jj % bw and b == jj // bw:
            srow[r + bh * a] = x
    return srow


def _sym(name, seed):
    rng = np.random.default_rng(seed)
    n = _n(name)
    A = rng.standard_normal((n, n))
    return A + A.T, rng


@pytest.mark.parametrize('name', list(LAYOUTS))
def test_slots_represent_the_matrix_once(name):
    A, _ = _sym(name, 1)
    R = _load(name, A)
    bh, bw = LAYOUTS[name][:2]
    assert len(R) == bh * bw * {'k_tridiag': 78, 'k_tridiag_w': 72}[name]
    np.testing.assert_allclose(_matrix(name, R), A, rtol=0, atol=1e-12)


@pytest.mark.parametrize('name', list(LAYOUTS))
def test_product_and_quadratic_form(name):
    A, rng = _sym(name, 2)
    v = rng.standard_normal(len(A))
    R = _load(name, A)
    y, vav = _product(name, R, v)
    np.testing.assert_allclose(y, A @ v, rtol=1e-12, atol=1e-10)
    assert abs(vav - v @ A @ v) <= 1e-10 * abs(v) @ abs(A) @ abs(v)
    # planted: column sums of the diagonal squares too count their entries twice
    yb, _ = _product(name, R, v, colsum_diag=True)
    assert np.abs(yb - A @ v).max() > 1e-3 * np.abs(A @ v).max()


@pytest.mark.parametrize('name', list(LAYOUTS))
def test_rank2_update_keeps_the_representation(name):
    A, rng = _sym(name, 3)
    v, w = rng.standard_normal(len(A)), rng.standard_normal(len(A))
    R = _load(name, A)
    _update(name, R, v, w)
    np.testing.assert_allclose(_matrix(name, R), A - np.outer(v, w) - np.outer(w, v), rtol=0, atol=1e-11)


@pytest.mark.parametrize('name', list(LAYOUTS))
def test_column_copy(name):
    A, _ = _sym(name, 4)
    R = _load(name, A)
    for jj in (0, 5, 31, 32, 100, len(A) - 1):
        srow = _column(name, R, jj)
        lo = (jj // LAYOUTS[name][1]) * LAYOUTS[name][1]  # rows from the column block's square down
        np.testing.assert_array_equal(srow[lo:], A[lo:, jj])
