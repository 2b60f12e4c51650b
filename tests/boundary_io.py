"""Serialisers for the native boundary tests: problem files for the sanitizer
harnesses (tests/native/problem_io.hpp) and MATLAB values for the mex.h
stand-in's driver (tests/native/mex/mex_shim.cpp)."""
import os
import struct
import subprocess

import numpy as np
import scipy.sparse as sps

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, 'tests', 'native')
BUILD = os.path.join(NATIVE, '_build')


def build_native(target='all'):
    subprocess.run(['make', '-s', '-C', NATIVE, target], check=True)


def _u64(f, v):
    f.write(struct.pack('<Q', int(v)))


def write_problem(path, I, qlist, grid, rho0, coil, dt, nsteps):
    dim = I.shape[0]
    with open(path, 'wb') as f:
        _u64(f, dim)
        for m in [I] + list(qlist):
            if m is None or m.nnz == 0:
                _u64(f, 0)
                continue
            c = sps.csc_matrix(m, dtype=np.complex128)
            c.sum_duplicates()
            c.sort_indices()
            _u64(f, dim)
            _u64(f, c.nnz)
            f.write(np.asarray(c.indptr, '<u8').tobytes())
            f.write(np.asarray(c.indices, '<u8').tobytes())
            f.write(np.ascontiguousarray(c.data, np.complex128).view('<f8').tobytes())
        n_or = len(grid['weights'])
        _u64(f, n_or)
        for k in ('alphas', 'betas', 'gammas', 'weights'):
            f.write(np.asarray(grid[k], '<f8').tobytes())
        f.write(np.asarray(rho0, np.complex128).view('<f8').tobytes())
        f.write(np.asarray(coil, np.complex128).view('<f8').tobytes())
        f.write(struct.pack('<d', float(dt)))
        _u64(f, nsteps)


# ---- MATLAB values: ('dense', a) ('sparse', m) ('cell', [..], (m, n)) ('struct', {..}) ('char', s) ('logical', b)

def mx_dense(a):
    return ('dense', np.atleast_2d(np.asarray(a)) if np.ndim(a) != 1 else np.asarray(a).reshape(-1, 1))


def mx_sparse(m):
    return ('sparse', sps.csc_matrix(m))


def mx_cell(items, shape):
    return ('cell', list(items), shape)


def mx_struct(fields):
    return ('struct', dict(fields))


def mx_char(s):
    return ('char', s)


def mx_logical(b):
    return ('logical', bool(b))


def _write_mx(f, v):
    kind = v[0]
    if kind == 'dense':
        a = v[1]
        cplx = np.iscomplexobj(a)
        m, n = a.shape
        f.write(struct.pack('<BQQB', 0, m, n, 1 if cplx else 0))
        data = np.asarray(a, np.complex128 if cplx else np.float64).reshape(-1, order='F')
        f.write(np.ascontiguousarray(data).view('<f8').tobytes() if cplx else data.astype('<f8').tobytes())
    elif kind == 'sparse':
        c = v[1].copy()
        c.sum_duplicates()
        c.sort_indices()
        cplx = np.iscomplexobj(c.data)
        m, n = c.shape
        f.write(struct.pack('<BQQB', 1, m, n, 1 if cplx else 0))
        _u64(f, c.nnz)
        f.write(np.asarray(c.indptr, '<u8').tobytes())
        f.write(np.asarray(c.indices, '<u8').tobytes())
        d = np.ascontiguousarray(c.data, np.complex128 if cplx else np.float64)
        f.write(d.view('<f8').tobytes() if cplx else d.astype('<f8').tobytes())
    elif kind == 'cell':
        items, (m, n) = v[1], v[2]
        f.write(struct.pack('<BQQB', 2, m, n, 0))
        for it in items:  # column-major order
            _write_mx(f, it)
    elif kind == 'struct':
        f.write(struct.pack('<BQQB', 3, 1, 1, 0))
        _u64(f, len(v[1]))
        for name, child in v[1].items():
            b = name.encode()
            _u64(f, len(b))
            f.write(b)
            _write_mx(f, child)
    elif kind == 'char':
        b = v[1].encode()
        f.write(struct.pack('<BQQB', 4, 1, len(b), 0))
        f.write(b)
    elif kind == 'logical':
        f.write(struct.pack('<BQQB', 5, 1, 1, 0))
        f.write(bytes([1 if v[1] else 0]))
    else:
        raise ValueError(kind)


def write_case(path, rhs):
    with open(path, 'wb') as f:
        _u64(f, len(rhs))
        for v in rhs:
            _write_mx(f, v)


def read_outputs(path):
    out = []
    with open(path, 'rb') as f:
        count = struct.unpack('<Q', f.read(8))[0]
        for _ in range(count):
            kind, m, n, cplx = struct.unpack('<BQQB', f.read(18))
            assert kind == 0
            k = m * n * (2 if cplx else 1)
            d = np.frombuffer(f.read(8 * k), '<f8')
            a = d.view(np.complex128) if cplx else d
            out.append(a.reshape((m, n), order='F'))
    return out


def powder_hip_rhs(I, qnested, grid, rho0, coil, dt, nsteps, tols=None, opts=None, sparse_states=True):
    """The right-hand sides of powder_hip(I,Q,grid,rho0,coil,dt,nsteps,tols,opts)
    as powder.m would pass them: Q as {3x3 cell, 5x5 cell}, cells column-major."""
    def cell_of(block, r):
        items = []
        for m in range(2 * r + 1):      # column-major: row index fastest
            for k in range(2 * r + 1):
                q = block[k][m]
                items.append(mx_sparse(q) if q is not None and q.nnz else mx_dense(np.zeros((0, 0))))
        return mx_cell(items, (2 * r + 1, 2 * r + 1))
    qcell = mx_cell([cell_of(qnested[0], 1), cell_of(qnested[1], 2)], (1, 2))
    g = mx_struct({k: mx_dense(np.asarray(grid[k], float).reshape(-1, 1)) for k in
                   ('alphas', 'betas', 'gammas', 'weights')})
    st = (lambda v: mx_sparse(sps.csc_matrix(np.asarray(v).reshape(-1, 1)))) if sparse_states else \
        (lambda v: mx_dense(np.asarray(v).reshape(-1, 1)))
    t = mx_struct({k: mx_dense(np.array([[float(v)]])) for k, v in (tols or {}).items()})
    o = mx_struct(opts or {})
    return [mx_sparse(I), qcell, g, st(rho0), st(coil), mx_dense(np.array([[float(dt)]])),
            mx_dense(np.array([[float(nsteps)]])), t, o]


def run_driver(exe, rhs, nlhs, tmp, env=None):
    case, out = os.path.join(tmp, 'case.bin'), os.path.join(tmp, 'out.bin')
    write_case(case, rhs)
    if os.path.exists(out):
        os.remove(out)
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe, case, out, str(nlhs)], capture_output=True, text=True, env=e, timeout=600)
    return r, (read_outputs(out) if r.stdout.startswith('OK') else None)
