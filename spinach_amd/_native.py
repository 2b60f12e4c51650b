"""ctypes binding of libspinach_powder.so (include/spinach_powder.h).

The library is the only compute path: if it is not built, or no GPU is
visible, every entry point raises. There is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os

import numpy as np
import scipy.sparse as sps

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libspinach_powder.so')

SP_MODE_FP64 = 0
SP_MODE_FP32 = 1
SP_MODE_COMPLEX = 16  # flag: keep complex arithmetic for a real generator
SP_MODE_PROPAGATOR = 32  # flag: propagator power scheme for every system (no spectral path)
SP_MODE_GRAPH = 64  # flag: record the pass once as a HIP graph and replay it
SP_MODE_KRYLOV = 128  # flag: time stepping (step() per time step) for every system
SP_NQ = 34

EXPORTS = ('sp_create', 'sp_destroy', 'sp_last_error', 'sp_powder_fid', 'sp_plan_powder',
           'sp_execute', 'sp_plan_result', 'sp_plan_destroy', 'sp_plan_powder_batch',
           'sp_plan_result_batch', 'sp_plan_mas', 'sp_mas_fid', 'sp_generator_is_real', 'sp_plan_set_dead_time',
           'sp_powder_fid_noalloc', 'sp_plan_autotune', 'sp_plan_get_tuning', 'sp_plan_set_tuning',
           'sp_plan_powder_relax', 'sp_plan_set_initial_states', 'sp_plan_propagate')
SP_MAS_FRAME_MAGNET = 0
SP_MAS_FRAME_ROTOR = 1


class SpCsc(C.Structure):
    _fields_ = [('n', C.c_int64), ('jc', C.POINTER(C.c_uint64)), ('ir', C.POINTER(C.c_uint64)),
                ('pr', C.POINTER(C.c_double)), ('is_complex', C.c_int32)]


class SpRelax(C.Structure):
    _fields_ = [('R0', C.POINTER(SpCsc)), ('diag', C.POINTER(C.c_double))]


class SpTols(C.Structure):
    _fields_ = [('liouv_zero', C.c_double), ('zte_tol', C.c_double), ('prop_chop', C.c_double),
                ('small_matrix', C.c_double), ('krylov_tol', C.c_double)]


class SpStats(C.Structure):
    _fields_ = [('elapsed_s', C.c_double), ('orient_steps_per_s', C.c_double), ('kernel_ms', C.c_double),
                ('full_dim', C.c_int64), ('reduced_dim', C.c_int64), ('reduced_nnz', C.c_int64),
                ('padded_dim', C.c_int32), ('block', C.c_int32), ('taylor_terms', C.c_int32),
                ('max_squarings', C.c_int32), ('ps_block', C.c_int32), ('sum_squarings_lo', C.c_int32),
                ('real_systems', C.c_int32), ('spectral_systems', C.c_int32), ('eig_max_sweeps', C.c_int32),
                ('krylov_systems', C.c_int32), ('stage_ms', C.c_double * 4), ('step_products', C.c_int64),
                ('plan_ms', C.c_double), ('eig_fallbacks', C.c_int32), ('mas_tiles', C.c_int32),
                ('mas_pairs', C.c_int32), ('mas_detected', C.c_int32)]

    def as_dict(self):
        out = {k: getattr(self, k) for k, _ in self._fields_}
        out['stage_ms'] = list(self.stage_ms)
        return out


class SpTuning(C.Structure):
    _fields_ = [('log2_block', C.c_int32), ('prop_wg_per_cu', C.c_int32), ('tri_wg_per_cu', C.c_int32),
                ('syn_wg_per_cu', C.c_int32), ('step_wg_per_cu', C.c_int32), ('candidates', C.c_int32),
                ('default_ms', C.c_double), ('best_ms', C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SpRotor(C.Structure):
    _fields_ = [('axis', C.c_double * 3), ('rate_hz', C.c_double), ('ticks', C.c_int32), ('frame', C.c_int32)]


class SpSystem(C.Structure):
    _fields_ = [('I', C.POINTER(SpCsc)), ('Q', C.POINTER(SpCsc)), ('alphas', C.POINTER(C.c_double)),
                ('betas', C.POINTER(C.c_double)), ('gammas', C.POINTER(C.c_double)),
                ('weights', C.POINTER(C.c_double)), ('n_or', C.c_int64), ('rho0', C.POINTER(C.c_double)),
                ('coil', C.POINTER(C.c_double)), ('fid_out', C.POINTER(C.c_double))]


_lib = None
_ctx = {}

_dp = C.POINTER(C.c_double)


class SpHamTerm(C.Structure):
    _fields_ = [('nspin', C.c_int32), ('spin', C.c_int32 * 2), ('ist', C.c_int32 * 2), ('pad', C.c_int32),
                ('coef', C.c_double), ('iso', C.c_double * 2), ('ist_coef', C.c_double * 16),
                ('irr', C.c_double * 16)]


_OPNAME = {'L+': (1, -2 ** 0.5), 'Lz': (2, 1.0), 'L-': (3, 2 ** 0.5)}


def build_hamiltonian(nspins, basis, descr, tol):
    """sp_ham_build: I and Q (nested [rank1 3x3, rank2 5x5], None where empty) of the
    Liouvillian commutation superoperator from the Hamiltonian's descriptors
    (ops, spins, iso, ist[8], irr[8]), natively. basis: (dim, nspins) IST indices, rows
    ascending."""
    import numpy as np
    import scipy.sparse as sps
    lib = load()
    if not getattr(lib, '_ham_ready', False):
        lib.sp_ham_build.restype = C.c_void_p
        lib.sp_ham_build.argtypes = [C.c_int32, C.POINTER(C.c_int8), C.c_int64, C.POINTER(SpHamTerm), C.c_int32,
                                     C.c_double]
        lib.sp_ham_matrix.restype = C.c_int
        lib.sp_ham_matrix.argtypes = [C.c_void_p, C.c_int32, C.POINTER(SpCsc)]
        lib.sp_ham_free.restype = None
        lib.sp_ham_free.argtypes = [C.c_void_p]
        lib._ham_ready = True
    b = np.ascontiguousarray(basis, dtype=np.int8)
    dim = b.shape[0]
    terms = (SpHamTerm * max(1, len(descr)))()
    for i, (ops, spins, iso, ist, irr) in enumerate(descr):
        t = terms[i]
        t.nspin = len(ops)
        coef = 1.0
        for a, (name, sp) in enumerate(zip(ops, spins)):
            idx, c = _OPNAME[name]
            t.spin[a] = int(sp)
            t.ist[a] = idx
            coef *= c
        t.coef = coef
        t.iso[0], t.iso[1] = complex(iso).real, complex(iso).imag
        for k in range(8):
            t.ist_coef[2 * k], t.ist_coef[2 * k + 1] = complex(ist[k]).real, complex(ist[k]).imag
            t.irr[2 * k], t.irr[2 * k + 1] = complex(irr[k]).real, complex(irr[k]).imag
    h = lib.sp_ham_build(int(nspins), b.ctypes.data_as(C.POINTER(C.c_int8)), int(dim), terms, len(descr),
                         float(tol))
    if not h:
        raise SpinachPowderError(-1, 'sp_ham_build: ' + lib.sp_last_error(None).decode())
    try:
        mats = []
        for w in range(35):
            m = SpCsc()
            if lib.sp_ham_matrix(h, w, C.byref(m)):
                raise SpinachPowderError(-1, 'sp_ham_matrix failed')
            nnz = int(m.jc[dim]) if dim > 0 else 0
            jc = np.ctypeslib.as_array(m.jc, shape=(dim + 1,)).astype(np.int64)
            if nnz:
                ir = np.ctypeslib.as_array(m.ir, shape=(nnz,)).astype(np.int64)
                pr = np.ctypeslib.as_array(m.pr, shape=(2 * nnz,)).copy().view(np.complex128)
                mats.append(sps.csc_matrix((pr, ir, jc), shape=(dim, dim)).tocsr())
            else:
                mats.append(None)
    finally:
        lib.sp_ham_free(h)
    I = mats[0] if mats[0] is not None else sps.csr_matrix((dim, dim), dtype=complex)
    Q = [[mats[1 + 3 * k + m] for m in range(3)] for k in range(3)], \
        [[mats[10 + 5 * k + m] for m in range(5)] for k in range(5)]
    return I, [Q[0], Q[1]]


def load():
    """Load the library and declare its signatures (no GPU call)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f'{LIB_PATH} is not built: run __graft_entry__.build() first')
    lib = C.CDLL(LIB_PATH)
    lib.sp_create.restype = C.c_void_p
    lib.sp_create.argtypes = [C.c_int]
    lib.sp_destroy.restype = None
    lib.sp_destroy.argtypes = [C.c_void_p]
    lib.sp_last_error.restype = C.c_char_p
    lib.sp_last_error.argtypes = [C.c_void_p]
    lib.sp_powder_fid.restype = C.c_int
    lib.sp_powder_fid.argtypes = [C.c_void_p, C.POINTER(SpCsc), C.POINTER(SpCsc), _dp, _dp, _dp, _dp, C.c_int64,
                                  _dp, _dp, C.c_double, C.c_int64, C.POINTER(SpTols), C.c_int, _dp, _dp,
                                  C.POINTER(SpStats)]
    lib.sp_plan_powder.restype = C.c_void_p
    lib.sp_plan_powder.argtypes = [C.c_void_p, C.POINTER(SpCsc), C.POINTER(SpCsc), _dp, _dp, _dp, _dp, C.c_int64,
                                   _dp, _dp, C.c_double, C.c_int64, C.POINTER(SpTols), C.c_int]
    lib.sp_plan_powder_relax.restype = C.c_void_p
    lib.sp_plan_powder_relax.argtypes = lib.sp_plan_powder.argtypes + [C.POINTER(SpRelax)]
    lib.sp_execute.restype = C.c_int
    lib.sp_execute.argtypes = [C.c_void_p, C.POINTER(SpStats)]
    lib.sp_plan_result.restype = C.c_int
    lib.sp_plan_result.argtypes = [C.c_void_p, _dp, _dp]
    lib.sp_plan_destroy.restype = None
    lib.sp_plan_destroy.argtypes = [C.c_void_p]
    lib.sp_plan_powder_batch.restype = C.c_void_p
    lib.sp_plan_powder_batch.argtypes = [C.c_void_p, C.POINTER(SpSystem), C.c_int64, C.c_double, C.c_int64,
                                         C.POINTER(SpTols), C.c_int]
    lib.sp_plan_result_batch.restype = C.c_int
    lib.sp_plan_result_batch.argtypes = [C.c_void_p]
    lib.sp_plan_mas.restype = C.c_void_p
    lib.sp_plan_mas.argtypes = [C.c_void_p, C.POINTER(SpCsc), C.POINTER(SpCsc), _dp, _dp, _dp, _dp, C.c_int64,
                                _dp, _dp, C.POINTER(SpRotor), C.c_int64, C.c_int]
    lib.sp_mas_fid.restype = C.c_int
    lib.sp_mas_fid.argtypes = [C.c_void_p, C.POINTER(SpCsc), C.POINTER(SpCsc), _dp, _dp, _dp, _dp, C.c_int64,
                               _dp, _dp, C.POINTER(SpRotor), C.c_int64, C.c_int, _dp, _dp, C.POINTER(SpStats)]
    lib.sp_plan_autotune.restype = C.c_int
    lib.sp_plan_autotune.argtypes = [C.c_void_p, C.c_int32, C.POINTER(SpTuning)]
    lib.sp_plan_get_tuning.restype = C.c_int
    lib.sp_plan_get_tuning.argtypes = [C.c_void_p, C.POINTER(SpTuning)]
    lib.sp_plan_set_tuning.restype = C.c_int
    lib.sp_plan_set_tuning.argtypes = [C.c_void_p, C.POINTER(SpTuning)]
    lib.sp_powder_fid_noalloc.restype = C.c_int
    lib.sp_powder_fid_noalloc.argtypes = lib.sp_powder_fid.argtypes
    lib.sp_plan_set_dead_time.restype = C.c_int
    lib.sp_plan_set_dead_time.argtypes = [C.c_void_p, C.c_double]
    lib.sp_plan_set_initial_states.restype = C.c_int
    lib.sp_plan_set_initial_states.argtypes = [C.c_void_p, _dp]
    lib.sp_plan_propagate.restype = C.c_int
    lib.sp_plan_propagate.argtypes = [C.c_void_p, C.c_double, C.c_int64, _dp]
    lib.sp_generator_is_real.restype = C.c_int
    lib.sp_generator_is_real.argtypes = [C.POINTER(SpCsc), C.POINTER(SpCsc), C.POINTER(C.c_int32)]
    _lib = lib
    return lib


class SpinachPowderError(RuntimeError):
    pass


def context(device=0):
    """Process-wide context for a device (created on first use)."""
    lib = load()
    if device not in _ctx:
        c = lib.sp_create(device)
        if not c:
            raise SpinachPowderError('sp_create failed: ' + lib.sp_last_error(None).decode())
        _ctx[device] = c
    return _ctx[device]


@atexit.register
def _release():
    if _lib is not None:
        for c in _ctx.values():
            _lib.sp_destroy(c)
        _ctx.clear()


def _err(ctx, rc, what):
    msg = load().sp_last_error(ctx).decode()
    raise SpinachPowderError(f'{what} failed (code {rc}): {msg}')


# ------------------------------------------------------------ marshalling
class _Keep:
    """Holds numpy buffers alive for the duration of a library call."""

    def __init__(self):
        self.refs = []

    def arr(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.refs.append(a)
        return a

    def dptr(self, a, n=None):
        a = self.arr(np.asarray(a, dtype=np.float64).ravel(), np.float64)
        if n is not None and a.size != n:
            raise ValueError(f'expected {n} values, got {a.size}')
        return a.ctypes.data_as(_dp)

    def cvec(self, v, n):
        v = self.arr(np.asarray(v, dtype=np.complex128).ravel(), np.complex128)
        if v.size != n:
            raise ValueError(f'state vector has {v.size} entries, the operators have dimension {n}')
        return v.view(np.float64).ctypes.data_as(_dp)

    def grid(self, g):
        """(alphas, betas, gammas, weights) pointers and the orientation count."""
        n = len(g['weights'])
        return tuple(self.dptr(g[key], n) for key in ('alphas', 'betas', 'gammas', 'weights')) + (n,)

    def csc(self, m):
        s = SpCsc()
        if m is None:
            s.n = 0
            return s
        m = sps.csc_matrix(m, dtype=np.complex128)
        if m.shape[0] != m.shape[1]:
            raise ValueError(f'sparse operand must be square, got shape {m.shape}')
        m.sum_duplicates()
        jc = self.arr(m.indptr, np.uint64)
        ir = self.arr(m.indices, np.uint64)
        pr = self.arr(m.data, np.complex128)
        s.n = m.shape[1]  # the library checks it against I's dimension
        s.jc = jc.ctypes.data_as(C.POINTER(C.c_uint64))
        s.ir = ir.ctypes.data_as(C.POINTER(C.c_uint64))
        s.pr = pr.view(np.float64).ctypes.data_as(_dp)
        s.is_complex = 1
        return s

    def qarray(self, qlist):
        arr = (SpCsc * SP_NQ)()
        for t in range(SP_NQ):
            arr[t] = self.csc(qlist[t])
        self.refs.append(arr)
        return arr


def _tols(t):
    t = t or {}
    return SpTols(t.get('liouv_zero', 1e-10), t.get('zte_tol', 1e-24), t.get('prop_chop', 1e-10),
                  t.get('small_matrix', 200), t.get('krylov_tol', 10000))


def mode_id(mode):
    """'fp64' / 'fp32' (or SP_MODE_*), with optional '+complex' (SP_MODE_COMPLEX:
    complex arithmetic even for a real generator), '+propagator'
    (SP_MODE_PROPAGATOR: no spectral path), '+graph' (SP_MODE_GRAPH: the pass
    recorded once as a HIP graph and replayed) and '+krylov' (SP_MODE_KRYLOV: time
    stepping for every system, sys.enable={'krylov'}) flags."""
    flags = {'complex': SP_MODE_COMPLEX, 'propagator': SP_MODE_PROPAGATOR, 'graph': SP_MODE_GRAPH,
             'krylov': SP_MODE_KRYLOV}
    if isinstance(mode, str) and '+' in mode:
        base, *extra = mode.split('+')
        out = mode_id(base)
        for f in extra:
            if f not in flags:
                raise ValueError(f'unknown mode flag {f!r}')
            out |= flags[f]
        return out
    mask = SP_MODE_COMPLEX | SP_MODE_PROPAGATOR | SP_MODE_GRAPH | SP_MODE_KRYLOV
    if isinstance(mode, (int, np.integer)) and not isinstance(mode, bool) and int(mode) & mask:
        return mode_id(int(mode) & ~mask) | (int(mode) & mask)
    if mode in (SP_MODE_FP64, 'fp64', 'c128'):
        return SP_MODE_FP64
    if mode in (SP_MODE_FP32, 'fp32', 'c64'):
        return SP_MODE_FP32
    raise ValueError(f'unknown precision mode {mode!r}')


class Plan:
    """Device-resident powder problem (sp_plan_powder / sp_plan_powder_batch)."""

    def __init__(self, handle, ctx, npts, n_or, batch_outputs=None, keep=None):
        self.h = handle
        self.ctx = ctx
        self.npts = npts
        self.n_or = n_or
        self.batch_outputs = batch_outputs
        self._keep = keep
        self.stats = None

    def execute(self):
        st = SpStats()
        rc = load().sp_execute(self.h, C.byref(st))
        if rc:
            _err(self.ctx, rc, 'sp_execute')
        self.stats = st.as_dict()
        return self.stats

    def autotune(self, passes=2):
        """sp_plan_autotune: time candidate launch configurations, keep the
        fastest (cached per plan shape); returns the chosen tuning."""
        t = SpTuning()
        rc = load().sp_plan_autotune(self.h, int(passes), C.byref(t))
        if rc:
            _err(self.ctx, rc, 'sp_plan_autotune')
        return t.as_dict()

    def tuning(self):
        t = SpTuning()
        rc = load().sp_plan_get_tuning(self.h, C.byref(t))
        if rc:
            _err(self.ctx, rc, 'sp_plan_get_tuning')
        return t.as_dict()

    def set_tuning(self, tuning):
        t = SpTuning(**{k: v for k, v in tuning.items() if k in dict(SpTuning._fields_)})
        rc = load().sp_plan_set_tuning(self.h, C.byref(t))
        if rc:
            _err(self.ctx, rc, 'sp_plan_set_tuning')

    def set_initial_states(self, states):
        """sp_plan_set_initial_states: states is (n_or, dim) complex (full
        dimension), or None to restore rho0."""
        if states is None:
            rc = load().sp_plan_set_initial_states(self.h, None)
        else:
            a = np.ascontiguousarray(states, np.complex128)
            if a.ndim != 2 or a.shape[0] != self.n_or:
                raise ValueError(f'initial states must have shape ({self.n_or}, dim), got {a.shape}')
            self._init = a
            rc = load().sp_plan_set_initial_states(self.h, a.ctypes.data_as(_dp))
        if rc:
            _err(self.ctx, rc, 'sp_plan_set_initial_states')

    def propagate(self, dt, nsteps, dim):
        """sp_plan_propagate: every orientation's state after nsteps steps of dt
        under its own generator, as an (n_or, dim) complex array."""
        out = np.zeros((self.n_or, int(dim)), np.complex128)
        rc = load().sp_plan_propagate(self.h, float(dt), int(nsteps), out.ctypes.data_as(_dp))
        if rc:
            _err(self.ctx, rc, 'sp_plan_propagate')
        return out

    def set_dead_time(self, dead_time):
        """sp_plan_set_dead_time: step(L, rho0, dead_time) before the acquisition
        (experiments/acquire.m:65-67); 0 removes it."""
        rc = load().sp_plan_set_dead_time(self.h, float(dead_time))
        if rc:
            _err(self.ctx, rc, 'sp_plan_set_dead_time')

    def result(self, per_orientation=False):
        fid = np.zeros(self.npts, np.complex128)
        per = np.zeros((self.n_or, self.npts), np.complex128) if per_orientation else None
        rc = load().sp_plan_result(self.h, fid.view(np.float64).ctypes.data_as(_dp),
                                   per.view(np.float64).ctypes.data_as(_dp) if per is not None else None)
        if rc:
            _err(self.ctx, rc, 'sp_plan_result')
        return (fid, per) if per_orientation else fid

    def result_batch(self):
        rc = load().sp_plan_result_batch(self.h)
        if rc:
            _err(self.ctx, rc, 'sp_plan_result_batch')
        return [o.copy() for o in self.batch_outputs]

    def close(self):
        if self.h:
            load().sp_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generator_is_real(I, qlist):
    """sp_generator_is_real: host-only test whether H(Omega) = I + sum D Q is
    real at every orientation (the planner then uses real products)."""
    lib = load()
    k = _Keep()
    dim = I.shape[0]
    out = C.c_int32(-1)
    rc = lib.sp_generator_is_real(C.byref(k.csc(I)), k.qarray(qlist), C.byref(out))
    if rc:
        raise SpinachPowderError(f'sp_generator_is_real failed (code {rc}): ' + lib.sp_last_error(None).decode())
    return bool(out.value)


def plan_powder(I, qlist, grid, rho0, coil, dt, nsteps, tols=None, mode='fp64', device=0, relax=None):
    """relax: None, or a dict with 'R0' (sparse dim x dim, or None) and 'diag'
    (7 x dim array, or None): the relaxation superoperator of sp_plan_powder_relax."""
    lib = load()
    ctx = context(device)
    k = _Keep()
    dim = I.shape[0]
    Ic = k.csc(I)
    Qa = k.qarray(qlist)
    n_or = len(grid['weights'])
    t = _tols(tols)
    args = (ctx, C.byref(Ic), Qa, *k.grid(grid), k.cvec(rho0, dim), k.cvec(coil, dim), float(dt), int(nsteps),
            C.byref(t), mode_id(mode))
    if relax is None:
        h = lib.sp_plan_powder(*args)
    else:
        r = SpRelax()
        if relax.get('R0') is not None:
            r0 = k.csc(relax['R0'])
            k.refs.append(r0)
            r.R0 = C.pointer(r0)
        if relax.get('diag') is not None:
            g = np.asarray(relax['diag'], np.float64)
            if g.shape != (7, dim):
                raise ValueError(f'relaxation diagonals must have shape (7, {dim}), got {g.shape}')
            r.diag = k.dptr(g)
        h = lib.sp_plan_powder_relax(*args, C.byref(r))
    if not h:
        _err(ctx, -1, 'sp_plan_powder_relax' if relax is not None else 'sp_plan_powder')
    return Plan(h, ctx, int(nsteps) + 1, n_or)


def plan_powder_batch(systems, dt, nsteps, tols=None, mode='fp64', device=0):
    """systems: list of dicts with keys I, Q (34-slot list), grid, rho0, coil."""
    lib = load()
    ctx = context(device)
    k = _Keep()
    npts = int(nsteps) + 1
    arr = (SpSystem * len(systems))()
    outs = []
    for s, y in enumerate(systems):
        dim = y['I'].shape[0]
        Ic = k.csc(y['I'])
        k.refs.append(Ic)
        Qa = k.qarray(y['Q'])
        out = np.zeros(npts, np.complex128)
        outs.append(out)
        g = y['grid']
        arr[s].I = C.pointer(Ic)
        arr[s].Q = C.cast(Qa, C.POINTER(SpCsc))
        arr[s].alphas, arr[s].betas, arr[s].gammas, arr[s].weights, arr[s].n_or = k.grid(g)
        arr[s].rho0 = k.cvec(y['rho0'], dim)
        arr[s].coil = k.cvec(y['coil'], dim)
        arr[s].fid_out = out.view(np.float64).ctypes.data_as(_dp)
    t = _tols(tols)
    h = lib.sp_plan_powder_batch(ctx, arr, len(systems), float(dt), int(nsteps), C.byref(t), mode_id(mode))
    if not h:
        _err(ctx, -1, 'sp_plan_powder_batch')
    k.refs.append(arr)
    return Plan(h, ctx, npts, sum(len(y['grid']['weights']) for y in systems), batch_outputs=outs, keep=k)


def powder_fid(I, qlist, grid, rho0, coil, dt, nsteps, tols=None, mode='fp64', per_orientation=False, device=0):
    """One-shot sp_powder_fid call. Returns the FID (and the per-orientation
    FIDs when requested) plus the stats dict."""
    lib = load()
    ctx = context(device)
    k = _Keep()
    dim = I.shape[0]
    Ic = k.csc(I)
    Qa = k.qarray(qlist)
    n_or = len(grid['weights'])
    npts = int(nsteps) + 1
    fid = np.zeros(npts, np.complex128)
    per = np.zeros((n_or, npts), np.complex128) if per_orientation else None
    st = SpStats()
    t = _tols(tols)
    rc = lib.sp_powder_fid(ctx, C.byref(Ic), Qa, *k.grid(grid), k.cvec(rho0, dim), k.cvec(coil, dim),
                           float(dt), int(nsteps), C.byref(t), mode_id(mode),
                           fid.view(np.float64).ctypes.data_as(_dp),
                           per.view(np.float64).ctypes.data_as(_dp) if per is not None else None, C.byref(st))
    if rc:
        _err(ctx, rc, 'sp_powder_fid')
    return fid, per, st.as_dict()


def powder_fid_noalloc(I, qlist, grid, rho0, coil, dt, nsteps, tols=None, mode='fp64', per_orientation=False,
                       device=0):
    """sp_powder_fid_noalloc: sp_powder_fid on the spectral path with all device
    memory taken from the library's static scratch (no allocation in the call)."""
    lib = load()
    ctx = context(device)
    k = _Keep()
    dim = I.shape[0]
    Ic = k.csc(I)
    Qa = k.qarray(qlist)
    n_or = len(grid['weights'])
    npts = int(nsteps) + 1
    fid = np.zeros(npts, np.complex128)
    per = np.zeros((n_or, npts), np.complex128) if per_orientation else None
    st = SpStats()
    t = _tols(tols)
    rc = lib.sp_powder_fid_noalloc(ctx, C.byref(Ic), Qa, *k.grid(grid), k.cvec(rho0, dim), k.cvec(coil, dim),
                                   float(dt), int(nsteps), C.byref(t), mode_id(mode),
                                   fid.view(np.float64).ctypes.data_as(_dp),
                                   per.view(np.float64).ctypes.data_as(_dp) if per is not None else None,
                                   C.byref(st))
    if rc:
        _err(ctx, rc, 'sp_powder_fid_noalloc')
    return fid, per, st.as_dict()


def _rotor(axis, rate_hz, ticks, frame):
    r = SpRotor()
    for i in range(3):
        r.axis[i] = float(axis[i])
    r.rate_hz = float(rate_hz)
    r.ticks = int(ticks)
    r.frame = SP_MAS_FRAME_MAGNET if frame in ('magnet', SP_MAS_FRAME_MAGNET) else SP_MAS_FRAME_ROTOR
    return r


def _dense_colmajor(m, keep, dim):
    m = np.asarray(m.toarray() if sps.issparse(m) else m, dtype=np.complex128)
    if m.shape != (dim, dim):
        raise ValueError(f'state matrix has shape {m.shape}, the operators have dimension {dim}')
    return keep.cvec(np.asfortranarray(m).ravel(order='F'), dim * dim)


def plan_mas(H, qlist, grid, rho0, coil, axis, rate_hz, ticks, nsteps, frame='magnet', mode='fp32', device=0):
    """Magic-angle-spinning Hilbert-space plan (sp_plan_mas). H, Q are the
    zeeman-hilb I and Q; rho0, coil dense N x N."""
    lib = load()
    ctx = context(device)
    k = _Keep()
    dim = H.shape[0]
    Hc = k.csc(H)
    Qa = k.qarray(qlist)
    n_or = len(grid['weights'])
    rot = _rotor(axis, rate_hz, ticks, frame)
    h = lib.sp_plan_mas(ctx, C.byref(Hc), Qa, *k.grid(grid), _dense_colmajor(rho0, k, dim), _dense_colmajor(coil, k, dim),
                        C.byref(rot), int(nsteps), mode_id(mode))
    if not h:
        _err(ctx, -1, 'sp_plan_mas')
    return Plan(h, ctx, int(nsteps) + 1, n_or)


def mas_fid(H, qlist, grid, rho0, coil, axis, rate_hz, ticks, nsteps, frame='magnet', mode='fp32',
            per_orientation=False, device=0):
    """One-shot sp_mas_fid call."""
    lib = load()
    ctx = context(device)
    k = _Keep()
    dim = H.shape[0]
    Hc = k.csc(H)
    Qa = k.qarray(qlist)
    n_or = len(grid['weights'])
    npts = int(nsteps) + 1
    fid = np.zeros(npts, np.complex128)
    per = np.zeros((n_or, npts), np.complex128) if per_orientation else None
    st = SpStats()
    rot = _rotor(axis, rate_hz, ticks, frame)
    rc = lib.sp_mas_fid(ctx, C.byref(Hc), Qa, *k.grid(grid), _dense_colmajor(rho0, k, dim), _dense_colmajor(coil, k, dim),
                        C.byref(rot), int(nsteps), mode_id(mode), fid.view(np.float64).ctypes.data_as(_dp),
                        per.view(np.float64).ctypes.data_as(_dp) if per is not None else None, C.byref(st))
    if rc:
        _err(ctx, rc, 'sp_mas_fid')
    return fid, per, st.as_dict()


# ------------------------------------------------------------- surrogate
def _load_surrogate():
    lib = load()
    if getattr(lib, '_sg_ready', False):
        return lib
    lib.sp_surrogate_create.restype = C.c_void_p
    lib.sp_surrogate_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                        C.c_uint64]
    lib.sp_surrogate_destroy.restype = None
    lib.sp_surrogate_destroy.argtypes = [C.c_void_p]
    lib.sp_surrogate_last_error.restype = C.c_char_p
    lib.sp_surrogate_last_error.argtypes = [C.c_void_p]
    lib.sp_surrogate_num_params.restype = C.c_int64
    lib.sp_surrogate_num_params.argtypes = [C.c_void_p]
    _fp = C.POINTER(C.c_float)
    _ip = C.POINTER(C.c_int32)
    lib.sp_surrogate_get_params.restype = C.c_int
    lib.sp_surrogate_get_params.argtypes = [C.c_void_p, _fp]
    lib.sp_surrogate_set_params.restype = C.c_int
    lib.sp_surrogate_set_params.argtypes = [C.c_void_p, _fp]
    lib.sp_surrogate_grad.restype = C.c_int
    lib.sp_surrogate_grad.argtypes = [C.c_void_p, _dp, _dp, _dp, _ip, C.c_int64, C.c_double, _fp, _fp, _fp]
    lib.sp_surrogate_train.restype = C.c_int
    lib.sp_surrogate_train.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int64, _dp, C.c_int64, C.c_double, C.c_int64,
                                       C.c_int64, C.c_double, C.c_uint64, _fp]
    lib.sp_surrogate_powder.restype = C.c_int
    lib.sp_surrogate_powder.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_int64, C.c_int64, C.c_double, _dp, _dp,
                                        _dp]
    lib._sg_ready = True
    return lib


class SurrogateNet:
    """Device-resident Fourier-feature MLP (sp_surrogate_*)."""

    def __init__(self, t_max, width=256, depth=4, ka=3, kt=22, seed=0, device=0, t_feat=None):
        """t_max: record length (the linear time feature t/t_max); t_feat: time
        scale of the harmonics cos/sin(pi m t/t_feat), m = 1..kt (default
        t_max; kt*dt spans the band up to the Nyquist frequency)."""
        lib = _load_surrogate()
        context(device)
        t_feat = float(t_max if t_feat is None else t_feat)
        self.h = lib.sp_surrogate_create(device, width, depth, ka, kt, float(t_max), t_feat, int(seed))
        self.t_feat = t_feat
        if not self.h:
            raise SpinachPowderError('sp_surrogate_create failed')
        self.width, self.depth, self.ka, self.kt, self.t_max = width, depth, ka, kt, float(t_max)
        self.nparam = lib.sp_surrogate_num_params(self.h)

    def _chk(self, rc, what):
        if rc:
            msg = _load_surrogate().sp_surrogate_last_error(self.h).decode()
            raise SpinachPowderError(f'{what} failed (code {rc}): {msg}')

    def params(self):
        out = np.zeros(self.nparam, np.float32)
        self._chk(_load_surrogate().sp_surrogate_get_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float))),
                  'get_params')
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32).ravel()
        if p.size != self.nparam:
            raise ValueError(f'expected {self.nparam} parameters, got {p.size}')
        self._chk(_load_surrogate().sp_surrogate_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))),
                  'set_params')

    def grad(self, a, b, g, tidx, dt, target):
        """Loss and gradient for explicit points (one angle triple per point)."""
        n = len(a)
        k = _Keep()
        t = np.ascontiguousarray(tidx, np.int32).ravel()
        if t.size != n:
            raise ValueError(f'expected {n} time indices, got {t.size}')
        tg = np.ascontiguousarray(np.asarray(target, np.float32).reshape(n, 2))
        loss = C.c_float()
        gr = np.zeros(self.nparam, np.float32)
        rc = _load_surrogate().sp_surrogate_grad(
            self.h, k.dptr(a, n), k.dptr(b, n), k.dptr(g, n), t.ctypes.data_as(C.POINTER(C.c_int32)), n, float(dt),
            tg.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss), gr.ctypes.data_as(C.POINTER(C.c_float)))
        self._chk(rc, 'grad')
        return loss.value, gr

    def train(self, grid, fids, dt, steps, batch=4096, lr=1e-3, seed=1):
        """Adam on random (orientation, time) points of the training FIDs."""
        k = _Keep()
        f = np.ascontiguousarray(fids, np.complex128)
        n_or, npts = f.shape
        trace = np.zeros(max(steps, 1), np.float32)
        rc = _load_surrogate().sp_surrogate_train(
            self.h, k.dptr(grid['alphas'], n_or), k.dptr(grid['betas'], n_or), k.dptr(grid['gammas'], n_or), n_or,
            f.view(np.float64).ctypes.data_as(_dp), npts, float(dt), int(steps), int(batch), float(lr), int(seed),
            trace.ctypes.data_as(C.POINTER(C.c_float)))
        self._chk(rc, 'train')
        return trace[:steps]

    def powder(self, grid, npts, dt, per_orientation=False):
        """Weighted powder sum of the network's FIDs over a grid (FP64 sum);
        returns (fid, per-orientation or None, device ms)."""
        k = _Keep()
        n_or = len(grid['weights'])
        fid = np.zeros(npts, np.complex128)
        per = np.zeros((n_or, npts), np.complex128) if per_orientation else None
        ms = C.c_double()
        rc = _load_surrogate().sp_surrogate_powder(
            self.h, *k.grid(grid), int(npts), float(dt), fid.view(np.float64).ctypes.data_as(_dp),
            per.view(np.float64).ctypes.data_as(_dp) if per is not None else None, C.byref(ms))
        self._chk(rc, 'powder')
        return fid, per, ms.value

    def close(self):
        if self.h:
            _load_surrogate().sp_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
