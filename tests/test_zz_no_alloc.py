"""sp_powder_fid_noalloc (the entry point that calls no allocator: all device
memory from the library's static scratch) against sp_powder_fid, through the
C-ABI. Two calls in a row, the second with a smaller input; results must be
identical bit for bit, since both run the same kernels on the same inputs and
add the 64-orientation chunk partials in the same order."""
import numpy as np
import pytest

from spinach_amd import _native, grids, workloads

pytestmark = pytest.mark.gpu


def _both(w, grid, nsteps):
    a = _native.powder_fid(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], nsteps, per_orientation=True)
    b = _native.powder_fid_noalloc(w['I'], w['Q'], grid, w['rho0'], w['coil'], w['dt'], nsteps,
                                   per_orientation=True)
    return a, b


def test_noalloc_matches_sp_powder_fid_twice_second_smaller():
    w8 = workloads.static_8spin(n_orient=4000)
    g8 = grids.subset(w8['grid'], np.arange(1100))  # several 64-aligned blocks of the scratch
    (f, per, st), (g, gper, gst) = _both(w8, g8, 4096)
    print(f'8-spin 1100 x 4096: noalloc vs sp_powder_fid max |diff| {np.max(np.abs(f - g)):.3e}, '
          f'{gst["kernel_ms"]:.2f} ms', flush=True)
    assert np.array_equal(f, g)
    assert np.array_equal(per, gper)
    w2 = workloads.static_2spin(npoints=256, n_orient=40)
    (f2, per2, _), (g2, gper2, _) = _both(w2, w2['grid'], 255)
    print(f'2-spin 40 x 255 after it: max |diff| {np.max(np.abs(f2 - g2)):.3e}', flush=True)
    assert np.array_equal(f2, g2)
    assert np.array_equal(per2, gper2)


def test_noalloc_rejects_what_it_does_not_run():
    w = workloads.static_2spin(npoints=64, n_orient=8)
    args = (w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], 63)
    with pytest.raises(_native.SpinachPowderError, match='spectral path'):
        _native.powder_fid_noalloc(*args, mode='fp32')
    with pytest.raises(_native.SpinachPowderError, match='spectral path'):
        _native.powder_fid_noalloc(*args, mode='fp64+complex')
    with pytest.raises(_native.SpinachPowderError, match='static scratch'):
        _native.powder_fid_noalloc(w['I'], w['Q'], w['grid'], w['rho0'], w['coil'], w['dt'], 200000)
