"""The C-ABI library loads and exports every entry point include/*.h
declares (no compute call: these run without a GPU)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from spinach_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        text = open(h).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        for m in re.finditer(r'\b(sp_\w+)\s*\(', text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_expected_entry_points():
    names = declared_functions()
    for n in ('sp_create', 'sp_destroy', 'sp_last_error', 'sp_powder_fid', 'sp_plan_powder', 'sp_execute',
              'sp_plan_result', 'sp_plan_destroy', 'sp_plan_powder_batch', 'sp_plan_result_batch'):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_native.EXPORTS) <= set(declared_functions())


def test_binding_declares_signatures():
    lib = _native.load()
    assert lib.sp_last_error.restype is ctypes.c_char_p
    assert lib.sp_powder_fid.argtypes is not None


STRUCTS = {'sp_csc': _native.SpCsc, 'sp_tols': _native.SpTols, 'sp_stats': _native.SpStats,
           'sp_system': _native.SpSystem, 'sp_rotor': _native.SpRotor, 'sp_tuning': _native.SpTuning,
           'sp_relax': _native.SpRelax}


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors have the size and field offsets a C compiler gives
    the header's structs (compiled here with gcc against include/)."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "spinach_powder.h"', 'int main(void) {']
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['  return 0;', '}']
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines) + '\n')
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-std=c99', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split('\n')
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in STRUCTS.items():
        assert got[(cname, 'size')] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_last_error_without_context_is_a_string():
    lib = _native.load()
    assert isinstance(lib.sp_last_error(None), bytes)


def test_graph_flag_in_mode_strings():
    assert _native.mode_id('fp64+graph') == _native.SP_MODE_FP64 | _native.SP_MODE_GRAPH
    assert _native.mode_id('fp32+propagator+graph') == (_native.SP_MODE_FP32 | _native.SP_MODE_PROPAGATOR
                                                          | _native.SP_MODE_GRAPH)
    assert _native.mode_id(_native.SP_MODE_GRAPH | _native.SP_MODE_FP32) == _native.SP_MODE_FP32 | _native.SP_MODE_GRAPH
