"""The C-ABI library loads and exports every entry point include/*.h
declares (no compute call: these run without a GPU)."""
import ctypes
import glob
import os
import re

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


def test_struct_layouts_match_header():
    # sp_csc: int64 n, 2 pointers, 1 pointer, int32 -> 40 bytes with padding
    assert ctypes.sizeof(_native.SpCsc) == 40
    assert ctypes.sizeof(_native.SpTols) == 5 * 8
    assert ctypes.sizeof(_native.SpStats) == 3 * 8 + 3 * 8 + 6 * 4
    assert ctypes.sizeof(_native.SpSystem) == 10 * 8


def test_last_error_without_context_is_a_string():
    lib = _native.load()
    assert isinstance(lib.sp_last_error(None), bytes)
