"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, struct layouts agree with the binding, and argument errors are
reported through mw_last_error without touching the device."""

import ctypes
import os
import re

import pytest

from witwin_maxwell_amd import build as mwbuild
from witwin_maxwell_amd.fdtd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    text = open(os.path.join(ROOT, "include", "maxwell_fdtd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mw_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(mwbuild.library_path()):
        mwbuild.build()
    return native.load_library()


def test_exports_every_declared_symbol(lib):
    names = _header_functions()
    assert "mw_run_steps" in names and "mw_update_electric" in names
    for name in names:
        assert hasattr(lib, name), name
    assert set(native.EXPORTS) <= set(names)


def test_struct_layouts(lib):
    sizes = (ctypes.c_int64 * 16)()
    n = lib.mw_abi_struct_sizes(sizes, 16)
    assert n == len(native.STRUCTS)
    for i, st in enumerate(native.STRUCTS):
        assert ctypes.sizeof(st) == sizes[i], st.__name__


def test_argument_errors_are_reported(lib):
    g = native.MwGrid()
    g.nx, g.ny, g.nz = 1, 4, 4
    rc = lib.mw_update_magnetic(ctypes.byref(g), None, None, None, 7, None)
    assert rc == -1
    assert "grid" in lib.mw_last_error().decode() or "field" in lib.mw_last_error().decode()
    rc = lib.mw_add_source_patch(ctypes.c_void_p(16), 4, 4, 4, ctypes.c_void_p(16), 2, 2, 2, 3, 0, 0, 1.0, None)
    assert rc == -1
    assert "outside" in lib.mw_last_error().decode()
    prog = native.MwProgram()
    rc = lib.mw_run_steps(ctypes.byref(prog), 0, 1, None)
    assert rc == -1


def test_adjoint_and_batch_argument_errors(lib):
    """The adjoint and batch entry points reject bad arguments before any HIP call."""
    vp = ctypes.c_void_p
    assert lib.mw_capture_steps(None, 0, 1, None, None) == -1
    assert lib.mw_reverse_steps(None, None, None, 0, 1, None) == -1
    prog = native.MwProgram()
    adj = native.MwAdjoint()
    assert lib.mw_reverse_steps(ctypes.byref(prog), ctypes.byref(adj), None, 0, 1, None) == -1
    handle = vp()
    assert lib.mw_batch_create(None, 0, 0, ctypes.byref(handle)) == -1
    progs = (ctypes.POINTER(native.MwProgram) * 1)(ctypes.POINTER(native.MwProgram)())
    assert lib.mw_batch_create(progs, 1, 0, ctypes.byref(handle)) == -1
    assert "NULL" in lib.mw_last_error().decode()
    assert lib.mw_batch_create(progs, 1, 999, ctypes.byref(handle)) == -1
    assert lib.mw_batch_run(None, 0, 1, 1, None) == -1
    assert lib.mw_batch_destroy(None) == 0
    assert lib.mw_adjoint_batch_create(None, None, None, None, 1, ctypes.byref(handle)) == -1
    assert lib.mw_adjoint_batch_capture(None, 0, 1, None) == -1
    assert lib.mw_adjoint_batch_reverse(None, 0, 1, None) == -1
