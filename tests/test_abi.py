"""The C-ABI library loads and exports every entry point include/uppe.h declares; the
host-only size queries answer without a GPU."""
import ctypes
import os
import re

import pytest

from project_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "uppe.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(uppe\w+)\s*\(", src, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from project_amd import build
        build.build()
    return _lib.load()


def test_every_declared_symbol_is_exported(lib):
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures must cover the header"


def test_workspace_sizes_without_gpu(lib):
    d = _lib.Uppe3DDesc(nx=128, ny=128, nt=1024, dt=1e-15, a_scale=1.0, lut_n=2000)
    nb = ctypes.c_size_t()
    assert lib.uppe3d_workspace_size(ctypes.byref(d), ctypes.byref(nb)) == 0
    field = 8 * 128 * 128 * 1024
    assert 7 * field <= nb.value < 7 * field + (1 << 22)
    d.nx = d.ny = 96
    assert lib.uppe3d_workspace_size(ctypes.byref(d), ctypes.byref(nb)) == -1
    assert b"unsupported" in lib.uppe_last_error()
    d1 = _lib.Uppe1DDesc(nt=4096)
    assert lib.uppe1d_workspace_size(ctypes.byref(d1), ctypes.byref(nb)) == 0
    assert nb.value > 4096 * 8 * 5


def test_t_axis_lengths_without_gpu(lib):
    """t axes: powers of two and other lengths with prime factors <= 31 (at most one above 13)
    in [128, 2048]; the adjoint takes powers of two only."""
    nb = ctypes.c_size_t()
    for nt, ok in ((2046, True), (1001, True), (250, True), (1536, True), (2048, True),
                   (2039, False), (1798, False), (4094, False), (126, False)):
        d = _lib.Uppe3DDesc(nx=32, ny=32, nt=nt, dt=1e-15, a_scale=1.0, lut_n=2000)
        rc = lib.uppe3d_workspace_size(ctypes.byref(d), ctypes.byref(nb))
        assert (rc == 0) == ok, nt
        if not ok:
            assert b"nt = %d unsupported" % nt in lib.uppe_last_error()
    d = _lib.Uppe3DDesc(nx=32, ny=32, nt=2046, dt=1e-15, a_scale=1.0, lut_n=2000)
    assert lib.uppe3d_adjoint_workspace_size(ctypes.byref(d), ctypes.byref(nb)) == -1
    assert b"power of two" in lib.uppe_last_error()
    d.nt = 2048
    assert lib.uppe3d_adjoint_workspace_size(ctypes.byref(d), ctypes.byref(nb)) == 0
