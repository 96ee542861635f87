"""The C-ABI library builds for gfx950, loads without a GPU, and exports every
entry point declared in include/giwaxsim_amd.h (no compute calls here)."""

import os
import re

import pytest

from giwaxsim_amd import _lib
from giwaxsim_amd.build import LIB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, 'include', 'giwaxsim_amd.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(gx_[a-z0-9_]+)\s*\(', src)))


def test_library_present():
    assert os.path.exists(LIB), "run __graft_entry__.build() first"


def test_exports_every_declared_symbol():
    lib = _lib.load()
    decl = declared_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(_lib.EXPORTS), set(decl) ^ set(_lib.EXPORTS)


def test_no_gpu_calls_needed_for_metadata():
    lib = _lib.load()
    assert lib.gx_abi_version() == _lib.GX_ABI_VERSION == 12
    assert [lib.gx_supported_grid(n) for n in (8, 128, 256, 1048, 2048, 2049, 4096)] == \
        [0, 1, 1, 1, 1, 0, 0]


def test_argument_errors_are_reported():
    """Validation runs before any launch, so error paths are testable on the CPU."""
    import ctypes
    with pytest.raises(_lib.GxError, match='NULL'):
        _lib.call('gx_rotate_points', (ctypes.c_double * 9)(), None, None, None, 3, None, None,
                  None, None)
    d = _lib.GxFrames(N=1000, r=0.2)
    with pytest.raises(_lib.GxError, match='chirp-z length'):
        _lib.call('gx_frames_rowpass', ctypes.byref(d), 0, 1, None)
    a = _lib.GxAdjoint()
    with pytest.raises(_lib.GxError, match='not 256/512/1024/2048'):
        _lib.call('gx_adj_rowpass', ctypes.byref(d), ctypes.byref(a), 0, 1, None)
    d = _lib.GxFrames(N=512, r=0.2, fill_bkg=1)
    with pytest.raises(_lib.GxError, match='fill_bkg needs col_nm'):
        _lib.call('gx_adj_backward', ctypes.byref(d), ctypes.byref(a), 0, 1, None)
    with pytest.raises(_lib.GxError, match='crop window'):
        _lib.call('gx_bins_finalize', ctypes.c_void_p(8), ctypes.c_void_p(8), 10, 1.0, 5, 12, 0.0,
                  0.0, 0.0, _lib.dbl_array([0] * 9), 6.0, None, None, None)
