"""The C-ABI library loads and exports every entry point include/slx.h declares (no GPU calls)."""
import ctypes
import os
import re

from slevolve_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "slx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(slx_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.lib_path())
    names = declared_symbols()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_abi_version_and_error_string():
    lib = _lib.load()
    assert lib.slx_abi_version() == 3
    assert isinstance(lib.slx_last_error(), bytes)


def test_bad_arguments_are_rejected_without_a_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.slx_sht_create(63, 128, 31, 0, 3, ctypes.byref(h)) != 0  # odd nlat
    assert b"even" in lib.slx_last_error()
    assert lib.slx_sht_create(64, 96, 31, 0, 3, ctypes.byref(h)) != 0   # nlon not a power of two
    assert lib.slx_sht_create(64, 128, 65, 0, 3, ctypes.byref(h)) != 0  # lmax > nlon/2
    assert lib.slx_sht_create(64, 128, 31, 2, 3, ctypes.byref(h)) != 0  # unknown precision
    assert b"precision" in lib.slx_last_error()


def test_struct_layouts_match_the_header():
    """The ctypes mirrors of slx_solver_params and slx_status end where the C structs do."""
    assert ctypes.sizeof(_lib.SlxStatus) % 8 == 0
    assert _lib.SlxStatus.sweeps32.offset + 8 == ctypes.sizeof(_lib.SlxStatus)
    assert _lib.SlxSolverParams.max_inner.offset + 8 == ctypes.sizeof(_lib.SlxSolverParams)
