"""The C-ABI library loads and exports every symbol include/gamdpy_amd.h declares
(no compute calls: runs without a GPU)."""
import ctypes

from gamdpy_amd import _native


def test_library_exports_header_symbols():
    lib = _native.load()
    names = _native.header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
        assert n in _native._SIGNATURES, f"{n} has no ctypes signature"


def test_abi_version_and_desc_layout():
    lib = _native.load()
    assert lib.gmd_abi_version() == _native.ABI_VERSION
    assert lib.gmd_desc_size() == ctypes.sizeof(_native.GmdDesc)


def test_null_arguments_are_errors():
    lib = _native.load()
    assert lib.gmd_create(None, None) != 0
    assert b"null" in lib.gmd_last_error()
    assert lib.gmd_run(None, ctypes.c_float(0.0), 1) != 0
    assert lib.gmd_destroy(None) == 0
