"""The C-ABI library loads and exports every symbol include/sphswe.h declares
(no compute call: this runs without a GPU)."""
import ctypes

from sph_swe_amd import capi


def test_library_exports_header_symbols():
    lib = capi.load()
    names = capi.header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_grid_auto_needs_no_gpu():
    assert capi.grid_auto(511) == (768, 1536)
    assert capi.grid_auto(64) == (98, 196)
    assert capi.grid_auto(127) == (192, 384)


def test_bad_arguments_are_rejected_without_gpu():
    lib = capi.load()
    h = ctypes.c_void_p()
    # mmax != lmax, odd nlat, too few longitudes: rejected before touching the device
    assert lib.sphswe_create(31, 30, 48, 96, 0, ctypes.byref(h)) == capi.ERR_ARG
    assert lib.sphswe_create(31, 31, 47, 96, 0, ctypes.byref(h)) == capi.ERR_ARG
    assert lib.sphswe_create(31, 31, 48, 60, 0, ctypes.byref(h)) == capi.ERR_ARG
    assert lib.sphswe_create(31, 31, 48, 11 * 13, 0, ctypes.byref(h)) == capi.ERR_ARG  # prime factors > 7
