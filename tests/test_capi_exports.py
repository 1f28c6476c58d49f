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


def test_rexi_coefficients_entry_matches_golden():
    # REXI::setup through the C ABI (host only) against the reference build's vectors
    import os
    import numpy as np
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for h, M, half in [(0.2, 64, True), (0.2, 188, True), (0.15, 32, False)]:
        arr = np.loadtxt(os.path.join(golden, "rexi_h%s_M%d_%s.txt" % (repr(h), M, "half" if half else "full")))
        a, br, bi = capi.rexi_coefficients(h, M, half, with_beta_im=True)
        assert a.size == arr.shape[0]
        assert np.max(np.abs(a - (arr[:, 0] + 1j * arr[:, 1]))) <= 1e-15 * np.max(np.abs(a))
        ref_br = arr[:, 2] + 1j * arr[:, 3]
        ref_bi = arr[:, 4] + 1j * arr[:, 5]
        assert np.max(np.abs(br - ref_br)) <= 1e-14 * np.max(np.abs(ref_br))
        assert np.max(np.abs(bi - ref_bi)) <= 1e-14 * np.max(np.abs(ref_bi))


def test_rexi_coefficients_rejects_bad_arguments():
    lib = capi.load()
    n = ctypes.c_int()
    assert lib.sphswe_rexi_coefficients(0.2, 0, 1, ctypes.byref(n), None, None, None) == capi.ERR_ARG
    assert lib.sphswe_rexi_coefficients(-0.2, 8, 1, ctypes.byref(n), None, None, None) == capi.ERR_ARG
    assert lib.sphswe_rexi_coefficients(0.2, 8, 1, ctypes.byref(n), None, None, None) == 0 and n.value == 20
