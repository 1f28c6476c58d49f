"""sphswe_sht_analysis_noalloc (the real analysis that allocates nothing)
against sphswe_sht_analysis through the C ABI: FP64 and the matrix-core modes
bit for bit, two calls in a row with the second on a smaller input, the
size limit rejected on the host, and no device allocation during a call."""
import ctypes

import numpy as np
import pytest

from sph_swe_amd import capi

pytestmark = pytest.mark.gpu


def _hip_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _both(ctx, spat, prec):
    nf = spat.shape[0]
    d_in = ctx.to_device(spat)
    d_a = ctx.empty(nf * ctx.nlm * 16)
    d_b = ctx.empty(nf * ctx.nlm * 16)
    ctx._check(ctx.lib.sphswe_sht_analysis(ctx.h, d_in.ptr, d_a.ptr, nf, prec))
    print("noalloc T%d nfields=%d prec=%d" % (ctx.L, nf, prec), flush=True)
    ctx._check(ctx.lib.sphswe_sht_analysis_noalloc(ctx.h, d_in.ptr, d_b.ptr, nf, prec))
    a = ctx.to_host(d_a, np.complex128, (nf, ctx.nlm))
    b = ctx.to_host(d_b, np.complex128, (nf, ctx.nlm))
    return a, b


@pytest.mark.parametrize("L,prec", [(511, capi.PREC_FP64), (127, capi.PREC_FP64), (127, capi.PREC_BF16X3),
                                    (127, capi.PREC_FP32), (31, capi.PREC_BF16)])
def test_noalloc_equals_analysis_twice(L, prec):
    ctx = capi.Context(L)
    try:
        rng = np.random.default_rng(L + prec)
        first = 3 if L == 511 else 9
        for nf in (first, 1):   # the second call on a smaller input
            spat = rng.standard_normal((nf, ctx.nspat))
            a, b = _both(ctx, spat, prec)
            assert np.array_equal(a, b), (L, prec, nf)
    finally:
        ctx.close()


def test_noalloc_fp32_tables_within_tolerance():
    # with the stored FP32 tables switched on, sphswe_sht_analysis takes the
    # table GEMM; the no-allocation form keeps the recurrence kernels (the same
    # FP32 roundings, another summation order)
    ctx = capi.Context(127)
    try:
        ctx.set_fp32_tables(1, 1)
        rng = np.random.default_rng(2)
        for nf in (16, 2):
            a, b = _both(ctx, rng.standard_normal((nf, ctx.nspat)), capi.PREC_FP32)
            for f in range(nf):
                assert np.linalg.norm(a[f] - b[f]) <= 2e-6 * np.linalg.norm(a[f])
    finally:
        ctx.close()


def test_noalloc_allocates_nothing():
    ctx = capi.Context(255)
    try:
        spat = np.random.default_rng(4).standard_normal((4, ctx.nspat))
        d_in = ctx.to_device(spat)
        d_out = ctx.empty(4 * ctx.nlm * 16)
        ctx.sync()
        before = _hip_free_bytes()
        ctx._check(ctx.lib.sphswe_sht_analysis_noalloc(ctx.h, d_in.ptr, d_out.ptr, 4, capi.PREC_FP64))
        ctx.sync()
        assert _hip_free_bytes() == before
        ref = ctx.analysis(spat)
        assert np.array_equal(ctx.to_host(d_out, np.complex128, (4, ctx.nlm)), ref)
    finally:
        ctx.close()


def test_noalloc_rejects_sizes_beyond_the_limit():
    lib = capi.load()
    limit = lib.sphswe_noalloc_scratch_bytes()
    assert limit == 96 << 20
    ctx = capi.Context(511)
    try:
        nf = 16   # 16 Fourier row sets of 6.3 MB and one operand group of 50 MB
        d_in = ctx.empty(8)
        d_out = ctx.empty(8)
        rc = ctx.lib.sphswe_sht_analysis_noalloc(ctx.h, d_in.ptr, d_out.ptr, nf, capi.PREC_FP64)
        assert rc == capi.ERR_ARG
        assert b"limit" in ctx.lib.sphswe_last_error(ctx.h)
    finally:
        ctx.close()
