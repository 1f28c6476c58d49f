"""Autotuner and the launch variants it chooses between: every variant of the
real FFT (sequences per workgroup) and the FP64 synthesis on the matrix cores
give the oracle's results to FP64 round-off, and tuning changes nothing but the
order of the sums."""
import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi
from tests._util import random_real_spec, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-13
# candidates of the compile-time FFT per truncation (grid nphi 96, 196, 384)
NSEQ = {31: [8, 16], 64: [4, 8], 127: [2, 4, 8]}


@pytest.fixture(scope="module", params=sorted(NSEQ))
def pair(request):
    L = request.param
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    yield ctx, orc
    ctx.close()


def test_every_fft_variant_matches_oracle(pair):
    ctx, orc = pair
    qs = np.stack([random_real_spec(ctx.L, seed=300 + s) for s in range(3)])
    refs = [orc.synthesis(q) for q in qs]
    for ns in NSEQ[ctx.L]:
        ctx.set_tuning(fft_nseq_r2c=ns, fft_nseq_c2r=ns)
        f = ctx.synthesis(qs)
        for k in range(3):
            assert rel_l2(f[k], refs[k]) < TOL, ns
        q = ctx.analysis(f)
        for k in range(3):
            assert rel_l2(q[k], orc.analysis(f[k])) < TOL, ns
    ctx.set_tuning()


def test_matrix_core_fp64_synthesis_matches_oracle(pair):
    ctx, orc = pair
    qs = np.stack([random_real_spec(ctx.L, seed=400 + s) for s in range(11)])
    ctx.set_tuning(synth_mfma_min=1)
    f = ctx.synthesis(qs)
    ctx.set_tuning()
    for k in range(11):
        assert rel_l2(f[k], orc.synthesis(qs[k])) < TOL


@pytest.mark.parametrize("nf", [1, 2, 3, 4, 5, 9, 12])
def test_legendre_shapes_match_oracle(pair, nf):
    # both FP64 synthesis shapes (8 x 3 in groups of three; 4 x 2 in groups of
    # four and three) and both analysis variants (4x4x4 kernel for a remainder
    # of at most four fields; 16x16x4 only), every field count class
    ctx, orc = pair
    qs = np.stack([random_real_spec(ctx.L, seed=600 + s) for s in range(nf)])
    refs = [orc.synthesis(q) for q in qs]
    for shape in (0, 1):
        for variant in (0, 1):
            ctx.set_legendre_tuning(synth_shape=shape, anal_variant=variant)
            f = ctx.synthesis(qs)
            q = ctx.analysis(f)
            for k in range(nf):
                assert rel_l2(f[k], refs[k]) < TOL, (shape, variant, k)
                assert rel_l2(q[k], orc.analysis(f[k])) < TOL, (shape, variant, k)
                assert rel_l2(q[k], qs[k]) < TOL, (shape, variant, k)
    ctx.set_legendre_tuning()


def test_autotune_choices_and_results(pair):
    ctx, _ = pair
    qs = np.stack([random_real_spec(ctx.L, seed=500 + s) for s in range(16)])
    f0 = ctx.synthesis(qs)
    q0 = ctx.analysis(f0)
    t = ctx.autotune()
    assert t["tuned"]
    assert t["fft_nseq_r2c"] in NSEQ[ctx.L] and t["fft_nseq_c2r"] in NSEQ[ctx.L]
    assert t["synth_mfma_min"] in (0, 1, 3, 8, 16)
    assert t["fp32_table_synth_min"] in (0, 16) and t["fp32_table_anal_min"] in (0, 16)
    assert ctx.legendre_tuning()["synth_shape"] in (0, 1) and ctx.legendre_tuning()["anal_variant"] in (0, 1)
    f1 = ctx.synthesis(qs)
    q1 = ctx.analysis(f1)
    for k in range(16):
        assert rel_l2(f1[k], f0[k]) < TOL
        assert rel_l2(q1[k], q0[k]) < TOL
    # FP32 transforms with the tuned choice against the recurrence form
    g1 = ctx.synthesis(qs, capi.PREC_FP32)
    r1 = ctx.analysis(f0, capi.PREC_FP32)
    ctx.set_fp32_tables(0)
    g0 = ctx.synthesis(qs, capi.PREC_FP32)
    r0 = ctx.analysis(f0, capi.PREC_FP32)
    for k in range(16):
        assert rel_l2(g1[k], g0[k]) < 2e-6
        assert rel_l2(r1[k], r0[k]) < 2e-6
    ctx.set_tuning()
    ctx.set_legendre_tuning()


def test_bad_tuning_is_rejected(pair):
    ctx, _ = pair
    with pytest.raises(capi.SphSweError):
        ctx.set_tuning(fft_nseq_r2c=-1)
    with pytest.raises(capi.SphSweError):
        ctx.set_legendre_tuning(synth_shape=2)
    with pytest.raises(capi.SphSweError):
        ctx.set_legendre_tuning(anal_variant=-1)
