"""Mixed-precision Legendre modes (FP32, BF16 hi/lo, BF16 hi/mid/lo) against the
CPU oracle and the FP64 GPU path.

The recurrence is FP64 in every mode; only the contraction is rounded.  The
tolerances are relative L2 / Linf over a field and follow from the unit
roundoff of the products: about 2^-24 for FP32 and for the three-term bf16
split (which represents FP32 values exactly and drops products of order
2^-24), about 2^-16 for the two-term bf16 split, times a slow growth with the
number of terms summed.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi
from tests._util import random_cplx_spec, random_real_spec, rel_l2, rel_linf

pytestmark = pytest.mark.gpu

TOL = {capi.PREC_FP32: 2e-6, capi.PREC_BF16: 3e-5, capi.PREC_BF16X3: 2e-6}
MODES = sorted(TOL)


@pytest.fixture(scope="module", params=[16, 31, 64, 127])
def pair(request):
    L = request.param
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    yield ctx, orc
    ctx.close()


def _spectra(L, n, seed):
    qs = np.stack([random_real_spec(L, seed=seed + s) for s in range(n)])
    qs[:, 0] += 3.0   # a dominant mean, as the geopotential has
    return qs


@pytest.mark.parametrize("prec", MODES)
def test_real_pair_vs_oracle(pair, prec):
    ctx, orc = pair
    qs = _spectra(ctx.L, 3, 40)
    f = ctx.synthesis(qs, prec)
    for k in range(3):
        ref = orc.synthesis(qs[k])
        assert rel_l2(f[k], ref) < TOL[prec]
        assert rel_linf(f[k], ref) < TOL[prec]
    fr = np.stack([orc.synthesis(q) for q in qs])
    q = ctx.analysis(fr, prec)
    for k in range(3):
        ref = orc.analysis(fr[k])
        assert rel_l2(q[k], ref) < TOL[prec]
        assert rel_linf(q[k], ref) < TOL[prec]


@pytest.mark.parametrize("prec", MODES)
def test_complex_pair_vs_oracle(pair, prec):
    ctx, orc = pair
    qs = np.stack([random_cplx_spec(ctx.L, seed=60 + s) for s in range(2)])
    f = ctx.synthesis_cplx(qs, prec)
    for k in range(2):
        assert rel_l2(f[k], orc.synthesis_cplx(qs[k])) < TOL[prec]
    q = ctx.analysis_cplx(f, prec)
    for k in range(2):
        assert rel_l2(q[k], orc.analysis_cplx(f[k])) < TOL[prec]


@pytest.mark.parametrize("prec", MODES)
def test_field_groups_and_padding(pair, prec):
    # 1..17 fields: groups of 8 slots with unused columns; every field matches
    # the FP64 path, and fields of a batch do not leak into each other
    ctx, _ = pair
    qs = _spectra(ctx.L, 17, 80)
    ref = ctx.synthesis(qs)
    for n in (1, 2, 5, 8, 9, 17):
        f = ctx.synthesis(qs[:n], prec)
        for k in range(n):
            assert rel_l2(f[k], ref[k]) < TOL[prec]
        q = ctx.analysis(ref[:n], prec)
        qref = ctx.analysis(ref[:n])
        for k in range(n):
            assert rel_l2(q[k], qref[k]) < TOL[prec]


def test_bf16_two_term_is_less_accurate_than_three_term(pair):
    # the two-term split really drops bits: a check that the modes differ
    ctx, _ = pair
    qs = _spectra(ctx.L, 2, 120)
    ref = ctx.synthesis(qs)
    e2 = rel_l2(ctx.synthesis(qs, capi.PREC_BF16), ref)
    e3 = rel_l2(ctx.synthesis(qs, capi.PREC_BF16X3), ref)
    assert e2 > 1e-7
    assert e3 < e2 / 5


def test_unknown_precision_is_rejected(pair):
    ctx, _ = pair
    qs = _spectra(ctx.L, 1, 1)
    with pytest.raises(capi.SphSweError):
        ctx.synthesis(qs, 7)


@pytest.mark.parametrize("prec", MODES)
def test_full_size_round_trip(prec):
    # T511 on the 768 x 1536 grid: synthesis in the mode against FP64, and the
    # analysis of the FP64 field in the mode against the exact coefficients
    ctx = capi.Context(511)
    try:
        qs = _spectra(511, 3, 200)
        ref = ctx.synthesis(qs)
        f = ctx.synthesis(qs, prec)
        q = ctx.analysis(ref, prec)
        for k in range(3):
            assert rel_l2(f[k], ref[k]) < TOL[prec]
            assert rel_l2(q[k], qs[k]) < TOL[prec]
    finally:
        ctx.close()


# nonlinear SWE (config 0: T64, Gaussian dam, dt = 120 s, 100 steps) with the
# Legendre contractions in a reduced mode against the FP64 oracle: relative L2
# of the geopotential (phi = g h) spectrum, and of u and v, which start at rest
# and are measured on their own (small) scale.  The two-term bf16 split keeps
# phi within a few 1e-6 but not the winds: the weighted Fourier rows that the
# analysis splits are dominated by the 10 km mean height, their 2^-16 relative
# splitting error differs from ring to ring and reaches every degree, and the
# gradients of that error drive winds of about 1e-2 of the true ones after 100
# steps.  FP32 and the three-term split hold the winds to about 1e-5.
SWE_TOL_PHI = {capi.PREC_FP32: 1e-6, capi.PREC_BF16: 2e-5, capi.PREC_BF16X3: 1e-6}
SWE_TOL_UV = {capi.PREC_FP32: 1e-4, capi.PREC_BF16: 3e-2, capi.PREC_BF16X3: 1e-4}


@pytest.mark.parametrize("prec", MODES)
def test_swe_config0_in_mode(prec):
    from tests._util import gaussian_dam, swe_prm
    L, dt, nsteps = 64, 120.0, 100
    ctx = capi.Context(L)
    try:
        orc = O.Oracle(L, ctx.nlat, ctx.nphi)
        h = gaussian_dam(orc.lon(), orc.ct)
        z = np.zeros_like(h)
        ref = orc.swe_rk4_run(swe_prm(dt), h, z, z, nsteps)
        state = ctx.to_device(ctx.analysis(np.stack([h, z, z])))
        ctx.swe_rk4(capi.swe_params(dt=dt, precision=prec), state, 1, nsteps)
        out = ctx.to_host(state, np.complex128, (3, ctx.nlm))
        e = {name: rel_l2(out[k], ref[name]) for k, name in enumerate("huv")}
        print(prec, e)
        assert e["h"] < SWE_TOL_PHI[prec]
        assert e["u"] < SWE_TOL_UV[prec] and e["v"] < SWE_TOL_UV[prec]
    finally:
        ctx.close()
