"""The stored-table form of the real Legendre transforms (SPHSWE_PREC_MIXED):
associated-Legendre values from the FP64 recurrence stored as bf16 hi + lo,
field operands split the same way, products hi*hi + hi*lo + lo*hi on the
double-K bf16 matrix cores, FP32 accumulation.

Against the CPU oracle and the FP64 GPU path.  The bound is the two-term bf16
split's: each operand carries 16 significant bits (relative error 2^-17) and
the dropped lo*lo term is 2^-16 of a product, so a contraction of positive and
cancelling terms stays within a few 2^-16 of the field's scale; 3e-5 as for the
recurrence form's BF16 mode.  Field counts cover every column-group width, a
partial last group and fields that must not leak into each other.
"""
import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi
from tests._util import random_real_spec, rel_l2, rel_linf

pytestmark = pytest.mark.gpu

TOL = 3e-5


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


def test_pair_vs_oracle(pair):
    ctx, orc = pair
    qs = _spectra(ctx.L, 3, 140)
    f = ctx.synthesis(qs, capi.PREC_MIXED)
    for k in range(3):
        ref = orc.synthesis(qs[k])
        assert rel_l2(f[k], ref) < TOL
        assert rel_linf(f[k], ref) < TOL
    fr = np.stack([orc.synthesis(q) for q in qs])
    q = ctx.analysis(fr, capi.PREC_MIXED)
    for k in range(3):
        ref = orc.analysis(fr[k])
        assert rel_l2(q[k], ref) < TOL
        assert rel_linf(q[k], ref) < TOL


@pytest.mark.parametrize("n", [1, 2, 7, 8, 9, 16, 17, 24, 33, 40, 65])
def test_field_counts_vs_fp64(pair, n):
    ctx, _ = pair
    qs = _spectra(ctx.L, n, 300)
    ref = ctx.synthesis(qs)
    f = ctx.synthesis(qs, capi.PREC_MIXED)
    for k in range(n):
        assert rel_l2(f[k], ref[k]) < TOL, (n, k)
    qref = ctx.analysis(ref)
    q = ctx.analysis(ref, capi.PREC_MIXED)
    for k in range(n):
        assert rel_l2(q[k], qref[k]) < TOL, (n, k)
        assert rel_linf(q[k], qref[k]) < TOL, (n, k)


def test_single_modes_t511():
    # single (l, m) modes at T511: the synthesis of a unit coefficient is the
    # table column itself, the analysis of it returns the unit coefficient
    ctx = capi.Context(511)
    try:
        nlm = ctx.nlm
        picks = [(0, 0), (511, 0), (300, 17), (511, 255), (256, 256), (511, 511), (400, 399)]
        qs = np.zeros((len(picks), nlm), np.complex128)
        for k, (l, m) in enumerate(picks):
            qs[k, m * (2 * 511 - m + 1) // 2 + l] = 1.0 + 0.5j if m > 0 else 1.0
        ref = ctx.synthesis(qs)
        f = ctx.synthesis(qs, capi.PREC_MIXED)
        for k in range(len(picks)):
            assert rel_l2(f[k], ref[k]) < TOL, picks[k]
        q = ctx.analysis(ref, capi.PREC_MIXED)
        for k in range(len(picks)):
            assert rel_l2(q[k], qs[k]) < TOL, picks[k]
    finally:
        ctx.close()


def test_many_fields_t511_vs_fp64():
    # the batched widths of the SWE stepper at T511: 9 x 16 syntheses and
    # 4 x 16 analyses in one call each
    ctx = capi.Context(511)
    try:
        qs = _spectra(511, 144, 700)
        ref = ctx.synthesis(qs)
        f = ctx.synthesis(qs, capi.PREC_MIXED)
        errs = [rel_l2(f[k], ref[k]) for k in range(144)]
        assert max(errs) < TOL
        qref = ctx.analysis(ref[:64])
        q = ctx.analysis(ref[:64], capi.PREC_MIXED)
        errs = [rel_l2(q[k], qref[k]) for k in range(64)]
        assert max(errs) < TOL
    finally:
        ctx.close()


# ---- the SWE stepper in the mixed form --------------------------------------
# Legendre contractions from the bf16 tables, FP32 Fourier rows, FFTs and
# pointwise terms, FP64 state and stage sums.  Config 0 (T64, Gaussian dam,
# dt = 120 s, 100 steps) against the FP64 oracle with the bounds of the
# two-term bf16 split (the FP32 Fourier and physical stages add rounding of
# 2^-24, far below it): phi within a few 1e-6, the winds (which start at rest)
# within a few 1e-3 of their own scale.
def _dam_state(ctx, orc):
    from tests._util import gaussian_dam
    h = gaussian_dam(orc.lon(), orc.ct)
    z = np.zeros_like(h)
    return h, z


def test_swe_config0_mixed_vs_oracle():
    from tests._util import swe_prm
    L, dt, nsteps = 64, 120.0, 100
    ctx = capi.Context(L)
    try:
        orc = O.Oracle(L, ctx.nlat, ctx.nphi)
        h, z = _dam_state(ctx, orc)
        ref = orc.swe_rk4_run(swe_prm(dt), h, z, z, nsteps)
        state = ctx.to_device(ctx.analysis(np.stack([h, z, z])))
        ctx.swe_rk4(capi.swe_params(dt=dt, precision=capi.PREC_MIXED), state, 1, nsteps)
        out = ctx.to_host(state, np.complex128, (3, ctx.nlm))
        e = {name: rel_l2(out[k], ref[name]) for k, name in enumerate("huv")}
        print(e)
        assert e["h"] < 2e-5
        assert e["u"] < 3e-2 and e["v"] < 3e-2
    finally:
        ctx.close()


@pytest.mark.parametrize("L", [31, 64, 127])
def test_swe_mixed_members_and_graph(L):
    # three members with different initial states: each equals its own
    # single-member run bitwise (columns of one contraction are independent,
    # the FFTs and pointwise terms are per field), and the captured step
    # replayed equals the eager one bitwise
    ctx = capi.Context(L)
    try:
        orc = O.Oracle(L, ctx.nlat, ctx.nphi)
        h, z = _dam_state(ctx, orc)
        states = []
        for k in range(3):
            hk = 10000.0 + (h - 10000.0) * (0.6 + 0.3 * k)
            states.append(ctx.analysis(np.stack([hk, z, z])))
        prm = capi.swe_params(dt=60.0, precision=capi.PREC_MIXED)
        both = ctx.to_device(np.stack(states))
        ctx.swe_rk4(prm, both, 3, 4)
        multi = ctx.to_host(both, np.complex128, (3, 3, ctx.nlm))
        for k in range(3):
            one = ctx.to_device(states[k])
            ctx.swe_rk4(prm, one, 1, 4)
            assert np.array_equal(ctx.to_host(one, np.complex128, (3, ctx.nlm)), multi[k])
        g = ctx.to_device(np.stack(states))
        ctx.swe_rk4_graph(prm, g, 3, 2)
        ctx.swe_rk4_graph(prm, g, 3, 2)
        assert np.array_equal(ctx.to_host(g, np.complex128, (3, 3, ctx.nlm)), multi)
    finally:
        ctx.close()


def test_swe_t511_mixed_vs_fp64():
    # the batched T511 form: 4 members of the dam at dt = 0.0005 a / nlat,
    # 20 steps, against the FP64 stepper: phi within 1e-6 (the north star's
    # bound after 100 steps)
    from tests._util import gaussian_dam
    ctx = capi.Context(511)
    try:
        lon = np.arange(ctx.nphi) * 2 * np.pi / ctx.nphi
        ct = ctx.ct
        dt = 0.0005 * 6.37122e6 / ctx.nlat
        states = []
        for k in range(4):
            h = gaussian_dam(lon, ct, lat_c=np.pi / 4 + 0.05 * k, amp=400.0 + 50 * k)
            z = np.zeros_like(h)
            states.append(ctx.analysis(np.stack([h, z, z])))
        s0 = np.stack(states)
        a = ctx.to_device(s0)
        ctx.swe_rk4(capi.swe_params(dt=dt), a, 4, 20)
        ref = ctx.to_host(a, np.complex128, (4, 3, ctx.nlm))
        b = ctx.to_device(s0)
        ctx.swe_rk4_graph(capi.swe_params(dt=dt, precision=capi.PREC_MIXED), b, 4, 20)
        out = ctx.to_host(b, np.complex128, (4, 3, ctx.nlm))
        for k in range(4):
            e = rel_l2(out[k, 0], ref[k, 0])
            print(k, e, rel_l2(out[k, 1], ref[k, 1]))
            assert e < 1e-6
    finally:
        ctx.close()
