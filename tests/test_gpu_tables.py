"""FP32 real Legendre transforms from the stored FP32 tables (the batched GEMM
form of SPHSWE_PREC_FP32) against the CPU oracle, against the on-the-fly
recurrence form of the same precision, and inside the RK4 stepper.

Both forms round the same FP64 recurrence values and the same operands to FP32
and accumulate in FP32; they differ in the order of the sums, so they agree to
the FP32 tolerance of the mode (unit roundoff 2^-24 times a slow growth with
the number of terms), not bitwise."""
import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi
from tests._util import random_real_spec, rel_l2, rel_linf

pytestmark = pytest.mark.gpu

TOL = 2e-6


@pytest.fixture(scope="module", params=[16, 31, 64, 127])
def pair(request):
    L = request.param
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    yield ctx, orc
    ctx.close()


@pytest.mark.parametrize("nf", [1, 3, 9])
def test_table_pair_vs_oracle_and_recurrence(pair, nf):
    ctx, orc = pair
    print(f"fp32 tables L={ctx.L} nf={nf}", flush=True)
    qs = np.stack([random_real_spec(ctx.L, seed=70 + s) for s in range(nf)])
    qs[:, 0] += 3.0
    ctx.set_fp32_tables(0)
    f_rec = ctx.synthesis(qs, capi.PREC_FP32)
    ctx.set_fp32_tables(1)
    f = ctx.synthesis(qs, capi.PREC_FP32)
    fr = np.stack([orc.synthesis(q) for q in qs])
    q = ctx.analysis(fr, capi.PREC_FP32)
    ctx.set_fp32_tables(0)
    q_rec = ctx.analysis(fr, capi.PREC_FP32)
    for k in range(nf):
        assert rel_l2(f[k], fr[k]) < TOL and rel_linf(f[k], fr[k]) < TOL
        assert rel_l2(f[k], f_rec[k]) < TOL
        ref = orc.analysis(fr[k])
        assert rel_l2(q[k], ref) < TOL and rel_linf(q[k], ref) < TOL
        assert rel_l2(q[k], q_rec[k]) < TOL


def test_table_m0_imaginary_part_ignored():
    # Im q_l0 is ignored by synthesis (SHTNS real-field convention)
    ctx = capi.Context(31)
    ctx.set_fp32_tables(1)
    q = random_real_spec(31, seed=5)
    q2 = q.copy()
    q2[:32] += 1j * np.linspace(1, 2, 32)   # m = 0 block
    a = ctx.synthesis(q[None], capi.PREC_FP32)
    b = ctx.synthesis(q2[None], capi.PREC_FP32)
    assert np.array_equal(a, b)
    ctx.close()


def test_table_swe_rk4_matches_recurrence_form():
    # config-0-like dam at T64, 20 steps, FP32 mode with and without tables
    L = 64
    ctx = capi.Context(L)
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    phi1 = np.arcsin(np.asarray(ctx.ct))[None, :]
    d = np.arccos(np.clip(np.sin(phi1) * np.sin(np.pi / 4) + np.cos(phi1) * np.cos(np.pi / 4)
                          * np.cos(lon[:, None] - np.pi), -1, 1))
    h = (10000.0 + 500.0 * np.exp(-(d / 0.1) ** 2)).reshape(-1)
    z = np.zeros_like(h)
    x0 = ctx.analysis(np.stack([h, z, z]))
    prm = capi.swe_params(dt=120.0, precision=capi.PREC_FP32)
    out = {}
    for tmin in (0, 1):
        ctx.set_fp32_tables(tmin)
        s = ctx.to_device(x0)
        ctx.swe_rk4(prm, s, 1, 20)
        out[tmin] = ctx.to_host(s, np.complex128, (3, ctx.nlm))
    s = ctx.to_device(x0)
    ctx.swe_rk4(capi.swe_params(dt=120.0), s, 1, 20)
    ref = ctx.to_host(s, np.complex128, (3, ctx.nlm))
    assert rel_l2(out[1][0], ref[0]) < 1e-6
    assert rel_l2(out[1][0], out[0][0]) < 1e-6
    # graph-captured steps with the tables equal the eager ones
    ctx.set_fp32_tables(1)
    s = ctx.to_device(x0)
    ctx.swe_rk4_graph(prm, s, 1, 20)
    g = ctx.to_host(s, np.complex128, (3, ctx.nlm))
    assert np.array_equal(g, out[1])
    ctx.close()
