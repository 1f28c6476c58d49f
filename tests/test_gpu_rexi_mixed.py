"""REXI step with the per-pole gradient operator on the FP32 matrix cores and/or
the band LU factors stored in FP32 (solves in FP64), against the oracle's FP64
restatement of AppSWE's REXI loop and SWERexiSPH::solve (src/AppSWE.hpp:415-568,
src/SWERexiSPH.hpp:137-244), and at the full T256 / 200-pole size against the
closed form of the pole sum on gravity modes and against the FP64 GPU step."""
import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi
from tests._util import A_EARTH, G, H0, OMEGA, random_real_spec, real_spec_index, rel_l2

pytestmark = pytest.mark.gpu

# relative L2 bounds per (gradient-operator precision, factor precision), FP64
# reference.  FP32 factors: the solve's error is about kappa * 2^-24 of the
# per-pole solution; after the pole sum 2e-8 .. 4e-8 measured (T31 .. T127).
# FP32 operator: operands rounded to 2^-24, FP32 accumulation over up to L/2
# terms; about 5e-7 measured at T256.
BOUNDS = {
    ("fp64", "fp32"): 1e-6,
    ("fp32", "fp64"): 2e-6,
    ("fp32", "fp32"): 2e-6,
}
# FP32 factors with one sweep of FP64 iterative refinement (residual with the
# FP64 band rebuilt from its pole-independent pieces): the first solve's error
# eps ~ 3e-8 becomes ~eps^2 plus FP64 round-off of the residual, i.e. the FP64
# path's agreement with the oracle (1e-12 there, ~1e-15 measured)
REFINED_BOUND = 1e-11


def _state(L, seed):
    h = random_real_spec(L, seed, power=-1.5) * 50.0
    h[0] = 0.0
    u = random_real_spec(L, seed + 1, power=-1.5)
    v = random_real_spec(L, seed + 2, power=-1.5)
    return h, u, v


def _run(ctx, params, state, nsteps):
    rx = ctx.rexi(params)
    d = ctx.to_device(np.stack(state))
    rx.step(d, nsteps)
    out = ctx.to_host(d, np.complex128, (3, ctx.nlm))
    rx.close()
    return out


@pytest.mark.parametrize("L,M", [(31, 8), (64, 32)])
@pytest.mark.parametrize("leg,band", sorted(BOUNDS))
def test_rexi_modes_vs_oracle(L, M, leg, band):
    print(f"rexi mixed vs oracle L={L} M={M} legendre={leg} factors={band}", flush=True)
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    h, u, v = _state(L, 11)
    dt, nsteps = 3600.0, 2
    prm = dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=1e5, rexi_h=0.2, rexi_M=M, half=1, dt=dt)
    ref = orc.swe_rexi_run(prm, h, u, v, nsteps)
    params = capi.rexi_params(omega=OMEGA, visc=1e5, dt=dt, rexi_h=0.2, rexi_M=M, half=1,
                              precision=capi.PRECISIONS[leg], band_precision=capi.PRECISIONS[band])
    out = _run(ctx, params, (h, u, v), nsteps)
    for k in range(3):
        e = rel_l2(out[k], ref[k])
        print(L, leg, band, "huv"[k], e)
        assert e < BOUNDS[(leg, band)]
    ctx.close()


@pytest.mark.parametrize("L,M", [(31, 8), (64, 32)])
def test_rexi_refined_fp32_factors_vs_oracle(L, M):
    print(f"rexi refined FP32 factors vs oracle L={L} M={M}", flush=True)
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    h, u, v = _state(L, 11)
    dt, nsteps = 3600.0, 2
    prm = dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=1e5, rexi_h=0.2, rexi_M=M, half=1, dt=dt)
    ref = orc.swe_rexi_run(prm, h, u, v, nsteps)
    params = capi.rexi_params(omega=OMEGA, visc=1e5, dt=dt, rexi_h=0.2, rexi_M=M, half=1,
                              band_precision=capi.PREC_FP32, refine=1)
    out = _run(ctx, params, (h, u, v), nsteps)
    unref = _run(ctx, capi.rexi_params(omega=OMEGA, visc=1e5, dt=dt, rexi_h=0.2, rexi_M=M, half=1,
                                       band_precision=capi.PREC_FP32), (h, u, v), nsteps)
    for k in range(3):
        e, e0 = rel_l2(out[k], ref[k]), rel_l2(unref[k], ref[k])
        print(L, "refined", "huv"[k], e, "unrefined", e0)
        assert e < REFINED_BOUND
        assert e < 1e-3 * e0   # the refinement removes the FP32 factor error
    ctx.close()


def test_rexi_rejects_bad_precision():
    ctx = capi.Context(16)
    for kw in (dict(band_precision=capi.PREC_BF16), dict(precision=capi.PREC_BF16X3), dict(precision=7),
               dict(refine=2)):
        with pytest.raises(RuntimeError):
            ctx.rexi(capi.rexi_params(dt=3600.0, rexi_M=8, **kw))
    ctx.close()


L_FULL, M_FULL = 256, 188


@pytest.fixture(scope="module")
def full():
    ctx = capi.Context(L_FULL)
    yield ctx
    ctx.close()


def test_rexi_full_size_fp32_factors_gravity_modes(full):
    # no rotation: the solves are the diagonal Helmholtz inverses, stored in FP32;
    # h(dt) = C(w_l dt) h per degree with C the cos-part of the pole sum
    ctx = full
    print("rexi full-size gravity modes, FP32 factors", flush=True)
    ls, _ = real_spec_index(L_FULL)
    h = np.zeros(ctx.nlm, complex)
    for (l, m) in [(1, 0), (10, 3), (100, 50), (200, 7), (256, 256)]:
        h[m * (2 * L_FULL - m + 1) // 2 + l] = 1.0 if m == 0 else 0.6 + 0.8j
    z = np.zeros_like(h)
    dt = 3600.0
    params = capi.rexi_params(omega=0.0, visc=0.0, dt=dt, rexi_h=0.2, rexi_M=M_FULL, half=1,
                              band_precision=capi.PREC_FP32)
    out = _run(ctx, params, (h, z, z), 1)
    al, be = O.rexi_coeffs(0.2, M_FULL, True)
    w = np.sqrt(G * H0 * ls * (ls + 1.0)) / A_EARTH
    x = (w * dt)[:, None]
    C = np.sum(be * al / (al * al + x * x), -1).real
    assert rel_l2(out[0], C * h) < 1e-6


@pytest.mark.parametrize("leg,band", [("fp64", "fp32"), ("fp32", "fp64"), ("fp32", "fp32")])
def test_rexi_full_size_modes_vs_fp64(full, leg, band):
    # T256, 200 poles, rotation and viscosity, two steps: each reduced form
    # against the FP64 GPU step (itself within 1e-12 of closed forms here)
    ctx = full
    print(f"rexi full-size legendre={leg} factors={band} vs fp64", flush=True)
    x = _state(L_FULL, 21)
    ref = _run(ctx, capi.rexi_params(omega=OMEGA, visc=1e5, dt=3600.0, rexi_h=0.2, rexi_M=M_FULL, half=1), x, 2)
    params = capi.rexi_params(omega=OMEGA, visc=1e5, dt=3600.0, rexi_h=0.2, rexi_M=M_FULL, half=1,
                              precision=capi.PRECISIONS[leg], band_precision=capi.PRECISIONS[band])
    out = _run(ctx, params, x, 2)
    for k in range(3):
        e = rel_l2(out[k], ref[k])
        print(leg, band, "huv"[k], e)
        assert e < 4 * BOUNDS[(leg, band)]


def test_rexi_full_size_refined_fp32_factors_vs_fp64(full):
    # T256, 200 poles, rotation and viscosity, two steps: FP32 factors refined once
    ctx = full
    print("rexi full-size refined FP32 factors vs fp64", flush=True)
    x = _state(L_FULL, 21)
    ref = _run(ctx, capi.rexi_params(omega=OMEGA, visc=1e5, dt=3600.0, rexi_h=0.2, rexi_M=M_FULL, half=1), x, 2)
    params = capi.rexi_params(omega=OMEGA, visc=1e5, dt=3600.0, rexi_h=0.2, rexi_M=M_FULL, half=1,
                              band_precision=capi.PREC_FP32, refine=1)
    out = _run(ctx, params, x, 2)
    for k in range(3):
        e = rel_l2(out[k], ref[k])
        print("refined", "huv"[k], e)
        assert e < REFINED_BOUND
