"""The largest configuration, T1023 on the 1536 x 3072 Gaussian grid, through
the C ABI: SHT round trips in FP64 and in the reduced-precision modes, single
modes against closed forms and scipy's orthonormal spherical harmonics
(Condon-Shortley phase, the convention of SPHConfig.hpp:338), REXI closed forms and linearity, and the
split step (REXI linear part + RK4 nonlinear remainder) at this size; the
200-pole rotating REXI step against the pointwise inertial oscillation.  The CPU
oracle would take minutes per transform here, so these are the size-independent
properties; the oracle comparisons run at the smaller truncations."""
import numpy as np
import pytest
from scipy.special import sph_harm_y

from oracle import oracle as O
from sph_swe_amd import capi, workloads
from tests._util import A_EARTH, G, H0, OMEGA, random_real_spec, real_spec_index, rel_l2

pytestmark = pytest.mark.gpu

L = 1023


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(L)
    yield c
    c.close()


def test_t1023_grid_and_round_trip(ctx):
    print("T1023 grid and FP64 round trip", flush=True)
    assert (ctx.nlat, ctx.nphi) == (1536, 3072)
    qs = np.stack([random_real_spec(L, seed=s, power=-1.5) for s in range(3)])
    q = ctx.analysis(ctx.synthesis(qs))
    for k in range(3):
        assert rel_l2(q[k], qs[k]) < 1e-12


@pytest.mark.parametrize("mode,tol", [(capi.PREC_FP32, 2e-6), (capi.PREC_BF16X3, 2e-6), (capi.PREC_BF16, 3e-5)])
def test_t1023_reduced_precision_modes(ctx, mode, tol):
    print(f"T1023 precision mode {mode}", flush=True)
    qs = np.stack([random_real_spec(L, seed=10 + s, power=-1.5) for s in range(2)])
    f64 = ctx.synthesis(qs)
    fm = ctx.synthesis(qs, precision=mode)
    qm = ctx.analysis(f64, precision=mode)
    for k in range(2):
        assert rel_l2(fm[k], f64[k]) < tol
        assert rel_l2(qm[k], qs[k]) < tol


def _ylm_ref(l, m, th, ph):
    """Y_l^m (orthonormal, Condon-Shortley phase) without scipy's overflow at
    high degree: m = 0 by numpy's Legendre series (Clenshaw), m = l by the closed
    form (-1)^l sqrt((2l+1)/(4 pi) (2l)!) / (2^l l!) sin^l(theta) e^{i l phi},
    other orders by scipy where it stays finite."""
    if m == 0:
        c = np.zeros(l + 1)
        c[l] = 1.0
        return np.sqrt((2 * l + 1) / (4 * np.pi)) * np.polynomial.legendre.legval(np.cos(th), c) + 0 * ph
    if m == l:
        from math import lgamma
        lc = 0.5 * (np.log(2 * l + 1) - np.log(4 * np.pi) + lgamma(2 * l + 1)) - l * np.log(2) - lgamma(l + 1)
        return (-1) ** l * np.exp(lc + l * np.log(np.sin(th))) * np.exp(1j * l * ph)
    return sph_harm_y(l, m, th, ph)


def test_t1023_single_modes_vs_closed_forms(ctx):
    # unit coefficients: f = Y_l^0, or 2 Re Y_l^m for m > 0 (real-field synthesis)
    print("T1023 single modes vs closed forms and scipy", flush=True)
    modes = [(1, 0), (1023, 0), (1023, 1023), (512, 1), (600, 100), (511, 300)]
    ls, ms = real_spec_index(L)
    q = np.zeros((len(modes), ctx.nlm), complex)
    for k, (l, m) in enumerate(modes):
        q[k, np.nonzero((ls == l) & (ms == m))[0][0]] = 1.0
    f = ctx.synthesis(q).reshape(len(modes), ctx.nphi, ctx.nlat)
    cols = np.array([0, 5, 777, 2048])
    th = np.arccos(np.asarray(ctx.ct))[None, :]
    ph = (2 * np.pi * cols / ctx.nphi)[:, None]
    for k, (l, m) in enumerate(modes):
        Y = _ylm_ref(l, m, th, ph)
        assert np.all(np.isfinite(Y)), (l, m)
        ref = Y.real if m == 0 else 2 * Y.real
        assert np.max(np.abs(f[k][cols] - ref)) < 1e-11 * np.max(np.abs(ref)), (l, m)


def test_t1023_rexi_gravity_modes():
    # no rotation, u = v = 0: h(dt) = C(w_l dt) h per degree, 200 poles
    print("T1023 REXI gravity modes (200 poles, no rotation)", flush=True)
    c = capi.Context(L)
    ls, ms = real_spec_index(L)
    h = np.zeros(c.nlm, complex)
    for (l, m) in [(1, 0), (100, 50), (511, 7), (1023, 1023)]:
        h[np.nonzero((ls == l) & (ms == m))[0][0]] = 1.0 if m == 0 else 0.6 + 0.8j
    z = np.zeros_like(h)
    dt = 300.0
    rx = c.rexi(capi.rexi_params(omega=0.0, visc=0.0, dt=dt, rexi_h=0.2, rexi_M=188, half=1))
    d = c.to_device(np.stack([h, z, z]))
    rx.step(d, 1)
    out = c.to_host(d, np.complex128, (3, c.nlm))
    rx.close()
    al, be = O.rexi_coeffs(0.2, 188, True)
    w = np.sqrt(G * H0 * ls * (ls + 1.0)) / A_EARTH
    x = (w * dt)[:, None]
    C = np.sum(be * al / (al * al + x * x), -1).real
    assert rel_l2(out[0], C * h) < 1e-12
    c.close()


def test_t1023_rexi_200_poles_with_rotation():
    # the rotating branch (band solves for phi and the winds at every order and
    # pole) with the 200 poles of the T1023 split configuration, dt = 300 s.
    # (1) h0 = 0, no viscosity: h stays 0 and (u, v) turn pointwise by the REXI
    # approximation of exp(dt f J), f = 2 Omega mu: per pole the wind solve is
    # (alpha^2 + f^2) u = alpha u0 - f v0, (alpha^2 + f^2) v = f u0 + alpha v0
    # (SWERexiSPH.hpp:222-229), so u(dt) = C u0 - S v0, v(dt) = S u0 + C v0 with
    # C(x) = Re sum beta alpha / (alpha^2 + x^2), S(x) = Re sum beta x / (alpha^2 + x^2),
    # x = f dt, the coefficients of REXI::setup(0.2, 188); winds of degree <= 64,
    # whose products with C and S are resolved at T1023 (C and S are analytic
    # for |mu| < 20).  C ~ cos x and S ~ -sin x: the physical rotation to 1e-6.
    print("T1023 REXI 200 poles with rotation: inertial oscillation and linearity", flush=True)
    c = capi.Context(L)
    dt, M = 300.0, 188
    ls, _ = real_spec_index(L)
    u0, v0 = (random_real_spec(L, seed=s, power=-1.5) for s in (41, 42))
    u0[ls > 64] = 0
    v0[ls > 64] = 0
    rx = c.rexi(capi.rexi_params(omega=OMEGA, h0=0.0, visc=0.0, dt=dt, rexi_h=0.2, rexi_M=M, half=1))
    assert rx.npoles == 200
    d = c.to_device(np.stack([np.zeros_like(u0), u0, v0]))
    rx.step(d, 1)
    out = c.to_host(d, np.complex128, (3, c.nlm))
    rx.close()
    uf, vf = c.synthesis(np.stack([u0, v0])).reshape(2, c.nphi, c.nlat)
    al, be = O.rexi_coeffs(0.2, M, True)
    x = (2 * OMEGA * dt * np.asarray(c.ct))[:, None]
    C = np.sum(be * al / (al * al + x * x), -1).real[None, :]
    S = np.sum(be * x / (al * al + x * x), -1).real[None, :]
    ref = c.analysis(np.stack([C * uf - S * vf, S * uf + C * vf]))
    eu, ev = rel_l2(out[1], ref[0]), rel_l2(out[2], ref[1])
    print("inertial oscillation rel L2", eu, ev, flush=True)
    assert eu < 1e-12 and ev < 1e-12
    assert np.max(np.abs(out[0])) == 0.0
    fx = 2 * OMEGA * dt * np.asarray(c.ct)[None, :]
    phys = c.analysis(np.stack([np.cos(fx) * uf + np.sin(fx) * vf, np.cos(fx) * vf - np.sin(fx) * uf]))
    assert rel_l2(out[1], phys[0]) < 2e-6 and rel_l2(out[2], phys[1]) < 2e-6
    # (2) the full operator (h0 = 10 km, viscosity 1e5, z1..z8): the step is linear
    rx = c.rexi(capi.rexi_params(omega=OMEGA, visc=1e5, dt=dt, rexi_h=0.2, rexi_M=M, half=1))
    x = np.stack([random_real_spec(L, 51 + s, power=-1.5) for s in range(3)])
    y = np.stack([random_real_spec(L, 61 + s, power=-1.5) for s in range(3)])

    def step(s):
        dd = c.to_device(s)
        rx.step(dd, 1)
        return c.to_host(dd, np.complex128, (3, c.nlm))

    sx, sy, sxy = step(x), step(y), step(x - 3 * y)
    for k in range(3):
        e = rel_l2(sxy[k], sx[k] - 3 * sy[k])
        print("linearity", "huv"[k], e, flush=True)
        assert e < 3e-12   # round-off of the 200-term pole sum (measured 5e-13 for h)
    rx.close()
    c.close()


def test_t1023_rexi_and_split_linearity():
    # with rotation and viscosity, 20 poles: the REXI step is linear
    # (step(x + 2y) = step(x) + 2 step(y)); then two split steps of the Galewsky
    # state with FP64 and with BF16X3 transforms in the RK4 part agree to the
    # reduced mode's rounding
    print("T1023 REXI linearity with rotation (20 poles) and split step", flush=True)
    c = capi.Context(L)
    dt = 120.0
    rx = c.rexi(capi.rexi_params(omega=OMEGA, visc=1e5, dt=dt, rexi_h=0.2, rexi_M=8, half=1))
    x = np.stack([random_real_spec(L, 21 + s, power=-1.5) for s in range(3)])
    y = np.stack([random_real_spec(L, 31 + s, power=-1.5) for s in range(3)])

    def step(s):
        d = c.to_device(s)
        rx.step(d, 1)
        return c.to_host(d, np.complex128, (3, c.nlm))

    sx, sy, sxy = step(x), step(y), step(x + 2 * y)
    for k in range(3):
        assert rel_l2(sxy[k], sx[k] + 2 * sy[k]) < 1e-12
    # the Galewsky state: two split steps in FP64 and with BF16X3 transforms in the RK4 part
    x0 = workloads.galewsky_state(c, seed=1023, pert_amp=1.0)
    outs = []
    for prec in (capi.PREC_FP64, capi.PREC_BF16X3):
        d = c.to_device(x0)
        rx.split_step(capi.swe_params(dt=dt, precision=prec), d, 2)
        outs.append(c.to_host(d, np.complex128, (3, c.nlm)))
    assert np.all(np.isfinite(outs[0]))
    # h and u against their own size; v (zero in the jet, small after two steps)
    # against the wind scale |u|, since the rounding of the transforms is relative
    # to the fields transformed (unit roundoff 2^-24 of the BF16X3 products)
    assert rel_l2(outs[1][0], outs[0][0]) < 1e-7
    assert rel_l2(outs[1][1], outs[0][1]) < 1e-7
    assert np.linalg.norm(outs[1][2] - outs[0][2]) < 1e-7 * np.linalg.norm(outs[0][1])
    rx.close()
    c.close()
