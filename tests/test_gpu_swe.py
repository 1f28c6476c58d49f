"""Fused RK4 SWE step on the GPU against the oracle's op-by-op restatement of
AppSWE::p_run_euler_timestep_update + TimesteppingRK (RK4)."""
import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi
from tests._util import gaussian_dam, rel_l2, swe_prm

pytestmark = pytest.mark.gpu

# FP64 tolerances (relative L2 per spectral field).  The fused step forms some
# sums in physical space before one analysis where the reference analyses each
# term and sums spectrally, and it skips the round trips A(S(x)) of band-limited
# fields that the lazy reference performs; both differ only by round-off, which
# the nonlinear dynamics amplifies slowly with the number of steps.  u and v
# start at zero, so their relative error is measured on small fields.
# Measured sensitivity of the oracle itself: a 1e-16 relative perturbation of
# the T64 initial height grows to rel-L2 3e-11 (u) and 1.2e-10 (v) after 100
# steps; round-off injected at every step accumulates to about 100x that.
TOL_SHORT = 1e-10   # <= 5 steps
TOL_LONG = 1e-7     # 100 steps


def _run(L, nsteps, dt, visc=1e5, omega=7.292e-5, nonlinear=1):
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    h = gaussian_dam(orc.lon(), orc.ct)
    z = np.zeros_like(h)
    ref = orc.swe_rk4_run(swe_prm(dt, visc=visc, omega=omega, nonlinear=nonlinear), h, z, z, nsteps)
    state = ctx.to_device(ctx.analysis(np.stack([h, z, z])))
    ctx.swe_rk4(capi.swe_params(dt=dt, visc=visc, omega=omega, nonlinear=nonlinear), state, 1, nsteps)
    out = ctx.to_host(state, np.complex128, (3, ctx.nlm))
    ctx.close()
    return out, ref


@pytest.mark.parametrize("nonlinear", [1, 0])
def test_rk4_t31(nonlinear):
    out, ref = _run(31, 5, 240.0, nonlinear=nonlinear)
    for k, name in enumerate("huv"):
        e = rel_l2(out[k], ref[name])
        print(name, e)
        assert e < TOL_SHORT, name


def test_rk4_no_rotation_no_viscosity():
    out, ref = _run(31, 5, 240.0, visc=0.0, omega=0.0)
    for k, name in enumerate("huv"):
        e = rel_l2(out[k], ref[name])
        print(name, e)
        assert e < TOL_SHORT, name


def _diagnostics(L, out, a=6.37122e6, g=9.80616):
    """phi = g h, vort = (grad_lon v - grad_lat u)/a and div = (div_lon u +
    div_lat v)/a of a spectral (h, u, v) on the GPU through the C ABI, with the
    reference's projection points (AppSWE.hpp:85-95, SPHOperators.hpp:100-119,
    339-439, 507-529)."""
    ctx = capi.Context(L)
    try:
        h, u, v = out

        def glon(x):
            return ctx.spat_op(capi.SPAT_DIV_COSPHI, ctx.synthesis(ctx.spec_op(capi.OP_DIFF_LON, x))[0])

        def glat(x):
            return ctx.spat_op(capi.SPAT_DIV_COSPHI, ctx.synthesis(ctx.spec_op(capi.OP_ONE_MINUS_MU2_DMU, x))[0])

        def dlat(x):
            xs = ctx.analysis(ctx.spat_op(capi.SPAT_MUL_ST, ctx.synthesis(x)[0]))[0]
            return ctx.spat_op(capi.SPAT_DIV_ST, glat(xs))

        vort = ctx.analysis(glon(v) - glat(u))[0] / a
        div = ctx.analysis(glon(u) + dlat(v))[0] / a
        return g * h, vort, div
    finally:
        ctx.close()


def test_rk4_config0_t64():
    # nonlinear SWE, T64 on 98 x 196, Gaussian dam at (pi/4, pi), dt = 120 s,
    # 100 steps; the configuration compares geopotential, vorticity and
    # divergence spectra
    out, ref = _run(64, 100, 120.0)
    for k, name in enumerate("huv"):
        e = rel_l2(out[k], ref[name])
        print(name, e)
        assert e < TOL_LONG, name
    for x, name in zip(_diagnostics(64, out), ("phi", "vort", "div")):
        e = rel_l2(x, ref[name])
        print(name, e)
        assert e < TOL_LONG, name


def test_ensemble_members_are_independent():
    L = 31
    ctx = capi.Context(L)
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    states = []
    for k in range(3):
        h = gaussian_dam(lon, ctx.ct, lat_c=np.pi / 4 + 0.1 * k, amp=300.0 + 100 * k)
        z = np.zeros_like(h)
        states.append(ctx.analysis(np.stack([h, z, z])))
    prm = capi.swe_params(dt=300.0)
    # with one analysis kernel for every field count the members equal single
    # runs bit for bit; the default picks the kernel by field count (4x4x4 for
    # at most four fields), and the two kernels agree to round-off
    for variant, tol in ((1, 0.0), (0, 1e-13)):
        ctx.set_legendre_tuning(anal_variant=variant)
        batch = ctx.to_device(np.concatenate(states))
        ctx.swe_rk4(prm, batch, 3, 4)
        outb = ctx.to_host(batch, np.complex128, (9, ctx.nlm))
        for k in range(3):
            single = ctx.to_device(states[k])
            ctx.swe_rk4(prm, single, 1, 4)
            outs = ctx.to_host(single, np.complex128, (3, ctx.nlm))
            if tol == 0.0:
                assert np.array_equal(outs, outb[3 * k:3 * k + 3])
            else:
                for f in range(3):
                    assert rel_l2(outs[f], outb[3 * k + f]) < tol, (k, f)
    ctx.close()


def _members(ctx, n, seed=3):
    rng = np.random.default_rng(seed)
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    out = []
    for _ in range(n):
        h = gaussian_dam(lon, ctx.ct, lat_c=np.pi / 4 + rng.uniform(-0.2, 0.2),
                         lon_c=np.pi + rng.uniform(-0.2, 0.2), amp=rng.uniform(300, 700))
        z = np.zeros_like(h)
        out.append(ctx.analysis(np.stack([h, z, z])))
    return np.concatenate(out)


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_BF16X3])
def test_graph_replays_match_eager(prec):
    # one captured RK4 step replayed: every replay equals the eager steps bit
    # for bit, with eager work and synchronisations between the replays, across
    # a change of dt (the graph is rebuilt) and of the member count
    L = 31
    ctx = capi.Context(L)
    try:
        x0 = _members(ctx, 3)
        eager = ctx.to_device(x0)
        graph = ctx.to_device(x0)
        for k, (dt, nsteps) in enumerate([(300.0, 1), (300.0, 4), (150.0, 2), (300.0, 3)]):
            prm = capi.swe_params(dt=dt, precision=prec)
            ctx.swe_rk4(prm, eager, 3, nsteps)
            ctx.swe_rk4_graph(prm, graph, 3, nsteps)
            a = ctx.to_host(eager, np.complex128, (9, ctx.nlm))
            b = ctx.to_host(graph, np.complex128, (9, ctx.nlm))
            assert np.array_equal(a, b), k
        # a different member count on another buffer: rebuilt, still identical
        one_e = ctx.to_device(x0[:3])
        one_g = ctx.to_device(x0[:3])
        prm = capi.swe_params(dt=300.0, precision=prec)
        ctx.swe_rk4(prm, one_e, 1, 2)
        ctx.swe_rk4_graph(prm, one_g, 1, 2)
        assert np.array_equal(ctx.to_host(one_e, np.complex128, (3, ctx.nlm)),
                              ctx.to_host(one_g, np.complex128, (3, ctx.nlm)))
    finally:
        ctx.close()


def test_ensemble_config3_members_match_single_runs():
    # config 3 at reduced length: 64 members at T127 (192 x 384), perturbed dam
    # centres and amplitudes, advanced together (graph path), dt the reference's
    # rule 0.002 a / nlat (the configuration's 600 s is outside the RK4
    # stability interval of the fastest gravity wave); each member equals its
    # own single-member run to round-off
    L = 127
    ctx = capi.Context(L)
    try:
        x0 = _members(ctx, 64, seed=11)
        prm = capi.swe_params(dt=0.002 * 6.37122e6 / ctx.nlat)
        batch = ctx.to_device(x0)
        ctx.swe_rk4_graph(prm, batch, 64, 5)
        out = ctx.to_host(batch, np.complex128, (64 * 3, ctx.nlm))
        assert np.all(np.isfinite(out))
        for k in (0, 17, 63):
            single = ctx.to_device(x0[3 * k:3 * k + 3])
            ctx.swe_rk4(prm, single, 1, 5)
            ref = ctx.to_host(single, np.complex128, (3, ctx.nlm))
            for f in range(3):
                assert rel_l2(out[3 * k + f], ref[f]) < 1e-13, (k, f)
    finally:
        ctx.close()


def test_graph_rebuilt_after_workspace_growth():
    # regression: a step captured for 1 member, then an eager call for 4
    # members (the workspace is reallocated), then the 1-member graph again
    # with unchanged arguments: the captured step must be rebuilt on the new
    # workspace, and equal the eager steps bit for bit
    L = 31
    ctx = capi.Context(L)
    try:
        x0 = _members(ctx, 4)
        prm = capi.swe_params(dt=300.0)
        g = ctx.to_device(x0[:3])
        e = ctx.to_device(x0[:3])
        ctx.swe_rk4_graph(prm, g, 1, 1)
        big = ctx.to_device(x0)
        ctx.swe_rk4(prm, big, 4, 1)
        ctx.swe_rk4_graph(prm, g, 1, 2)
        ctx.swe_rk4(prm, e, 1, 3)
        assert np.array_equal(ctx.to_host(g, np.complex128, (3, ctx.nlm)), ctx.to_host(e, np.complex128, (3, ctx.nlm)))
    finally:
        ctx.close()


def test_graph_rebuilt_after_scratch_growth():
    # regression: a step captured for 1 member, then library calls that
    # reallocate the context's transform scratch (a 64-field analysis; a REXI
    # object, whose gradient operator runs the pole-layout analysis; the
    # autotuner), then the 1-member graph again with unchanged arguments: the
    # captured step must be rebuilt on the new buffers and equal the eager
    # steps bit for bit
    L = 31
    ctx = capi.Context(L)
    try:
        x0 = _members(ctx, 1)
        prm = capi.swe_params(dt=300.0)
        g = ctx.to_device(x0)
        e = ctx.to_device(x0)
        ctx.swe_rk4_graph(prm, g, 1, 1)
        ctx.swe_rk4(prm, e, 1, 1)
        rng = np.random.default_rng(5)
        ctx.analysis(rng.standard_normal((64, ctx.nspat)))
        ctx.swe_rk4_graph(prm, g, 1, 2)
        ctx.swe_rk4(prm, e, 1, 2)
        assert np.array_equal(ctx.to_host(g, np.complex128, (3, ctx.nlm)), ctx.to_host(e, np.complex128, (3, ctx.nlm)))
        rx = ctx.rexi(capi.rexi_params(dt=600.0, rexi_M=32))
        ctx.swe_rk4_graph(prm, g, 1, 1)
        ctx.swe_rk4(prm, e, 1, 1)
        assert np.array_equal(ctx.to_host(g, np.complex128, (3, ctx.nlm)), ctx.to_host(e, np.complex128, (3, ctx.nlm)))
        ctx.autotune()
        ctx.swe_rk4_graph(prm, g, 1, 2)
        ctx.swe_rk4(prm, e, 1, 2)
        assert np.array_equal(ctx.to_host(g, np.complex128, (3, ctx.nlm)), ctx.to_host(e, np.complex128, (3, ctx.nlm)))
        rx.close()
    finally:
        ctx.close()


def test_objects_outlive_their_context():
    # a REXI object and a band solver whose context is destroyed first: their
    # destroy calls free their own memory and touch no freed context; a
    # context created afterwards works and starts with an empty block pool
    ctx = capi.Context(15)
    rx = ctx.rexi(capi.rexi_params(dt=600.0, rexi_M=16))
    bd = ctx.band(2)
    bd.add(0, 1.0)   # solver_component_scalar_phi
    buf = ctx.empty(4096)
    buf.free()
    lib = ctx.lib
    ctx.close()
    assert lib.sphswe_rexi_destroy(rx.h) == 0
    rx.h = None
    assert lib.sphswe_band_destroy(bd.h) == 0
    bd.h = None
    ctx2 = capi.Context(15)
    try:
        q = np.zeros((1, ctx2.nlm), np.complex128)
        q[0, 1] = 1.0
        back = ctx2.analysis(ctx2.synthesis(q))
        assert rel_l2(back, q) < 1e-13
    finally:
        ctx2.close()


def test_nonlinear_remainder_vs_oracle():
    # nonlinear = 2: RK4 of the full tendency minus the linear one (the part a
    # split scheme leaves to RK4), against RK4 on the host with both tendencies
    # from the oracle; a jet with a height perturbation, 3 steps of 300 s
    from sph_swe_amd import workloads
    L = 31
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    x0 = workloads.galewsky_state(ctx, seed=7, pert_amp=50.0)
    # non-zonal winds, so that every advection term is active (a zonal jet with
    # v = 0 has no advection of u, and the increment would be pure round-off)
    x0[1] += 5.0 * workloads.seeded_spectra(L, 1, 8)[0]
    x0[2] += 5.0 * workloads.seeded_spectra(L, 1, 9)[0]
    dt, nsteps = 300.0, 3
    prm = dict(a=6.37122e6, g=9.80616, omega=7.292e-5, h0=10000.0, visc=0.0)

    def N(x):
        full = orc.swe_tendency(dict(prm, nonlinear=1), *x)
        lin = orc.swe_tendency(dict(prm, nonlinear=0), *x)
        return [f - l for f, l in zip(full, lin)]

    x = [c.copy() for c in x0]
    for _ in range(nsteps):
        k1 = N(x)
        k2 = N([a + 0.5 * dt * k for a, k in zip(x, k1)])
        k3 = N([a + 0.5 * dt * k for a, k in zip(x, k2)])
        k4 = N([a + dt * k for a, k in zip(x, k3)])
        x = [a + dt / 6.0 * (p + 2 * q + 2 * r + s) for a, p, q, r, s in zip(x, k1, k2, k3, k4)]
    st = ctx.to_device(x0)
    ctx.swe_rk4(capi.swe_params(dt=dt, visc=0.0, nonlinear=2), st, 1, nsteps)
    out = ctx.to_host(st, np.complex128, (3, ctx.nlm))
    # The increments are compared (the state itself is dominated by h0 and the
    # jet).  The host reference forms N as the difference of two tendencies whose
    # gravity and Coriolis terms are 10-100x the remainder, so its own round-off
    # relative to the increment is ~1e-11 per step; measured GPU-host agreement
    # 1e-10 (v) and below.
    for k in range(3):
        ref_step = x[k] - x0[k]
        assert rel_l2(out[k] - x0[k], ref_step) < 1e-9, k
    with pytest.raises(capi.SphSweError):
        ctx.swe_rk4(capi.swe_params(dt=dt, nonlinear=3), st, 1, 1)
    ctx.close()


@pytest.mark.parametrize("L,omega,visc", [(31, 7.292e-5, 1e5), (64, 7.292e-5, 0.0), (31, 0.0, 0.0)])
def test_rk4_robert_functions_vs_oracle(L, omega, visc):
    # linear equations with Robert-function velocities (AppSWE.hpp:606-620):
    # h tendency complete after the first analysis level, no metric division in
    # the u, v tendencies; 5 steps from a height bump with U, V bumps
    print(f"rk4 Robert functions vs oracle L={L} omega={omega} visc={visc}", flush=True)
    from tests.test_oracle_robert import _bump
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    h = (1e4 + 50.0 * _bump(orc, 0.2, 1.0, 0.3)).reshape(-1)
    U = (10.0 * _bump(orc, -0.1, 2.5, 0.35)).reshape(-1)
    V = (8.0 * _bump(orc, 0.3, 4.0, 0.3)).reshape(-1)
    dt, nsteps = 120.0, 5
    ref = orc.swe_rk4_run(dict(swe_prm(dt, visc=visc, omega=omega, nonlinear=0), robert=1), h, U, V, nsteps)
    x0 = ctx.analysis(np.stack([h, U, V]))
    st = ctx.to_device(x0)
    ctx.swe_rk4(capi.swe_params(dt=dt, visc=visc, omega=omega, nonlinear=0, robert=1), st, 1, nsteps)
    out = ctx.to_host(st, np.complex128, (3, ctx.nlm))
    for k, f in enumerate("huv"):
        e = rel_l2(out[k], ref[f])
        print("robert", f, e)
        assert e < TOL_SHORT, f
    # the standard formulation differs (the flag is not ignored) ...
    st = ctx.to_device(x0)
    ctx.swe_rk4(capi.swe_params(dt=dt, visc=visc, omega=omega, nonlinear=0), st, 1, nsteps)
    assert rel_l2(ctx.to_host(st, np.complex128, (3, ctx.nlm))[0], ref["h"]) > 1e-9
    # ... and the nonlinear equations read no flag, as in the reference
    runs = []
    for rb in (0, 1):
        st = ctx.to_device(x0)
        ctx.swe_rk4(capi.swe_params(dt=dt, visc=visc, omega=omega, nonlinear=1, robert=rb), st, 1, 2)
        runs.append(ctx.to_host(st, np.complex128, (3, ctx.nlm)))
    assert np.array_equal(runs[0], runs[1])
    with pytest.raises(capi.SphSweError):
        ctx.swe_rk4(capi.swe_params(dt=dt, nonlinear=0, robert=2), st, 1, 1)
    ctx.close()
