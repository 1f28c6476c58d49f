"""Benchmark of the spectral SWE hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one SHT round trip (synthesis + analysis) of 3 fields (phi, u, v) at
T511 on the 768 x 1536 Gaussian grid, FP64, spectral inputs already resident
in HBM, seeded coefficients with power ~ l^-3.  The JSON line also carries the
same round trip with the Legendre contractions in FP32, BF16 (hi/lo), BF16
(hi/mid/lo) and the mixed form (bf16 hi/lo products from stored tables; in the
SWE stepper also FP32 Fourier rows, FFTs and pointwise terms) with per-field
errors, the nonlinear SWE RK4 rate at T511
(Gaussian dam, dt = 0.001 a / nlat = 8.296 s) in FP64 and in each reduced mode
with the geopotential error after 100 steps,
the linear REXI rate at T256 (200 poles, dt = 3600 s, seeded (phi, vort, div)
start state) in FP64 and its precision forms, the T127 ensemble and the batched
T511 form, the split scheme at T1023 (REXI linear part, RK4 nonlinear
remainder, Galewsky jet), the roofline of the dominant kernel (the Legendre analysis)
measured with HIP events on the library's stream, and the CPU oracle timed on a
bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from sph_swe_amd import capi, workloads  # noqa: E402

L_BENCH = 511
NFIELDS = 3
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def seeded_spec(L, nf, seed=1234):
    rng = np.random.default_rng(seed)
    out = []
    for f in range(nf):
        qs = []
        for m in range(L + 1):
            l = np.arange(m, L + 1)
            amp = np.where(l > 0, np.maximum(l, 1) ** -1.5, 0.0)
            q = amp * np.exp(1j * rng.uniform(0, 2 * np.pi, l.size))
            if m == 0:
                q = q.real.astype(complex)
            qs.append(q)
        out.append(np.concatenate(qs))
    return np.stack(out)


def gaussian_dam(lon, ct, lat_c=np.pi / 4, lon_c=np.pi, amp=500.0, width=0.1, h0=10000.0):
    phi1 = np.arcsin(ct)[None, :]
    c = np.sin(phi1) * np.sin(lat_c) + np.cos(phi1) * np.cos(lat_c) * np.cos(np.asarray(lon)[:, None] - lon_c)
    d = np.arccos(np.clip(c, -1, 1))
    return (h0 + amp * np.exp(-(d / width) ** 2)).reshape(-1)


def bench_sht(ctx, steps, warmup):
    lib, h = ctx.lib, ctx.h
    spec = ctx.to_device(seeded_spec(ctx.L, NFIELDS))
    spat = ctx.empty(NFIELDS * ctx.nspat * 8)
    spec2 = ctx.empty(NFIELDS * ctx.nlm * 16)

    def round_trip():
        ctx._check(lib.sphswe_sht_synthesis(h, spec.ptr, spat.ptr, NFIELDS, capi.PREC_FP64))
        ctx._check(lib.sphswe_sht_analysis(h, spat.ptr, spec2.ptr, NFIELDS, capi.PREC_FP64))

    for _ in range(warmup):
        round_trip()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.mark(0)
    for _ in range(steps):
        round_trip()
    ctx.mark(1)
    ctx.sync()
    wall = time.perf_counter() - t0
    dev_ms = ctx.elapsed_ms(0, 1)
    # accuracy of the round trip (band-limited input -> identity)
    q_in = ctx.to_host(spec, np.complex128, (NFIELDS, ctx.nlm))
    q_out = ctx.to_host(spec2, np.complex128, (NFIELDS, ctx.nlm))
    err = max(np.linalg.norm(q_out[k] - q_in[k]) / np.linalg.norm(q_in[k]) for k in range(NFIELDS))
    return wall, dev_ms, err


def bench_legendre_kernel(ctx, reps=20):
    """Average duration of one Legendre synthesis launch (3 fields), HIP events
    on the library stream, and its algorithmic FLOP count."""
    lib, h = ctx.lib, ctx.h
    spec = ctx.to_device(seeded_spec(ctx.L, NFIELDS, seed=7))
    four = ctx.empty(NFIELDS * (ctx.L + 1) * ctx.nlat * 16)
    for _ in range(3):
        ctx._check(lib.sphswe_legendre_synthesis(h, spec.ptr, four.ptr, NFIELDS, capi.PREC_FP64))
    ctx.mark(2)
    for _ in range(reps):
        ctx._check(lib.sphswe_legendre_synthesis(h, spec.ptr, four.ptr, NFIELDS, capi.PREC_FP64))
    ctx.mark(3)
    ms = ctx.elapsed_ms(2, 3) / reps
    flops = 4.0 * ctx.nlm * (ctx.nlat // 2) * NFIELDS
    # analysis kernel as well
    ctx.mark(4)
    for _ in range(reps):
        ctx._check(lib.sphswe_legendre_analysis(h, four.ptr, spec.ptr, NFIELDS, capi.PREC_FP64))
    ctx.mark(5)
    ms_a = ctx.elapsed_ms(4, 5) / reps
    return ms, ms_a, flops


def bench_bt_contractions(ctx, reps=10):
    """The mixed form's Legendre contractions at the batched T511 widths
    (144-field synthesis, 64-field analysis: 16 members at level 0 and 1),
    through the FP64-row entry points: the operand split and the GEMM on the
    stored bf16 hi/lo tables, average time with HIP events.  achieved counts the
    bf16 MFMA work issued, three products per FP64-equivalent multiply-add
    (2 * 3 * nlm * nlat/2 * 2 * nf), against the 2.5 PFLOP/s dense bf16 peak."""
    lib, h = ctx.lib, ctx.h
    out = {}
    for name, nf, synth in (("synthesis_144", 144, True), ("analysis_64", 64, False)):
        spec = ctx.empty(nf * ctx.nlm * 16)
        four = ctx.empty(nf * (ctx.L + 1) * ctx.nlat * 16)
        ctx._check(lib.sphswe_fill_zero(h, spec.ptr, nf * ctx.nlm * 16))
        ctx._check(lib.sphswe_fill_zero(h, four.ptr, nf * (ctx.L + 1) * ctx.nlat * 16))
        if synth:
            fn = lambda: lib.sphswe_legendre_synthesis(h, spec.ptr, four.ptr, nf, capi.PREC_MIXED)  # noqa: E731
        else:
            fn = lambda: lib.sphswe_legendre_analysis(h, four.ptr, spec.ptr, nf, capi.PREC_MIXED)  # noqa: E731
        for _ in range(3):
            ctx._check(fn())
        ctx.mark(2)
        for _ in range(reps):
            ctx._check(fn())
        ctx.mark(3)
        ms = ctx.elapsed_ms(2, 3) / reps
        fl = 4.0 * ctx.nlm * (ctx.nlat // 2) * nf
        out[name] = {"avg_ms": ms, "fp64_equivalent_tflops": fl / (ms * 1e-3) / 1e12,
                     "achieved": 3 * fl / (ms * 1e-3) / 1e12, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": 3 * fl / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS}
        spec.free()
        four.free()
    return dict(out, engine="bf16 matrix cores (v_mfma_f32_16x16x32_bf16), hi/lo operands, 3 products",
                includes="the operand split pass from FP64 spectra / Fourier rows, FP64 Fourier rows")


def bench_fft_kernels(ctx, nf=9, reps=20):
    """Longitude FFTs of nf fields at the context's grid (the reference's
    [i * nlat + j] layout, through the C ABI): average launch time with HIP
    events and achieved HBM bandwidth on the algorithmic bytes, 8 nlat nphi
    (physical) + 16 (L + 1) nlat (Fourier rows) per field."""
    lib, h = ctx.lib, ctx.h
    spat = ctx.empty(nf * ctx.nspat * 8)
    four = ctx.empty(nf * (ctx.L + 1) * ctx.nlat * 16)
    ctx._check(lib.sphswe_fill_zero(h, spat.ptr, nf * ctx.nspat * 8))
    ctx._check(lib.sphswe_fill_zero(h, four.ptr, nf * (ctx.L + 1) * ctx.nlat * 16))
    bytes_per = float(nf * (8 * ctx.nspat + 16 * (ctx.L + 1) * ctx.nlat))
    out = {"fields": nf, "bytes_per_launch": bytes_per, "peak_gbs": HBM_PEAK_GBS}
    for name, fn in (("r2c", lambda: lib.sphswe_fft_analysis(h, spat.ptr, four.ptr, nf)),
                     ("c2r", lambda: lib.sphswe_fft_synthesis(h, four.ptr, spat.ptr, nf))):
        for _ in range(3):
            ctx._check(fn())
        ctx.mark(2)
        for _ in range(reps):
            ctx._check(fn())
        ctx.mark(3)
        ms = ctx.elapsed_ms(2, 3) / reps
        gbs = bytes_per / (ms * 1e-3) / 1e9
        out[name] = {"avg_ms": ms, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS}
    return out


def bench_swe(ctx, steps, warmup):
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    hgt = gaussian_dam(lon, ctx.ct)
    z = np.zeros_like(hgt)
    spec = ctx.analysis(np.stack([hgt, z, z]))
    state = ctx.to_device(spec)
    dt = 0.001 * 6.37122e6 / ctx.nlat
    prm = capi.swe_params(dt=dt)
    ctx.swe_rk4(prm, state, 1, warmup)
    ctx.sync()
    ctx.mark(6)
    ctx.swe_rk4(prm, state, 1, steps)
    ctx.mark(7)
    ms = ctx.elapsed_ms(6, 7)
    out = ctx.to_host(state, np.complex128, (3, ctx.nlm))
    if not np.all(np.isfinite(out)):
        raise RuntimeError("SWE state became non-finite")
    return steps / (ms / 1000.0), dt


def bench_rexi(steps, L=256, M=188):
    """Linear SWE by REXI: seeded start state, 200 poles, dt = 3600 s."""
    ctx = capi.Context(L)
    tuning = ctx.autotune()
    state = ctx.to_device(workloads.rexi_state(ctx))
    prm = capi.rexi_params(dt=3600.0, rexi_h=0.2, rexi_M=M, half=1, visc=1e5)
    t0 = time.perf_counter()
    rx = ctx.rexi(prm)
    ctx.sync()
    setup_s = time.perf_counter() - t0
    rx.step(state, 1)
    ctx.sync()
    ctx.mark(8)
    rx.step(state, steps)
    ctx.mark(9)
    ms = ctx.elapsed_ms(8, 9)
    out = ctx.to_host(state, np.complex128, (3, ctx.nlm))
    if not np.all(np.isfinite(out)):
        raise RuntimeError("REXI state became non-finite")
    cfg = {"lmax": L, "nlat": ctx.nlat, "nphi": ctx.nphi, "rexi_M": M, "rexi_h": 0.2, "poles": rx.npoles,
           "dt": 3600.0, "steps": steps, "viscosity": 1e5, "setup_s": setup_s, "autotune": tuning}
    rx.close()
    # the same workload with Robert-function velocities (SWERexiSPHRobert): the
    # gradients are spectral, so the per-pole gradient operator is not applied
    rxr = ctx.rexi(capi.rexi_params(dt=3600.0, rexi_h=0.2, rexi_M=M, half=1, visc=1e5, robert=1))
    state = ctx.to_device(workloads.rexi_state(ctx))
    rxr.step(state, 1)
    ctx.sync()
    ctx.mark(8)
    rxr.step(state, steps)
    ctx.mark(9)
    cfg["robert_functions_steps_per_s"] = steps / (ctx.elapsed_ms(8, 9) / 1000.0)
    if not np.all(np.isfinite(ctx.to_host(state, np.complex128, (3, ctx.nlm)))):
        raise RuntimeError("Robert REXI state became non-finite")
    rxr.close()
    ctx.close()
    return steps / (ms / 1000.0), cfg


MODES = ("fp32", "bf16", "bf16x3", "mixed")
BF16_PEAK_TFLOPS = 2500.0

# REXI precision forms: (precision of the per-pole gradient operator, storage
# precision of the band LU factors)
REXI_MODES = (("fp64", "fp32"), ("fp64", "fp32+refine"), ("fp32", "fp64"), ("fp32", "fp32"),
              ("fp32", "fp32+refine"))


def bench_rexi_modes(steps, L=256, M=188):
    """Config 1 (T256, 200 poles, dt = 3600 s, seeded start state) with the
    per-pole gradient operator on the FP32 matrix cores and/or the band LU
    factors stored in FP32 (solves in FP64): steps/s and the relative L2 error
    of h, u and v after the run against the FP64 run from the same start.
    "+refine": FP32 factors with one sweep of FP64 iterative refinement."""
    ctx = capi.Context(L)
    ctx.autotune()
    x0 = workloads.rexi_state(ctx)
    res, finals = {}, {}
    for leg, band in (("fp64", "fp64"),) + REXI_MODES:
        name = leg + "_gradient_" + band + "_factors"
        bp = band.split("+")[0]
        prm = capi.rexi_params(dt=3600.0, rexi_h=0.2, rexi_M=M, half=1, visc=1e5,
                               precision=capi.PRECISIONS[leg], band_precision=capi.PRECISIONS[bp],
                               refine=int(band.endswith("+refine")))
        rx = ctx.rexi(prm)
        state = ctx.to_device(x0)
        rx.step(state, 1)
        state = ctx.to_device(x0)
        ctx.sync()
        ctx.mark(8)
        rx.step(state, steps)
        ctx.mark(9)
        ms = ctx.elapsed_ms(8, 9)
        rx.close()
        finals[name] = ctx.to_host(state, np.complex128, (3, ctx.nlm))
        if not np.all(np.isfinite(finals[name])):
            raise RuntimeError("REXI state became non-finite (%s)" % name)
        res[name] = {"steps_per_s": steps / (ms / 1000.0)}
    ref = "fp64_gradient_fp64_factors"
    for name in res:
        if name == ref:
            continue
        for k, f in enumerate("huv"):
            res[name][f + "_rel_l2"] = _rel(finals[name][k], finals[ref][k])[0]
        res[name]["speedup_vs_fp64"] = res[name]["steps_per_s"] / res[ref]["steps_per_s"]
    ctx.close()
    return {"lmax": L, "poles_M": M, "steps": steps, "dt": 3600.0, "modes": res}


def _rel(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref)), float(np.abs(x - ref).max() / np.abs(ref).max())


def bench_sht_modes(ctx, steps, warmup):
    """SHT round trip of 3 fields at T511 with the Legendre contractions in each
    reduced mode: round trips/s (HIP events) and, per field, the relative L2 and
    Linf errors of the synthesis and of the analysis against the FP64 GPU path
    (itself within 1e-13 of the FP64 oracle)."""
    lib, h = ctx.lib, ctx.h
    q = seeded_spec(ctx.L, NFIELDS)
    spec = ctx.to_device(q)
    spat = ctx.empty(NFIELDS * ctx.nspat * 8)
    spec2 = ctx.empty(NFIELDS * ctx.nlm * 16)
    ref_f = ctx.synthesis(q)
    ref_q = ctx.analysis(ref_f)
    out = {}
    for name in MODES:
        prec = capi.PRECISIONS[name]

        def round_trip():
            ctx._check(lib.sphswe_sht_synthesis(h, spec.ptr, spat.ptr, NFIELDS, prec))
            ctx._check(lib.sphswe_sht_analysis(h, spat.ptr, spec2.ptr, NFIELDS, prec))

        for _ in range(warmup):
            round_trip()
        ctx.sync()
        ctx.mark(10)
        for _ in range(steps):
            round_trip()
        ctx.mark(11)
        ms = ctx.elapsed_ms(10, 11)
        f = ctx.synthesis(q, prec)
        qa = ctx.analysis(ref_f, prec)
        es = [_rel(f[k], ref_f[k]) for k in range(NFIELDS)]
        ea = [_rel(qa[k], ref_q[k]) for k in range(NFIELDS)]
        out[name] = {"round_trips_per_s": steps / (ms / 1000.0),
                     "synthesis_rel_l2": [e[0] for e in es], "synthesis_rel_linf": [e[1] for e in es],
                     "analysis_rel_l2": [e[0] for e in ea], "analysis_rel_linf": [e[1] for e in ea]}
    return out


def bench_swe_modes(ctx, nsteps=100):
    """Nonlinear SWE at T511 (Gaussian dam, 100 steps) in FP64 and with the
    Legendre contractions in each reduced mode: steps/s and the relative L2
    error of the geopotential spectrum (phi = g h) after the run against the
    FP64 GPU run.  dt = 0.0005 a / nlat = 4.148 s, half the reference's rule:
    at 8.296 s the m = 1 polar mode of the scalar-wind formulation grows by
    about 1.6 per step and the FP64 run itself overflows before step 100
    (stable up to about 7.5 s), so errors there would measure that growth."""
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    hgt = gaussian_dam(lon, ctx.ct)
    z = np.zeros_like(hgt)
    x0 = ctx.analysis(np.stack([hgt, z, z]))
    dt = 0.0005 * 6.37122e6 / ctx.nlat
    res, finals = {}, {}
    for name in ("fp64",) + MODES:
        prm = capi.swe_params(dt=dt, precision=capi.PRECISIONS[name])
        warm = ctx.to_device(x0)
        ctx.swe_rk4(prm, warm, 1, 2)
        state = ctx.to_device(x0)
        ctx.sync()
        ctx.mark(12)
        ctx.swe_rk4(prm, state, 1, nsteps)
        ctx.mark(13)
        ms = ctx.elapsed_ms(12, 13)
        finals[name] = ctx.to_host(state, np.complex128, (3, ctx.nlm))
        if not np.all(np.isfinite(finals[name])):
            raise RuntimeError("SWE state became non-finite (%s)" % name)
        res[name] = {"steps_per_s": nsteps / (ms / 1000.0)}
    for name in MODES:
        res[name]["phi_rel_l2"] = _rel(finals[name][0], finals["fp64"][0])[0]
        res[name]["u_rel_l2"] = _rel(finals[name][1], finals["fp64"][1])[0]
        res[name]["v_rel_l2"] = _rel(finals[name][2], finals["fp64"][2])[0]
        res[name]["speedup_vs_fp64"] = res[name]["steps_per_s"] / res["fp64"]["steps_per_s"]
    return {"steps": nsteps, "dt": dt, "modes": res}


def bench_swe_batched(ctx, nmem=16, nsteps=100):
    """The north star's batched form at T511: nmem nonlinear SWE members
    (Gaussian dams, centres offset by U(-0.2, 0.2) rad, amplitudes U(300, 700) m,
    seeded) advanced together by the captured RK4 step, in FP64 and in each
    reduced mode, against the plain form: one FP64 state stepped eagerly.
    Reports member-steps/s, the speedup over the plain form's steps/s, and the
    largest member's relative L2 geopotential error after nsteps against the
    batched FP64 run (member 0 of which equals the plain run).  dt = 0.0005 a /
    nlat = 4.148 s, as the mixed T511 leg."""
    dt = 0.0005 * 6.37122e6 / ctx.nlat
    rng = np.random.default_rng(4242)
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    x0 = []
    for _ in range(nmem):
        h = gaussian_dam(lon, ctx.ct, lat_c=np.pi / 4 + rng.uniform(-0.2, 0.2), lon_c=np.pi + rng.uniform(-0.2, 0.2),
                         amp=rng.uniform(300, 700))
        z = np.zeros_like(h)
        x0.append(ctx.analysis(np.stack([h, z, z])))
    x0 = np.concatenate(x0)
    # plain form: member 0 alone, FP64, eager
    prm = capi.swe_params(dt=dt, precision=capi.PREC_FP64)
    one = ctx.to_device(x0[:3])
    ctx.swe_rk4(prm, one, 1, 2)
    one = ctx.to_device(x0[:3])
    ctx.sync()
    ctx.mark(12)
    ctx.swe_rk4(prm, one, 1, nsteps)
    ctx.mark(13)
    plain = nsteps / (ctx.elapsed_ms(12, 13) / 1000.0)
    plain_final = ctx.to_host(one, np.complex128, (3, ctx.nlm))
    del one
    res, finals = {}, {}
    for name in ("fp64",) + MODES:
        prm = capi.swe_params(dt=dt, precision=capi.PRECISIONS[name])
        state = ctx.to_device(x0)
        ctx.swe_rk4_graph(prm, state, nmem, 2)
        ctx._check(ctx.lib.sphswe_memcpy_h2d(ctx.h, state.ptr, np.ascontiguousarray(x0).ctypes.data, x0.nbytes))
        ctx.sync()
        ctx.mark(12)
        ctx.swe_rk4_graph(prm, state, nmem, nsteps)
        ctx.mark(13)
        ms = ctx.elapsed_ms(12, 13)
        finals[name] = ctx.to_host(state, np.complex128, (nmem * 3, ctx.nlm))
        del state
        if not np.all(np.isfinite(finals[name])):
            raise RuntimeError("batched SWE state became non-finite (%s)" % name)
        rate = nmem * nsteps / (ms / 1000.0)
        res[name] = {"member_steps_per_s": rate, "ms_per_step": ms / nsteps, "speedup_vs_plain_fp64": rate / plain}
    ref = finals["fp64"]
    res["fp64"]["member0_vs_plain_phi_rel_l2"] = _rel(ref[0], plain_final[0])[0]
    for name in MODES:
        res[name]["phi_rel_l2_max_member"] = max(_rel(finals[name][3 * k], ref[3 * k])[0] for k in range(nmem))
    return {"members": nmem, "steps": nsteps, "dt": dt, "plain_fp64_steps_per_s": plain, "runs": res}


def bench_ensemble(nsteps=200, nmem=64, L=127):
    """Config 3: 64 nonlinear SWE members at T127 (192 x 384), 200 steps, each
    a Gaussian dam with centre offsets U(-0.2, 0.2) rad and amplitude
    U(300, 700) m (seeded), all members in the same launches.  FP64 stepped
    eagerly and with the captured step; the reduced modes with the captured
    step, with the largest relative L2 error of a member's geopotential
    spectrum against its FP64 run.  dt is the reference's own rule at this
    grid, 0.002 a / nlat = 66.4 s: with the 600 s of the configuration the
    fastest gravity wave has omega dt = 3.8, outside the RK4 stability interval
    (2.83), and every dt from 120 s up overflows within 200 steps (the m = 1
    polar mode of the scalar-wind formulation)."""
    ctx = capi.Context(L)
    tuning = ctx.autotune()
    dt = 0.002 * 6.37122e6 / ctx.nlat
    rng = np.random.default_rng(2024)
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    x0 = []
    for _ in range(nmem):
        h = gaussian_dam(lon, ctx.ct, lat_c=np.pi / 4 + rng.uniform(-0.2, 0.2), lon_c=np.pi + rng.uniform(-0.2, 0.2),
                         amp=rng.uniform(300, 700))
        z = np.zeros_like(h)
        x0.append(ctx.analysis(np.stack([h, z, z])))
    x0 = np.concatenate(x0)
    res, finals = {}, {}
    runs = [("fp64_eager", "fp64", False), ("fp64", "fp64", True)] + [(m, m, True) for m in MODES]
    for key, name, graph in runs:
        prm = capi.swe_params(dt=dt, precision=capi.PRECISIONS[name])
        step = ctx.swe_rk4_graph if graph else ctx.swe_rk4
        warm = ctx.to_device(x0)
        step(prm, warm, nmem, 2)
        state = ctx.to_device(x0) if not graph else warm
        if graph:
            # the captured step is bound to its buffer: restart it from x0
            ctx._check(ctx.lib.sphswe_memcpy_h2d(ctx.h, state.ptr, np.ascontiguousarray(x0).ctypes.data, x0.nbytes))
        ctx.sync()
        ctx.mark(14)
        step(prm, state, nmem, nsteps)
        ctx.mark(15)
        ms = ctx.elapsed_ms(14, 15)
        finals[key] = ctx.to_host(state, np.complex128, (nmem * 3, ctx.nlm))
        if not np.all(np.isfinite(finals[key])):
            raise RuntimeError("ensemble state became non-finite (%s)" % key)
        res[key] = {"member_steps_per_s": nmem * nsteps / (ms / 1000.0), "ms_per_step": ms / nsteps,
                    "hip_graph": graph}
    ref = finals["fp64_eager"]
    for key in res:
        if key == "fp64_eager":
            continue
        res[key]["phi_rel_l2_max_member"] = max(
            float(np.linalg.norm(finals[key][3 * k] - ref[3 * k]) / np.linalg.norm(ref[3 * k])) for k in range(nmem))
        res[key]["speedup_vs_fp64_eager"] = res[key]["member_steps_per_s"] / res["fp64_eager"]["member_steps_per_s"]
    cfg = {"lmax": L, "nlat": ctx.nlat, "nphi": ctx.nphi, "members": nmem, "dt": dt, "steps": nsteps,
           "viscosity": 1e5, "autotune": tuning}
    ctx.close()
    return {"config": cfg, "runs": res}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(L, ctx_nlat, ctx_nphi, runs=5):
    """The oracle (the C restatement of the reference's SHT path, OpenMP over
    rings and orders as SHTNS runs over all cores) on the host: the median of
    `runs` timed SHT round trips with all the cores this process may use
    (3 fields) and with one thread (1 field, scaled to 3), after one warm-up."""
    from oracle import oracle as O
    orc = O.Oracle(L, ctx_nlat, ctx_nphi)
    spec = seeded_spec(L, NFIELDS)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))

    def timed(nf):
        t0 = time.perf_counter()
        for k in range(nf):
            orc.analysis(orc.synthesis(spec[k]))
        return time.perf_counter() - t0

    O.set_threads(cores)
    timed(1)
    multi = sorted(timed(NFIELDS) for _ in range(runs))[runs // 2]
    O.set_threads(1)
    single = sorted(timed(1) for _ in range(runs))[runs // 2] * NFIELDS
    O.set_threads(cores)
    model = _cpu_model()
    return {"value": 1.0 / multi, "unit": "round_trips/s", "cores": cores, "kind": "port",
            "cpu_model": model,
            "sample": "median of %d round trips of 3 fields at T%d (%dx%d), oracle/sphswe_oracle.c with %d OpenMP "
                      "threads, %.2f s each" % (runs, L, ctx_nphi, ctx_nlat, cores, multi),
            "single_thread": {"value": 1.0 / single, "unit": "round_trips/s", "cores": 1,
                              "sample": "median of %d one-field round trips x 3, %.2f s per 3-field round trip"
                                        % (runs, single)}}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "leg_anal_T511_3fields_hbm.json")
FFT_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fft_T511_9fields_hbm.json")


def pmc_kernels():
    """Counter-collection workload: 5 Legendre syntheses and 5 analyses (3 fields,
    T511), 5 real FFTs each way of 9 fields (the reference layout, default tile
    width) and 5 launches of a copy-shaped kernel of known size (k_axpby with
    y = 0: 64 MiB read, 64 MiB written) that calibrates FETCH_SIZE / WRITE_SIZE."""
    ctx = capi.Context(L_BENCH)
    spec = ctx.to_device(seeded_spec(ctx.L, NFIELDS, seed=7))
    four = ctx.empty(NFIELDS * (ctx.L + 1) * ctx.nlat * 16)
    nff = 9
    fspat = ctx.to_device(np.random.default_rng(8).standard_normal(nff * ctx.nspat))
    ffour = ctx.empty(nff * (ctx.L + 1) * ctx.nlat * 16)
    for _ in range(5):
        ctx._check(ctx.lib.sphswe_fft_analysis(ctx.h, fspat.ptr, ffour.ptr, nff))
    for _ in range(5):
        ctx._check(ctx.lib.sphswe_fft_synthesis(ctx.h, ffour.ptr, fspat.ptr, nff))
    ncopy = 1 << 22
    src = ctx.to_device(np.ones(ncopy, np.complex128))
    dst = ctx.empty(ncopy * 16)
    for _ in range(5):
        ctx._check(ctx.lib.sphswe_legendre_synthesis(ctx.h, spec.ptr, four.ptr, NFIELDS, capi.PREC_FP64))
    for _ in range(5):
        ctx._check(ctx.lib.sphswe_legendre_analysis(ctx.h, four.ptr, spec.ptr, NFIELDS, capi.PREC_FP64))
    for _ in range(5):
        ctx._check(ctx.lib.sphswe_spec_axpby(ctx.h, ncopy, 1.0, 0.0, src.ptr, 0.0, 0.0, None, dst.ptr))
    ctx.sync()
    ctx.close()
    print("pmc kernels done", flush=True)


def mixed_swe_kernels(nmem=16, nsteps=3):
    """Trace and counter workload of the mixed form: nmem T511 members (the
    batched leg's Gaussian dams) stepped eagerly (no captured graph) for one
    warm-up and nsteps RK4 steps in SPHSWE_PREC_MIXED, then 5 launches of the
    copy-shaped calibration kernel (k_axpby, 64 MiB read, 64 MiB written)."""
    ctx = capi.Context(L_BENCH)
    dt = 0.0005 * 6.37122e6 / ctx.nlat
    rng = np.random.default_rng(4242)
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    x0 = []
    for _ in range(nmem):
        h = gaussian_dam(lon, ctx.ct, lat_c=np.pi / 4 + rng.uniform(-0.2, 0.2), lon_c=np.pi + rng.uniform(-0.2, 0.2),
                         amp=rng.uniform(300, 700))
        z = np.zeros_like(h)
        x0.append(ctx.analysis(np.stack([h, z, z])))
    state = ctx.to_device(np.concatenate(x0))
    prm = capi.swe_params(dt=dt, precision=capi.PREC_MIXED)
    ctx.swe_rk4(prm, state, nmem, 1)
    ctx.sync()
    ctx.swe_rk4(prm, state, nmem, nsteps)
    ncopy = 1 << 22
    src = ctx.to_device(np.ones(ncopy, np.complex128))
    dst = ctx.empty(ncopy * 16)
    for _ in range(5):
        ctx._check(ctx.lib.sphswe_spec_axpby(ctx.h, ncopy, 1.0, 0.0, src.ptr, 0.0, 0.0, None, dst.ptr))
    ctx.sync()
    ctx.close()
    print("mixed swe kernels done", flush=True)


def measured_traffic(path=TRAFFIC_FILE, key="bytes_per_launch"):
    """HBM bytes per launch from a committed counter summary (FETCH_SIZE and
    WRITE_SIZE scaled by the copy calibration), or None."""
    try:
        with open(path) as f:
            return json.load(f)[key]
    except (OSError, ValueError, KeyError):
        return None


# split-scheme forms at T1023: (RK4 Legendre precision, REXI gradient-operator
# precision, REXI factor storage, FP64 refinement sweep)
SPLIT_FORMS = (("fp64", "fp64", "fp64", "fp64", 0),
               ("mixed", "bf16x3", "fp32", "fp32", 0),
               ("mixed_refined", "bf16x3", "fp32", "fp32", 1))


def bench_split(nsteps=50, L=1023, M=188, dt=300.0, seg=2):
    """Config 4: nonlinear SWE at T1023 on 1536 x 3072, Galewsky balanced jet
    (no bump) plus a seeded height perturbation of 1 m with spectrum ~ l^-2,
    dt = 300 s, split scheme per step: the linear equations by REXI (200 poles,
    M = 188, h = 0.2, implicit viscosity 1e5) and the nonlinear remainder by one
    RK4 step.  FP64 form, and mixed forms (BF16X3 Legendre contractions in the
    RK4 part, FP32 REXI gradient operator, FP32 band factors with and without
    one sweep of FP64 refinement).

    The formulation diverges at this size: a straight FP64 run overflows within
    a few steps (reported as finite_steps), from an m = 1 mode that the REXI
    step of the scalar-wind equations amplifies (about 80x per step at T1023;
    the oracle's REXI shows the same top-degree m = 1 growth at T63), and at
    300 s the jet's advection of m > 530 is outside RK4's stability interval.
    The nsteps timed steps therefore run as nsteps / seg segments of seg steps,
    each from the initial state (a device-to-device copy of 8 MB inside the
    timed region, against ~60 ms of work per step), so that every timed step
    works on finite values.  Errors: relative L2 of h (= phi / g), u and v
    after seg steps against the FP64 form."""
    ctx = capi.Context(L)
    tuning = ctx.autotune()
    x0 = np.ascontiguousarray(workloads.galewsky_state(ctx, seed=1023, pert_amp=1.0))
    d_x0 = ctx.to_device(x0)
    res, finals = {}, {}
    for name, sht, grad, band, refine in SPLIT_FORMS:
        t0 = time.perf_counter()
        rx = ctx.rexi(capi.rexi_params(dt=dt, rexi_h=0.2, rexi_M=M, half=1, visc=1e5,
                                       precision=capi.PRECISIONS[grad], band_precision=capi.PRECISIONS[band],
                                       refine=refine))
        ctx.sync()
        setup_s = time.perf_counter() - t0
        prm = capi.swe_params(dt=dt, precision=capi.PRECISIONS[sht])
        state = ctx.to_device(x0)
        rx.split_step(prm, state, 1)
        ctx.sync()
        ctx.mark(12)
        for _ in range(nsteps // seg):
            ctx._check(ctx.lib.sphswe_memcpy_d2d(ctx.h, state.ptr, d_x0.ptr, x0.nbytes))
            rx.split_step(prm, state, seg)
        ctx.mark(13)
        ms = ctx.elapsed_ms(12, 13)
        steps = seg * (nsteps // seg)
        finals[name] = ctx.to_host(state, np.complex128, (3, ctx.nlm))
        if not np.all(np.isfinite(finals[name])):
            raise RuntimeError("split SWE state became non-finite within %d steps (%s)" % (seg, name))
        res[name] = {"steps_per_s": steps / (ms / 1000.0), "ms_per_step": ms / steps, "rexi_setup_s": setup_s,
                     "rk4_legendre": sht, "rexi_gradient": grad, "rexi_factors": band, "refine": refine}
        if name == "fp64":
            # the two parts of a step, timed separately from the initial state
            nl = capi.swe_params(dt=dt, visc=0.0, nonlinear=2, precision=capi.PREC_FP64)
            ctx._check(ctx.lib.sphswe_memcpy_d2d(ctx.h, state.ptr, d_x0.ptr, x0.nbytes))
            ctx.mark(12)
            rx.step(state, seg)
            ctx.mark(13)
            ctx._check(ctx.lib.sphswe_memcpy_d2d(ctx.h, state.ptr, d_x0.ptr, x0.nbytes))
            ctx.mark(14)
            ctx.swe_rk4(nl, state, 1, seg)
            ctx.mark(15)
            res[name]["rexi_ms_per_step"] = ctx.elapsed_ms(12, 13) / seg
            res[name]["rk4_ms_per_step"] = ctx.elapsed_ms(14, 15) / seg
            # a straight run: steps until the state is non-finite
            ctx._check(ctx.lib.sphswe_memcpy_d2d(ctx.h, state.ptr, d_x0.ptr, x0.nbytes))
            finite = 0
            for n in range(nsteps):
                rx.split_step(prm, state, 1)
                if not np.all(np.isfinite(ctx.to_host(state, np.complex128, (3, ctx.nlm)))):
                    break
                finite = n + 1
            res[name]["finite_steps"] = finite
        del state
        rx.close()
    ref = finals["fp64"]
    for name in res:
        if name == "fp64":
            continue
        for k, f in enumerate("huv"):
            res[name][f + "_rel_l2_after_%d_steps" % seg] = _rel(finals[name][k], ref[k])[0]
        # v is zero in the jet and small after seg steps: also against the wind scale |u|
        res[name]["v_err_over_u_l2_after_%d_steps" % seg] = float(
            np.linalg.norm(finals[name][2] - ref[2]) / np.linalg.norm(ref[1]))
        res[name]["speedup_vs_fp64"] = res[name]["steps_per_s"] / res["fp64"]["steps_per_s"]
    cfg = {"lmax": L, "nlat": ctx.nlat, "nphi": ctx.nphi, "poles": M + 12, "rexi_M": M, "rexi_h": 0.2, "dt": dt,
           "steps": nsteps, "segment_steps": seg, "viscosity": 1e5,
           "scheme": "Lie split: REXI (linear) then RK4 (nonlinear remainder)",
           "initial": "Galewsky jet without bump + 1 m seeded perturbation ~ l^-2", "autotune": tuning}
    ctx.close()
    return {"config": cfg, "forms": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--swe-steps", type=int, default=50)
    ap.add_argument("--rexi-steps", type=int, default=10)
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--split-steps", type=int, default=50)
    ap.add_argument("--batched-members", type=int, default=16,
                    help="members of the batched T511 leg (0: skip the leg)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (built-in launch shapes)")
    ap.add_argument("--pmc-kernels", action="store_true",
                    help="only launch the kernels used for counter collection")
    ap.add_argument("--mixed-swe-kernels", action="store_true",
                    help="only step the 16-member T511 mixed form eagerly (kernel trace and counter workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (see DESIGN.md, sharding).",
              file=sys.stderr)
        sys.exit(2)
    if args.pmc_kernels:
        pmc_kernels()
        return
    if args.mixed_swe_kernels:
        mixed_swe_kernels()
        return

    ctx = capi.Context(L_BENCH)
    t0 = time.perf_counter()
    tuning = ctx.autotune() if not args.no_tune else ctx.tuning()
    tune_s = time.perf_counter() - t0
    wall, dev_ms, rt_err = bench_sht(ctx, args.steps, args.warmup)
    ms_step = dev_ms / args.steps
    value = args.steps / (dev_ms / 1000.0)
    ms_syn, ms_ana, flops = bench_legendre_kernel(ctx)
    fft = bench_fft_kernels(ctx)
    swe_rate, swe_dt = bench_swe(ctx, args.swe_steps, 2)
    achieved = flops / (ms_syn * 1e-3) / 1e12
    achieved_ana = flops / (ms_ana * 1e-3) / 1e12
    mixed_sht = bench_sht_modes(ctx, max(args.steps // 4, 10), 3)
    bt = bench_bt_contractions(ctx)
    mixed_swe = bench_swe_modes(ctx) if args.swe_steps > 0 else None
    batched_swe = (bench_swe_batched(ctx, nmem=args.batched_members)
                   if args.swe_steps > 0 and args.batched_members > 0 else None)
    ensemble = bench_ensemble(args.ensemble_steps) if args.ensemble_steps > 0 else None
    rexi_rate, rexi_cfg = bench_rexi(args.rexi_steps) if args.rexi_steps > 0 else (None, None)
    rexi_modes = bench_rexi_modes(args.rexi_steps) if args.rexi_steps > 0 else None
    split = bench_split(args.split_steps) if args.split_steps > 0 else None
    cpu = None if args.no_cpu else cpu_baseline(L_BENCH, ctx.nlat, ctx.nphi)
    line = {
        "metric": "SHT round-trips/sec at T511 (3 fields, FP64); SWE steps/sec at T511 in swe_steps_per_s",
        "value": value,
        "unit": "round_trips/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded spectra, power ~ l^-3; Gaussian dam for the SWE rate)",
        "config": {"workload": "sht_round_trip_T511_3fields", "lmax": L_BENCH, "nlat": ctx.nlat,
                   "nphi": ctx.nphi, "fields": NFIELDS, "precision": "fp64"},
        "wall_s": wall,
        "autotune": dict(tuning, setup_s=tune_s),
        "round_trip_rel_l2": rt_err,
        "swe_steps_per_s": swe_rate,
        "swe_config": {"lmax": L_BENCH, "dt": swe_dt, "scheme": "RK4 nonlinear", "steps": args.swe_steps,
                       "viscosity": 1e5},
        "mixed_sht_T511_3fields": mixed_sht,
        "mixed_swe_T511_dam_100steps": mixed_swe,
        "batched_mixed_swe_T511_100steps": batched_swe,
        "ensemble_T127_64members": ensemble,
        "rexi_steps_per_s": rexi_rate,
        "rexi_config": rexi_cfg,
        "mixed_rexi_T256_200poles": rexi_modes,
        "split_T1023_galewsky": split,
        "roofline": {"bound": "mfma", "engine": "fp64 matrix cores (v_mfma_f64_4x4x4f64)",
                     "kernel": "k_leg_anal4 (Legendre analysis, 3 fields, T511)",
                     "achieved": achieved_ana, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_ana / FP64_PEAK_TFLOPS, "traffic": measured_traffic(),
                     "avg_ms": ms_ana, "flops_per_launch": flops,
                     "legendre_tuning": ctx.legendre_tuning(),
                     "synthesis": {"kernel": "k_leg_synth (FP64 VALU)", "avg_ms": ms_syn, "achieved": achieved,
                                   "frac": achieved / FP64_PEAK_TFLOPS},
                     "bf16_tables_T511": bt,
                     "fft_T511_9fields": dict(fft, bound="hbm", kernel="k_fft_wide (reference layout)",
                                              traffic={"r2c": measured_traffic(FFT_TRAFFIC_FILE, "r2c_bytes_per_launch"),
                                                       "c2r": measured_traffic(FFT_TRAFFIC_FILE, "c2r_bytes_per_launch")})},
        "cpu_baseline": cpu,
    }
    ctx.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
