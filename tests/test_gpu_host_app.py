"""The C++ host program (host/main.cpp with the SPHConfig / SPHData /
SPHOperators / TimesteppingRK / REXI / SWERexiSPH / AppSWE classes over the C
ABI) against the oracle: RK4 and REXI runs through both the one-call-per-step
path (G1) and the reference's class-by-class loop (G0); the self-test
programs P1 (real operators against analytic derivatives), P2 (complex
operators), P3 (the banded solver on real spectra), P4 (the known-answer tests
of the complex banded solver) and P5 (one CSV per spherical harmonic)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi, workloads
from tests._util import A_EARTH, G, H0, OMEGA, rel_l2

pytestmark = pytest.mark.gpu

APP = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph_swe_amd", "lib", "sphswe_app")


def _run(args, cwd):
    print("sphswe_app", " ".join(args), flush=True)
    r = subprocess.run([APP] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _load(path, nlm):
    return np.fromfile(path, dtype=np.complex128).reshape(3, nlm)


def _dam(orc):
    lon = orc.lon()
    return workloads.dam_height(lon, orc.ct)


@pytest.mark.parametrize("fused", [1, 0])
def test_app_rk4_vs_oracle(tmp_path, fused):
    L, dt, nsteps = 31, 120.0, 4
    nlat, nphi = capi.grid_auto(L)
    orc = O.Oracle(L, nlat, nphi)
    dump = str(tmp_path / "state.bin")
    _run(["T%d" % L, "P0", "Q4", "N1", "R%g" % dt, "E%g" % (dt * nsteps), "O1e12", "V100000", "G%d" % fused,
          "Z" + dump], str(tmp_path))
    out = _load(dump, orc.nlm)
    h = _dam(orc)
    z = np.zeros_like(h)
    ref = orc.swe_rk4_run(dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=1e5, nonlinear=1, dt=dt), h, z, z, nsteps)
    for k, name in enumerate("huv"):
        e = rel_l2(out[k], ref[name])
        print("rk4", fused, name, e)
        assert e < 1e-10
    # the reference's output files: initial and final state, four fields each;
    # the final height file holds the grid values of the final spectrum (8
    # significant digits), rows south to north after the longitude header
    assert len([f for f in os.listdir(tmp_path) if f.endswith(".csv")]) == 8
    names = sorted(f for f in os.listdir(tmp_path) if f.startswith("prog_h"))
    with open(tmp_path / names[0]) as f:
        assert f.readline().startswith("#TI")
    with open(tmp_path / names[-1]) as f:
        lines = f.read().splitlines()
    rows = np.array([r.split("\t") for r in lines[4:]], float)[::-1]
    lat = rows[:, 0]
    assert np.allclose(lat, np.degrees(np.arcsin(np.asarray(orc.ct))), rtol=1e-7, atol=1e-6)
    grid = orc.synthesis(out[0]).reshape(orc.nphi, orc.nlat).T   # [ring][longitude]
    assert np.max(np.abs(rows[:, 1:] - grid)) < 1e-7 * np.max(np.abs(grid))


@pytest.mark.parametrize("fused", [1, 0])
def test_app_rk4_robert_functions_vs_oracle(tmp_path, fused):
    # N0 S1: the linear equations with Robert-function velocities (AppSWE.hpp:606-620)
    L, dt, nsteps = 31, 120.0, 4
    nlat, nphi = capi.grid_auto(L)
    orc = O.Oracle(L, nlat, nphi)
    dump = str(tmp_path / "state.bin")
    _run(["T%d" % L, "P0", "Q4", "N0", "S1", "R%g" % dt, "E%g" % (dt * nsteps), "O1e12", "V100000",
          "G%d" % fused, "Z" + dump], str(tmp_path))
    out = _load(dump, orc.nlm)
    h = _dam(orc)
    z = np.zeros_like(h)
    ref = orc.swe_rk4_run(dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=1e5, nonlinear=0, robert=1, dt=dt),
                          h, z, z, nsteps)
    for k, name in enumerate("huv"):
        e = rel_l2(out[k], ref[name])
        print("rk4 robert", fused, name, e)
        assert e < 1e-10
    # REXI with Robert functions: the one-call step (G1) and the class-by-class
    # loop of SWERexiSPHRobert terms (G0) match the oracle's SWERexiSPHRobert
    M, dtr = 8, 3600.0
    _run(["T%d" % L, "P0", "Q4", "N0", "S1", "M%d" % M, "H0.2", "I1", "R%g" % dtr, "E%g" % (2 * dtr), "O1e12", "V100000",
          "G%d" % fused, "Z" + dump], str(tmp_path))
    out = _load(dump, orc.nlm)
    hs = orc.analysis(h)
    zs = np.zeros_like(hs)
    ref = orc.swe_rexi_run(dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=1e5, rexi_h=0.2, rexi_M=M, half=1, dt=dtr,
                                robert=1), hs, zs, zs, 2)
    for k in range(3):
        e = rel_l2(out[k], ref[k])
        print("rexi robert", fused, "huv"[k], e)
        assert e < 1e-11, k


@pytest.mark.parametrize("fused,omega", [(1, OMEGA), (0, OMEGA), (0, 0.0)])
def test_app_rexi_vs_oracle(tmp_path, fused, omega):
    L, dt, nsteps, M = 31, 3600.0, 2, 8
    nlat, nphi = capi.grid_auto(L)
    orc = O.Oracle(L, nlat, nphi)
    dump = str(tmp_path / "state.bin")
    _run(["T%d" % L, "P0", "Q4", "M%d" % M, "H0.2", "I1", "F%g" % omega, "R%g" % dt, "E%g" % (dt * nsteps),
          "O1e12", "V100000", "G%d" % fused, "Z" + dump], str(tmp_path))
    out = _load(dump, orc.nlm)
    hs = orc.analysis(_dam(orc))
    z = np.zeros_like(hs)
    ref = orc.swe_rexi_run(dict(a=A_EARTH, g=G, omega=omega, h0=H0, visc=1e5, rexi_h=0.2, rexi_M=M, half=1, dt=dt),
                           hs, z, z, nsteps)
    for k in range(3):
        e = rel_l2(out[k], ref[k])
        print("rexi", fused, omega, "huv"[k], e)
        assert e < 1e-11


def _copy_modes(spec, L_src, L_dst):
    """SPHData::spec_copyToDifferentModes on m-major real spectra (numpy)."""
    out = np.zeros((L_dst + 1) * (L_dst + 2) // 2, complex)
    for m in range(min(L_src, L_dst) + 1):
        n = min(L_src, L_dst) - m + 1
        od, os_ = m * (2 * L_dst - m + 1) // 2 + m, m * (2 * L_src - m + 1) // 2 + m
        out[od:od + n] = spec[os_:os_ + n]
    return out


@pytest.mark.parametrize("fused", [1, 0])
def test_app_rexi_extended_modes_vs_oracle(tmp_path, fused):
    # X4 (AppSWE.hpp:338-351,434-445,524-537): each step copies the state into
    # the T(L+4) configuration, runs the REXI terms there and copies back; the
    # oracle does the same with its own REXI step at T(L+4)
    L, X, dt, nsteps, M = 31, 4, 3600.0, 2, 8
    nlat, nphi = capi.grid_auto(L)
    orc = O.Oracle(L, nlat, nphi)
    nl2, np2 = capi.grid_auto(L + X)
    orc2 = O.Oracle(L + X, nl2, np2)
    dump = str(tmp_path / "state.bin")
    _run(["T%d" % L, "P0", "Q4", "M%d" % M, "H0.2", "I1", "X%d" % X, "R%g" % dt, "E%g" % (dt * nsteps), "O1e12",
          "V100000", "G%d" % fused, "Z" + dump], str(tmp_path))
    out = _load(dump, orc.nlm)
    x = [orc.analysis(_dam(orc)), np.zeros(orc.nlm, complex), np.zeros(orc.nlm, complex)]
    prm = dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=1e5, rexi_h=0.2, rexi_M=M, half=1, dt=dt)
    for _ in range(nsteps):
        y = orc2.swe_rexi_run(prm, *[_copy_modes(f, L, L + X) for f in x], 1)
        x = [_copy_modes(f, L + X, L) for f in y]
    plain = orc.swe_rexi_run(prm, orc.analysis(_dam(orc)), np.zeros(orc.nlm, complex), np.zeros(orc.nlm, complex),
                             nsteps)
    for k in range(3):
        e = rel_l2(out[k], x[k])
        print("rexi X%d" % X, fused, "huv"[k], e, "vs plain", rel_l2(x[k], plain[k]))
        assert e < 1e-11
    # the extra modes change the result (else the test would not see them)
    assert rel_l2(x[2], plain[2]) > 1e-8


def test_app_galewsky_initial_state(tmp_path):
    # scenario Q1 (AppSWE.hpp:199-207): the balanced jet (pinned on the CPU in
    # tests/test_host_galewsky.py) plus the bump, analysed on the GPU; against
    # the same grid values computed here (the jet of sph_swe_amd.workloads with
    # the bump in numpy) and the oracle's analysis
    from sph_swe_amd import workloads
    L = 63
    nlat, nphi = capi.grid_auto(L)
    orc = O.Oracle(L, nlat, nphi)
    ctx = capi.Context(L)
    jet = workloads.galewsky_state(ctx, seed=None)
    hj = ctx.synthesis(jet[0:1])[0].reshape(nphi, nlat)
    ctx.close()
    lam = orc.lon()[:, None]
    phi = np.arcsin(np.asarray(orc.ct))[None, :]
    bump = 120.0 * np.cos(phi) * np.exp(-((lam - np.pi) * 3.0) ** 2) * np.exp(-((np.pi / 4 - phi) * 15.0) ** 2)
    dump = str(tmp_path / "state.bin")
    _run(["T%d" % L, "P0", "Q1", "N1", "R1", "E0", "O1e12", "G1", "Z" + dump], str(tmp_path))
    out = _load(dump, orc.nlm)
    ref_h = orc.analysis((hj + bump).ravel())
    e = rel_l2(out[0], ref_h)
    print("galewsky h", e, "u", rel_l2(out[1], jet[1]))
    assert e < 1e-11
    assert rel_l2(out[1], jet[1]) < 1e-12
    assert np.max(np.abs(out[2])) == 0.0


def test_app_p4_solver_known_answers(tmp_path):
    # AppTestSPHSolverComplex.hpp:29-517 at T32: for the 76 poles of
    # REXI::setup(0.2, 64), scalar, mu, (1-mu^2) d/dmu, z1..z8 and Helmholtz
    # solves through the GPU band solver reproduce the Gaussian bump to 1e-10
    out = _run(["T32", "P4"], str(tmp_path))
    errs = [float(line.rsplit(":", 1)[1]) for line in out.splitlines() if "||| Error:" in line]
    print("P4 largest error", max(errs))
    assert len(errs) == 76 * 12
    assert max(errs) < 1e-10
    assert "P4: all 76 poles within" in out


def _gauss_div(lon, mu):
    """The Gaussian bump of the self-test programs (centre (pi/5, pi/4),
    f = exp(-10 d^2)) with div_lat f = d(cos(phi) f)/dphi / cos(phi) and
    div_lon f = grad_lon f = (df/dlam) / cos(phi), grad_lat f = df/dphi."""
    phi = np.arcsin(mu)
    lon0, phi0 = np.pi / 5, np.pi / 4
    c = np.clip(np.sin(phi) * np.sin(phi0) + np.cos(phi) * np.cos(phi0) * np.cos(lon - lon0), -1, 1)
    d = np.arccos(c)
    s = np.sqrt(1 - c * c)
    q = np.where(s > 0, d / np.where(s > 0, s, 1), 1.0)
    f = np.exp(-10 * d * d)
    dlam = 20 * q * f * (-np.cos(phi) * np.cos(phi0) * np.sin(lon - lon0))
    dphi = 20 * q * f * (np.cos(phi) * np.sin(phi0) - np.sin(phi) * np.cos(phi0) * np.cos(lon - lon0))
    return f, dphi - np.tan(phi) * f, dlam / np.cos(phi), dphi


def test_app_p1_operators_vs_analytic(tmp_path):
    # AppTestSPHOperators.hpp:25-520 at T64: every real operator on the GPU
    # against the analytic derivatives of the Gaussian bump (the reference only
    # prints the errors).  Where the operator is exact on band-limited input the
    # error is the bump's spectral tail (exp(-l^2 / 40) at l = 64, times the
    # derivative); div_lat multiplies by cos(phi) before its projection, which
    # is not band-limited where the bump reaches the pole (f = 2e-3 there), so
    # DIV LAT and the sums of two operator results (DIVERGENCE, VORTICITY,
    # div.grad; projected by SPHData's spectral + and -) are checked against the
    # same errors computed by the oracle's operators
    L = 64
    errs = _p_errors(_run(["T%d" % L, "P1"], str(tmp_path)))
    print(errs)
    assert len(errs) == 18, errs
    nlat, nphi = capi.grid_auto(L)
    orc = O.Oracle(L, nlat, nphi)
    lon = np.repeat(orc.lon(), nlat).reshape(nphi, nlat)
    mu = np.tile(np.asarray(orc.ct), nphi).reshape(nphi, nlat)
    f, div_mu, div_lam, dphi = _gauss_div(lon, mu)
    f = f.ravel()
    grad_lon = orc.phys_op(0, f)
    grad_lat = orc.phys_op(1, f)
    # SPHData's + and - work on the spectra (SPHData.hpp:231-314): a sum of two
    # operator results is projected before the pointwise comparison
    proj = lambda x: orc.synthesis(orc.analysis(x))
    inexact = {
        "TEST DIV LAT": np.max(np.abs(orc.phys_op(2, f) - div_mu.ravel())),
        "TEST DIVERGENCE": np.max(np.abs(proj(grad_lon + orc.phys_op(2, f)) - (div_mu + div_lam).ravel())),
        "TEST VORTICITY": np.max(np.abs(proj(grad_lon - grad_lat) - (div_lam - dphi).ravel())),
        "TEST LAPLACE (div.grad vs. sph laplace)":
            np.max(np.abs(proj(orc.phys_op(0, grad_lon) + orc.phys_op(2, grad_lat)) - orc.phys_op(3, f))),
    }
    for name, e in errs.items():
        if name in inexact:
            print(name, "app", e, "oracle", inexact[name])
            assert abs(e - inexact[name]) < 1e-9 + 1e-6 * inexact[name], name
            assert e < 0.2
        else:
            assert 0 <= e < 1e-10, (name, e)


def _p_errors(out):
    import re
    errs = {}
    for line in out.splitlines():
        mt = (re.match(r"^(.*?)[: -]+max error: ([-+0-9.eE]+)$", line) or re.match(r"^ERROR (.*): ([-+0-9.eE]+)$", line)
              or re.match(r"^(Test .*) \|\|\| Error: ([-+0-9.eE]+)$", line))
        if mt:
            errs[mt.group(1).strip()] = float(mt.group(2))
    return errs


def test_app_p2_complex_operators_vs_analytic(tmp_path):
    # AppTestSPHOperatorsComplex.hpp:26-606 at T64: the complex operators on a
    # real-valued bump; the exact cases to 1e-10, div_lat and div.grad equal the
    # errors of the real operators computed by the oracle (the complex analysis
    # of a real field has the real analysis' coefficients)
    L = 64
    errs = _p_errors(_run(["T%d" % L, "P2"], str(tmp_path)))
    print(errs)
    assert len(errs) == 17, errs
    nlat, nphi = capi.grid_auto(L)
    orc = O.Oracle(L, nlat, nphi)
    lon = np.repeat(orc.lon(), nlat).reshape(nphi, nlat)
    mu = np.tile(np.asarray(orc.ct), nphi).reshape(nphi, nlat)
    f, div_mu, _, _ = _gauss_div(lon, mu)
    f = f.ravel()
    proj = lambda x: orc.synthesis(orc.analysis(x))
    inexact = {
        "TEST DIV LAT": np.max(np.abs(orc.phys_op(2, f) - div_mu.ravel())),
        "TEST LAPLACE": np.max(np.abs(proj(orc.phys_op(0, orc.phys_op(0, f)) + orc.phys_op(2, orc.phys_op(1, f)))
                                      - orc.phys_op(3, f))),
    }
    for name, e in errs.items():
        if name in inexact:
            print(name, "app", e, "oracle", inexact[name])
            assert abs(e - inexact[name]) < 1e-9 + 1e-6 * inexact[name], name
        else:
            assert 0 <= e < 1e-10, (name, e)


def test_app_p3_real_solver_known_answers(tmp_path):
    # AppTestSPHSolver.hpp:34-482 at T32: the banded solver on real spectra
    # (alpha = 3 and the REXI scalars), each solution against the bump
    L = 32
    errs = _p_errors(_run(["T%d" % L, "P3"], str(tmp_path)))
    print(errs)
    assert len(errs) == 11, errs
    for name, e in errs.items():
        assert 0 <= e < 1e-10, (name, e)   # measured <= 2.4e-11 (z3)


def test_app_p5_writes_every_harmonic(tmp_path):
    # AppOutputSphericalHarmonics.hpp: one CSV per (n, m); the grid values of
    # a unit coefficient are Y_n^0, or 2 Re Y_n^m for m > 0 (orthonormal,
    # Condon-Shortley phase), written with 8 significant digits
    from scipy.special import sph_harm_y
    L = 8
    _run(["T%d" % L, "P5"], str(tmp_path))
    names = sorted(f for f in os.listdir(tmp_path) if f.startswith("SPH_n"))
    assert len(names) == (L + 1) * (L + 2) // 2
    for n, m in [(0, 0), (2, 1), (5, 3), (8, 8)]:
        with open(tmp_path / ("SPH_n%d_m%d.csv" % (n, m))) as f:
            lines = f.read().splitlines()
        assert lines[0].startswith("#TI") and lines[1] == "#TX Longitude" and lines[2] == "#TY Latitude"
        lon = np.radians(np.array(lines[3].split("\t")[1:], float))
        rows = np.array([r.split("\t") for r in lines[4:]], float)
        lat = np.radians(rows[:, 0])
        Y = sph_harm_y(n, m, (np.pi / 2 - lat)[:, None], lon[None, :])
        ref = Y.real if m == 0 else 2 * Y.real
        assert np.max(np.abs(rows[:, 1:] - ref)) < 1e-7 * max(1.0, np.max(np.abs(ref))), (n, m)


def test_app_reports_missing_programs(tmp_path):
    r = subprocess.run([APP, "T16", "P9"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "not available" in r.stderr
