"""Failure paths and checks of the C ABI on the GPU, forced from the host:
NaN/Inf detection (SPHData::isAnyNaNorInf, SPHData.hpp:560-569), the
physical reductions (SPHData.hpp:800-875), the zero-pivot status of the band
factorisation (zgbsv INFO > 0, BandedMatrixSolver.hpp:416-431) and the host
program's exit on an unstable run (AppSWE.hpp:323-329)."""
import os
import subprocess

import numpy as np
import pytest

from sph_swe_amd import capi

pytestmark = pytest.mark.gpu

APP = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph_swe_amd", "lib", "sphswe_app")


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(15)
    yield c
    c.close()


@pytest.mark.parametrize("bad", [None, np.nan, np.inf, -np.inf])
@pytest.mark.parametrize("where", [0, 4097, -1])
def test_any_nan(ctx, bad, where):
    x = np.random.default_rng(1).standard_normal(10001)
    if bad is not None:
        x[where] = bad
    d = ctx.to_device(x)
    assert ctx.any_nan(d, x.size) == (0 if bad is None else 1)


def test_any_nan_empty_and_prefix(ctx):
    x = np.zeros(300)
    x[200] = np.nan
    d = ctx.to_device(x)
    assert ctx.any_nan(d, 0) == 0
    assert ctx.any_nan(d, 200) == 0
    assert ctx.any_nan(d, 201) == 1


@pytest.mark.parametrize("n", [1, 255, 256, 70001])
def test_reduce_minmax(ctx, n):
    x = np.random.default_rng(n).standard_normal(n) * 3.0 - 0.5
    d = ctx.to_device(x)
    mn, mx, ab = ctx.minmax(d, n)
    assert mn == x.min() and mx == x.max() and ab == np.abs(x).max()


def test_band_zero_pivot_reports_singular(ctx):
    # an empty matrix: every pivot is zero, the factorisation reports it with
    # SPHSWE_ERR_SINGULAR and a message; adding the identity makes it solvable
    b = ctx.band(2)
    rhs = (np.arange(ctx.ncplx) + 1j).astype(np.complex128)
    with pytest.raises(capi.SphSweError) as e:
        b.solve(rhs)
    assert e.value.code == capi.ERR_SINGULAR
    assert "zero pivot" in str(e.value)
    b.add(0, 2.0)   # solver_component_scalar_phi
    x = b.solve(rhs)
    assert np.allclose(x, rhs / 2.0, rtol=0, atol=1e-15 * np.abs(rhs).max())
    b.close()


def test_rk4_propagates_nan(ctx):
    # a NaN planted in one spectral coefficient reaches the state after a step
    # and is reported by the check the host program runs after every step
    q = np.zeros((3, ctx.nlm), np.complex128)
    q[0, 0] = 1e4 * np.sqrt(4 * np.pi)
    d = ctx.to_device(q)
    ctx.swe_rk4(capi.swe_params(dt=60.0), d, 1, 1)
    assert ctx.any_nan(d, 6 * ctx.nlm) == 0
    q[1, 5] = np.nan
    d = ctx.to_device(q)
    ctx.swe_rk4(capi.swe_params(dt=60.0), d, 1, 1)
    assert ctx.any_nan(d, 6 * ctx.nlm) == 1


def test_host_program_exits_on_instability(tmp_path):
    # RK4 far outside its stability interval (dt = 1e6 s at T15) overflows
    # within a few tens of steps; the program stops with the reference's
    # message and status 1 instead of running on
    args = ["T15", "P0", "N1", "R1000000", "E100000000", "O1e12"]
    print("sphswe_app", " ".join(args), flush=True)
    r = subprocess.run([APP] + args, cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 1, r.stdout[-1000:] + r.stderr[-1000:]
    assert "Instability detected" in r.stderr
    # it stopped early: fewer step dots than the 100 requested
    assert r.stdout.count(".") < 100
