"""The C++ host program (host/main.cpp with the SPHConfig / SPHData /
SPHOperators / TimesteppingRK / REXI / SWERexiSPH / AppSWE classes over the C
ABI) against the oracle: RK4 and REXI runs through both the one-call-per-step
path (G1) and the reference's class-by-class loop (G0); program P4, the
reference's known-answer tests of the complex banded solver."""
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
    # the reference's output files: initial and final state, four fields each
    assert len([f for f in os.listdir(tmp_path) if f.endswith(".csv")]) == 8
    with open(tmp_path / sorted(f for f in os.listdir(tmp_path) if f.startswith("prog_h"))[0]) as f:
        assert f.readline().startswith("#TI")


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
    # REXI with Robert functions: the one-call step (G1) matches the oracle's
    # SWERexiSPHRobert; the class-by-class loop (G0) is refused, as documented
    M, dtr = 8, 3600.0
    _run(["T%d" % L, "P0", "Q4", "N0", "S1", "M%d" % M, "H0.2", "I1", "R%g" % dtr, "E%g" % (2 * dtr), "O1e12", "V100000",
          "G1", "Z" + dump], str(tmp_path))
    out = _load(dump, orc.nlm)
    hs = orc.analysis(h)
    zs = np.zeros_like(hs)
    ref = orc.swe_rexi_run(dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=1e5, rexi_h=0.2, rexi_M=M, half=1, dt=dtr,
                                robert=1), hs, zs, zs, 2)
    for k in range(3):
        assert rel_l2(out[k], ref[k]) < 1e-11, k
    r = subprocess.run([APP, "T%d" % L, "P0", "N0", "S1", "M8", "R3600", "E3600", "O1e12", "G0"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0


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


def test_app_reports_missing_programs(tmp_path):
    r = subprocess.run([APP, "T16", "P5"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "not available" in r.stderr
