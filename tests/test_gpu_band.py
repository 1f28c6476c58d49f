"""SPHSolverComplex on the GPU (sphswe_band_*): per-m pivoted LU solves of the
band components against the oracle's restatement (itself pinned against LAPACK
zgbsv and the reference's known-answer test), and the single-pole REXI solve
(SWERexiSPH for one pole) through sphswe_rexi_apply against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from sph_swe_amd import capi
from tests._util import A_EARTH, G, H0, OMEGA, random_cplx_spec, random_real_spec, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,halo,comps", [
    (31, 2, [(0, 2.0 - 1.0j), (1, 0.7), (2, 0.3 + 0.1j)]),
    (31, 4, [(11, (1.5 + 0.5j) ** 4), (12, 0.8j), (13, 0.3), (14, 0.2), (16, 0.4), (17, -0.1 + 0.05j), (18, 0.25)]),
    (64, 4, [(11, 3.0 + 1.0j), (13, 1.0), (15, 0.5 - 0.5j), (18, 0.7)]),
])
def test_band_solve_vs_oracle(L, halo, comps):
    print(f"band solve L={L} halo={halo}", flush=True)
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    r = 2.0       # a radius of order one keeps z4..z8 visible next to z1
    Mo = O.Matrix(orc, halo)
    band = ctx.band(halo)
    for comp, s in comps:
        Mo.add(comp, s, r)
        band.add(comp, s, r)
    rhs = random_cplx_spec(L, 3)
    x_ref, info = Mo.solve(rhs)
    assert info == 0
    x = band.solve(rhs)
    assert rel_l2(x, x_ref) < 1e-12
    # a second solve after a change refactorises
    band.add(0, 0.5, r)
    Mo.add(0, 0.5, r)
    x_ref2, _ = Mo.solve(rhs)
    assert rel_l2(band.solve(rhs), x_ref2) < 1e-12
    band.close()
    ctx.close()


def test_band_rejects_component_wider_than_halo():
    ctx = capi.Context(16)
    band = ctx.band(2)
    with pytest.raises(capi.SphSweError):
        band.add(13, 1.0, 1.0)      # mu^4 needs halo 4
    band.close()
    ctx.close()


@pytest.mark.parametrize("omega", [OMEGA, 0.0])
def test_single_pole_apply_vs_oracle(omega):
    # one SWERexiSPH pole (alpha, beta) applied through sphswe_rexi_apply equals the
    # oracle's REXI run restricted to that pole (h = phi/g in and out, no viscosity)
    L = 31
    print(f"single pole omega={omega}", flush=True)
    ctx = capi.Context(L)
    orc = O.Oracle(L, ctx.nlat, ctx.nphi)
    dt = 3600.0
    al, be = O.rexi_coeffs(0.2, 8, True)
    k = 5
    prm = capi.rexi_params(omega=omega, visc=0.0, dt=dt)
    rx = ctx.rexi(prm, alpha=al[k:k + 1], beta=be[k:k + 1])
    phi = random_real_spec(L, 1) * 1e3
    u, v = random_real_spec(L, 2), random_real_spec(L, 3)
    d_in = ctx.to_device(np.stack([phi, u, v]))
    d_out = ctx.empty(3 * ctx.nlm * 16)
    rx.apply(d_in, d_out)
    out = ctx.to_host(d_out, np.complex128, (3, ctx.nlm))
    # oracle: the pole's complex solution, times beta, real part, analysed
    shape = (ctx.nphi, ctx.nlat)
    sol = orc.rexi_pole_solve(dict(a=A_EARTH, g=G, omega=omega, h0=H0), al[k] / dt, phi, u, v)
    for q in range(3):
        ref = orc.analysis((be[k] / dt * sol[q]).real.reshape(-1))
        assert rel_l2(out[q], ref) < 1e-12, q
    rx.close()
    ctx.close()
