"""Pins of the oracle's SWERexiSPHRobert (SWERexiSPHRobert.hpp) and of the
Robert band components (SPHSolverComplex.hpp:241-370):
- z4robert against its continuous operator (1 - mu^2) i m phi / r^2;
- z5robert, which the reference assembles with all five entries of
  (mu^2 - mu^4) i m / r^2 on the -2 diagonal (its defect B3), against that sum
  formed from the z2 (mu^2) and z3 (mu^4) components, on interior rows;
- the REXI step with Robert functions to first order in dt against the oracle's
  Robert linear tendency (itself pinned against the standard tendency), by the
  Richardson pair used for the standard step, with and without rotation;
- the gravity-mode closed form without rotation: h(dt) = C(w dt) h and the
  Robert winds (g/a) dt S(w dt) (i m h, D' h).
z6robert has no independent closed form here: it enters only through the
first-order check with rotation."""
import numpy as np
import pytest
import sympy as sp

from oracle import oracle as O
from tests._util import A_EARTH, G, H0, OMEGA, random_real_spec, rel_l2
from tests.test_oracle_rexi import _C_S
from tests.test_oracle_rexi_pole import MU, S, _matvec, _nc


@pytest.fixture(scope="module")
def orc31():
    return O.Oracle(31)


@pytest.mark.parametrize("m", [3, -2, 5])
def test_z4robert_vs_continuous(orc31, m):
    orc, L = orc31, 31
    rng = np.random.default_rng(20 + m)
    c = rng.standard_normal(6)
    P = sum(float(c[i]) * sp.legendre(i, MU) for i in range(6))
    phi = S ** abs(m) * P
    lon = orc.lon()[:, None]
    mu = orc.ct[None, :]
    e = np.exp(1j * m * lon)
    x = orc.analysis_cplx((sp.lambdify(MU, phi, "numpy")(mu) * e).reshape(-1))
    Mx = O.Matrix(orc, 2)
    Mx.add(19, 1.0, 1.0)
    y = _matvec(L, Mx.data(), Mx.halo, x)
    fn = sp.lambdify(MU, sp.expand(sp.I * m * (1 - MU ** 2) * phi), "numpy")
    yr = orc.analysis_cplx((np.broadcast_to(fn(mu), mu.shape) * e + 0j).reshape(-1))
    assert np.linalg.norm(y - yr) < 1e-12 * np.linalg.norm(yr)


def test_z5robert_is_the_collapsed_band(orc31):
    orc, L = orc31, 31
    s = 0.7 - 0.4j
    z5 = O.Matrix(orc, 4)
    z5.add(20, s, 1.0)
    z2 = O.Matrix(orc, 4)
    z2.add(12, 1.0, 0.0)
    z3 = O.Matrix(orc, 4)
    z3.add(13, 1.0, 0.0)
    d5, d2, d3 = z5.data(), z2.data(), z3.data()
    checked = 0
    for m in range(-L, L + 1):
        for n in range(abs(m) + 2, L - 3):
            r = _nc(L, n, m)
            want = s * 1j * m * np.sum(d2[r] - d3[r])
            assert abs(d5[r][2] - want) <= 1e-13 * max(1.0, abs(want)), (n, m)
            assert np.all(d5[r][[0, 1, 3, 4, 5, 6, 7, 8]] == 0), (n, m)
            checked += 1
    assert checked > 300


@pytest.mark.parametrize("omega", [OMEGA, 0.0])
def test_robert_rexi_first_order_matches_robert_tendency(orc31, omega):
    orc = orc31
    h = random_real_spec(31, 3, power=-2.0) * 50
    h[0] = 0
    u, v = random_real_spec(31, 4, power=-2.0), random_real_spec(31, 5, power=-2.0)
    hh, M = 0.2, 64
    al, be = O.rexi_coeffs(hh, M, True)
    C0, S0 = np.sum(be / al).real, np.sum(be / al ** 2).real
    T = orc.swe_tendency(dict(a=A_EARTH, g=G, omega=omega, h0=H0, visc=0.0, nonlinear=0, dt=0.0, robert=1),
                         h, u, v)

    def D(dt):
        prm = dict(a=A_EARTH, g=G, omega=omega, h0=H0, visc=0.0, rexi_h=hh, rexi_M=M, half=1, dt=dt, robert=1)
        x = orc.swe_rexi_run(prm, h, u, v, 1)
        return [(x[k] - C0 * y) / dt for k, y in enumerate((h, u, v))]

    errs = []
    for dt in (4.0, 8.0):
        d1, d2 = D(dt), D(2 * dt)
        errs.append([rel_l2(2 * d1[k] - d2[k], -S0 * T[k]) for k in range(3)])
    print(errs)
    assert max(errs[0]) < 1e-5, errs
    assert max(errs[0]) < 0.4 * max(errs[1]), errs


@pytest.mark.parametrize("l,m", [(1, 0), (5, 3), (20, 0), (31, 31)])
def test_robert_rexi_gravity_mode_closed_form(orc31, l, m):
    orc = orc31
    h = np.zeros(orc.nlm, complex)
    h[orc.idx(l, m)] = 1.0 if m == 0 else 0.6 + 0.8j
    z = np.zeros_like(h)
    dt, hh, M = 3600.0, 0.2, 128
    prm = dict(a=A_EARTH, g=G, omega=0.0, h0=H0, visc=0.0, rexi_h=hh, rexi_M=M, half=1, dt=dt, robert=1)
    x = orc.swe_rexi_run(prm, h, z, z, 1)
    al, be = O.rexi_coeffs(hh, M, True)
    w = np.sqrt(G * H0 * l * (l + 1)) / A_EARTH
    C, Sx = _C_S(al, be, w * dt)
    assert rel_l2(x[0], C * h) < 1e-12
    for k, op in ((1, 0), (2, 1)):   # the oracle's diff_lon and (1 - mu^2) d/dmu
        ref = (G / A_EARTH) * dt * Sx * orc.spec_op(op, h)
        scale = max(np.linalg.norm(ref), 1e-3 * np.linalg.norm(x[0]) * G * dt / A_EARTH)
        assert np.linalg.norm(x[k] - ref) < 1e-12 * scale, k
