"""Pins of the oracle's REXI coefficients, banded solver and band assembly.

- REXI alpha/beta against the reference's own REXI.hpp / GaussianApproximation.hpp /
  ExponentialApproximation.hpp, compiled from /root/reference into oracle/_ref;
- the rational approximation property sum Re(beta/(ix+alpha)) ~ cos x;
- the banded LU with partial pivoting against scipy's LAPACK zgbsv;
- the reference's only asserting test, AppTestSPHSolverComplex (src/AppTestSPHSolverComplex.hpp:
  29-517): band assembly z1..z8 + solve against closed-form Gaussian-bump
  solutions, 1e-10 max error, at T32;
- the REXI time step through closed forms of the pole sum: a gravity mode without
  rotation, the inertial oscillation, and first-order agreement with the linear
  tendency with rotation.
"""
import os

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from oracle import oracle as O
from tests._util import A_EARTH, G, H0, OMEGA, random_real_spec, rel_l2


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference tree)")
@pytest.mark.parametrize("h,M,half", [(0.2, 64, True), (0.2, 128, True), (0.2, 188, True), (0.15, 32, False)])
def test_rexi_coefficients_vs_reference_build(h, M, half):
    a_ref, b_ref = O.ref_rexi(h, M, half)
    a, b = O.rexi_coeffs(h, M, half)
    assert a.shape == a_ref.shape
    assert np.max(np.abs(a - a_ref)) <= 1e-15 * np.max(np.abs(a_ref))
    assert np.max(np.abs(b - b_ref)) <= 1e-14 * np.max(np.abs(b_ref))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("h,M,half", [(0.2, 64, True), (0.2, 128, True), (0.2, 188, True), (0.15, 32, False)])
def test_rexi_coefficients_vs_golden(h, M, half):
    # vectors written by tests/golden/make_golden.py from the reference's REXI.hpp build
    arr = np.loadtxt(os.path.join(GOLDEN, "rexi_h%s_M%d_%s.txt" % (repr(h), M, "half" if half else "full")))
    a_ref, b_ref = arr[:, 0] + 1j * arr[:, 1], arr[:, 2] + 1j * arr[:, 3]
    a, b = O.rexi_coeffs(h, M, half)
    assert a.shape == a_ref.shape
    assert np.max(np.abs(a - a_ref)) <= 1e-15 * np.max(np.abs(a_ref))
    assert np.max(np.abs(b - b_ref)) <= 1e-14 * np.max(np.abs(b_ref))


def test_identities_vs_golden():
    ref = np.loadtxt(os.path.join(GOLDEN, "identities_L20.txt"))
    for row in ref:
        n, m = row[0], row[1]
        d = O.identities(n, m)
        for k, name in enumerate("RSABCDEGH"):
            want = row[2 + k]
            if np.isnan(want) and name == "H":
                continue   # H(0, m): the reference asserts on E(-1, m); z6robert uses H(n >= 1, m)
            if np.isnan(want):
                assert np.isnan(d[name]), (n, m, name)
            else:
                assert abs(d[name] - want) <= 1e-15 * max(1.0, abs(want)), (n, m, name)


def test_pole_counts():
    # full: 2(M+11)+1 poles, half: M+12; M = 188 half gives the 200-pole configuration
    assert O.rexi_coeffs(0.2, 128, True)[0].size == 140
    assert O.rexi_coeffs(0.2, 188, True)[0].size == 200
    assert O.rexi_coeffs(0.2, 64, False)[0].size == 151


@pytest.mark.parametrize("M", [32, 64])
def test_rexi_approximates_cosine(M):
    h = 0.2
    a, b = O.rexi_coeffs(h, M, False)
    xs = np.linspace(-0.7 * h * M, 0.7 * h * M, 401)
    approx = np.array([np.sum((b / (1j * x + a)).real) for x in xs])
    assert np.max(np.abs(approx - np.cos(xs))) < 1e-10


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_identities_vs_reference_build():
    ref = O.ref_identities(20)
    for row in ref[::7]:
        n, m = row[0], row[1]
        d = O.identities(n, m)
        for k, name in enumerate("RSABCDEGH"):
            if np.isnan(row[2 + k]) and name == "H":
                continue
            assert abs(d[name] - row[2 + k]) <= 1e-15 * max(1.0, abs(row[2 + k])), (n, m, name)


@pytest.mark.parametrize("n,kl", [(40, 4), (57, 2), (9, 4), (1, 2)])
def test_gbsv_vs_scipy(n, kl):
    rng = np.random.default_rng(n + kl)
    ku = kl
    ldab = 2 * kl + ku + 1
    AB = np.zeros((ldab, n), complex)
    # random band, weak diagonal so that pivoting matters
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            AB[kl + ku + i - j, j] = rng.standard_normal() + 1j * rng.standard_normal()
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, ipiv, info = O.gbsv(AB, kl, ku, b)
    lu, piv, xs, info_s = lapack.zgbsv(kl, ku, AB.copy(), b.copy())
    assert info == 0 and info_s == 0
    assert np.max(np.abs(x - xs)) < 1e-12 * np.max(np.abs(xs))
    assert np.array_equal(ipiv, piv)   # scipy reports 0-based pivots


def test_gbsv_reports_singular_pivot():
    n, kl = 6, 2
    AB = np.zeros((3 * kl + 1, n), complex)
    x, ipiv, info = O.gbsv(AB, kl, kl, np.ones(n, complex))
    assert info == 1


# --- AppTestSPHSolverComplex (P4) ------------------------------------------------

class Gauss:
    """Closed forms of SphereTestSolutions_Gaussian (exp(-10 d^2) bump at (pi/5, pi/4))."""
    f, lon_c, lat_c = 10.0, np.pi / 5, np.pi / 4

    def _d(self, lon, mu):
        p1 = np.arcsin(mu)
        c = np.sin(p1) * np.sin(self.lat_c) + np.cos(p1) * np.cos(self.lat_c) * np.cos(lon - self.lon_c)
        return p1, c, np.arccos(np.clip(c, -1, 1))

    def fun(self, lon, mu):
        _, _, d = self._d(lon, mu)
        return np.exp(-d * d * self.f)

    def dlambda(self, lon, mu):
        p1, c, d = self._d(lon, mu)
        return -(self.f * 2 * np.cos(p1) * np.cos(self.lat_c) * np.sin(lon - self.lon_c)) * d * np.exp(-d * d * self.f) \
            / np.sqrt(1 - c * c)

    def dphi(self, lon, mu):
        p1, c, d = self._d(lon, mu)
        return self.f * 2 * d * (np.cos(p1) * np.sin(self.lat_c) - np.sin(p1) * np.cos(self.lat_c) * np.cos(lon - self.lon_c)) \
            * np.exp(-d * d * self.f) / np.sqrt(1 - c * c)

    def grad_lambda(self, lon, mu):
        return self.dlambda(lon, mu) / np.sqrt(1 - mu * mu)

    def one_minus_mu2_dmu(self, lon, mu):
        return self.dphi(lon, mu) / np.sqrt(1 - mu * mu) * (1 - mu * mu)


@pytest.fixture(scope="module")
def p4():
    orc = O.Oracle(32)
    lon = orc.lon()[:, None]
    mu = orc.ct[None, :]
    return orc, lon, mu


def _solve(orc, halo, comps, b_phys):
    Mx = O.Matrix(orc, halo)
    for comp, sc in comps:
        Mx.add(comp, sc, A_EARTH)
    x, info = Mx.solve(orc.analysis_cplx(b_phys.astype(complex).reshape(-1)))
    assert info == 0
    return orc.synthesis_cplx(x).reshape(orc.nphi, orc.nlat)


def test_p4_known_answers(p4):
    orc, lon, mu = p4
    g = Gauss()
    fun = g.fun(lon, mu)
    r, two_omega = A_EARTH, 2 * OMEGA
    alphas, _ = O.rexi_coeffs(0.2, 64, True)
    worst = 0.0
    for alpha in alphas[::5]:
        cases = [
            (2, [(0, alpha)], fun * alpha),
            (2, [(0, alpha), (1, 1.0)], mu * fun + alpha * fun),
            (2, [(0, alpha), (2, 1.0)], g.one_minus_mu2_dmu(lon, mu) + alpha * fun),
            (2, [(11, alpha ** 4)], fun * alpha ** 4),
            (2, [(12, alpha * alpha * two_omega ** 2)], alpha * alpha * two_omega ** 2 * mu * mu * fun),
            (4, [(13, two_omega ** 4)], two_omega ** 4 * mu ** 4 * fun),
            (2, [(14, -alpha * alpha * two_omega), (0, alpha)],
             (-alpha * alpha * two_omega) / (r * r) * g.dlambda(lon, mu) + fun * alpha),
            (2, [(15, two_omega ** 3), (0, alpha)],
             two_omega ** 3 / (r * r) * np.sqrt(1 - mu * mu) * mu * mu * g.grad_lambda(lon, mu) + fun * alpha),
            (2, [(16, 2 * alpha * two_omega ** 2), (0, alpha)],
             2 * alpha * two_omega ** 2 / (r * r) * np.sqrt(1 - mu * mu) * mu * g.dphi(lon, mu) + fun * alpha),
        ]
        for halo, comps, b in cases:
            x = _solve(orc, halo, comps, b)
            worst = max(worst, np.max(np.abs(x.real - fun)), np.max(np.abs(x.imag)))
        # z7 / z8: right-hand sides through the oracle's own spectral operators
        fs = orc.analysis_cplx(fun.astype(complex).reshape(-1))
        lap = orc.spec_op(O.OP_LAPLACE, fs)
        for comp, rhs in ((17, lap / (r * r) + fs * alpha), (18, orc.spec_op(O.OP_MU2, lap) / (r * r) + fs * alpha)):
            Mx = O.Matrix(orc, 2)
            Mx.add(comp, 1.0, r)
            Mx.add(0, alpha, r)
            x, info = Mx.solve(rhs)
            xp = orc.synthesis_cplx(x).reshape(orc.nphi, orc.nlat)
            worst = max(worst, np.max(np.abs(xp.real - fun)), np.max(np.abs(xp.imag)))
    assert worst < 1e-10, worst


def test_band_rows_quirk_is_harmless(p4):
    # rowElement_add tests n < m with the signed m (SPHMatrixComplex.hpp:193-197), so
    # for m < 0 it keeps entries whose column falls below the block start.  The
    # recurrence identities vanish there, so every such entry is an exact zero and
    # the band-to-LAPACK conversion may drop them; checked for every component.
    orc, _, _ = p4
    L = orc.L

    def ncompact(n, m):
        if m < 0:
            minc = L + m
            return ((minc - 1) * minc >> 1) + minc + (n + m)
        return (L * (L + 1) >> 1) + (m * (2 * L - m + 1) >> 1) + n

    for halo, comp in ((2, 1), (2, 2), (2, 12), (4, 13), (2, 14), (2, 15), (2, 16), (2, 17), (2, 18)):
        Mx = O.Matrix(orc, halo)
        Mx.add(comp, 1.0 + 0.5j, A_EARTH)
        d = Mx.data()
        assert np.all(np.isfinite(d)), comp
        for m in range(-L, L + 1):
            for n in range(abs(m), L + 1):
                for rel in range(-halo, halo + 1):
                    if n + rel < abs(m):
                        assert d[ncompact(n, m), rel + halo] == 0, (comp, n, m, rel)


# --- REXI time step: closed forms of the pole sum ---------------------------------
#
# One REXI step applies f(dt L) = sum_n beta_n (alpha_n + dt L)^-1 to the real state,
# L the linear SWE operator.  AppSWE uses the beta_re weights, whose sum reproduces
# Re e^{ix} = cos x but only approximates sin x (relative error ~5e-4 at h = 0.2),
# so the closed forms below use the scalar sums of the same poles,
#   C(x) = Re sum beta alpha / (alpha^2 + x^2)   (~ cos x)
#   S(x) = Re sum beta / (alpha^2 + x^2)         (~ -sin(x) / x),
# rather than cos and sin.  A dropped or mis-signed term of SWERexiSPH::solve changes
# the operator that f is applied to and fails one of these.

def _C_S(alpha, beta, x):
    x = np.asarray(x, dtype=float)[..., None]
    d = alpha * alpha + x * x
    return np.sum(beta * alpha / d, -1).real, np.sum(beta / d, -1).real


def _lowdeg(L, seed, lmax):
    from tests._util import real_spec_index
    ls, _ = real_spec_index(L)
    q = random_real_spec(L, seed, power=-1.0)
    q[ls > lmax] = 0
    return q


@pytest.fixture(scope="module")
def orc31():
    return O.Oracle(31)


@pytest.mark.parametrize("l,m", [(1, 0), (5, 3), (20, 0), (31, 31)])
def test_rexi_gravity_mode_closed_form(orc31, l, m):
    # F0 (no rotation), u = v = 0, h a single mode: h(dt) = C(w dt) h,
    # (u, v)(dt) = (g/a) dt S(w dt) grad(h), w = sqrt(g h0 l(l+1)) / a
    orc = orc31
    h = np.zeros(orc.nlm, complex)
    h[orc.idx(l, m)] = 1.0 if m == 0 else 0.6 + 0.8j
    z = np.zeros_like(h)
    dt, hh, M = 3600.0, 0.2, 128
    prm = dict(a=A_EARTH, g=G, omega=0.0, h0=H0, visc=0.0, rexi_h=hh, rexi_M=M, half=1, dt=dt)
    x = orc.swe_rexi_run(prm, h, z, z, 1)
    al, be = O.rexi_coeffs(hh, M, True)
    w = np.sqrt(G * H0 * l * (l + 1)) / A_EARTH
    C, S = _C_S(al, be, w * dt)
    hp = orc.synthesis(h)
    assert rel_l2(x[0], C * h) < 1e-12
    for k, which in ((1, 0), (2, 1)):
        ref = (G / A_EARTH) * dt * S * orc.analysis(orc.phys_op(which, hp))
        scale = max(np.linalg.norm(ref), 1e-3 * np.linalg.norm(x[0]) * G * dt / A_EARTH)
        assert np.linalg.norm(x[k] - ref) < 1e-12 * scale, (k, np.linalg.norm(x[k] - ref) / scale)


@pytest.mark.parametrize("dt", [600.0, 3600.0])
def test_rexi_inertial_oscillation_closed_form(orc31, dt):
    # rotation only (h0 = 0, h = 0): u_t = f v, v_t = -f u with f = 2 Omega mu, so
    # pointwise (u, v)(dt) = C(f dt)(u0, v0) + f dt S(f dt)(-v0, u0)
    orc = orc31
    u, v = _lowdeg(31, 1, 8), _lowdeg(31, 2, 8)
    z = np.zeros_like(u)
    hh, M = 0.2, 64
    prm = dict(a=A_EARTH, g=G, omega=OMEGA, h0=0.0, visc=0.0, rexi_h=hh, rexi_M=M, half=1, dt=dt)
    x = orc.swe_rexi_run(prm, z, u, v, 1)
    al, be = O.rexi_coeffs(hh, M, True)
    xf = 2 * OMEGA * orc.ct * dt
    C, S = _C_S(al, be, xf)
    shape = (orc.nphi, orc.nlat)
    u0, v0 = orc.synthesis(u).reshape(shape), orc.synthesis(v).reshape(shape)
    ue = C * u0 - xf * S * v0
    ve = C * v0 + xf * S * u0
    assert np.max(np.abs(x[0])) == 0.0
    assert rel_l2(orc.synthesis(x[1]).reshape(shape), ue) < 1e-12
    assert rel_l2(orc.synthesis(x[2]).reshape(shape), ve) < 1e-12


@pytest.mark.parametrize("omega", [OMEGA, 0.0])
def test_rexi_first_order_matches_linear_tendency(orc31, omega):
    # f(dt L) = C0 - dt L S0 + O(dt^2) with C0 = sum beta/alpha, S0 = sum beta/alpha^2:
    # the Richardson pair of (f(dt L) U - C0 U) / dt tends to -S0 L U, L the linear
    # tendency (pinned against the analytic operators in test_oracle_swe.py),
    # including every Coriolis term
    orc = orc31
    h = random_real_spec(31, 3, power=-2.0) * 50
    h[0] = 0
    u, v = random_real_spec(31, 4, power=-2.0), random_real_spec(31, 5, power=-2.0)
    hh, M = 0.2, 64
    al, be = O.rexi_coeffs(hh, M, True)
    C0, S0 = np.sum(be / al).real, np.sum(be / al ** 2).real
    T = orc.swe_tendency(dict(a=A_EARTH, g=G, omega=omega, h0=H0, visc=0.0, nonlinear=0, dt=0.0), h, u, v)

    def D(dt):
        prm = dict(a=A_EARTH, g=G, omega=omega, h0=H0, visc=0.0, rexi_h=hh, rexi_M=M, half=1, dt=dt)
        x = orc.swe_rexi_run(prm, h, u, v, 1)
        return [(x[k] - C0 * y) / dt for k, y in enumerate((h, u, v))]

    errs = []
    for dt in (4.0, 8.0):
        d1, d2 = D(dt), D(2 * dt)
        errs.append([rel_l2(2 * d1[k] - d2[k], -S0 * T[k]) for k in range(3)])
    # second-order Richardson: the error falls ~4x per halving of dt
    assert max(errs[0]) < 1e-5, errs
    assert max(errs[0]) < 0.4 * max(errs[1]), errs
