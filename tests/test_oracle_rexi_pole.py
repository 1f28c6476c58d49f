"""Pins of the oracle's per-pole REXI solve (SWERexiSPH::setup/solve,
src/SWERexiSPH.hpp:68-130, 137-244) against continuous calculus.

The oracle assembles the phi system from the band components z1..z8 of
SPHSolverComplex (src/include/sph/SPHSolverComplex.hpp:88-420) and builds the
right-hand side through spectral operators and physical products.  Here the same
operator and right-hand side are written as differential expressions in mu, and
for a state confined to an equatorial band (so that every product and metric
factor of the discrete path is exact at T63) the per-m ODE is solved by an
independent Legendre collocation.  The symbolic checks record how the reference's
formulation relates to the exact elimination of (u, v) from the linear SWE: it
agrees except for z8 and the sign of the Fc term (both reference behaviour, kept).
"""
import numpy as np
import pytest
import sympy as sp
from numpy.polynomial import legendre as Lg

from oracle import oracle as O
from tests._util import A_EARTH, G, H0, OMEGA

MU, AL, TO, GB, R = sp.symbols("mu alpha to Gbar r")
S = sp.sqrt(1 - MU ** 2)


def _lap(x, m):
    return sp.diff((1 - MU ** 2) * sp.diff(x, MU), MU) - m * m * x / (1 - MU ** 2)


def _components(phi, m):
    """Continuous form of the band components on phi(mu) e^{i m lam} (unit sphere;
    the 1/r^2 of z4..z8 included).  z8 is what solver_component_rexi_z8 assembles
    (SPHSolverComplex.hpp:400-420): 2 phi - lap(mu^2 phi) - m^2 mu^2 phi, not the
    mu^2 lap(phi) its comment names."""
    return {
        11: phi, 12: MU ** 2 * phi, 13: MU ** 4 * phi,
        14: sp.I * m * phi / R ** 2,
        15: sp.I * m * MU ** 2 * phi / R ** 2,
        16: MU * (1 - MU ** 2) * sp.diff(phi, MU) / R ** 2,
        17: _lap(phi, m) / R ** 2,
        18: (2 * phi - _lap(MU ** 2 * phi, m) - m * m * MU ** 2 * phi) / R ** 2,
    }


def _phi_operator(phi, m, mu2lap_z8=False):
    """The phi system of SWERexiSPH::setup (SWERexiSPH.hpp:88-113)."""
    z = _components(phi, m)
    if mu2lap_z8:
        z[18] = MU ** 2 * _lap(phi, m) / R ** 2
    coef = {11: AL ** 4, 12: 2 * TO ** 2 * AL ** 2, 13: TO ** 4, 14: -GB * AL * TO,
            15: GB / AL * TO ** 3, 16: GB * 2 * TO ** 2, 17: -GB * AL ** 2, 18: -GB * TO ** 2}
    return sum(coef[k] * z[k] for k in coef)


def _rhs(phi0, u0, v0, m):
    """Right-hand side of SWERexiSPH::solve (SWERexiSPH.hpp:160-200): div0 and eta0
    from the reference operators (vort = grad_lon(v) - grad_lat(u), no metric term)."""
    glon = lambda x: sp.I * m * x / S
    glat = lambda x: S * sp.diff(x, MU)
    div0 = (glon(u0) + sp.diff(v0 * S, MU)) / R
    eta0 = (glon(v0) - glat(u0)) / R
    foo = GB * (div0 - TO / AL * MU * eta0) + (AL * phi0 + TO ** 2 / AL * MU ** 2 * phi0)
    fc = (TO / R) * S * (-(AL ** 2 * u0 - TO ** 2 * MU ** 2 * u0) + 2 * AL * TO * MU * v0)
    return AL ** 2 * foo + TO ** 2 * MU ** 2 * foo + fc * GB / AL, fc


def _elimination(phi, phi0, u0, v0, m):
    """Exact elimination of (u, v) from alpha x + L x = x0 (L the linear SWE
    operator of AppSWE.hpp:591-603), multiplied by (alpha^2 + f^2)^2: returns
    (operator on phi, right-hand side)."""
    f = TO * MU
    den = AL ** 2 + f ** 2
    glon = lambda x: sp.I * m * x / S
    div = lambda u, v: glon(u) + sp.diff(v * S, MU)
    A, B = glon(phi) / R, S * sp.diff(phi, MU) / R
    op = den ** 2 * (AL * phi - GB * div((AL * A - f * B) / den, (f * A + AL * B) / den) / R)
    rhs = den ** 2 * phi0 + GB / R * den ** 2 * div((AL * u0 - f * v0) / den, (f * u0 + AL * v0) / den)
    return op, rhs


def test_phi_operator_equals_elimination_except_z8():
    psi = sp.Function("psi")(MU)
    m = sp.Symbol("m")
    op_el, _ = _elimination(psi, 0, 0, 0, m)
    # with mu^2 lap in place of the assembled z8, alpha * Mphi is the exact elimination
    d = sp.simplify(sp.expand(AL * _phi_operator(psi, m, mu2lap_z8=True) - op_el))
    assert d == 0
    # the assembled z8 differs from mu^2 lap by -2 mu^2 lap - 4 mu (1-mu^2) d/dmu + (6 - m^2) mu^2
    z8 = _components(psi, m)[18] * R ** 2
    expect = (-MU ** 2 * _lap(psi, m) - 4 * MU * (1 - MU ** 2) * sp.diff(psi, MU)
              + (6 - m * m) * MU ** 2 * psi)
    assert sp.simplify(sp.expand(z8 - expect)) == 0


def test_rhs_relation_to_elimination():
    phi0, u0, v0 = (sp.Function(n)(MU) for n in ("phi0", "u0", "v0"))
    m = sp.Symbol("m")
    rhs_ref, fc = _rhs(phi0, u0, v0, m)
    # exact elimination with the true vorticity (1/cos) d(u cos)/dphi in the Coriolis
    # term reproduces alpha * rhs_ref up to twice the Fc term: the reference's Fc has
    # the opposite sign of the exact elimination
    glon = lambda x: sp.I * m * x / S
    div0 = (glon(u0) + sp.diff(v0 * S, MU)) / R
    eta_true = (glon(v0) - sp.diff(u0 * S, MU)) / R
    foo = GB * (div0 - TO / AL * MU * eta_true) + (AL * phi0 + TO ** 2 / AL * MU ** 2 * phi0)
    rhs_true_vort = AL ** 2 * foo + TO ** 2 * MU ** 2 * foo - fc * GB / AL
    _, rhs_el = _elimination(sp.Integer(0), phi0, u0, v0, m)
    assert sp.simplify(sp.expand(AL * rhs_true_vort - rhs_el)) == 0


# --- band components against their continuous operators (r = 1) -------------

def _nc(L, n, m):
    if m < 0:
        minc = L + m
        return ((minc - 1) * minc >> 1) + minc + (n + m)
    return (L * (L + 1) >> 1) + (m * (2 * L - m + 1) >> 1) + n


def _matvec(L, d, halo, x):
    y = np.zeros_like(x)
    for m in range(-L, L + 1):
        for n in range(abs(m), L + 1):
            row = d[_nc(L, n, m)]
            for rel in range(-halo, halo + 1):
                nn = n + rel
                if abs(m) <= nn <= L:
                    y[n * n + n + m] += row[rel + halo] * x[nn * nn + nn + m]
    return y


@pytest.fixture(scope="module")
def orc31():
    return O.Oracle(31)


@pytest.mark.parametrize("m", [0, 3, -2])
def test_band_components_vs_continuous(orc31, m):
    # z-components at r = 1 (at the Earth radius z4..z8 are 1e-14 of z1, which is why
    # the reference's own known-answer test cannot see them)
    orc, L = orc31, 31
    rng = np.random.default_rng(10 + m)
    c = rng.standard_normal(6)
    k = abs(m)
    # phi = s^|m| P(mu), P a degree-5 Legendre series: degree |m| + 5 <= L - 4
    P = sum(float(c[i]) * sp.legendre(i, MU) for i in range(6))
    phi = S ** k * P
    lon = orc.lon()[:, None]
    mu = orc.ct[None, :]
    e = np.exp(1j * m * lon)
    x = orc.analysis_cplx((sp.lambdify(MU, phi, "numpy")(mu) * e).reshape(-1))
    comps = _components(phi, m)
    comps[1] = MU * phi
    comps[2] = (1 - MU ** 2) * sp.diff(phi, MU)
    for comp, expr in comps.items():
        Mx = O.Matrix(orc, 4 if comp == 13 else 2)
        Mx.add(comp, 1.0, 1.0)
        y = _matvec(L, Mx.data(), Mx.halo, x)
        fn = sp.lambdify(MU, sp.simplify(expr.subs(R, 1)), "numpy")
        yr = orc.analysis_cplx((np.broadcast_to(fn(mu), mu.shape) * e + 0j).reshape(-1))
        # round-off leakage of the analysis (~1e-16 of x) is amplified up to L^2 by z7, z8
        scale = max(np.linalg.norm(yr), np.linalg.norm(x))
        assert np.linalg.norm(y - yr) < 1e-11 * scale, comp


# --- one pole: discrete solve vs per-m ODE of the same formulation ------------

SIGMA = 0.25


def _band(mu):
    """Equatorial band exp(-(lat/sigma)^2) and its mu-derivative: below 1e-17 at the
    poles and band-limited at T63."""
    lat = np.arcsin(mu)
    b = np.exp(-lat ** 2 / SIGMA ** 2)
    return b, b * (-2 * lat / SIGMA ** 2) / np.sqrt(1 - mu * mu)


def _mode_coefficients(m):
    """Lambdified ODE coefficients for psi (phi = s^|m| psi) and the right-hand side in
    terms of phi0, u0, u0', v0, v0' at mu; arguments (mu, alpha, to, Gbar, r, ...)."""
    k = abs(m)
    psi = sp.Function("psi")(MU)
    op = (_phi_operator(S ** k * psi, m) / S ** k).doit()
    y0, y1, y2 = sp.symbols("y0 y1 y2")
    lin = op.subs(sp.diff(psi, MU, 2), y2).subs(sp.diff(psi, MU), y1).subs(psi, y0)
    args = (MU, AL, TO, GB, R)
    cs = [sp.lambdify(args, sp.diff(lin, y), "numpy") for y in (y2, y1, y0)]
    fp, fu, fv = (sp.Function(n)(MU) for n in ("fp", "fu", "fv"))
    rhs, _ = _rhs(fp, fu, fv, m)
    P, U, Ud, V, Vd = sp.symbols("P U Ud V Vd")
    rhs = (rhs / S ** k).doit()
    rhs = rhs.subs(sp.diff(fu, MU), Ud).subs(sp.diff(fv, MU), Vd).subs({fu: U, fv: V, fp: P})
    return cs, sp.lambdify(args + (P, U, Ud, V, Vd), rhs, "numpy")


def _solve_mode(m, prm, alpha, cp, cu, cv, N=110):
    """phi_m(mu) of Mphi phi = rhs for phi0, u0, v0 = band(mu) * (cp, cu, cv) e^{i m lam},
    by Legendre collocation at N + 1 Gauss points."""
    k = abs(m)
    (f2, f1, f0), fr = _mode_coefficients(m)
    x, _ = Lg.leggauss(N + 1)
    a = (x, alpha, 2 * prm["omega"], prm["g"] * prm["h0"], prm["a"])
    b, db = _band(x)
    a2, a1, a0 = (np.broadcast_to(f(*a), x.shape) for f in (f2, f1, f0))
    V = np.zeros((N + 1, N + 1), complex)
    for j in range(N + 1):
        cj = np.zeros(N + 1)
        cj[j] = 1
        V[:, j] = (a2 * Lg.legval(x, Lg.legder(cj, 2)) + a1 * Lg.legval(x, Lg.legder(cj))
                   + a0 * Lg.legval(x, cj))
    rhs = np.broadcast_to(fr(*a, cp * b, cu * b, cu * db, cv * b, cv * db), x.shape)
    coef = np.linalg.solve(V, rhs.astype(complex))
    assert np.max(np.abs(coef[-6:])) < 1e-11 * np.max(np.abs(coef))    # resolved
    return lambda mu: (1 - mu * mu) ** (k / 2) * Lg.legval(mu, coef)


@pytest.mark.parametrize("n", [0, -7, -40])
def test_pole_phi_vs_continuous(n):
    L = 63
    orc = O.Oracle(L)
    prm = dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0)
    dt = 3600.0
    alpha = 0.2 * (-4.315321510875024 + 1j * n) / dt
    rng = np.random.default_rng(40 + n)
    ms = [0, 1, 2, 3]
    coeffs = {m: rng.standard_normal(3) + 1j * rng.standard_normal(3) for m in ms}
    coeffs[0] = coeffs[0].real + 0j
    for m in ms:
        coeffs[m] *= np.array([1e3, 1.0, 1.0])          # phi0 [m^2/s^2], u0, v0 [m/s]
    lon = orc.lon()[:, None]
    mu = orc.ct[None, :]
    bnum = _band(mu)[0]
    fields = []
    for q in range(3):
        f = np.zeros((orc.nphi, orc.nlat))
        for m in ms:
            w = 1.0 if m == 0 else 2.0
            f = f + w * (coeffs[m][q] * bnum * np.exp(1j * m * lon)).real
        fields.append(orc.analysis(f))
    phi, _, _ = orc.rexi_pole_solve(prm, alpha, *fields)
    ref = np.zeros((orc.nphi, orc.nlat), complex)
    for m in ms:
        for sgn in ((1, -1) if m else (1,)):
            c = coeffs[m] if sgn == 1 else np.conj(coeffs[m])
            fm = _solve_mode(sgn * m, prm, alpha, *c)
            ref += fm(mu) * np.exp(1j * sgn * m * lon)
    err = np.linalg.norm(phi - ref) / np.linalg.norm(ref)
    print("pole", n, "rel err", err)
    assert err < 1e-7
