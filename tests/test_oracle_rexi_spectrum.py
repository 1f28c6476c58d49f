"""Spectrum of the linear SWE operator and of the REXI step, per order m, on
the CPU (oracle only).

The linear tendency of the scalar-wind equations (AppSWE.hpp:593-606, no
viscosity) couples degrees only within one order m, so its restriction to
order m is a dense (3n x 3n) matrix, n = L - m + 1, built column by column
from the oracle's tendency applied to unit coefficients.  Two facts about the
reference's method follow from it:

* the spatial operator is neutral: every eigenvalue is imaginary to round-off,
  so the exact flow exp(dt T) has spectral radius 1, and classic RK4
  (TimesteppingRK.hpp:314-327) is stable exactly for dt <= 2 sqrt(2) / rho(T),
  the imaginary-axis limit of its stability polynomial;
* the REXI step (SWERexiSPH.hpp:137-244 summed over the poles of
  REXI::setup) is not exp(dt T) at finite truncation: its matrix, built the
  same way from the oracle's REXI step, has spectral radius above 1 with the
  growing vector in the upper degrees of the order.  The step agrees with
  exp(dt T) to first order in dt only: its dt^2 term is not T^2 / 2, since
  the per-pole solve eliminates the winds with operator identities of the
  continuous equations, which the truncated scalar-wind fields do not
  satisfy.  The excess of the spectral radius over 1 shrinks like dt^2.  That growth therefore belongs to the
  reference's REXI formulation, not to the tendency and not to an
  implementation of either (the GPU reproduces the oracle's growth value by
  value at T63).
"""
import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O
from tests._util import A_EARTH, G, H0, OMEGA

L, M_ORD = 16, 1


@pytest.fixture(scope="module")
def blocks():
    orc = O.Oracle(L)
    idx = [orc.idx(l, M_ORD) for l in range(M_ORD, L + 1)]
    n = len(idx)

    def unit(k):
        f, j = divmod(k, n)
        x = [np.zeros(orc.nlm, complex) for _ in range(3)]
        x[f][idx[j]] = 1.0
        return x

    def block(apply):
        B = np.zeros((3 * n, 3 * n), complex)
        for k in range(3 * n):
            B[:, k] = np.concatenate([o[idx] for o in apply(unit(k))])
        return B

    prm = dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=0.0, nonlinear=0)
    T = block(lambda x: orc.swe_tendency(prm, *x))
    # the block is the whole operator on order m: nothing leaks to other orders
    x = unit(5)
    full = orc.swe_tendency(prm, *x)
    leak = [np.delete(full[f], idx) for f in range(3)]
    assert max(np.max(np.abs(y)) for y in leak) < 1e-12 * np.max(np.abs(T))

    def rexi(dt):
        pr = dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=0.0, rexi_h=0.2, rexi_M=16, half=1, dt=dt)
        return block(lambda x: orc.swe_rexi_run(pr, *x, 1))

    return T, rexi, n


def test_linear_tendency_is_neutral_and_rk4_limit(blocks):
    T, _, n = blocks
    ev = np.linalg.eigvals(T)
    rho = np.max(np.abs(ev))
    assert np.max(np.abs(ev.real)) < 1e-12 * rho
    # the exact flow keeps every mode's amplitude
    dt = 300.0
    assert abs(np.max(np.abs(np.linalg.eigvals(sl.expm(dt * T)))) - 1.0) < 1e-10
    # classic RK4: R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24, |R(iy)| <= 1 iff |y| <= 2 sqrt 2
    dt_max = 2 * np.sqrt(2) / rho
    for fac, stable in ((0.95, True), (1.05, False)):
        z = fac * dt_max * ev
        amp = np.max(np.abs(1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24))
        assert (amp <= 1 + 1e-12) == stable, (fac, amp)
    print("T%d m=%d: rho(T) = %.4e 1/s, RK4 limit dt = %.1f s" % (L, M_ORD, rho, dt_max))


def test_rexi_step_amplifies_and_is_first_order_only(blocks):
    T, rexi, n = blocks
    S = {dt: rexi(dt) for dt in (300.0, 150.0)}
    I = np.eye(3 * n)
    # fields in common units: h by sqrt(g / h0), winds by 1
    w = np.concatenate([np.full(n, np.sqrt(G / H0)), np.ones(2 * n)])
    sc = lambda A: np.linalg.norm(w[:, None] * A / w[None, :])
    rho, r1, r2 = {}, {}, {}
    for dt, Sd in S.items():
        es, V = np.linalg.eig(Sd)
        j = np.argmax(np.abs(es))
        rho[dt] = abs(es[j])
        print("T%d m=%d dt=%g: rho(REXI step) = %.6f" % (L, M_ORD, dt, rho[dt]))
        if dt == 300.0:
            # the growing vector lives in the upper half of the degrees
            e = np.abs(w * V[:, j]).reshape(3, n) ** 2
            assert e[:, n // 2:].sum() > 0.9 * e.sum()
        r1[dt] = sc(Sd - I - dt * T)                       # first-order residual
        r2[dt] = sc(Sd - I - dt * T - (dt * T) @ (dt * T) / 2)  # second-order residual
    # spectral radius above 1, the excess shrinking like dt^2 (so the growth per
    # unit time vanishes as dt -> 0: consistent, but unstable at a fixed dt)
    assert rho[300.0] > 1.02
    assert 3.0 < (rho[300.0] - 1) / (rho[150.0] - 1) < 5.0
    # S = I + dt T + O(dt^2): the first-order residual falls like dt^2 ...
    assert 3.0 < r1[300.0] / r1[150.0] < 5.0
    # ... but the second-order residual also falls only like dt^2 (an exact
    # exponential would give dt^3): the step's dt^2 term is not T^2 / 2
    assert 3.0 < r2[300.0] / r2[150.0] < 5.0
    assert r2[300.0] > 0.5 * r1[300.0]
