"""Pin of the oracle's Robert-function tendency (AppSWE.hpp:606-620,
SPHOperators.hpp:124-208): the linear equations for U = u cos(phi),
V = v cos(phi).  In the continuum, with U, V the Robert functions of (u, v),
  h_t(Robert)     = h_t(standard),
  U_t(Robert)     = cos(phi) u_t(standard),
  V_t(Robert)     = cos(phi) v_t(standard),
so the Robert branch is checked against the oracle's standard linear tendency
(pinned through the SHT and operator tests) on smooth, well-resolved fields
that vanish near the poles.  A dropped term, a wrong sign or a missing
division by cos^2 phi changes the result at O(1)."""
import numpy as np

from oracle import oracle as O
from tests._util import A_EARTH, G, H0, OMEGA, rel_l2


def _bump(orc, lat0, lon0, w):
    lat = np.arcsin(orc.ct)[None, :]
    lon = orc.lon()[:, None]
    c = np.sin(lat) * np.sin(lat0) + np.cos(lat) * np.cos(lat0) * np.cos(lon - lon0)
    d = np.arccos(np.clip(c, -1, 1))
    return np.exp(-(d / w) ** 2)


def test_robert_tendency_equals_cos_weighted_standard():
    L = 127
    orc = O.Oracle(L)
    cosphi = np.asarray(orc.st)[None, :]
    h = (50.0 * _bump(orc, 0.2, 1.0, 0.3)).reshape(-1)
    u = (10.0 * _bump(orc, -0.1, 2.5, 0.35)).reshape(orc.nphi, orc.nlat)
    v = (8.0 * _bump(orc, 0.3, 4.0, 0.3)).reshape(orc.nphi, orc.nlat)
    prm = dict(a=A_EARTH, g=G, omega=OMEGA, h0=H0, visc=0.0, nonlinear=0)
    hs, us, vs = (orc.analysis(x.reshape(-1)) for x in (h, u, v))
    Us, Vs = (orc.analysis((x * cosphi).reshape(-1)) for x in (u, v))
    th, tu, tv = orc.swe_tendency(prm, hs, us, vs)
    rh, rU, rV = orc.swe_tendency(dict(prm, robert=1), hs, Us, Vs)
    shape = (orc.nphi, orc.nlat)
    # measured 8.5e-10 (h), 4.4e-11 (U), 6.7e-11 (V): truncation of the products
    # with cos(phi); passing u for U, or dropping the Coriolis term, gives 3e-2
    # and 0.8
    assert rel_l2(rh, th) < 1e-8
    assert rel_l2(orc.synthesis(rU).reshape(shape), cosphi * orc.synthesis(tu).reshape(shape)) < 1e-8
    assert rel_l2(orc.synthesis(rV).reshape(shape), cosphi * orc.synthesis(tv).reshape(shape)) < 1e-8
    # the flag belongs to the linear branch only (the reference reads it there)
    nl = dict(prm, nonlinear=1)
    a = orc.swe_tendency(nl, hs, us, vs)
    b = orc.swe_tendency(dict(nl, robert=1), hs, us, vs)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
