"""The Galewsky jet of host program scenario Q1
(host/include/benchmarks/BenchmarkGalewsky.hpp, after src/include/benchmarks/
BenchmarkGalewsky.hpp:13-246) and the composite Gauss rule it integrates with
(host/include/libmath/GaussQuadrature.hpp, after src/include/libmath/
GaussQuadrature.hpp:15-91), on the CPU: the 5-point rule is exact for degree 9
and not for degree 10; the balanced height on the T63 Gauss latitudes equals an
independent computation with SciPy's adaptive quadrature; the area mean is
10 km; geostrophic balance holds between rings."""
import os
import subprocess

import numpy as np
import pytest
from scipy.integrate import quad

HOST_INC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph_swe_amd", "host", "include")

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "benchmarks/BenchmarkGalewsky.hpp"
int main() {
	int mode;
	std::scanf("%d", &mode);
	if (mode == 0) {   // integrate x^k over [a, b] with one strip, k = 0..10
		double a, b;
		std::scanf("%lf %lf", &a, &b);
		for (int k = 0; k <= 10; k++)
			std::printf("%.17g\n", GaussQuadrature::integrate5(a, b, [k](double x) { return std::pow(x, k); }));
		std::printf("%.17g\n", GaussQuadrature::integrate5_intervals_adaptive(0.0, 1.5707963267948966,
			[](double x) { return std::cos(x); }));
		return 0;
	}
	int n;
	std::scanf("%d", &n);
	std::vector<double> lat(n);
	for (auto &x : lat) std::scanf("%lf", &x);
	SimVars sv;
	BenchmarkGalewsky g(sv);
	for (double h : g.balanced_h(lat)) std::printf("%.17g\n", h);
	for (double p : lat) std::printf("%.17g\n", g.initial_condition_u(0.0, p));
	std::printf("%.17g\n", g.bump(3.0, 0.8));
}
"""

A, G, OMEGA = 6.37122e6, 9.80616, 7.292e-5


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("galewsky")
    (d / "g.cpp").write_text(DRIVER)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", HOST_INC, str(d / "g.cpp"), "-o", str(d / "g")], check=True)
    return str(d / "g")


def _run(exe, text):
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout
    return np.array([float(x) for x in out.split()])


def test_gauss5_exactness(exe):
    a, b = -0.3, 1.7
    v = _run(exe, "0 %r %r\n" % (a, b))
    exact = np.array([(b ** (k + 1) - a ** (k + 1)) / (k + 1) for k in range(11)])
    assert np.max(np.abs(v[:10] - exact[:10]) / np.abs(exact[:10])) < 1e-14
    assert abs(v[10] - exact[10]) / abs(exact[10]) > 1e-8       # degree 10 is not exact
    assert abs(v[11] - 1.0) < 1e-13                             # adaptive, int cos over [0, pi/2]


def test_balanced_jet_vs_scipy(exe):
    x, w = np.polynomial.legendre.leggauss(96)
    lat = np.arcsin(x[::-1])            # north first
    v = _run(exe, "1 %d\n%s\n" % (len(lat), " ".join("%.17g" % p for p in lat)))
    h, u, bump = v[:96], v[96:192], v[192]
    phi0, phi1 = np.pi / 7, np.pi / 2 - np.pi / 7
    en = np.exp(-4 / (phi1 - phi0) ** 2)
    uo = lambda p: 0.0 if (p <= phi0 or p >= phi1) else 80.0 / en * np.exp(1 / ((p - phi0) * (p - phi1)))
    F = lambda p: A * uo(p) * (2 * OMEGA * np.sin(p) + np.tan(p) / A * uo(p))

    def integ(s, e):
        s2, e2 = max(s, phi0), min(e, phi1)
        return 0.0 if e2 <= s2 else quad(F, s2, e2, epsabs=1e-12, epsrel=1e-13, limit=200)[0]

    hg, start, prev, hsum, area = np.empty(96), -np.pi / 2, 0.0, 0.0, 0.0
    for j in range(95, -1, -1):
        prev -= integ(start, lat[j])
        hg[j] = prev
        wgt = (np.sin(lat[j]) - np.sin(start)) * 2 * np.pi
        hsum += prev * wgt
        area += wgt
        start = lat[j]
    last = prev - integ(start, np.pi / 2)
    wgt = (1 - np.sin(start)) * 2 * np.pi
    hsum += last * wgt
    area += wgt
    ref = hg / G + (10000.0 - hsum / G / area)
    assert np.max(np.abs(h - ref)) < 1e-8
    assert np.allclose(u, [uo(p) for p in lat], rtol=1e-14, atol=1e-300)
    assert np.max(u) > 79.0 and np.min(h) < 9300 and np.max(h) > 10000   # the jet is there
    # geostrophic balance between rings, g dh = -F dphi (midpoint rule, second order)
    dh = np.diff(h[::-1]) * G
    mid = 0.5 * (lat[::-1][1:] + lat[::-1][:-1])
    dphi = np.diff(lat[::-1])
    Fm = np.array([F(p) for p in mid])
    assert np.max(np.abs(dh + Fm * dphi)) < 1.5e-2 * np.max(np.abs(Fm * dphi))
    # the bump's closed form at one point
    assert abs(bump - 120 * np.cos(0.8) * np.exp(-((3.0 - np.pi) * 3) ** 2) * np.exp(-((np.pi / 4 - 0.8) * 15) ** 2)) < 1e-12
