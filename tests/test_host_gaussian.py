"""The Gaussian bump of the self-test programs
(host/include/benchmarks/SphereTestSolutions_Gaussian.hpp, after
src/include/benchmarks/SphereTestSolutions_Gaussian.hpp) against central
finite differences of its own test function: every analytic derivative the P4
program feeds the GPU band solver is pinned by the definition of the
derivative, not by a second copy of the formula."""
import os
import subprocess

import numpy as np
import pytest

HOST_INC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph_swe_amd", "host", "include")

DRIVER = r"""
#include <cstdio>
#include "benchmarks/SphereTestSolutions_Gaussian.hpp"
int main() {
	SphereTestSolutions_Gaussian t;
	double lon, mu;
	while (std::scanf("%lf %lf", &lon, &mu) == 2) {
		double v[10];
		t.test_function__grid_gaussian(lon, mu, v[0]);
		t.correct_result_diff_lambda__grid_gaussian(lon, mu, v[1]);
		t.correct_result_diff_phi__grid_gaussian(lon, mu, v[2]);
		t.correct_result_diff_mu__grid_gaussian(lon, mu, v[3]);
		t.correct_result_grad_lambda__grid_gaussian(lon, mu, v[4]);
		t.correct_result_mu__grid_gaussian(lon, mu, v[5]);
		t.correct_result_one_minus_mu_squared_diff_lat_mu__grid_gaussian(lon, mu, v[6]);
		t.correct_result_grad_phi__grid_gaussian(lon, mu, v[7]);
		t.correct_result_div_lambda__grid_gaussian(lon, mu, v[8]);
		t.correct_result_div_mu__grid_gaussian(lon, mu, v[9]);
		for (int k = 0; k < 10; k++) std::printf("%.17g ", v[k]);
		std::printf("\n");
	}
}
"""


@pytest.fixture(scope="module")
def gauss(tmp_path_factory):
    d = tmp_path_factory.mktemp("gauss")
    src, exe = d / "g.cpp", d / "g"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-O1", "-I", HOST_INC, str(src), "-o", str(exe)], check=True)

    def ev(lon, mu):
        inp = "".join("%.17g %.17g\n" % (a, b) for a, b in zip(np.ravel(lon), np.ravel(mu)))
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout
        return np.array([[float(x) for x in line.split()] for line in out.splitlines()])

    return ev


def test_gaussian_bump_peak_and_decay(gauss):
    # f = 1 at the centre (pi/5, pi/4), exp(-10 d^2) at great-circle angle d
    v = gauss([np.pi / 5, np.pi / 5 + np.pi], [np.sin(np.pi / 4), np.sin(-np.pi / 4)])
    assert abs(v[0, 0] - 1.0) < 1e-15
    assert abs(v[1, 0] - np.exp(-10 * np.pi ** 2)) < 1e-18
    # every derivative vanishes at the peak
    assert np.all(np.abs(v[0, [1, 2, 3, 4, 6, 7, 8]]) < 1e-12)


def test_gaussian_derivatives_match_finite_differences(gauss):
    rng = np.random.default_rng(5)
    lon = rng.uniform(0, 2 * np.pi, 40)
    phi = rng.uniform(-1.3, 1.3, 40)
    mu = np.sin(phi)
    h = 1e-5
    v = gauss(lon, mu)
    f = lambda a, p: gauss(a, np.sin(p))[:, 0]
    d_lon = (f(lon + h, phi) - f(lon - h, phi)) / (2 * h)
    d_phi = (f(lon, phi + h) - f(lon, phi - h)) / (2 * h)
    d_mu = (gauss(lon, mu + h)[:, 0] - gauss(lon, mu - h)[:, 0]) / (2 * h)
    cphi = np.cos(phi)
    # d(cos(phi) f)/dphi / cos(phi), the divergence of (0, f)
    div_mu = ((np.cos(phi + h) * f(lon, phi + h) - np.cos(phi - h) * f(lon, phi - h)) / (2 * h)) / cphi
    tol = 1e-7
    assert np.max(np.abs(v[:, 1] - d_lon)) < tol
    assert np.max(np.abs(v[:, 2] - d_phi)) < tol
    assert np.max(np.abs(v[:, 3] - d_mu)) < tol
    assert np.max(np.abs(v[:, 4] - d_lon / cphi)) < tol
    assert np.max(np.abs(v[:, 5] - mu * v[:, 0])) < 1e-15
    assert np.max(np.abs(v[:, 6] - (1 - mu ** 2) * d_mu)) < tol
    assert np.max(np.abs(v[:, 7] - d_phi)) < tol
    assert np.max(np.abs(v[:, 8] - d_lon / cphi)) < tol
    assert np.max(np.abs(v[:, 9] - div_mu)) < tol
    # the derivatives are not trivially small where the bump is
    assert np.max(np.abs(d_lon)) > 0.1 and np.max(np.abs(d_phi)) > 0.1


SPH_DRIVER = r"""
#include <cstdio>
#include "benchmarks/SphereTestSolutions_SPH.hpp"
int main() {
	int n, m;
	double lon, mu;
	while (std::scanf("%d %d %lf %lf", &n, &m, &lon, &mu) == 4) {
		SphereTestSolutions_SPH t(n, m);
		double f, d;
		t.test_function__grid_gaussian(lon, mu, f);
		t.correct_result_diff_mu__grid_gaussian(lon, mu, d);
		std::printf("%.17g %.17g\n", f, d);
	}
}
"""


def test_sph_test_solution_normalisation_and_derivative(tmp_path):
    # host/include/benchmarks/SphereTestSolutions_SPH.hpp: the norm on the
    # sphere (Gauss-Legendre in mu, exact for these polynomial degrees), the
    # closed form the reference lists for n = 2, m = 1, and (1 - mu^2) d/dmu
    # against central differences
    src, exe = tmp_path / "s.cpp", tmp_path / "s"
    src.write_text(SPH_DRIVER)
    subprocess.run(["g++", "-O1", "-I", HOST_INC, str(src), "-o", str(exe)], check=True)

    def ev(n, m, lon, mu):
        inp = "".join("%d %d %.17g %.17g\n" % (n, m, a, b) for a, b in zip(np.ravel(lon), np.ravel(mu)))
        out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout
        return np.array([[float(x) for x in line.split()] for line in out.splitlines()])

    x, w = np.polynomial.legendre.leggauss(24)
    lon = np.linspace(0, 2 * np.pi, 48, endpoint=False)
    for n, m in [(0, 0), (1, 0), (2, 1), (3, 3), (6, 4)]:
        v = ev(n, m, np.repeat(lon, len(x)), np.tile(x, len(lon)))[:, 0].reshape(len(lon), len(x))
        norm = np.sum((v ** 2) * w[None, :]) * (2 * np.pi / len(lon))
        # cos(m lon) / sqrt(2 pi) carries half the norm of e^{i m lon} for m > 0
        assert abs(norm - (1.0 if m == 0 else 0.5)) < 1e-12, (n, m, norm)
    mu = np.linspace(-0.95, 0.95, 21)
    v = ev(2, 1, np.zeros_like(mu), mu)
    assert np.max(np.abs(v[:, 0] - (-mu / 2 * np.sqrt(15 * (1 - mu * mu)) / np.sqrt(2 * np.pi)))) < 1e-15
    h = 1e-6
    for n, m in [(2, 1), (5, 3), (7, 0)]:
        d = (ev(n, m, 0.3 + 0 * mu, mu + h)[:, 0] - ev(n, m, 0.3 + 0 * mu, mu - h)[:, 0]) / (2 * h)
        assert np.max(np.abs(ev(n, m, 0.3 + 0 * mu, mu)[:, 1] - (1 - mu * mu) * d)) < 1e-8, (n, m)
