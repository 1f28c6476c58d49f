// Composite 5-point Gauss-Legendre quadrature (src/include/libmath/
// GaussQuadrature.hpp:15-91): the rule on one interval, on n equal strips, and
// the adaptive form that adds strips (1, 2, 3, ...) until two successive sums
// differ by no more than a threshold.  The nodes are the roots of P_5, 0 and
// +-(1/3) sqrt(5 -+ 2 sqrt(10/7)), with weights 128/225 and (322 +- 13 sqrt(70))/900.
#ifndef SPHSWE_HOST_GAUSSQUADRATURE_HPP
#define SPHSWE_HOST_GAUSSQUADRATURE_HPP

#include <cmath>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <limits>

class GaussQuadrature
{
public:
	static double integrate5(double i_start, double i_end, std::function<double(double)> i_fun)
	{
		const double s10 = std::sqrt(10.0 / 7.0), s70 = std::sqrt(70.0);
		const double xa = std::sqrt(5.0 - 2.0 * s10) / 3.0, xb = std::sqrt(5.0 + 2.0 * s10) / 3.0;
		const double x[5] = {0.0, -xa, xa, -xb, xb};
		const double w[5] = {128.0 / 225.0, (322.0 + 13.0 * s70) / 900.0, (322.0 + 13.0 * s70) / 900.0,
				     (322.0 - 13.0 * s70) / 900.0, (322.0 - 13.0 * s70) / 900.0};
		const double half = 0.5 * (i_end - i_start), mid = 0.5 * (i_end + i_start);
		double acc = 0.0;
		for (int i = 0; i < 5; i++) acc += w[i] * i_fun(half * x[i] + mid);
		return acc * half;
	}

	static double integrate5_intervals(double i_start, double i_end, std::function<double(double)> i_fun, int n_strips)
	{
		const double d = (i_end - i_start) / (double)n_strips;
		double acc = 0.0;
		for (int i = 0; i < n_strips; i++) acc += integrate5(i_start + d * (double)i, i_start + d * (double)(i + 1), i_fun);
		return acc;
	}

	static double integrate5_intervals_adaptive(double i_start, double i_end, std::function<double(double)> i_fun,
						    double i_error_threshold = 10e-13)
	{
		double prev = std::numeric_limits<double>::infinity();
		for (int n = 1; n < 1024 * 128; n++) {
			const double v = integrate5_intervals(i_start, i_end, i_fun, n);
			if (std::abs(prev - v) <= i_error_threshold) return v;
			prev = v;
		}
		std::cerr << "No convergence reached for integrate5_intervals_adaptive" << std::endl;
		exit(1);
	}
};

#endif
