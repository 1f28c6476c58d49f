// REXI coefficients (src/include/rexi/REXI.hpp:52-166): poles alpha and the
// weights beta_re / beta_im of the rational approximation
// cos x ~ sum Re(beta_re / (ix + alpha)), sin x ~ sum Re(beta_im / (ix + alpha)),
// computed by the library (sphswe_rexi_coefficients, no GPU needed).
#ifndef SPHSWE_HOST_REXI_HPP
#define SPHSWE_HOST_REXI_HPP

#include <complex>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "sphswe.h"

class REXI
{
	typedef std::complex<double> complex;

public:
	std::vector<complex> alpha, beta_re, beta_im;

	// i_L: order of the Gaussian approximation; 0 selects the reference's default (11)
	void setup(double i_h, int i_M, int i_L = 0, bool i_reduce_to_half = true)
	{
		if (i_L != 0 && i_L != 11) {
			std::cerr << "REXI: only the published L = 11 Gaussian approximation is available" << std::endl;
			exit(1);
		}
		int n = 0;
		if (sphswe_rexi_coefficients(i_h, i_M, i_reduce_to_half, &n, nullptr, nullptr, nullptr) != SPHSWE_OK) {
			std::cerr << "REXI: invalid h or M" << std::endl;
			exit(1);
		}
		alpha.assign(n, 0.0);
		beta_re.assign(n, 0.0);
		beta_im.assign(n, 0.0);
		sphswe_rexi_coefficients(i_h, i_M, i_reduce_to_half, &n, (double *)alpha.data(), (double *)beta_re.data(),
					 (double *)beta_im.data());
	}

	complex eval_e_ix(double i_x) { return std::exp(complex(0, i_x)); }

	complex approx_e_ix(double i_x)
	{
		double sum_re = 0, sum_im = 0;
		for (std::size_t n = 0; n < alpha.size(); n++) {
			const complex denom = complex(0, i_x) + alpha[n];
			sum_re += (beta_re[n] / denom).real();
			sum_im += (beta_im[n] / denom).real();
		}
		return complex(sum_re, sum_im);
	}
};

#endif
