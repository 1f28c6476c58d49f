// A Gaussian bump on the sphere and its analytic derivatives, the known answers
// of the self-test programs (src/include/benchmarks/SphereTestSolutions_Gaussian.hpp:
// f = exp(-A d^2), d the great-circle angle to (center_lon, center_lat) =
// (pi/5, pi/4), A = 10).  With c = cos d = sin(phi) sin(phi0) +
// cos(phi) cos(phi0) cos(lam - lam0):
//   df/dx = 2 A d f (dc/dx) / sin d,
//   dc/dlam = -cos(phi) cos(phi0) sin(lam - lam0),
//   dc/dphi = cos(phi) sin(phi0) - sin(phi) cos(phi0) cos(lam - lam0),
// and d / sin d -> 1 at the centre.
#ifndef SPHSWE_HOST_SPHERETESTSOLUTIONS_GAUSSIAN_HPP
#define SPHSWE_HOST_SPHERETESTSOLUTIONS_GAUSSIAN_HPP

#include <cmath>

class SphereTestSolutions_Gaussian
{
	double exp_fac = 10.0;
	double center_lon = M_PI / 5;
	double center_lat = M_PI / 4;

	// (cos d, d / sin d, f) at (lon, phi)
	void geometry(double lon, double phi, double &c, double &d_over_sin, double &f) const
	{
		c = std::sin(phi) * std::sin(center_lat) + std::cos(phi) * std::cos(center_lat) * std::cos(lon - center_lon);
		if (c > 1.0) c = 1.0;
		if (c < -1.0) c = -1.0;
		const double d = std::acos(c), s = std::sqrt(1.0 - c * c);
		d_over_sin = s > 0.0 ? d / s : 1.0;
		f = std::exp(-exp_fac * d * d);
	}
	double dc_dlon(double lon, double phi) const
	{
		return -std::cos(phi) * std::cos(center_lat) * std::sin(lon - center_lon);
	}
	double dc_dphi(double lon, double phi) const
	{
		return std::cos(phi) * std::sin(center_lat) - std::sin(phi) * std::cos(center_lat) * std::cos(lon - center_lon);
	}

public:
	// f at (lon, mu)
	void test_function__grid_gaussian(double lon, double mu, double &o_data) const
	{
		test_function_phi__grid_phi(lon, std::asin(mu), o_data);
	}
	// f at (lon, phi)
	void test_function_phi__grid_phi(double lon, double phi, double &o_data) const
	{
		double c, q, f;
		geometry(lon, phi, c, q, f);
		o_data = f;
	}
	// df/dlam
	void correct_result_diff_lambda__grid_gaussian(double lon, double mu, double &o_data) const
	{
		const double phi = std::asin(mu);
		double c, q, f;
		geometry(lon, phi, c, q, f);
		o_data = 2.0 * exp_fac * q * f * dc_dlon(lon, phi);
	}
	// df/dphi
	void correct_result_diff_phi__grid_gaussian(double lon, double mu, double &o_data) const
	{
		const double phi = std::asin(mu);
		double c, q, f;
		geometry(lon, phi, c, q, f);
		o_data = 2.0 * exp_fac * q * f * dc_dphi(lon, phi);
	}
	// df/dmu = df/dphi / cos(phi)
	void correct_result_diff_mu__grid_gaussian(double lon, double mu, double &o_data) const
	{
		correct_result_diff_phi__grid_gaussian(lon, mu, o_data);
		o_data /= std::sqrt(1.0 - mu * mu);
	}
	// df/dlam / cos(phi)
	void correct_result_grad_lambda__grid_gaussian(double lon, double mu, double &o_data) const
	{
		correct_result_diff_lambda__grid_gaussian(lon, mu, o_data);
		o_data /= std::sqrt(1.0 - mu * mu);
	}
	void correct_result_mu__grid_gaussian(double lon, double mu, double &o_data) const
	{
		test_function__grid_gaussian(lon, mu, o_data);
		o_data *= mu;
	}
	// (1 - mu^2) df/dmu
	void correct_result_one_minus_mu_squared_diff_lat_mu__grid_gaussian(double lon, double mu, double &o_data) const
	{
		correct_result_diff_mu__grid_gaussian(lon, mu, o_data);
		o_data *= 1.0 - mu * mu;
	}
	void correct_result_grad_phi__grid_gaussian(double lon, double mu, double &o_data) const
	{
		correct_result_diff_phi__grid_gaussian(lon, mu, o_data);
	}
	void correct_result_div_lambda__grid_gaussian(double lon, double mu, double &o_data) const
	{
		correct_result_grad_lambda__grid_gaussian(lon, mu, o_data);
	}
	// d(cos(phi) f)/dphi / cos(phi) = df/dphi - tan(phi) f
	void correct_result_div_mu__grid_gaussian(double lon, double mu, double &o_data) const
	{
		const double phi = std::asin(mu);
		double c, q, f;
		geometry(lon, phi, c, q, f);
		o_data = 2.0 * exp_fac * q * f * dc_dphi(lon, phi) - std::tan(phi) * f;
	}
};

#endif
