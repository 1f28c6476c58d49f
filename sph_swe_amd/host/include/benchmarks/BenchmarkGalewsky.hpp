// The Galewsky et al. (2004) barotropic jet with its height bump, benchmark
// scenario 1 of the shallow-water program (src/include/benchmarks/
// BenchmarkGalewsky.hpp:13-246, used at AppSWE.hpp:199-207):
//   u(phi) = u_max / e_n exp(1 / ((phi - phi0)(phi - phi1))) on phi0 < phi < phi1,
//     u_max = 80 m/s, phi0 = pi/7, phi1 = pi/2 - phi0, e_n = exp(-4 / (phi1 - phi0)^2), v = 0;
//   g h(lat_j) = -integral from -pi/2 to lat_j of a u (2 Omega sin phi + tan(phi) u / a),
//     accumulated ring by ring from the south pole with the adaptive 5-point
//     Gauss rule (threshold 1e-13), then shifted so that the area mean (weights
//     2 pi (sin(end) - sin(start)) per interval, the value at the interval's
//     end, one more interval up to the north pole) is h_avg = 10 km;
//   bump: h += h_hat cos(phi) exp(-((lambda - pi) / alpha)^2) exp(-((phi2 - phi) / beta)^2),
//     h_hat = 120 m, alpha = 1/3, beta = 1/15, phi2 = pi/4.
#ifndef SPHSWE_HOST_BENCHMARKGALEWSKY_HPP
#define SPHSWE_HOST_BENCHMARKGALEWSKY_HPP

#include <cmath>
#include <vector>

#include "../SimVars.hpp"
#include "../libmath/GaussQuadrature.hpp"

class BenchmarkGalewsky
{
	const double phi0 = M_PI / 7.0;
	const double phi1 = M_PI / 2.0 - phi0;
	const double u_max = 80.0;
	const double h_avg = 10000.0;
	const double en = std::exp(-4.0 / ((phi1 - phi0) * (phi1 - phi0)));
	const double phi2 = M_PI / 4.0;
	const double alpha = 1.0 / 3.0;
	const double beta = 1.0 / 15.0;
	const double h_hat = 120.0;
	const double error_threshold = 1.e-13;
	SimVars &simVars;

public:
	explicit BenchmarkGalewsky(SimVars &i_simVars) : simVars(i_simVars) {}

	double initial_condition_u(double, double phi) const
	{
		if (phi <= phi0 || phi >= phi1) return 0.0;
		return u_max / en * std::exp(1.0 / ((phi - phi0) * (phi - phi1)));
	}

	// the integrand of g h: a u (2 Omega sin phi + tan(phi) u / a)
	double balance_integrand(double phi) const
	{
		const double u = initial_condition_u(0.0, phi), a = simVars.earth_radius;
		return a * u * (2.0 * simVars.coriolis_omega * std::sin(phi) + (std::tan(phi) / a) * u);
	}

	// h on rings of latitude lat[j] (north first, all inside (-pi/2, pi/2))
	std::vector<double> balanced_h(const std::vector<double> &lat) const
	{
		const int nlat = (int)lat.size();
		auto integral = [&](double s, double e) {
			return GaussQuadrature::integrate5_intervals_adaptive(
				s, e, [this](double phi) { return balance_integrand(phi); }, error_threshold);
		};
		std::vector<double> hg(nlat);
		double hg_sum = 0.0, area = 0.0, start = -M_PI * 0.5, prev = 0.0;
		for (int j = nlat - 1; j >= 0; j--) {
			prev -= integral(start, lat[j]);
			hg[j] = prev;
			const double w = (std::sin(lat[j]) - std::sin(start)) * 2.0 * M_PI;
			hg_sum += prev * w;
			area += w;
			start = lat[j];
		}
		const double last = prev - integral(start, M_PI * 0.5);
		const double w = (1.0 - std::sin(start)) * 2.0 * M_PI;
		hg_sum += last * w;
		area += w;
		const double shift = h_avg - (hg_sum / simVars.gravitation) / area;
		for (int j = 0; j < nlat; j++) hg[j] = hg[j] / simVars.gravitation + shift;
		return hg;
	}

	double bump(double lambda, double phi) const
	{
		const double x = (lambda - M_PI) / alpha, y = (phi2 - phi) / beta;
		return h_hat * std::cos(phi) * std::exp(-x * x) * std::exp(-y * y);
	}

#ifdef SPHSWE_HOST_SPHDATA_HPP
	void setup_initial_h(SPHData &o_h) const
	{
		const SPHConfig *c = o_h.sphConfig;
		const std::vector<double> h = balanced_h(c->lat);
		o_h.data_spat.resize(c->spat_num_elems);
		for (int i = 0; i < c->spat_num_lon; i++)
			for (int j = 0; j < c->spat_num_lat; j++) o_h.data_spat[(std::size_t)i * c->spat_num_lat + j] = h[j];
		o_h.spat_upload();
	}
	void setup_initial_h_add_bump(SPHData &o_h) const
	{
		o_h.spat_update_lambda([&](double lambda, double phi, double &io) { io += bump(lambda, phi); });
	}
	void setup_initial_u(SPHData &o_u) const
	{
		o_u.spat_update_lambda([&](double lon, double phi, double &o) { o = initial_condition_u(lon, phi); });
	}
	void setup_initial_v(SPHData &o_v) const { o_v.spat_set_zero(); }
#endif
};

#endif
