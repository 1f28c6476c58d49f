// Program P1 (src/AppTestSPHOperators.hpp:25-520): the real spectral and
// composite operators of SPHOperators, run on the GPU, against the analytic
// derivatives of a Gaussian bump and of one spherical harmonic.  As in the
// reference, each case prints its largest grid-point error; nothing is
// asserted here (the program's test harness compares the printed errors with
// bounds).
#ifndef SPHSWE_HOST_APPTESTSPHOPERATORS_HPP
#define SPHSWE_HOST_APPTESTSPHOPERATORS_HPP

#include <functional>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "benchmarks/SphereTestSolutions_Gaussian.hpp"
#include "benchmarks/SphereTestSolutions_SPH.hpp"
#include "sph/SPHConfig.hpp"
#include "sph/SPHData.hpp"
#include "sph/SPHOperators.hpp"

class AppTestSPHOperators
{
	typedef SphereTestSolutions_Gaussian TS;
	typedef void (TS::*Fn)(double, double, double &) const;

	SPHOperators op;
	SphereTestSolutions_Gaussian ts;

	SPHData gauss(SPHConfig *c, std::function<double(double, double)> g) const
	{
		SPHData d(c);
		d.spat_update_lambda_gaussian_grid([&](double lon, double mu, double &o) { o = g(lon, mu); });
		return d;
	}
	std::function<double(double, double)> fn(Fn f) const
	{
		return [this, f](double lon, double mu) {
			double v;
			(ts.*f)(lon, mu, v);
			return v;
		};
	}
	static void report(const std::string &name, const SPHData &a, const SPHData &b)
	{
		std::cout << name << " - max error: " << a.spat_reduce_error_max(b) << std::endl;
	}

public:
	void run(SPHConfig *c)
	{
		std::cout << std::setprecision(10);
		// one spherical harmonic, n = 2, m = 1: (1 - mu^2) d/dmu
		{
			SphereTestSolutions_SPH sph(2, 1);
			SPHData h(c), r(c);
			h.spat_update_lambda_gaussian_grid([&](double lon, double mu, double &o) { sph.test_function__grid_gaussian(lon, mu, o); });
			r.spat_update_lambda_gaussian_grid(
				[&](double lon, double mu, double &o) { sph.correct_result_diff_mu__grid_gaussian(lon, mu, o); });
			report("TEST SPHbasis (1-mu*mu)*d/dmu", op.spec_one_minus_mu_squared_diff_lat_mu(h), r);
		}
		// scalar arithmetic on the latitude field (both sides truncated alike)
		{
			auto lat = [&](std::function<double(double)> g) {
				SPHData d(c);
				d.spat_update_lambda([&](double, double phi, double &o) { o = g(phi); });
				return d;
			};
			SPHData y = lat([](double p) { return p; });
			std::cout << "ERROR operator*123.0: "
				  << (y * 123.0).spec_truncate().spat_reduce_error_max(lat([](double p) { return 123.0 * p; }).spec_truncate())
				  << std::endl;
			std::cout << "ERROR operator+(double): "
				  << (y + 100.0).spec_truncate().spat_reduce_error_max(lat([](double p) { return p + 100.0; }).spec_truncate())
				  << std::endl;
		}
		SPHData f = gauss(c, fn(&TS::test_function__grid_gaussian));
		{
			SPHData fphi(c);
			fphi.spat_update_lambda([&](double lon, double phi, double &o) { ts.test_function_phi__grid_phi(lon, phi, o); });
			std::cout << "TEST PHI vs. MU: max error: " << f.spat_reduce_error_max(fphi) << std::endl;
			std::cout << "TEST IDENTITY: max error: "
				  << f.spat_reduce_error_max(gauss(c, fn(&TS::test_function__grid_gaussian))) << std::endl;
		}
		// single operators against the analytic results
		struct Case {
			const char *name;
			std::function<SPHData(const SPHData &)> apply;
			std::function<double(double, double)> exact;
		};
		const std::vector<Case> cases = {
			{"TEST DIFF LON", [&](const SPHData &x) { return op.diff_lon(x); }, fn(&TS::correct_result_diff_lambda__grid_gaussian)},
			{"TEST DIFF PHI", [&](const SPHData &x) { return op.diff_lat_phi(x); }, fn(&TS::correct_result_diff_phi__grid_gaussian)},
			{"TEST DIFF LAT MU", [&](const SPHData &x) { return op.diff_lat_mu(x); }, fn(&TS::correct_result_diff_mu__grid_gaussian)},
			{"TEST GRAD LON", [&](const SPHData &x) { return op.grad_lon(x); }, fn(&TS::correct_result_grad_lambda__grid_gaussian)},
			{"TEST mu*()", [&](const SPHData &x) { return op.mu(x); }, fn(&TS::correct_result_mu__grid_gaussian)},
			{"TEST mu*mu*()", [&](const SPHData &x) { return op.mu2(x); },
			 [&](double lon, double mu) { return mu * mu * fn(&TS::test_function__grid_gaussian)(lon, mu); }},
			{"TEST (1-mu*mu)*d/dmu", [&](const SPHData &x) { return op.spec_one_minus_mu_squared_diff_lat_mu(x); },
			 fn(&TS::correct_result_one_minus_mu_squared_diff_lat_mu__grid_gaussian)},
			{"TEST GRAD LAT", [&](const SPHData &x) { return op.grad_lat(x); }, fn(&TS::correct_result_grad_phi__grid_gaussian)},
			{"TEST DIV LON ", [&](const SPHData &x) { return op.div_lon(x); }, fn(&TS::correct_result_div_lambda__grid_gaussian)},
			{"TEST DIV LAT ", [&](const SPHData &x) { return op.div_lat(x); }, fn(&TS::correct_result_div_mu__grid_gaussian)},
			{"TEST DIVERGENCE", [&](const SPHData &x) { return op.div(x, x); },
			 [&](double lon, double mu) {
				 return fn(&TS::correct_result_div_mu__grid_gaussian)(lon, mu) +
					fn(&TS::correct_result_div_lambda__grid_gaussian)(lon, mu);
			 }},
			{"TEST VORTICITY", [&](const SPHData &x) { return op.vort(x, x); },
			 [&](double lon, double mu) {
				 return fn(&TS::correct_result_grad_lambda__grid_gaussian)(lon, mu) -
					fn(&TS::correct_result_grad_phi__grid_gaussian)(lon, mu);
			 }},
		};
		for (const Case &k : cases) report(k.name, k.apply(f), gauss(c, k.exact));
		// div(grad f) against the spectral Laplacian
		report("TEST LAPLACE (div.grad vs. sph laplace)", op.div_lon(op.grad_lon(f)) + op.div_lat(op.grad_lat(f)),
		       op.laplace(f));
	}
};

#endif
