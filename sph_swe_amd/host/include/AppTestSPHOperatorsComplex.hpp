// Program P2 (src/AppTestSPHOperatorsComplex.hpp:26-606): the complex
// spectral and composite operators of SPHOperatorsComplex, run on the GPU,
// against the analytic derivatives of a Gaussian bump and of one spherical
// harmonic (real-valued complex fields), and the complex scalar arithmetic.
// As in the reference, each case prints its largest error; nothing is
// asserted here.
#ifndef SPHSWE_HOST_APPTESTSPHOPERATORSCOMPLEX_HPP
#define SPHSWE_HOST_APPTESTSPHOPERATORSCOMPLEX_HPP

#include <complex>
#include <functional>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "benchmarks/SphereTestSolutions_Gaussian.hpp"
#include "benchmarks/SphereTestSolutions_SPH.hpp"
#include "sph/SPHConfig.hpp"
#include "sph/SPHDataComplex.hpp"
#include "sph/SPHOperatorsComplex.hpp"

class AppTestSPHOperatorsComplex
{
	typedef std::complex<double> cplx;
	typedef SphereTestSolutions_Gaussian TS;
	typedef void (TS::*Fn)(double, double, double &) const;

	SPHOperatorsComplex opComplex;
	TS ts;

	SPHDataComplex gauss(SPHConfig *c, std::function<double(double, double)> g) const
	{
		SPHDataComplex d(c);
		d.spat_update_lambda_gaussian_grid([&](double lon, double mu, cplx &o) { o = g(lon, mu); });
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
	static void report(const std::string &name, const SPHDataComplex &a, const SPHDataComplex &b)
	{
		std::cout << name << " - max error: " << a.spat_reduce_error_max(b) << std::endl;
	}

public:
	void run(SPHConfig *c)
	{
		std::cout << std::setprecision(10);
		{
			SphereTestSolutions_SPH sph(2, 1);
			SPHDataComplex h(c), r(c);
			h.spat_update_lambda_gaussian_grid([&](double lon, double mu, cplx &o) {
				double v;
				sph.test_function__grid_gaussian(lon, mu, v);
				o = v;
			});
			r.spat_update_lambda_gaussian_grid([&](double lon, double mu, cplx &o) {
				double v;
				sph.correct_result_diff_mu__grid_gaussian(lon, mu, v);
				o = v;
			});
			report("TEST SPHbasis (1-mu*mu)*d/dmu", opComplex.spec_one_minus_mu_squared_diff_lat_mu(h), r);
		}
		// scalar arithmetic on fields of the latitude (both sides truncated alike)
		auto lat = [&](std::function<cplx(double, double)> g) {
			SPHDataComplex d(c);
			d.spat_update_lambda([&](double lon, double phi, cplx &o) { o = g(lon, phi); });
			return d;
		};
		{
			const cplx sc(12.0, 34.0);
			SPHDataComplex y = lat([](double, double p) { return cplx(p); });
			std::cout << "ERROR operator*" << sc << ": "
				  << (y * sc).spec_truncate().spat_reduce_error_max(lat([&](double, double p) { return p * sc; }).spec_truncate())
				  << std::endl;
			std::cout << "ERROR operator+(double): "
				  << (y + 100.0).spec_truncate().spat_reduce_error_max(
					     lat([](double, double p) { return cplx(p + 100.0); }).spec_truncate())
				  << std::endl;
			const cplx off(200.0, 100.0);
			SPHDataComplex xy = lat([](double x, double p) { return cplx(x * p, x + p); });
			std::cout << "ERROR operator+(cplx): "
				  << (xy + off).spec_truncate().spat_reduce_error_max(
					     lat([&](double x, double p) { return cplx(x * p + off.real(), x + p + off.imag()); }).spec_truncate())
				  << std::endl;
		}
		SPHDataComplex f = gauss(c, fn(&TS::test_function__grid_gaussian));
		{
			SPHDataComplex fphi(c);
			fphi.spat_update_lambda([&](double lon, double phi, cplx &o) {
				double v;
				ts.test_function_phi__grid_phi(lon, phi, v);
				o = v;
			});
			std::cout << "TEST PHI vs. MU: max error: " << f.spat_reduce_error_max(fphi) << std::endl;
			std::cout << "TEST IDENTITY: max error: "
				  << f.spat_reduce_error_max(gauss(c, fn(&TS::test_function__grid_gaussian))) << std::endl;
		}
		struct Case {
			const char *name;
			std::function<SPHDataComplex(const SPHDataComplex &)> apply;
			std::function<double(double, double)> exact;
		};
		const std::vector<Case> cases = {
			{"TEST DIFF LON", [&](const SPHDataComplex &x) { return opComplex.diff_lon(x); },
			 fn(&TS::correct_result_diff_lambda__grid_gaussian)},
			{"TEST DIFF PHI", [&](const SPHDataComplex &x) { return opComplex.diff_lat_phi(x); },
			 fn(&TS::correct_result_diff_phi__grid_gaussian)},
			{"TEST DIFF LAT MU", [&](const SPHDataComplex &x) { return opComplex.diff_lat_mu(x); },
			 fn(&TS::correct_result_diff_mu__grid_gaussian)},
			{"TEST GRAD LON", [&](const SPHDataComplex &x) { return opComplex.grad_lon(x); },
			 fn(&TS::correct_result_grad_lambda__grid_gaussian)},
			{"TEST mu*()", [&](const SPHDataComplex &x) { return opComplex.mu(x); }, fn(&TS::correct_result_mu__grid_gaussian)},
			{"TEST mu*mu*()", [&](const SPHDataComplex &x) { return opComplex.mu2(x); },
			 [&](double lon, double mu) { return mu * mu * fn(&TS::test_function__grid_gaussian)(lon, mu); }},
			{"TEST (1-mu*mu)*d/dmu", [&](const SPHDataComplex &x) { return opComplex.spec_one_minus_mu_squared_diff_lat_mu(x); },
			 fn(&TS::correct_result_one_minus_mu_squared_diff_lat_mu__grid_gaussian)},
			{"TEST GRAD LAT", [&](const SPHDataComplex &x) { return opComplex.grad_lat(x); },
			 fn(&TS::correct_result_grad_phi__grid_gaussian)},
			{"TEST DIV LON ", [&](const SPHDataComplex &x) { return opComplex.div_lon(x); },
			 fn(&TS::correct_result_div_lambda__grid_gaussian)},
			{"TEST DIV LAT ", [&](const SPHDataComplex &x) { return opComplex.div_lat(x); },
			 fn(&TS::correct_result_div_mu__grid_gaussian)},
		};
		for (const Case &k : cases) report(k.name, k.apply(f), gauss(c, k.exact));
		report("TEST LAPLACE ",
		       opComplex.div_lon(opComplex.grad_lon(f)) + opComplex.div_lat(opComplex.grad_lat(f)), opComplex.laplace(f));
	}
};

#endif
