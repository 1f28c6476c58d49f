// Program P4: known-answer tests of the complex banded solver
// (src/AppTestSPHSolverComplex.hpp:29-517).  For every pole alpha of
// REXI::setup(0.2, 64) (76 poles), each solver component is set up with the
// scalars the REXI operator uses, the right-hand side is the component applied
// analytically to a Gaussian bump f, and the solution must reproduce f to
// 1e-10 in the largest real / imaginary difference; the program exits with
// status 1 at the first excess, as the reference does.  Also checked: the
// direct Helmholtz solve spec_solve_helmholtz(a, b, r).
#ifndef SPHSWE_HOST_APPTESTSPHSOLVERCOMPLEX_HPP
#define SPHSWE_HOST_APPTESTSPHSOLVERCOMPLEX_HPP

#include <cmath>
#include <complex>
#include <cstdlib>
#include <functional>
#include <iomanip>
#include <iostream>
#include <string>

#include "SimVars.hpp"
#include "benchmarks/SphereTestSolutions_Gaussian.hpp"
#include "rexi/REXI.hpp"
#include "sph/SPHConfig.hpp"
#include "sph/SPHDataComplex.hpp"
#include "sph/SPHOperatorsComplex.hpp"
#include "sph/SPHSolverComplex.hpp"

class AppTestSPHSolversComplex
{
	typedef std::complex<double> cplx;
	typedef SPHSolverComplex<cplx> Solver;

public:
	SimVars simVars;
	SPHOperatorsComplex opComplex;
	SPHConfig *sphConfig = nullptr;
	double threshold = 1e-10;
	double worst = 0.0;   // largest error over every case (reported at the end)

	void test_threshold(double i_value)
	{
		worst = std::max(worst, std::abs(i_value));
		if (std::abs(i_value) > threshold) {
			std::cout << "!!!!!!!!!!!!!!!! ERROR !!!!!!!!!!! threshold exceeded !!!!!!!!!!!!!!!!" << std::endl;
			exit(1);
		}
	}

	void run(SPHConfig *i_sphConfig)
	{
		std::cout << std::setprecision(10);
		sphConfig = i_sphConfig;
		if (sphConfig->spec_n_max < 32) std::cerr << "WARNING: AT LEAST 32 MODES REQUIRED for proper accuracy!!!" << std::endl;

		SphereTestSolutions_Gaussian ts;
		// the expected solution of every case: f itself
		SPHDataComplex x_result(sphConfig);
		x_result.spat_update_lambda_gaussian_grid([&](double lon, double mu, cplx &o) {
			double v;
			ts.test_function__grid_gaussian(lon, mu, v);
			o = v;
		});
		const double r = simVars.earth_radius;
		const double two_omega = 2.0 * simVars.coriolis_omega;

		// right-hand side g(lon, mu, f) on the Gaussian grid
		auto rhs = [&](std::function<cplx(double, double, double)> g) {
			SPHDataComplex b(sphConfig);
			b.spat_update_lambda_gaussian_grid([&](double lon, double mu, cplx &o) {
				double f;
				ts.test_function__grid_gaussian(lon, mu, f);
				o = g(lon, mu, f);
			});
			return b;
		};
		auto check = [&](const std::string &name, const SPHDataComplex &x) {
			std::cout << name;
			const double e = x.spat_reduce_error_max(x_result);
			std::cout << " ||| Error: " << e << std::endl;
			test_threshold(e);
		};
		auto derivative = [&](void (SphereTestSolutions_Gaussian::*fn)(double, double, double &) const, double lon,
				      double mu) {
			double v;
			(ts.*fn)(lon, mu, v);
			return v;
		};

		REXI rexi;
		rexi.setup(0.2, 64);
		for (std::size_t i = 0; i < rexi.alpha.size(); i++) {
			const cplx alpha = rexi.alpha[i];
			std::cout << std::endl << "***************************************************" << std::endl;
			std::cout << "Testing for alpha " << alpha << std::endl;
			std::cout << "***************************************************" << std::endl;

			{
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_scalar_phi(alpha);
				check("Test Zx = c*Phi(mu)", s.solve(rhs([&](double, double, double f) { return f * alpha; })));
			}
			{
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_scalar_phi(alpha);
				s.solver_component_mu_phi();
				check("Test Zx = mu*Phi(lam,mu) + a*Phi(lam,mu)",
				      s.solve(rhs([&](double, double mu, double f) { return mu * f + alpha * f; })));
			}
			{
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_scalar_phi(alpha);
				s.solver_component_one_minus_mu_mu_diff_mu_phi();
				check("Test Zx = (1-mu*mu)*d/dmu Phi(lam,mu) + a*Phi(lam,mu)",
				      s.solve(rhs([&](double lon, double mu, double f) {
					      return derivative(&SphereTestSolutions_Gaussian::
									correct_result_one_minus_mu_squared_diff_lat_mu__grid_gaussian,
								lon, mu) + alpha * f;
				      })));
			}
			std::cout << "*** TESTING REXI COMPONENTS ********************************" << std::endl;
			{
				const cplx sc = (alpha * alpha) * (alpha * alpha);
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_rexi_z1(sc, r);
				check("Test Z1 = alpha^4*Phi(mu)", s.solve(rhs([&](double, double, double f) { return f * sc; })));
			}
			{
				const cplx sc = alpha * alpha * two_omega * two_omega;
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_rexi_z2(sc, r);
				check("Test Z2 = mu^2*Phi(lam,mu)",
				      s.solve(rhs([&](double, double mu, double f) { return sc * (mu * mu) * f; })));
			}
			{
				const cplx sc = two_omega * two_omega * two_omega * two_omega;
				Solver s;
				s.setup(sphConfig, 4);
				s.solver_component_rexi_z3(sc, r);
				check("Test Z3 = mu^4*Phi(lam,mu)",
				      s.solve(rhs([&](double, double mu, double f) { return sc * (mu * mu) * (mu * mu) * f; })));
			}
			// z4 .. z8 with the scalar term alpha*Phi added (the operators alone are singular)
			{
				const cplx sc = -alpha * alpha * two_omega;
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_rexi_z4(sc, r);
				s.solver_component_scalar_phi(alpha);
				check("Test Z4 = grad_j(mu) grad_i(Phi(lam,mu)) = d/dlambda Phi(lam,mu)",
				      s.solve(rhs([&](double lon, double mu, double f) {
					      return sc / (r * r) *
							     derivative(&SphereTestSolutions_Gaussian::correct_result_diff_lambda__grid_gaussian,
									lon, mu) + f * alpha;
				      })));
			}
			{
				const cplx sc = two_omega * two_omega * two_omega;
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_rexi_z5(sc, r);
				s.solver_component_scalar_phi(alpha);
				check("Test Z5 = grad_j(mu) mu^2 grad_i(Phi(lam,mu))",
				      s.solve(rhs([&](double lon, double mu, double f) {
					      return sc / (r * r) * std::sqrt(1.0 - mu * mu) * mu * mu *
							     derivative(&SphereTestSolutions_Gaussian::correct_result_grad_lambda__grid_gaussian,
									lon, mu) + f * alpha;
				      })));
			}
			{
				const cplx sc = 2.0 * alpha * two_omega * two_omega;
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_rexi_z6(sc, r);
				s.solver_component_scalar_phi(alpha);
				check("Test Z6 = grad_j(mu) mu grad_j(Phi(lam,mu))",
				      s.solve(rhs([&](double lon, double mu, double f) {
					      return sc / (r * r) * std::sqrt(1.0 - mu * mu) * mu *
							     derivative(&SphereTestSolutions_Gaussian::correct_result_grad_phi__grid_gaussian,
									lon, mu) + f * alpha;
				      })));
			}
			{
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_rexi_z7(1.0, r);
				s.solver_component_scalar_phi(alpha);
				SPHDataComplex b = rhs([&](double, double, double f) { return cplx(f); });
				b = opComplex.laplace(b) / (r * r) + b * alpha;
				check("Test Z7 = laplace(Phi(lam,mu))", s.solve(b));
			}
			{
				Solver s;
				s.setup(sphConfig, 2);
				s.solver_component_rexi_z8(1.0, r);
				s.solver_component_scalar_phi(alpha);
				SPHDataComplex b = rhs([&](double, double, double f) { return cplx(f); });
				b = opComplex.mu2(opComplex.laplace(b)) / (r * r) + b * alpha;
				check("Test Z8 = mu*mu*laplace(Phi(lam,mu))", s.solve(b));
			}
			{
				// (a + b laplace / r^2) x = rhs, solved directly per degree
				const cplx a = alpha, bb = alpha * 2.0 - 1.0;
				SPHDataComplex t = rhs([&](double, double, double f) { return cplx(f); });
				t = a * t + (bb / (r * r)) * opComplex.laplace(t);
				check("Test Zx = a + b*laplace", t.spec_solve_helmholtz(a, bb, r));
			}
		}
		std::cout << std::endl << "P4: all " << rexi.alpha.size() << " poles within " << threshold
			  << ", largest error " << worst << std::endl;
	}
};

#endif
