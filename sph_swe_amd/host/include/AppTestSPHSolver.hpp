// Program P3 (src/AppTestSPHSolver.hpp:34-482): the banded solver on real
// fields.  With alpha = 3 and the scalars of the REXI operator, each component
// is set up, the right-hand side is the component applied analytically to the
// Gaussian bump f (real parts of the scalars), and the largest grid-point
// error of the solution against f is printed.  As in the reference nothing is
// asserted here; the reference's "DOUBLE MODES" case is compiled out there and
// is left out here.
#ifndef SPHSWE_HOST_APPTESTSPHSOLVER_HPP
#define SPHSWE_HOST_APPTESTSPHSOLVER_HPP

#include <cmath>
#include <complex>
#include <functional>
#include <iomanip>
#include <iostream>
#include <string>

#include "SimVars.hpp"
#include "benchmarks/SphereTestSolutions_Gaussian.hpp"
#include "sph/SPHConfig.hpp"
#include "sph/SPHData.hpp"
#include "sph/SPHOperators.hpp"
#include "sph/SPHSolver.hpp"

class AppTestSPHSolvers
{
	typedef std::complex<double> cplx;
	typedef SphereTestSolutions_Gaussian TS;
	typedef void (TS::*Fn)(double, double, double &) const;

public:
	SimVars simVars;
	SPHOperators op;
	SPHConfig *sphConfig = nullptr;

	void run(SPHConfig *i_sphConfig)
	{
		std::cout << std::setprecision(10);
		sphConfig = i_sphConfig;
		if (sphConfig->spec_n_max < 32) std::cerr << "WARNING: AT LEAST 32 MODES REQUIRED for proper accuracy!!!" << std::endl;
		TS ts;
		SPHData x_result(sphConfig);
		x_result.spat_update_lambda_gaussian_grid([&](double lon, double mu, double &o) { ts.test_function__grid_gaussian(lon, mu, o); });
		const cplx alpha(3.0, 0.0);
		const double r = simVars.earth_radius, two_omega = 2.0 * simVars.coriolis_omega;
		const double a = alpha.real();
		auto val = [&](Fn f, double lon, double mu) {
			double v;
			(ts.*f)(lon, mu, v);
			return v;
		};
		auto rhs = [&](std::function<double(double, double, double)> g) {
			SPHData b(sphConfig);
			b.spat_update_lambda_gaussian_grid([&](double lon, double mu, double &o) {
				o = g(lon, mu, val(&TS::test_function__grid_gaussian, lon, mu));
			});
			return b;
		};
		auto check = [&](const std::string &name, int halo, std::function<void(SPHSolver<cplx> &)> comps, const SPHData &b) {
			std::cout << name;
			SPHSolver<cplx> s;
			s.setup(sphConfig, halo);
			comps(s);
			std::cout << " ||| Error: " << s.solve(b).spat_reduce_error_max(x_result) << std::endl;
		};

		check("Test Zx = c*Phi(mu)", 2, [&](SPHSolver<cplx> &s) { s.solver_component_scalar_phi(alpha); },
		      rhs([&](double, double, double f) { return f * a; }));
		check("Test Zx = mu*Phi(lam,mu) + a*Phi(lam,mu)", 2,
		      [&](SPHSolver<cplx> &s) {
			      s.solver_component_scalar_phi(alpha);
			      s.solver_component_mu_phi();
		      },
		      rhs([&](double, double mu, double f) { return mu * f + a * f; }));
		check("Test Zx = (1-mu*mu)*d/dmu Phi(lam,mu) + a*Phi(lam,mu)", 2,
		      [&](SPHSolver<cplx> &s) {
			      s.solver_component_scalar_phi(alpha);
			      s.solver_component_one_minus_mu_mu_diff_mu_phi();
		      },
		      rhs([&](double lon, double mu, double f) {
			      return val(&TS::correct_result_one_minus_mu_squared_diff_lat_mu__grid_gaussian, lon, mu) + a * f;
		      }));
		{
			const cplx sc = (alpha * alpha) * (alpha * alpha);
			check("Test Z1 = alpha^4*Phi(mu)", 2, [&](SPHSolver<cplx> &s) { s.solver_component_rexi_z1(sc, r); },
			      rhs([&](double, double, double f) { return f * sc.real(); }));
		}
		{
			const cplx sc = alpha * alpha * two_omega * two_omega;
			check("Test Z2 = mu^2*Phi(lam,mu)", 2, [&](SPHSolver<cplx> &s) { s.solver_component_rexi_z2(sc, r); },
			      rhs([&](double, double mu, double f) { return sc.real() * (mu * mu) * f; }));
		}
		{
			const cplx sc = two_omega * two_omega * two_omega * two_omega;
			check("Test Z3 = mu^4*Phi(lam,mu)", 4, [&](SPHSolver<cplx> &s) { s.solver_component_rexi_z3(sc, r); },
			      rhs([&](double, double mu, double f) { return sc.real() * (mu * mu) * (mu * mu) * f; }));
		}
		{
			const cplx sc = -alpha * alpha * two_omega;
			check("Test Z4 = grad_j(mu) grad_i(Phi(lam,mu)) = d/dlambda Phi(lam,mu)", 2,
			      [&](SPHSolver<cplx> &s) {
				      s.solver_component_rexi_z4(sc, r);
				      s.solver_component_scalar_phi(alpha);
			      },
			      rhs([&](double lon, double mu, double f) {
				      return sc.real() / (r * r) * val(&TS::correct_result_diff_lambda__grid_gaussian, lon, mu) + f * a;
			      }));
		}
		{
			const cplx sc = two_omega * two_omega * two_omega;
			check("Test Z5 = grad_j(mu) mu^2 grad_i(Phi(lam,mu))", 2,
			      [&](SPHSolver<cplx> &s) {
				      s.solver_component_rexi_z5(sc, r);
				      s.solver_component_scalar_phi(alpha);
			      },
			      rhs([&](double lon, double mu, double f) {
				      return sc.real() / (r * r) * std::sqrt(1.0 - mu * mu) * mu * mu *
						     val(&TS::correct_result_grad_lambda__grid_gaussian, lon, mu) + f * a;
			      }));
		}
		{
			const cplx sc = 2.0 * alpha * two_omega * two_omega;
			check("Test Z6 = grad_j(mu) mu grad_j(Phi(lam,mu))", 2,
			      [&](SPHSolver<cplx> &s) {
				      s.solver_component_rexi_z6(sc, r);
				      s.solver_component_scalar_phi(alpha);
			      },
			      rhs([&](double lon, double mu, double f) {
				      return sc.real() / (r * r) * std::sqrt(1.0 - mu * mu) * mu *
						     val(&TS::correct_result_grad_phi__grid_gaussian, lon, mu) + f * a;
			      }));
		}
		{
			SPHData f = rhs([&](double, double, double v) { return v; });
			check("Test Z7 = laplace(Phi(lam,mu))", 2,
			      [&](SPHSolver<cplx> &s) {
				      s.solver_component_rexi_z7(1.0, r);
				      s.solver_component_scalar_phi(alpha);
			      },
			      op.laplace(f) / (r * r) + f * a);
			check("Test Z8 = mu*mu*laplace(Phi(lam,mu))", 2,
			      [&](SPHSolver<cplx> &s) {
				      s.solver_component_rexi_z8(1.0, r);
				      s.solver_component_scalar_phi(alpha);
			      },
			      op.mu2(op.laplace(f)) / (r * r) + f * a);
		}
	}
};

#endif
