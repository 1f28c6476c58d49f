// The shallow-water application (src/AppSWE.hpp): initial conditions, the RK4 and
// REXI time loops with output cadence and NaN checks, and the tendency
// p_run_euler_timestep_update written with the class API as in the reference
// (:574-659).  By default every time step is one library call; with
// simVars.gpu_fused_step = 0 the loops run class by class as the reference does.
#ifndef SPHSWE_HOST_APPSWE_HPP
#define SPHSWE_HOST_APPSWE_HPP

#include <cassert>
#include <cstdio>
#include <vector>

#include "SWERexiSPH.hpp"
#include "SimVars.hpp"
#include "rexi/REXI.hpp"
#include "sph/SPHDataComplex.hpp"
#include "sph/SPHOperators.hpp"
#include "sph/SPHOperatorsComplex.hpp"
#include "sweet/TimesteppingRK.hpp"
#include "benchmarks/BenchmarkGalewsky.hpp"

class AppSWE
{
public:
	SimVars &simVars;
	SPHOperators op;
	SPHOperatorsComplex opComplex;
	SPHConfig *sphConfig;
	TimesteppingRK timestepping;
	SPHData prog_h, prog_u, prog_v, fdata;
	REXI rexi;

	AppSWE(SPHConfig *i_sphConfig, SimVars &i_simVars)
	    : simVars(i_simVars), sphConfig(i_sphConfig), prog_h(i_sphConfig), prog_u(i_sphConfig), prog_v(i_sphConfig),
	      fdata(i_sphConfig)
	{
	}

	void write_output()
	{
		char buffer[1024];
		const double th = simVars.timecontrol.current_simulation_time / (60 * 60);
		const bool shift = simVars.benchmark_scenario_id == 0;
		std::snprintf(buffer, sizeof(buffer), "prog_h_t%020.8f.csv", th);
		shift ? prog_h.spat_write_file_lon_pi_shifted(buffer) : prog_h.spat_write_file(buffer);
		std::cout << buffer << " (min: " << prog_h.spat_reduce_min() << ", max: " << prog_h.spat_reduce_max() << ")"
			  << std::endl;
		std::snprintf(buffer, sizeof(buffer), "prog_u_t%020.8f.csv", th);
		shift ? prog_u.spat_write_file_lon_pi_shifted(buffer) : prog_u.spat_write_file(buffer);
		std::cout << buffer << std::endl;
		std::snprintf(buffer, sizeof(buffer), "prog_v_t%020.8f.csv", th);
		shift ? prog_v.spat_write_file_lon_pi_shifted(buffer) : prog_v.spat_write_file(buffer);
		std::cout << buffer << std::endl;
		std::snprintf(buffer, sizeof(buffer), "prog_eta_t%020.8f.csv", th);
		SPHData vort = op.vort(prog_u, prog_v) / simVars.earth_radius;
		shift ? vort.spat_write_file_lon_pi_shifted(buffer, "vorticity, lon pi shifted") : vort.spat_write_file(buffer);
		std::cout << buffer << std::endl;
	}

	// h = h0 (1 + 0.1 exp(-10 d^2)), d the great-circle angle to the centre (AppSWE.hpp:99-128)
	void setup_initial_conditions_gaussian(double i_center_lat = M_PI / 3, double i_center_lon = M_PI / 3)
	{
		const double h0 = simVars.h0;
		prog_h.spat_update_lambda_gaussian_grid([&](double lon, double mu, double &o) {
			const double phi1 = std::asin(mu);
			const double d = std::acos(std::sin(phi1) * std::sin(i_center_lat) +
						   std::cos(phi1) * std::cos(i_center_lat) * std::cos(lon - i_center_lon));
			o = std::exp(-d * d * 10.0) * 0.1 * h0 + h0;
		});
		prog_u.spat_set_zero();
		prog_v.spat_set_zero();
	}

	// scenario 4 (an addition): h = h0 + amp exp(-(d / width)^2) at (pi/4, pi), at rest
	void setup_initial_conditions_dam(double amp = 500.0, double width = 0.1, double i_center_lat = M_PI / 4,
					  double i_center_lon = M_PI)
	{
		const double h0 = simVars.h0;
		prog_h.spat_update_lambda_gaussian_grid([&](double lon, double mu, double &o) {
			const double phi1 = std::asin(mu);
			double c = std::sin(phi1) * std::sin(i_center_lat) +
				   std::cos(phi1) * std::cos(i_center_lat) * std::cos(lon - i_center_lon);
			c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
			const double d = std::acos(c);
			o = h0 + amp * std::exp(-(d / width) * (d / width));
		});
		prog_u.spat_set_zero();
		prog_v.spat_set_zero();
	}

	SPHData f(SPHData i_sphData) { return fdata * i_sphData * 2.0 * simVars.coriolis_omega; }

	void run()
	{
		if (simVars.timecontrol.max_simulation_time == -1) simVars.timecontrol.max_simulation_time = 200 * 60 * 60;
		simVars.next_output_dt = 0;
		if (simVars.timecontrol.current_timestep_size <= 0) {
			if (sphConfig->spat_num_lat < 256)
				simVars.timecontrol.current_timestep_size = 0.002 * simVars.earth_radius / (double)sphConfig->spat_num_lat;
			else
				simVars.timecontrol.current_timestep_size = 0.001 * simVars.earth_radius / (double)sphConfig->spat_num_lat;
		}
		if (simVars.output_dt <= 0) simVars.output_dt = 60 * 30;
		if (simVars.use_rexi) {
			if (simVars.timecontrol.current_timestep_size <= 0) {
				std::cout << "Timestep size not positive" << std::endl;
				exit(1);
			}
			simVars.use_nonlinear_equations = 0;
		}

		switch (simVars.benchmark_scenario_id) {
		case 0: setup_initial_conditions_gaussian(); break;
		case 1: {
			// the Galewsky jet and bump (AppSWE.hpp:199-207)
			BenchmarkGalewsky galewsky(simVars);
			galewsky.setup_initial_h(prog_h);
			galewsky.setup_initial_h_add_bump(prog_h);
			galewsky.setup_initial_u(prog_u);
			galewsky.setup_initial_v(prog_v);
			break;
		}
		case 2: case 3:
			simVars.h0 = 1;
			simVars.gravitation = 1;
			simVars.earth_radius = 1;
			simVars.use_nonlinear_equations = 0;
			setup_initial_conditions_gaussian(simVars.benchmark_scenario_id == 2 ? 0.0 : M_PI / 3.0);
			break;
		case 4: setup_initial_conditions_dam(); break;
		default:
			std::cerr << "Benchmark scenario " << simVars.benchmark_scenario_id << " is not available in this build"
				  << std::endl;
			exit(1);
		}
		const bool with_coriolis = simVars.coriolis_omega != 0;

		std::cout << "Using time step size dt = " << simVars.timecontrol.current_timestep_size << std::endl;
		std::cout << "Running simulation until t_end = " << simVars.timecontrol.max_simulation_time << std::endl;
		std::cout << "Parameters:" << std::endl;
		std::cout << " + gravity: " << simVars.gravitation << std::endl;
		std::cout << " + earth_radius: " << simVars.earth_radius << std::endl;
		std::cout << " + average height: " << simVars.h0 << std::endl;
		std::cout << " + coriolis_omega: " << simVars.coriolis_omega << std::endl;
		std::cout << " + viscosity D2: " << simVars.viscosity2 << std::endl;
		std::cout << " + use_nonlinear: " << simVars.use_nonlinear_equations << std::endl;
		std::cout << " + Coriolis: " << with_coriolis << std::endl;
		std::cout << " + Benchmark scenario id: " << simVars.benchmark_scenario_id << std::endl;
		std::cout << " + Use REXI: " << simVars.use_rexi << std::endl;
		std::cout << " + REXI h: " << simVars.rexi_h << std::endl;
		std::cout << " + REXI M: " << simVars.rexi_M << std::endl;
		std::cout << " + REXI use half poles: " << simVars.rexi_use_half_poles << std::endl;
		std::cout << " + timestep size: " << simVars.timecontrol.current_timestep_size << std::endl;
		std::cout << " + output timestep size: " << simVars.output_dt << std::endl;
		std::cout << " + one library call per step: " << simVars.gpu_fused_step << std::endl << std::endl;

		fdata.spat_update_lambda_gaussian_grid([](double, double mu, double &o) { o = mu; });

		if (!simVars.use_rexi) run_rk();
		else run_rexi(with_coriolis);
		dump_state();
	}

	// the tendency of AppSWE.hpp:574-659, written with the class API
	void p_run_euler_timestep_update(const SPHData &i_h, const SPHData &i_u, const SPHData &i_v, SPHData &o_h_t,
					 SPHData &o_u_t, SPHData &o_v_t, double &o_dt, double i_fixed_dt = 0,
					 double i_simulation_timestamp = -1)
	{
		(void)i_fixed_dt;
		(void)i_simulation_timestamp;
		o_dt = simVars.timecontrol.current_timestep_size;
		const double a = simVars.earth_radius, g = simVars.gravitation;
		if (!simVars.use_nonlinear_equations && simVars.use_robert_functions) {
			// linear equations, velocities as Robert functions (AppSWE.hpp:606-620)
			o_h_t = -(op.robert_div_lon(i_u) + op.robert_div_lat(i_v)) * (simVars.h0 / a);
			o_u_t = -op.robert_grad_lon(i_h) * (g / a);
			o_v_t = -op.robert_grad_lat(i_h) * (g / a);
			if (simVars.coriolis_omega != 0) {
				o_u_t += f(i_v);
				o_v_t -= f(i_u);
			}
		} else if (!simVars.use_nonlinear_equations) {
			o_h_t = -(op.div_lon(i_u) + op.div_lat(i_v)) * (simVars.h0 / a);
			o_u_t = -op.grad_lon(i_h) * (g / a);
			o_v_t = -op.grad_lat(i_h) * (g / a);
			if (simVars.coriolis_omega != 0) {
				o_u_t += f(i_v);
				o_v_t -= f(i_u);
			}
		} else {
			o_h_t = -(op.div_lon(i_h * i_u) + op.div_lat(i_h * i_v)) * (1.0 / a);
			o_u_t = -op.grad_lon(i_h) * (g / a);
			o_v_t = -op.grad_lat(i_h) * (g / a);
			if (simVars.coriolis_omega != 0) {
				o_u_t += f(i_v);
				o_v_t -= f(i_u);
			}
			o_u_t -= (i_u * op.grad_lon(i_u) + i_v * op.grad_lat(i_u)) * (1.0 / a);
			o_v_t -= (i_u * op.grad_lon(i_v) + i_v * op.grad_lat(i_v)) * (1.0 / a);
		}
		if (simVars.viscosity2 != 0) {
			const double scalar = simVars.viscosity2 / (a * a);
			o_h_t += op.laplace(i_h) * scalar;
			o_u_t += op.laplace(i_u) * scalar;
			o_v_t += op.laplace(i_v) * scalar;
		}
	}

private:
	std::vector<double> state_dump;

	// (h, u, v) spectra <-> one device buffer of 3 nlm complex
	void pack(double *d_state)
	{
		const std::size_t n = sizeof(std::complex<double>) * sphConfig->spec_num_elems;
		SPHData *p[3] = {&prog_h, &prog_u, &prog_v};
		for (int k = 0; k < 3; k++) {
			p[k]->request_data_spectral();
			sphConfig->check(sphswe_memcpy_d2d(sphConfig->gpu, (char *)d_state + k * n, p[k]->d_spec, n), "pack");
		}
	}
	void unpack(const double *d_state)
	{
		const std::size_t n = sizeof(std::complex<double>) * sphConfig->spec_num_elems;
		SPHData *p[3] = {&prog_h, &prog_u, &prog_v};
		for (int k = 0; k < 3; k++) {
			sphConfig->check(sphswe_memcpy_d2d(sphConfig->gpu, p[k]->d_spec, (const char *)d_state + k * n, n), "unpack");
			p[k]->data_spec_valid = true;
			p[k]->data_spat_valid = false;
		}
	}

	void output_if_due()
	{
		if (simVars.timecontrol.current_simulation_time >= simVars.next_output_dt) {
			std::cout << std::endl;
			write_output();
			simVars.next_output_dt += simVars.output_dt;
			if (simVars.next_output_dt < simVars.timecontrol.current_simulation_time)
				simVars.next_output_dt = simVars.timecontrol.current_simulation_time;
		}
	}

	void check_nan()
	{
		if (prog_h.isAnyNaNorInf()) {
			std::cerr << "Instability detected (NaN value in H)" << std::endl;
			exit(1);
		}
	}

	void run_rk()
	{
		double *d_state = nullptr;
		sphConfig->check(sphswe_malloc(sphConfig->gpu, 3 * sizeof(std::complex<double>) * sphConfig->spec_num_elems,
					       (void **)&d_state), "state");
		sphswe_swe_params p = {simVars.earth_radius, simVars.gravitation, simVars.coriolis_omega, simVars.h0,
				       simVars.viscosity2, simVars.timecontrol.current_timestep_size,
				       simVars.use_nonlinear_equations ? 1 : 0, SPHSWE_PREC_FP64,
				       simVars.use_robert_functions ? 1 : 0};
		simVars.timecontrol.current_simulation_time = 0;
		while (simVars.timecontrol.current_simulation_time < simVars.timecontrol.max_simulation_time) {
			output_if_due();
			if (simVars.gpu_fused_step) {
				if (simVars.timecontrol.timestepping_runge_kutta_order != 4) {
					std::cerr << "the one-call step is RK4" << std::endl;
					exit(1);
				}
				pack(d_state);
				sphConfig->check(sphswe_swe_rk4(sphConfig->gpu, &p, d_state, 1, 1), "sphswe_swe_rk4");
				unpack(d_state);
			} else {
				double o_dt;
				timestepping.run_rk_timestep(this, &AppSWE::p_run_euler_timestep_update, prog_h, prog_u, prog_v, o_dt,
							     simVars.timecontrol.current_timestep_size,
							     simVars.timecontrol.timestepping_runge_kutta_order,
							     simVars.timecontrol.current_simulation_time,
							     simVars.timecontrol.max_simulation_time);
			}
			std::cout << "." << std::flush;
			check_nan();
			simVars.timecontrol.current_simulation_time += simVars.timecontrol.current_timestep_size;
		}
		std::cout << std::endl;
		write_output();
		std::cout << std::endl;
		sphswe_free(sphConfig->gpu, d_state);
	}

	void run_rexi(bool with_coriolis)
	{
		// X > 0: the REXI terms work in a configuration with X more modes
		// (AppSWE.hpp:338-351); the state is copied into it and back every step
		SPHConfig sphConfigRexiAddedModes;
		SPHConfig *sphConfigRexi = sphConfig;
		const int X = simVars.rexi_use_extended_modes;
		if (X > 0) {
			sphConfigRexiAddedModes.setupAdditionalModes(sphConfig, X, X);
			sphConfigRexi = &sphConfigRexiAddedModes;
			std::cout << "REXI with " << X << " additional modes: T" << sphConfigRexi->spec_n_max << ", "
				  << sphConfigRexi->spat_num_lon << " x " << sphConfigRexi->spat_num_lat << std::endl;
		}
		rexi.setup(simVars.rexi_h, simVars.rexi_M, 0, simVars.rexi_use_half_poles);
		std::cout << "REXI poles: " << rexi.alpha.size() << std::endl;
		const double dt = simVars.timecontrol.current_timestep_size;
		sphswe_rexi *batched = nullptr;
		std::vector<SWERexiSPH> terms;
		std::vector<SWERexiSPHRobert> terms_robert;   // Robert-function velocities (AppSWE.hpp:372-390)
		double *d_state = nullptr;
		if (simVars.gpu_fused_step) {
			// the implicit viscosity is diagonal in the degree, so applying it in
			// the larger configuration before the copy back is the same as after
			sphswe_rexi_params p = {simVars.earth_radius, simVars.gravitation, with_coriolis ? simVars.coriolis_omega : 0.0,
						simVars.h0, simVars.viscosity2, dt, simVars.rexi_h, simVars.rexi_M,
						simVars.rexi_use_half_poles ? 1 : 0, SPHSWE_PREC_FP64, SPHSWE_PREC_FP64, 0,
						simVars.use_robert_functions ? 1 : 0};
			sphConfigRexi->check(sphswe_rexi_create(sphConfigRexi->gpu, &p, &batched), "sphswe_rexi_create");
			sphConfigRexi->check(sphswe_malloc(sphConfigRexi->gpu,
							   3 * sizeof(std::complex<double>) * sphConfigRexi->spec_num_elems,
							   (void **)&d_state), "state");
		} else {
			if (simVars.use_robert_functions) terms_robert.resize(rexi.alpha.size());
			else terms.resize(rexi.alpha.size());
			for (std::size_t i = 0; i < rexi.alpha.size(); i++) {
				if (simVars.use_robert_functions)
					terms_robert[i].setup(sphConfigRexi, rexi.alpha[i], rexi.beta_re[i], simVars.earth_radius,
							      simVars.coriolis_omega, simVars.h0 * simVars.gravitation, dt, with_coriolis);
				else
					terms[i].setup(sphConfigRexi, rexi.alpha[i], rexi.beta_re[i], simVars.earth_radius,
						       simVars.coriolis_omega, simVars.h0 * simVars.gravitation, dt, with_coriolis);
			}
		}
		SPHData tmp_phi(sphConfigRexi), tmp_u(sphConfigRexi), tmp_v(sphConfigRexi);
		SPHData acc_phi(sphConfigRexi), acc_u(sphConfigRexi), acc_v(sphConfigRexi);
		simVars.timecontrol.current_simulation_time = 0;
		while (simVars.timecontrol.current_simulation_time < simVars.timecontrol.max_simulation_time) {
			output_if_due();
			if (batched) {
				if (X > 0) {
					SPHData *p[3] = {&prog_h, &prog_u, &prog_v};
					const std::size_t n = sphConfigRexi->spec_num_elems;
					for (int k = 0; k < 3; k++) {
						p[k]->request_data_spectral();
						sphConfig->check(sphswe_spec_copy_modes(sphConfig->gpu, p[k]->d_spec, sphConfigRexi->gpu,
											(double *)((std::complex<double> *)d_state + k * n), 1),
								 "copy to the REXI modes");
					}
					sphConfigRexi->check(sphswe_rexi_step(sphConfigRexi->gpu, batched, d_state, 1), "sphswe_rexi_step");
					for (int k = 0; k < 3; k++) {
						sphConfig->check(sphswe_spec_copy_modes(sphConfigRexi->gpu,
											(const double *)((std::complex<double> *)d_state + k * n),
											sphConfig->gpu, p[k]->d_spec, 1),
								 "copy from the REXI modes");
						p[k]->data_spec_valid = true;
						p[k]->data_spat_valid = false;
					}
				} else {
					pack(d_state);
					sphConfig->check(sphswe_rexi_step(sphConfig->gpu, batched, d_state, 1), "sphswe_rexi_step");
					unpack(d_state);
				}
			} else {
				SPHData phi(sphConfigRexi), u(sphConfigRexi), v(sphConfigRexi);
				if (X == 0) {
					phi = prog_h * simVars.gravitation;
					u = prog_u;
					v = prog_v;
				} else {
					(prog_h * simVars.gravitation).spec_copyToDifferentModes(phi);
					prog_u.spec_copyToDifferentModes(u);
					prog_v.spec_copyToDifferentModes(v);
				}
				acc_phi.spec_set_zero();
				acc_u.spec_set_zero();
				acc_v.spec_set_zero();
				auto add_term = [&](auto &t) {
					t.solve(phi, u, v, tmp_phi, tmp_u, tmp_v, opComplex);
					acc_phi += tmp_phi;
					acc_u += tmp_u;
					acc_v += tmp_v;
				};
				for (auto &t : terms) add_term(t);
				for (auto &t : terms_robert) add_term(t);
				if (X == 0) {
					prog_h = acc_phi * (1.0 / simVars.gravitation);
					prog_u = acc_u;
					prog_v = acc_v;
				} else {
					(acc_phi * (1.0 / simVars.gravitation)).spec_copyToDifferentModes(prog_h);
					acc_u.spec_copyToDifferentModes(prog_u);
					acc_v.spec_copyToDifferentModes(prog_v);
				}
				if (simVars.viscosity2 != 0) {
					const double scalar = simVars.viscosity2 * dt, r = simVars.earth_radius;
					prog_h = prog_h.spec_solve_helmholtz(1.0, -scalar, r);
					prog_u = prog_u.spec_solve_helmholtz(1.0, -scalar, r);
					prog_v = prog_v.spec_solve_helmholtz(1.0, -scalar, r);
				}
			}
			std::cout << "." << std::flush;
			check_nan();
			simVars.timecontrol.current_simulation_time += dt;
		}
		write_output();
		std::cout << std::endl;
		if (batched) sphswe_rexi_destroy(batched);
		if (d_state) sphswe_free(sphConfigRexi->gpu, d_state);
	}

	void dump_state()
	{
		if (simVars.dump_file.empty()) return;
		FILE *fp = std::fopen(simVars.dump_file.c_str(), "wb");
		if (!fp) {
			std::cerr << "cannot write " << simVars.dump_file << std::endl;
			exit(1);
		}
		SPHData *p[3] = {&prog_h, &prog_u, &prog_v};
		for (int k = 0; k < 3; k++) {
			p[k]->spec_download();
			std::fwrite(p[k]->data_spec.data(), sizeof(std::complex<double>), p[k]->data_spec.size(), fp);
		}
		std::fclose(fp);
	}
};

#endif
