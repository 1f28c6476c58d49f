// Simulation variables with the reference's defaults (src/include/SimVars.hpp:13-84).
// Two fields are additions of this build: gpu_fused_step and dump_file.
#ifndef SPHSWE_HOST_SIMVARS_HPP
#define SPHSWE_HOST_SIMVARS_HPP

#include <string>

struct SimVars
{
	int spat_res_lon = -1;
	int spat_res_lat = -1;
	int spec_res_m = -1;
	int spec_res_n = -1;

	double coriolis_omega = 7.292e-5;
	double gravitation = 9.80616;
	double earth_radius = 6.37122e6;

	int program_id = 0;
	double viscosity2 = 1e5;

	double output_dt = 0.2;
	double next_output_dt = 0;

	double rexi_h = 0.2;
	int rexi_M = 128;
	bool rexi_use_half_poles = true;
	int use_robert_functions = 0;
	int rexi_use_extended_modes = 0;
	bool use_rexi = false;

	int benchmark_scenario_id = 0;
	bool use_nonlinear_equations = true;
	double h0 = 10000.0;

	struct TimeControl
	{
		int timestepping_runge_kutta_order = 4;
		double current_timestep_size = 0;
		double current_simulation_time = 0;
		double max_simulation_time = -1;
	} timecontrol;

	// 1: one library call per time step (sphswe_swe_rk4 / sphswe_rexi_step);
	// 0: the class-by-class path of the reference (TimesteppingRK with
	// p_run_euler_timestep_update, one SWERexiSPH per pole)
	int gpu_fused_step = 1;
	// if set, the final (h, u, v) spectra are written there as raw complex doubles
	std::string dump_file;
};

#endif
