// Command-line entry of the shallow-water program (src/main.cpp:33-262): the
// same flags and programs: P0 (AppSWE), P1 (AppTestSPHOperators), P2
// (AppTestSPHOperatorsComplex), P3 (AppTestSPHSolvers), P4
// (AppTestSPHSolversComplex) and P5 (AppOutputSphericalHarmonics).  Two letters have no counterpart in the reference:
// G<0|1> (one library call per time step, default 1) and Z<file> (write the
// final (h, u, v) spectra as raw complex doubles).
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "AppSWE.hpp"
#include "AppOutputSphericalHarmonics.hpp"
#include "AppTestSPHOperators.hpp"
#include "AppTestSPHOperatorsComplex.hpp"
#include "AppTestSPHSolver.hpp"
#include "AppTestSPHSolverComplex.hpp"
#include "SimVars.hpp"
#include "sph/SPHConfig.hpp"

SimVars simVars;

static int setup(int argc, const char *argv[], SPHConfig *cfg)
{
	if (argc < 2) {
		std::cerr << "Usage: " << argv[0] << " T[n]                         # Use T notation" << std::endl;
		std::cerr << "   OR: " << argv[0] << " [n] [m]                      # set modes" << std::endl;
		std::cerr << "   OR: " << argv[0] << " [n] [m] [res_lon] [res_lat]  # set modes and physical resolution" << std::endl;
		return -1;
	}
	if (argv[1][0] == 'T') {
		const int order = std::atoi(&argv[1][1]);
		std::cout << "Using T=" << order << std::endl;
		simVars.spec_res_n = order;
		simVars.spec_res_m = order;
		for (int i = 2; i < argc; i++) {
			const char *v = &argv[i][1];
			switch (argv[i][0]) {
			case 'F': simVars.coriolis_omega = std::atof(v); break;
			case 'P': simVars.program_id = std::atoi(v); break;
			case 'N': simVars.use_nonlinear_equations = std::atoi(v); break;
			case 'M':
				simVars.rexi_M = std::atoi(v);
				simVars.use_rexi = simVars.rexi_M > 0;
				break;
			case 'H': simVars.rexi_h = std::atof(v); break;
			case 'I': simVars.rexi_use_half_poles = std::atoi(v); break;
			case 'R': simVars.timecontrol.current_timestep_size = std::atof(v); break;
			case 'S': simVars.use_robert_functions = std::atoi(v); break;
			case 'X': simVars.rexi_use_extended_modes = std::atoi(v); break;
			case 'Q': simVars.benchmark_scenario_id = std::atoi(v); break;
			case 'E': simVars.timecontrol.max_simulation_time = std::atof(v); break;
			case 'V': simVars.viscosity2 = std::atof(v); break;
			case 'O': simVars.output_dt = std::atof(v); break;
			case 'W': simVars.rexi_M = std::atoi(v); break;
			case 'G': simVars.gpu_fused_step = std::atoi(v); break;
			case 'Z': simVars.dump_file = v; break;
			}
		}
	} else {
		if (argc > 1) simVars.spec_res_n = std::atoi(argv[1]);
		if (argc > 2) simVars.spec_res_m = std::atoi(argv[2]);
		if (argc > 3) simVars.spat_res_lon = std::atoi(argv[3]);
		if (argc > 4) simVars.spat_res_lat = std::atoi(argv[4]);
	}
	if (simVars.spec_res_m == -1) simVars.spec_res_m = simVars.spec_res_n;

	std::cout << "Running setup..." << std::endl;
	if (simVars.spat_res_lon <= 0)
		cfg->setupAutoPhysicalSpace(simVars.spec_res_n, simVars.spec_res_m, &simVars.spat_res_lon, &simVars.spat_res_lat);
	else
		cfg->setup(simVars.spec_res_n, simVars.spec_res_m, simVars.spat_res_lon, simVars.spat_res_lat);
	std::cout << "Using spectral modes m=" << simVars.spec_res_m << " and n=" << simVars.spec_res_n << std::endl;
	std::cout << "Requested physical resolution " << simVars.spat_res_lon << " x " << simVars.spat_res_lat << std::endl;
	std::cout << "Using physical resolution " << cfg->spat_num_lon << " x " << cfg->spat_num_lat << std::endl;
	std::cout << "Using scalar for Coriolis " << simVars.coriolis_omega << std::endl;
	return 0;
}

int main(int argc, const char *argv[])
{
	SPHConfig sphConfig;
	const int rc = setup(argc, argv, &sphConfig);
	if (rc != 0) return rc;
	if (simVars.program_id == 0) {
		std::cout << "************************************************" << std::endl;
		std::cout << "AppSWE" << std::endl;
		std::cout << "************************************************" << std::endl;
		AppSWE app(&sphConfig, simVars);
		app.run();
		return 0;
	}
	if (simVars.program_id == 1) {
		std::cout << "************************************************" << std::endl;
		std::cout << "AppTestSPHOperators" << std::endl;
		std::cout << "************************************************" << std::endl;
		AppTestSPHOperators test_operators;
		test_operators.run(&sphConfig);
		return 0;
	}
	if (simVars.program_id == 2) {
		std::cout << "************************************************" << std::endl;
		std::cout << "AppTestSPHOperatorsComplex" << std::endl;
		std::cout << "************************************************" << std::endl;
		AppTestSPHOperatorsComplex test_operators;
		test_operators.run(&sphConfig);
		return 0;
	}
	if (simVars.program_id == 3) {
		std::cout << "************************************************" << std::endl;
		std::cout << "AppTestSPHSolvers" << std::endl;
		std::cout << "************************************************" << std::endl;
		AppTestSPHSolvers test_solvers;
		test_solvers.run(&sphConfig);
		return 0;
	}
	if (simVars.program_id == 5) {
		std::cout << "************************************************" << std::endl;
		std::cout << "AppOutputSphericalHarmonics" << std::endl;
		std::cout << "************************************************" << std::endl;
		AppOutputSphericalHarmonics output_spherical_harmonics;
		output_spherical_harmonics.run(&sphConfig);
		return 0;
	}
	if (simVars.program_id == 4) {
		std::cout << "************************************************" << std::endl;
		std::cout << "AppTestSPHSolversComplex" << std::endl;
		std::cout << "************************************************" << std::endl;
		AppTestSPHSolversComplex test_solvers;
		test_solvers.run(&sphConfig);
		return 0;
	}
	std::cerr << "Program " << simVars.program_id << " is not available in this build" << std::endl;
	return 1;
}
