// Explicit Runge-Kutta time stepping on (h, u, v) SPHData fields
// (src/include/sweet/TimesteppingRK.hpp:82-120, 160-392): orders 1 to 4 with the
// reference's tableaux, stage inputs io + dt (sum a_j k_j) and the update
// io += dt (sum b_j k_j), all in spectral space.
#ifndef SPHSWE_HOST_TIMESTEPPINGRK_HPP
#define SPHSWE_HOST_TIMESTEPPINGRK_HPP

#include <cstdlib>
#include <iostream>
#include <limits>
#include <memory>
#include <vector>

#include "../sph/SPHData.hpp"

class TimesteppingRK
{
	std::vector<std::unique_ptr<SPHData>> RK_h_t, RK_u_t, RK_v_t;
	int runge_kutta_order = -1;

public:
	void setupBuffers(const SPHData &i_test_buffer, int i_rk_order)
	{
		if (!RK_h_t.empty()) return;
		runge_kutta_order = i_rk_order;
		for (int i = 0; i < i_rk_order; i++) {
			RK_h_t.emplace_back(new SPHData(i_test_buffer.sphConfig));
			RK_u_t.emplace_back(new SPHData(i_test_buffer.sphConfig));
			RK_v_t.emplace_back(new SPHData(i_test_buffer.sphConfig));
		}
	}

	template <class BaseClass>
	void run_rk_timestep(BaseClass *i_baseClass,
			     void (BaseClass::*i_compute_euler_timestep_update)(const SPHData &, const SPHData &, const SPHData &,
									      SPHData &, SPHData &, SPHData &, double &, double,
									      double),
			     SPHData &io_h, SPHData &io_u, SPHData &io_v, double &o_dt, double i_use_fixed_dt = 0,
			     int i_runge_kutta_order = 1, double i_simulation_time = -1,
			     double i_max_simulation_time = std::numeric_limits<double>::infinity())
	{
		// Butcher tableaux of the reference (TimesteppingRK.hpp:191-334)
		static const double A2[2][2] = {{0, 0}, {0.5, 0}};
		static const double B2[2] = {0.0, 1.0}, C2[2] = {0, 0.5};
		static const double A3[3][3] = {{0, 0, 0}, {1.0 / 3.0, 0, 0}, {0.0, 2.0 / 3.0, 0}};
		static const double B3[3] = {1.0 / 4.0, 0.0, 3.0 / 4.0}, C3[3] = {0, 1.0 / 3.0, 2.0 / 3.0};
		static const double A4[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0.0, 0.5, 0, 0}, {0.0, 0.0, 1.0, 0}};
		static const double B4[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0}, C4[4] = {0, 0.5, 0.5, 1.0};
		const int s = i_runge_kutta_order;
		if (s < 1 || s > 4) {
			std::cerr << "This order of the Runge-Kutta time stepping is not supported!" << std::endl;
			exit(1);
		}
		setupBuffers(io_h, s);
		auto a = [&](int i, int j) { return s == 2 ? A2[i][j] : s == 3 ? A3[i][j] : s == 4 ? A4[i][j] : 0.0; };
		auto b = [&](int j) { return s == 1 ? 1.0 : s == 2 ? B2[j] : s == 3 ? B3[j] : B4[j]; };
		auto c = [&](int i) { return s == 2 ? C2[i] : s == 3 ? C3[i] : s == 4 ? C4[i] : 0.0; };

		double &dt = o_dt;
		double dummy_dt = 0;
		(i_baseClass->*i_compute_euler_timestep_update)(io_h, io_u, io_v, *RK_h_t[0], *RK_u_t[0], *RK_v_t[0], dt,
								 i_use_fixed_dt, i_simulation_time);
		if (i_max_simulation_time >= 0 && dt + i_simulation_time > i_max_simulation_time)
			dt = i_max_simulation_time - i_simulation_time;
		for (int i = 1; i < s; i++) {
			// sum of the non-zero a_ij k_j of the stage
			auto stage = [&](SPHData &io, std::vector<std::unique_ptr<SPHData>> &K) {
				SPHData acc(io.sphConfig);
				bool first = true;
				for (int j = 0; j < i; j++) {
					if (a(i, j) == 0.0) continue;
					if (first) acc = a(i, j) * (*K[j]);
					else acc = acc + a(i, j) * (*K[j]);
					first = false;
				}
				return io + dt * acc;
			};
			(i_baseClass->*i_compute_euler_timestep_update)(stage(io_h, RK_h_t), stage(io_u, RK_u_t), stage(io_v, RK_v_t),
									 *RK_h_t[i], *RK_u_t[i], *RK_v_t[i], dummy_dt, dt,
									 i_simulation_time + c(i) * dt);
		}
		auto update = [&](SPHData &io, std::vector<std::unique_ptr<SPHData>> &K) {
			if (s == 1) {
				io += dt * (*K[0]);
				return;
			}
			SPHData acc = b(0) * (*K[0]);
			for (int j = 1; j < s; j++) acc = acc + b(j) * (*K[j]);
			io += dt * acc;
		};
		update(io_h, RK_h_t);
		update(io_u, RK_u_t);
		update(io_v, RK_v_t);
	}
};

#endif
