// SPHSolverComplex<T> over the C ABI: per-m banded matrices in the reference's
// NCompact row order, components z1..z8 (src/include/sph/SPHSolverComplex.hpp:
// setup :50-86, components :88-420, solve :462-511), factorised with partial
// pivoting as zgbsv and solved on the device.  The Robert-function components
// and apply() (a debugging helper) are not provided.
#ifndef SPHSWE_HOST_SPHSOLVERCOMPLEX_HPP
#define SPHSWE_HOST_SPHSOLVERCOMPLEX_HPP

#include "SPHDataComplex.hpp"

template <typename T = std::complex<double>>
class SPHSolverComplex
{
public:
	SPHConfig *sphConfig = nullptr;
	sphswe_band *band = nullptr;
	int halosize_off_diagonal = 0;

	SPHSolverComplex() {}
	SPHSolverComplex(const SPHSolverComplex &) = delete;
	SPHSolverComplex &operator=(const SPHSolverComplex &) = delete;
	~SPHSolverComplex() { shutdown(); }

	void setup(SPHConfig *i_sphConfig, int i_halosize_offdiagonal)
	{
		shutdown();
		sphConfig = i_sphConfig;
		halosize_off_diagonal = i_halosize_offdiagonal;
		sphConfig->check(sphswe_band_create(sphConfig->gpu, i_halosize_offdiagonal, &band), "SPHSolverComplex::setup");
	}

	void shutdown()
	{
		if (band) sphswe_band_destroy(band);
		band = nullptr;
	}

	void solver_component_scalar_phi(const T &s) { add(SPHSWE_BAND_SCALAR, s, 1.0); }
	void solver_component_mu_phi(const T &s = T(1.0)) { add(SPHSWE_BAND_MU, s, 1.0); }
	void solver_component_one_minus_mu_mu_diff_mu_phi(const T &s = T(1.0)) { add(SPHSWE_BAND_ONE_MINUS_MU2_DMU, s, 1.0); }
	void solver_component_rexi_z1(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 0, s, r); }
	void solver_component_rexi_z2(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 1, s, r); }
	void solver_component_rexi_z3(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 2, s, r); }
	void solver_component_rexi_z4(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 3, s, r); }
	void solver_component_rexi_z5(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 4, s, r); }
	void solver_component_rexi_z6(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 5, s, r); }
	void solver_component_rexi_z7(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 6, s, r); }
	void solver_component_rexi_z8(const T &s, double r) { add(SPHSWE_BAND_REXI_Z1 + 7, s, r); }

	// x with A x = rhs for every m; exits on a zero pivot like the reference's
	// zgbsv INFO check (BandedMatrixSolver.hpp:420-430)
	SPHDataComplex solve(const SPHDataComplex &i_rhs)
	{
		i_rhs.request_data_spectral();
		SPHDataComplex out(sphConfig);
		sphConfig->check(sphswe_band_solve(sphConfig->gpu, band, i_rhs.d_spec, out.d_spec), "SPHSolverComplex::solve");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

private:
	void add(int comp, const T &s, double r)
	{
		const std::complex<double> v(s);
		sphConfig->check(sphswe_band_add(band, comp, v.real(), v.imag(), r), "SPHSolverComplex component");
	}
};

#endif
