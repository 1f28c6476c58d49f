// SPHSolver<T> over the C ABI: the banded solver on real-field spectra
// (src/include/sph/SPHSolver.hpp: setup :46-62, components :64-300, solve
// :346-375).  The reference factorises the m >= 0 blocks of the real spectrum;
// here the m >= 0 coefficients are placed in the complex spectral layout
// (index l^2 + l + m, zero for m < 0), solved by the device band solver of
// SPHSolverComplex (the same per-m band rows, which depend on the signed m
// only) and taken back.  The layout change runs on the host: the class serves
// the self-test program P3.
#ifndef SPHSWE_HOST_SPHSOLVER_HPP
#define SPHSWE_HOST_SPHSOLVER_HPP

#include <complex>
#include <vector>

#include "SPHData.hpp"

template <typename T = std::complex<double>>
class SPHSolver
{
public:
	SPHConfig *sphConfig = nullptr;
	sphswe_band *band = nullptr;
	int halosize_off_diagonal = 0;

	SPHSolver() {}
	SPHSolver(const SPHSolver &) = delete;
	SPHSolver &operator=(const SPHSolver &) = delete;
	~SPHSolver() { shutdown(); }

	void setup(SPHConfig *i_sphConfig, int i_halosize_offdiagonal)
	{
		shutdown();
		sphConfig = i_sphConfig;
		halosize_off_diagonal = i_halosize_offdiagonal;
		sphConfig->check(sphswe_band_create(sphConfig->gpu, i_halosize_offdiagonal, &band), "SPHSolver::setup");
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

	// x with A x = rhs for every m >= 0; exits on a zero pivot like the
	// reference's zgbsv INFO check
	SPHData solve(const SPHData &i_rhs)
	{
		typedef std::complex<double> cplx;
		const int L = sphConfig->spec_n_max;
		const std::size_t nc = (std::size_t)(L + 1) * (L + 1);
		i_rhs.spec_download();
		std::vector<cplx> c(nc, cplx(0, 0));
		for (int m = 0; m <= sphConfig->spec_m_max; m++)
			for (int n = m; n <= L; n++)
				c[(std::size_t)n * n + n + m] = i_rhs.data_spec[sphConfig->getArrayIndexByModes(n, m)];
		double *d_in = nullptr, *d_out = nullptr;
		sphConfig->check(sphswe_malloc(sphConfig->gpu, nc * sizeof(cplx), (void **)&d_in), "SPHSolver::solve");
		sphConfig->check(sphswe_malloc(sphConfig->gpu, nc * sizeof(cplx), (void **)&d_out), "SPHSolver::solve");
		sphConfig->check(sphswe_memcpy_h2d(sphConfig->gpu, d_in, c.data(), nc * sizeof(cplx)), "SPHSolver::solve");
		sphConfig->check(sphswe_band_solve(sphConfig->gpu, band, d_in, d_out), "SPHSolver::solve");
		sphConfig->check(sphswe_memcpy_d2h(sphConfig->gpu, c.data(), d_out, nc * sizeof(cplx)), "SPHSolver::solve");
		sphswe_free(sphConfig->gpu, d_in);
		sphswe_free(sphConfig->gpu, d_out);
		SPHData out(sphConfig);
		out.data_spec.assign(sphConfig->spec_num_elems, cplx(0, 0));
		for (int m = 0; m <= sphConfig->spec_m_max; m++)
			for (int n = m; n <= L; n++)
				out.data_spec[sphConfig->getArrayIndexByModes(n, m)] = c[(std::size_t)n * n + n + m];
		out.spec_upload();
		return out;
	}

private:
	void add(int comp, const T &s, double r)
	{
		const std::complex<double> v(s);
		sphConfig->check(sphswe_band_add(band, comp, v.real(), v.imag(), r), "SPHSolver component");
	}
};

#endif
