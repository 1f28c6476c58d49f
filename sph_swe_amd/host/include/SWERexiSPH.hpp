// One REXI term of the linear SWE on the sphere (src/SWERexiSPH.hpp:68-244): for
// the pole alpha and weight beta, solve
//   (alpha / dt + L)(phi, u, v) = (phi0, u0, v0) / dt,  return Re(beta / dt (phi, u, v)),
// with the per-m band solves on the device.  The state carries the geopotential
// phi = g h.  The Coriolis forcing product covers the whole grid.
// SWERexiSPHRobert (src/SWERexiSPHRobert.hpp:68-258) is the same term for
// Robert-function velocities: z4..z6robert bands, robert_div / robert_vort
// right-hand sides and the spectral Robert gradients.
#ifndef SPHSWE_HOST_SWEREXISPH_HPP
#define SPHSWE_HOST_SWEREXISPH_HPP

#include "sph/SPHData.hpp"
#include "sph/SPHOperatorsComplex.hpp"

template <int ROBERT>
class SWERexiSPHTerm
{
	SPHConfig *sphConfig = nullptr;
	sphswe_rexi *term = nullptr;
	double *d_io = nullptr;       // (phi, u, v) real spectra

public:
	SWERexiSPHTerm() {}
	SWERexiSPHTerm(const SWERexiSPHTerm &) = delete;
	SWERexiSPHTerm &operator=(const SWERexiSPHTerm &) = delete;
	SWERexiSPHTerm(SWERexiSPHTerm &&o) noexcept : sphConfig(o.sphConfig), term(o.term), d_io(o.d_io)
	{
		o.term = nullptr;
		o.d_io = nullptr;
	}
	~SWERexiSPHTerm() { release(); }

	void setup(SPHConfig *i_sphConfig, const std::complex<double> &i_alpha, const std::complex<double> &i_beta,
		   double i_radius, double i_coriolis_omega, double i_avg_geopotential, double i_timestep_size,
		   bool i_include_coriolis_effect = true)
	{
		release();
		sphConfig = i_sphConfig;
		// the term sees only Phi = g h0: pass it as h0 with g = 1
		sphswe_rexi_params p = {i_radius, 1.0, i_include_coriolis_effect ? i_coriolis_omega : 0.0,
					i_avg_geopotential, 0.0, i_timestep_size, 0.2, 1, 1, SPHSWE_PREC_FP64, SPHSWE_PREC_FP64, 0,
					ROBERT};
		const double a[2] = {i_alpha.real(), i_alpha.imag()}, b[2] = {i_beta.real(), i_beta.imag()};
		sphConfig->check(sphswe_rexi_create_poles(sphConfig->gpu, &p, 1, a, b, &term), "SWERexiSPH::setup");
		sphConfig->check(sphswe_malloc(sphConfig->gpu, 3 * sizeof(std::complex<double>) * sphConfig->spec_num_elems,
					       (void **)&d_io), "SWERexiSPH::setup");
	}

	void solve(const SPHData &i_phi0, const SPHData &i_u0, const SPHData &i_v0, SPHData &o_phi, SPHData &o_u,
		   SPHData &o_v, SPHOperatorsComplex &)
	{
		const std::size_t n = sizeof(std::complex<double>) * sphConfig->spec_num_elems;
		const SPHData *in[3] = {&i_phi0, &i_u0, &i_v0};
		SPHData *out[3] = {&o_phi, &o_u, &o_v};
		for (int k = 0; k < 3; k++) {
			in[k]->request_data_spectral();
			sphConfig->check(sphswe_memcpy_d2d(sphConfig->gpu, (char *)d_io + k * n, in[k]->d_spec, n), "SWERexiSPH::solve");
		}
		sphConfig->check(sphswe_rexi_apply(sphConfig->gpu, term, d_io, d_io), "SWERexiSPH::solve");
		for (int k = 0; k < 3; k++) {
			sphConfig->check(sphswe_memcpy_d2d(sphConfig->gpu, out[k]->d_spec, (char *)d_io + k * n, n), "SWERexiSPH::solve");
			out[k]->data_spec_valid = true;
			out[k]->data_spat_valid = false;
		}
	}

private:
	void release()
	{
		if (term) sphswe_rexi_destroy(term);
		if (d_io && sphConfig && sphConfig->gpu) sphswe_free(sphConfig->gpu, d_io);
		term = nullptr;
		d_io = nullptr;
	}
};

typedef SWERexiSPHTerm<0> SWERexiSPH;
typedef SWERexiSPHTerm<1> SWERexiSPHRobert;

#endif
