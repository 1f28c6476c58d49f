// SPHDataComplex over the C ABI: a complex field on the sphere with the full
// (l, m = -l..l) spectrum, device-resident like SPHData.
// Mirrors src/include/sph/SPHDataComplex.hpp: conversions from / to SPHData
// (:67-77, 95-175), lazy representations (:179-216), arithmetic (:218-460),
// spec_solve_helmholtz (:523-545) and lambdas (:549-729).  The product of two
// complex fields covers the whole grid; the reference's operator* covers only
// the first (L+1)^2 points (:321-340).
#ifndef SPHSWE_HOST_SPHDATACOMPLEX_HPP
#define SPHSWE_HOST_SPHDATACOMPLEX_HPP

#include "SPHData.hpp"

class SPHDataComplex
{
public:
	typedef std::complex<double> cplx;

	SPHConfig *sphConfig = nullptr;
	double *d_spat = nullptr;     // device, spat_num_elems complex
	double *d_spec = nullptr;     // device, cplx_spec_num_elems complex
	mutable bool data_spat_valid = false;
	mutable bool data_spec_valid = false;
	mutable std::vector<cplx> data_spat;
	mutable std::vector<cplx> data_spec;

	explicit SPHDataComplex(SPHConfig *i_sphConfig) { setup(i_sphConfig); }
	SPHDataComplex(const SPHDataComplex &i) { setup(i.sphConfig); operator=(i); }
	SPHDataComplex(const SPHData &i) { setup(i.sphConfig); operator=(i); }
	~SPHDataComplex() { release(); }

	void setup(SPHConfig *i_sphConfig)
	{
		release();
		sphConfig = i_sphConfig;
		sphConfig->check(sphswe_malloc(gpu(), spat_bytes(), (void **)&d_spat), "SPHDataComplex: allocation");
		sphConfig->check(sphswe_malloc(gpu(), spec_bytes(), (void **)&d_spec), "SPHDataComplex: allocation");
		data_spat_valid = data_spec_valid = false;
	}

	SPHDataComplex &operator=(const SPHDataComplex &i)
	{
		if (&i == this) return *this;
		if (i.data_spat_valid) sphConfig->check(sphswe_memcpy_d2d(gpu(), d_spat, i.d_spat, spat_bytes()), "copy");
		if (i.data_spec_valid) sphConfig->check(sphswe_memcpy_d2d(gpu(), d_spec, i.d_spec, spec_bytes()), "copy");
		data_spat_valid = i.data_spat_valid;
		data_spec_valid = i.data_spec_valid;
		return *this;
	}

	SPHDataComplex &operator=(const SPHData &i)
	{
		spat_fromSPHData(i);
		return *this;
	}

	// complex physical field = real physical field of i (SPHDataComplex.hpp:162-175)
	void spat_fromSPHData(const SPHData &i)
	{
		i.request_data_physical();
		sphConfig->check(sphswe_spat_real_to_cplx(gpu(), i.d_spat, d_spat, 1), "spat_fromSPHData");
		data_spat_valid = true;
		data_spec_valid = false;
	}
	void spat_RealToSPHData(SPHData &o) const { part_to(o, 0); }
	void spat_ImagToSPHData(SPHData &o) const { part_to(o, 1); }

	void request_data_spectral() const
	{
		if (data_spec_valid) return;
		sphConfig->check(sphswe_sht_analysis_cplx(gpu(), d_spat, d_spec, 1, SPHSWE_PREC_FP64), "spat_cplx_to_SH");
		data_spec_valid = true;
		data_spat_valid = false;
	}
	void request_data_physical() const
	{
		if (data_spat_valid) return;
		sphConfig->check(sphswe_sht_synthesis_cplx(gpu(), d_spec, d_spat, 1, SPHSWE_PREC_FP64), "SH_to_spat_cplx");
		data_spat_valid = true;
		data_spec_valid = false;
	}

	SPHDataComplex operator+(const SPHDataComplex &i) const { return combine(i, 1.0); }
	SPHDataComplex operator-(const SPHDataComplex &i) const { return combine(i, -1.0); }
	SPHDataComplex &operator+=(const SPHDataComplex &i) { return accumulate(i, 1.0); }
	SPHDataComplex &operator-=(const SPHDataComplex &i) { return accumulate(i, -1.0); }
	SPHDataComplex operator-() const { return scaled(cplx(-1, 0)); }
	SPHDataComplex operator*(double v) const { return scaled(cplx(v, 0)); }
	SPHDataComplex operator*(const cplx &v) const { return scaled(v); }
	SPHDataComplex operator/(double v) const { return scaled(cplx(1.0 / v, 0)); }
	const SPHDataComplex &operator*=(double v) const { return scale_in_place(cplx(v, 0)); }
	const SPHDataComplex &operator*=(const cplx &v) const { return scale_in_place(v); }

	SPHDataComplex operator+(double v) const { return plus_scalar(cplx(v, 0)); }
	SPHDataComplex operator+(const cplx &v) const { return plus_scalar(v); }

	// complex physical product over the whole grid
	SPHDataComplex operator*(const SPHDataComplex &i) const
	{
		request_data_physical();
		i.request_data_physical();
		SPHDataComplex out(sphConfig);
		sphConfig->check(sphswe_spat_op_cplx(gpu(), SPHSWE_SPAT_MUL, d_spat, i.d_spat, out.d_spat, 1.0, 1),
				 "SPHDataComplex *");
		out.data_spat_valid = true;
		out.data_spec_valid = false;
		return out;
	}

	// largest |real| and |imaginary| difference to i over the grid
	// (SPHDataComplex.hpp:744-771), -1 for an empty grid; on the host
	double spat_reduce_error_max(const SPHDataComplex &i) const
	{
		spat_download();
		i.spat_download();
		double e = -1.0;
		for (std::size_t k = 0; k < data_spat.size(); k++) {
			const cplx d = data_spat[k] - i.data_spat[k];
			e = std::max(e, std::max(std::abs(d.real()), std::abs(d.imag())));
		}
		return e;
	}

	// the field projected on the spectral modes and back (synthesis, analysis)
	SPHDataComplex spec_truncate() const
	{
		request_data_spectral();
		SPHDataComplex out(sphConfig);
		sphConfig->check(sphswe_sht_synthesis_cplx(gpu(), d_spec, out.d_spat, 1, SPHSWE_PREC_FP64), "spec_truncate");
		sphConfig->check(sphswe_sht_analysis_cplx(gpu(), out.d_spat, out.d_spec, 1, SPHSWE_PREC_FP64), "spec_truncate");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	SPHDataComplex spec_solve_helmholtz(const cplx &i_a, const cplx &i_b, double r) const
	{
		request_data_spectral();
		SPHDataComplex out(sphConfig);
		sphConfig->check(sphswe_spec_op_cplx(gpu(), SPHSWE_OP_HELMHOLTZ, d_spec, out.d_spec, 1, i_a.real(), i_a.imag(),
						     i_b.real(), r, i_b.imag()), "spec_solve_helmholtz");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	void spec_download() const
	{
		request_data_spectral();
		data_spec.resize(sphConfig->cplx_spec_num_elems);
		sphConfig->check(sphswe_memcpy_d2h(gpu(), data_spec.data(), d_spec, spec_bytes()), "spec_download");
	}
	void spat_download() const
	{
		request_data_physical();
		data_spat.resize(sphConfig->spat_num_elems);
		sphConfig->check(sphswe_memcpy_d2h(gpu(), data_spat.data(), d_spat, spat_bytes()), "spat_download");
	}
	void spec_upload()
	{
		sphConfig->check(sphswe_memcpy_h2d(gpu(), d_spec, data_spec.data(), spec_bytes()), "spec_upload");
		data_spec_valid = true;
		data_spat_valid = false;
	}
	void spat_upload()
	{
		sphConfig->check(sphswe_memcpy_h2d(gpu(), d_spat, data_spat.data(), spat_bytes()), "spat_upload");
		data_spat_valid = true;
		data_spec_valid = false;
	}

	void spec_update_lambda(std::function<void(int, int, cplx &)> fn)
	{
		if (data_spat_valid || data_spec_valid) spec_download();
		else data_spec.assign(sphConfig->cplx_spec_num_elems, cplx(0, 0));
		for (int n = 0; n <= sphConfig->spec_n_max; n++)
			for (int m = -n; m <= n; m++) fn(n, m, data_spec[sphConfig->getArrayIndexByModes_Complex(n, m)]);
		spec_upload();
	}
	void spat_update_lambda(std::function<void(double, double, cplx &)> fn) { update_grid(fn, sphConfig->lat); }
	void spat_update_lambda_gaussian_grid(std::function<void(double, double, cplx &)> fn)
	{
		update_grid(fn, sphConfig->lat_gaussian);
	}
	void spat_update_lambda_cogaussian_grid(std::function<void(double, double, cplx &)> fn)
	{
		update_grid(fn, sphConfig->lat_cogaussian);
	}

	void spat_set_zero()
	{
		sphConfig->check(sphswe_fill_zero(gpu(), d_spat, spat_bytes()), "spat_set_zero");
		data_spat_valid = true;
		data_spec_valid = false;
	}
	void spec_set_zero()
	{
		sphConfig->check(sphswe_fill_zero(gpu(), d_spec, spec_bytes()), "spec_set_zero");
		data_spec_valid = true;
		data_spat_valid = false;
	}

	sphswe_ctx *gpu() const { return sphConfig->gpu; }

private:
	std::size_t spat_bytes() const { return sizeof(cplx) * (std::size_t)sphConfig->spat_num_elems; }
	std::size_t spec_bytes() const { return sizeof(cplx) * (std::size_t)sphConfig->cplx_spec_num_elems; }

	void release()
	{
		if (sphConfig && sphConfig->gpu) {
			if (d_spat) sphswe_free(sphConfig->gpu, d_spat);
			if (d_spec) sphswe_free(sphConfig->gpu, d_spec);
		}
		d_spat = d_spec = nullptr;
	}

	void part_to(SPHData &o, int part) const
	{
		request_data_physical();
		sphConfig->check(sphswe_spat_cplx_part(gpu(), d_spat, o.d_spat, part, 1), "spat_RealToSPHData");
		o.data_spat_valid = true;
		o.data_spec_valid = false;
	}

	SPHDataComplex combine(const SPHDataComplex &i, double sgn) const
	{
		request_data_spectral();
		i.request_data_spectral();
		SPHDataComplex out(sphConfig);
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->cplx_spec_num_elems, 1.0, 0.0, d_spec, sgn, 0.0, i.d_spec,
						   out.d_spec), "SPHDataComplex +/-");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	SPHDataComplex &accumulate(const SPHDataComplex &i, double sgn)
	{
		request_data_spectral();
		i.request_data_spectral();
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->cplx_spec_num_elems, 1.0, 0.0, d_spec, sgn, 0.0, i.d_spec,
						   d_spec), "SPHDataComplex +=/-=");
		data_spec_valid = true;
		data_spat_valid = false;
		return *this;
	}

	SPHDataComplex scaled(cplx v) const
	{
		request_data_spectral();
		SPHDataComplex out(sphConfig);
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->cplx_spec_num_elems, v.real(), v.imag(), d_spec, 0.0, 0.0,
						   nullptr, out.d_spec), "SPHDataComplex scale");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	const SPHDataComplex &scale_in_place(cplx v) const
	{
		request_data_spectral();
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->cplx_spec_num_elems, v.real(), v.imag(), d_spec, 0.0, 0.0,
						   nullptr, d_spec), "SPHDataComplex *=");
		return *this;
	}

	// (0,0) += v sqrt(4 pi); index 0 of the complex layout is (0,0) as well
	SPHDataComplex plus_scalar(cplx v) const
	{
		request_data_spectral();
		SPHDataComplex out(*this);
		sphConfig->check(sphswe_spec_add_scalar(gpu(), out.d_spec, v.real(), v.imag()), "SPHDataComplex + scalar");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	void update_grid(std::function<void(double, double, cplx &)> &fn, const std::vector<double> &coord)
	{
		if (data_spec_valid || data_spat_valid) spat_download();
		else data_spat.assign(sphConfig->spat_num_elems, cplx(0, 0));
		for (int i = 0; i < sphConfig->spat_num_lon; i++) {
			const double lon = ((double)i / (double)sphConfig->spat_num_lon) * 2.0 * M_PI;
			for (int j = 0; j < sphConfig->spat_num_lat; j++) fn(lon, coord[j], data_spat[i * sphConfig->spat_num_lat + j]);
		}
		spat_upload();
	}
};

inline SPHDataComplex operator*(const std::complex<double> &v, const SPHDataComplex &i) { return i * v; }
inline SPHDataComplex operator*(double v, const SPHDataComplex &i) { return i * v; }

#endif
