// SPHData over the C ABI: a real scalar field on the sphere whose physical and
// spectral representations live in device memory (pooled, sphswe_malloc).
// Mirrors src/include/sph/SPHData.hpp: the lazy dual representation with
// exactly one representation valid after a conversion (:190-227), spectral
// + - and scalar * / (:231-443), the physical product (:334-353),
// spec_solve_helmholtz (:475-497), lambdas (:536-556, 645-766), reductions
// (:800-875) and the CSV writers (:935-1037).  Host copies (`data_spat`,
// `data_spec`) are filled on demand by the methods that need them.
#ifndef SPHSWE_HOST_SPHDATA_HPP
#define SPHSWE_HOST_SPHDATA_HPP

#include <complex>
#include <cstdio>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <vector>

#include "SPHConfig.hpp"

class SPHData
{
public:
	typedef std::complex<double> cplx;

	SPHConfig *sphConfig = nullptr;
	double *d_spat = nullptr;     // device, spat_num_elems doubles
	double *d_spec = nullptr;     // device, spec_num_elems complex (re, im)
	mutable bool data_spat_valid = false;
	mutable bool data_spec_valid = false;
	// host copies, valid after spat_download() / spec_download()
	mutable std::vector<double> data_spat;
	mutable std::vector<cplx> data_spec;

	explicit SPHData(SPHConfig *i_sphConfig) { setup(i_sphConfig); }
	SPHData(const SPHData &i_sph_data) { setup(i_sph_data.sphConfig); operator=(i_sph_data); }
	~SPHData() { release(); }

	void setup(SPHConfig *i_sphConfig)
	{
		release();
		sphConfig = i_sphConfig;
		sphConfig->check(sphswe_malloc(gpu(), sizeof(double) * sphConfig->spat_num_elems, (void **)&d_spat),
				 "SPHData: allocation");
		sphConfig->check(sphswe_malloc(gpu(), sizeof(cplx) * sphConfig->spec_num_elems, (void **)&d_spec),
				 "SPHData: allocation");
		data_spat_valid = data_spec_valid = false;
	}

	SPHData &operator=(const SPHData &i_sph_data)
	{
		if (&i_sph_data == this) return *this;
		if (i_sph_data.data_spat_valid)
			sphConfig->check(sphswe_memcpy_d2d(gpu(), d_spat, i_sph_data.d_spat, spat_bytes()), "SPHData: copy");
		if (i_sph_data.data_spec_valid)
			sphConfig->check(sphswe_memcpy_d2d(gpu(), d_spec, i_sph_data.d_spec, spec_bytes()), "SPHData: copy");
		data_spat_valid = i_sph_data.data_spat_valid;
		data_spec_valid = i_sph_data.data_spec_valid;
		return *this;
	}

	// ---- representation changes (analysis / synthesis on the device) --------
	void request_data_spectral() const
	{
		if (data_spec_valid) return;
		sphConfig->check(sphswe_sht_analysis(gpu(), d_spat, d_spec, 1, SPHSWE_PREC_FP64), "spat_to_SH");
		data_spec_valid = true;
		data_spat_valid = false;
	}

	void request_data_physical() const
	{
		if (data_spat_valid) return;
		sphConfig->check(sphswe_sht_synthesis(gpu(), d_spec, d_spat, 1, SPHSWE_PREC_FP64), "SH_to_spat");
		data_spat_valid = true;
		data_spec_valid = false;
	}

	// ---- spectral arithmetic --------------------------------------------------
	SPHData operator+(const SPHData &i) const { return spec_combine(i, 1.0); }
	SPHData operator-(const SPHData &i) const { return spec_combine(i, -1.0); }

	SPHData &operator+=(const SPHData &i) { return spec_accumulate(i, 1.0); }
	SPHData &operator-=(const SPHData &i) { return spec_accumulate(i, -1.0); }

	SPHData operator-() const { return scaled(cplx(-1.0, 0.0)); }
	SPHData operator*(double v) const { return scaled(cplx(v, 0.0)); }
	SPHData operator/(double v) const { return scaled(cplx(1.0 / v, 0.0)); }

	const SPHData &operator*=(double v) const { return scale_in_place(cplx(v, 0.0)); }
	const SPHData &operator*=(const cplx &v) const { return scale_in_place(v); }

	// adds v to the mean: (0,0) coefficient += v sqrt(4 pi) (SPHData.hpp:429-443)
	SPHData operator+(double v) const
	{
		request_data_spectral();
		SPHData out(*this);
		sphConfig->check(sphswe_spec_add_scalar(gpu(), out.d_spec, v, 0.0), "SPHData + double");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	// physical product (SPHData.hpp:334-353)
	SPHData operator*(const SPHData &i) const
	{
		request_data_physical();
		i.request_data_physical();
		SPHData out(sphConfig);
		sphConfig->check(sphswe_spat_op(gpu(), SPHSWE_SPAT_MUL, d_spat, i.d_spat, out.d_spat, 1.0, 1), "SPHData *");
		out.data_spat_valid = true;
		out.data_spec_valid = false;
		return out;
	}

	// (a + b D^2) x = this, D^2 the Laplacian on the sphere of radius r
	SPHData spec_solve_helmholtz(const double &i_a, const double &i_b, double r) const
	{
		request_data_spectral();
		SPHData out(sphConfig);
		sphConfig->check(sphswe_spec_op(gpu(), SPHSWE_OP_HELMHOLTZ, d_spec, out.d_spec, 1, i_a, 0.0, i_b, r),
				 "spec_solve_helmholtz");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	SPHData spat_truncate() const
	{
		request_data_physical();
		SPHData out(sphConfig);
		sphConfig->check(sphswe_sht_analysis(gpu(), d_spat, out.d_spec, 1, SPHSWE_PREC_FP64), "spat_truncate");
		sphConfig->check(sphswe_sht_synthesis(gpu(), out.d_spec, out.d_spat, 1, SPHSWE_PREC_FP64), "spat_truncate");
		out.data_spat_valid = true;
		out.data_spec_valid = false;
		return out;
	}

	SPHData spec_truncate() const
	{
		request_data_spectral();
		SPHData out(sphConfig);
		sphConfig->check(sphswe_sht_synthesis(gpu(), d_spec, out.d_spat, 1, SPHSWE_PREC_FP64), "spec_truncate");
		sphConfig->check(sphswe_sht_analysis(gpu(), out.d_spat, out.d_spec, 1, SPHSWE_PREC_FP64), "spec_truncate");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	// the spectrum in o's truncation (SPHData.hpp:108-187): orders and degrees
	// beyond the smaller truncation are dropped or zero-filled
	const SPHData &spec_copyToDifferentModes(SPHData &o) const
	{
		if (o.sphConfig == sphConfig) {
			std::cerr << "spec_copyToDifferentModes: same configuration" << std::endl;
			exit(1);
		}
		request_data_spectral();
		sphConfig->check(sphswe_spec_copy_modes(gpu(), d_spec, o.gpu(), o.d_spec, 1), "spec_copyToDifferentModes");
		o.data_spec_valid = true;
		o.data_spat_valid = false;
		return *this;
	}

	// ---- host-side element access -----------------------------------------------
	void spec_download() const
	{
		request_data_spectral();
		data_spec.resize(sphConfig->spec_num_elems);
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

	void spec_update_lambda(std::function<void(int, int, cplx &)> i_lambda)
	{
		if (data_spat_valid || data_spec_valid) spec_download();
		else data_spec.assign(sphConfig->spec_num_elems, cplx(0, 0));
		for (int m = 0; m <= sphConfig->spec_m_max; m++) {
			std::size_t idx = sphConfig->getArrayIndexByModes(m, m);
			for (int n = m; n <= sphConfig->spec_n_max; n++) i_lambda(n, m, data_spec[idx++]);
		}
		spec_upload();
	}

	// lambda(longitude in [0, 2pi), latitude phi, value)
	void spat_update_lambda(std::function<void(double, double, double &)> i_lambda)
	{
		spat_update_grid(i_lambda, sphConfig->lat);
	}
	// lambda(longitude, mu = sin(phi), value)
	void spat_update_lambda_gaussian_grid(std::function<void(double, double, double &)> i_lambda)
	{
		spat_update_grid(i_lambda, sphConfig->lat_gaussian);
	}
	// lambda(longitude, cos(phi), value)
	void spat_update_lambda_cogaussian_grid(std::function<void(double, double, double &)> i_lambda)
	{
		spat_update_grid(i_lambda, sphConfig->lat_cogaussian);
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
	void spat_set_value(double v)
	{
		data_spat.assign(sphConfig->spat_num_elems, v);
		spat_upload();
	}

	// NaN/Inf in the representation that is valid (the reference scans the
	// physical array whatever its state, SPHData.hpp:560-569)
	bool isAnyNaNorInf() const
	{
		int flag = 0;
		if (data_spat_valid)
			sphConfig->check(sphswe_any_nan(gpu(), d_spat, sphConfig->spat_num_elems, &flag), "isAnyNaNorInf");
		else
			sphConfig->check(sphswe_any_nan(gpu(), d_spec, 2 * (size_t)sphConfig->spec_num_elems, &flag),
					 "isAnyNaNorInf");
		return flag != 0;
	}

	double spat_reduce_min() const { double mn, mx, ab; minmax(mn, mx, ab); return mn; }
	double spat_reduce_max() const { double mn, mx, ab; minmax(mn, mx, ab); return mx; }
	double spat_reduce_abs_max() const { double mn, mx, ab; minmax(mn, mx, ab); return ab; }
	// largest grid-point difference to i, -1 for an empty grid
	// (SPHData.hpp:800-821): pointwise on the physical values, on the host
	double spat_reduce_error_max(const SPHData &i) const
	{
		request_data_physical();
		i.request_data_physical();
		spat_download();
		i.spat_download();
		double e = -1.0;
		for (std::size_t k = 0; k < data_spat.size(); k++) e = std::max(e, std::abs(data_spat[k] - i.data_spat[k]));
		return e;
	}

	void spec_print(int i_precision = 8) const
	{
		spec_download();
		std::cout << std::setprecision(i_precision);
		for (int n = 0; n <= sphConfig->spec_n_max; n++) {
			for (int m = 0; m <= n; m++) std::cout << data_spec[sphConfig->getArrayIndexByModes(n, m)] << "\t";
			std::cout << std::endl;
		}
	}

	void spat_print(int i_precision = 8) const
	{
		spat_download();
		std::cout << std::setprecision(i_precision);
		for (int j = sphConfig->spat_num_lat - 1; j >= 0; j--) {
			for (int i = 0; i < sphConfig->spat_num_lon; i++)
				std::cout << data_spat[i * sphConfig->spat_num_lat + j] << "\t";
			std::cout << std::endl;
		}
	}

	void spat_write_file(const char *i_filename, const char *i_title = "", int i_precision = 8) const
	{
		write_csv(i_filename, i_title, i_precision, false);
	}
	void spat_write_file_lon_pi_shifted(const char *i_filename, std::string i_title = "", int i_precision = 8) const
	{
		write_csv(i_filename, i_title.c_str(), i_precision, true);
	}

	sphswe_ctx *gpu() const { return sphConfig->gpu; }

private:
	std::size_t spat_bytes() const { return sizeof(double) * (std::size_t)sphConfig->spat_num_elems; }
	std::size_t spec_bytes() const { return sizeof(cplx) * (std::size_t)sphConfig->spec_num_elems; }

	void release()
	{
		if (sphConfig && sphConfig->gpu) {
			if (d_spat) sphswe_free(sphConfig->gpu, d_spat);
			if (d_spec) sphswe_free(sphConfig->gpu, d_spec);
		}
		d_spat = d_spec = nullptr;
	}

	SPHData spec_combine(const SPHData &i, double sgn) const
	{
		request_data_spectral();
		i.request_data_spectral();
		SPHData out(sphConfig);
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->spec_num_elems, 1.0, 0.0, d_spec, sgn, 0.0, i.d_spec,
						   out.d_spec), "SPHData +/-");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	SPHData &spec_accumulate(const SPHData &i, double sgn)
	{
		request_data_spectral();
		i.request_data_spectral();
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->spec_num_elems, 1.0, 0.0, d_spec, sgn, 0.0, i.d_spec,
						   d_spec), "SPHData +=/-=");
		data_spec_valid = true;
		data_spat_valid = false;
		return *this;
	}

	SPHData scaled(cplx v) const
	{
		request_data_spectral();
		SPHData out(sphConfig);
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->spec_num_elems, v.real(), v.imag(), d_spec, 0.0, 0.0,
						   nullptr, out.d_spec), "SPHData scale");
		out.data_spec_valid = true;
		out.data_spat_valid = false;
		return out;
	}

	const SPHData &scale_in_place(cplx v) const
	{
		request_data_spectral();
		sphConfig->check(sphswe_spec_axpby(gpu(), sphConfig->spec_num_elems, v.real(), v.imag(), d_spec, 0.0, 0.0,
						   nullptr, d_spec), "SPHData *=");
		return *this;
	}

	void spat_update_grid(std::function<void(double, double, double &)> &fn, const std::vector<double> &coord)
	{
		if (data_spec_valid || data_spat_valid) spat_download();
		else data_spat.assign(sphConfig->spat_num_elems, 0.0);
		for (int i = 0; i < sphConfig->spat_num_lon; i++) {
			const double lon = ((double)i / (double)sphConfig->spat_num_lon) * 2.0 * M_PI;
			for (int j = 0; j < sphConfig->spat_num_lat; j++) fn(lon, coord[j], data_spat[i * sphConfig->spat_num_lat + j]);
		}
		spat_upload();
	}

	void minmax(double &mn, double &mx, double &ab) const
	{
		request_data_physical();
		sphConfig->check(sphswe_reduce_minmax(gpu(), d_spat, sphConfig->spat_num_elems, &mn, &mx, &ab), "spat_reduce");
	}

	void write_csv(const char *name, const char *title, int prec, bool shifted) const
	{
		spat_download();
		std::ofstream file(name, std::ios_base::trunc);
		file << std::setprecision(prec);
		file << "#TI " << title << std::endl;
		file << "#TX Longitude" << std::endl;
		file << "#TY Latitude" << std::endl;
		file << "0\t";
		const int nlon = sphConfig->spat_num_lon, nlat = sphConfig->spat_num_lat;
		for (int i = 0; i < nlon; i++) {
			double lon = ((double)i / (double)nlon) * 2.0 * M_PI;
			lon = shifted ? (lon - M_PI) / M_PI * 180.0 : lon / M_PI * 180.0;
			file << lon;
			if (i < nlon - 1) file << "\t";
		}
		file << std::endl;
		for (int j = nlat - 1; j >= 0; j--) {
			file << sphConfig->lat[j] / M_PI * 180.0 << "\t";
			for (int i = 0; i < nlon; i++) {
				int ia = i;
				if (shifted) {
					ia = i + nlon / 2;
					if (ia >= nlon) ia -= nlon;
				}
				file << data_spat[ia * nlat + j];
				if (i < nlon - 1) file << "\t";
			}
			file << std::endl;
		}
	}
};

inline SPHData operator*(const double i_value, const SPHData &i_array_data) { return i_array_data * i_value; }
inline SPHData operator-(const double i_value, const SPHData &i_array_data) { return -i_array_data + i_value; }
inline SPHData operator+(const double i_value, const SPHData &i_array_data) { return i_array_data + i_value; }

#endif
