// SPHConfig over the C ABI: one GPU context per configuration.
// Mirrors src/include/sph/SPHConfig.hpp (fields :20-105, index maps :106-156,
// setup :324-352, setupAutoPhysicalSpace :366-399, setupAdditionalModes
// :403-417, shutdown :421-446).
#ifndef SPHSWE_HOST_SPHCONFIG_HPP
#define SPHSWE_HOST_SPHCONFIG_HPP

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "sphswe.h"

class SPHConfig
{
public:
	sphswe_ctx *gpu = nullptr;     // context on device `device`
	int device = 0;

	int spat_num_lon = -1;
	int spat_num_lat = -1;
	int spat_num_elems = -1;
	int spec_n_max = -1;
	int spec_m_max = -1;
	int spec_num_elems = -1;        // (L+1)(L+2)/2
	int cplx_spec_num_elems = -1;   // (L+1)^2

	// latitude phi of each ring, mu = sin(phi) (shtns->ct), cos(phi) (shtns->st),
	// Gauss weights; ring 0 is the northernmost
	std::vector<double> lat, lat_gaussian, lat_cogaussian, gauss_weights;

	SPHConfig() {}
	SPHConfig(const SPHConfig &) = delete;
	SPHConfig &operator=(const SPHConfig &) = delete;
	~SPHConfig() { shutdown(); }

	std::size_t getArrayIndexByModes(int n, int m) const { return (m * (2 * spec_n_max - m + 1) >> 1) + n; }
	std::size_t getArrayIndexByModes_Complex(int n, int m) const { return n * n + (m + n); }
	std::size_t getArrayIndexByModes_Complex_NCompact(int n, int m) const
	{
		if (m < 0) {
			const int minc = spec_m_max + m;
			return (((minc - 1) * minc) >> 1) + minc + (n + m);
		}
		return ((spec_n_max * (spec_n_max + 1)) >> 1) + (m * (2 * spec_n_max - m + 1) >> 1) + n;
	}

	// argument order of the reference: (mmax, nmax, nphi, nlat)
	void setup(int mmax, int nmax, int nphi, int nlat)
	{
		shutdown();
		const int rc = sphswe_create(nmax, mmax, nlat, nphi, device, &gpu);
		if (rc != SPHSWE_OK) {
			std::cerr << "SPHConfig::setup: " << (gpu ? sphswe_last_error(gpu) : "context creation failed") << std::endl;
			if (gpu) sphswe_destroy(gpu);
			gpu = nullptr;
			exit(1);
		}
		spec_m_max = mmax;
		spec_n_max = nmax;
		spat_num_lon = nphi;
		spat_num_lat = nlat;
		spat_num_elems = nphi * nlat;
		spec_num_elems = (nmax + 1) * (nmax + 2) / 2;
		cplx_spec_num_elems = (nmax + 1) * (nmax + 1);
		lat_gaussian.assign(nlat, 0.0);
		lat_cogaussian.assign(nlat, 0.0);
		gauss_weights.assign(nlat, 0.0);
		sphswe_get_grid(gpu, lat_gaussian.data(), lat_cogaussian.data(), gauss_weights.data());
		lat.assign(nlat, 0.0);
		for (int j = 0; j < nlat; j++) lat[j] = M_PI * 0.5 - std::acos(lat_gaussian[j]);   // SPHConfig.hpp:303-319
	}

	// a configuration with more modes than i (SPHConfig.hpp:403-417): the
	// automatic grid of the larger truncation; this configuration must be empty
	void setupAdditionalModes(SPHConfig *i, int i_additional_modes_longitude, int i_additional_modes_latitude)
	{
		if (gpu) {
			std::cerr << "SPHConfig::setupAdditionalModes: configuration already set up" << std::endl;
			exit(1);
		}
		device = i->device;
		setupAutoPhysicalSpace(i->spec_m_max + i_additional_modes_longitude, i->spec_n_max + i_additional_modes_latitude,
				       &spat_num_lon, &spat_num_lat);
	}

	void setupAutoPhysicalSpace(int mmax, int nmax, int *nphi, int *nlat)
	{
		int la = 0, lo = 0;
		if (sphswe_grid_auto(nmax, 2, &la, &lo) != SPHSWE_OK) {
			std::cerr << "SPHConfig::setupAutoPhysicalSpace: no grid for T" << nmax << std::endl;
			exit(1);
		}
		setup(mmax, nmax, lo, la);
		*nphi = lo;
		*nlat = la;
	}

	void shutdown()
	{
		if (gpu) sphswe_destroy(gpu);
		gpu = nullptr;
	}

	// checked call: the reference exits on unrecoverable errors
	void check(int rc, const char *what) const
	{
		if (rc != SPHSWE_OK) {
			std::cerr << what << ": " << sphswe_last_error(gpu) << std::endl;
			exit(1);
		}
	}
};

#endif
