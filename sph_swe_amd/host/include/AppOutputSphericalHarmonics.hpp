// Program P5 (src/AppOutputSphericalHarmonics.hpp): every spherical harmonic
// of the configuration, one unit spectral coefficient at a time, synthesised
// on the GPU and written to SPH_n<n>_m<m>.csv in the grid CSV format of
// SPHData::spat_write_file.
#ifndef SPHSWE_HOST_APPOUTPUTSPHERICALHARMONICS_HPP
#define SPHSWE_HOST_APPOUTPUTSPHERICALHARMONICS_HPP

#include <algorithm>
#include <cstdio>
#include <iostream>

#include "sph/SPHConfig.hpp"
#include "sph/SPHData.hpp"

class AppOutputSphericalHarmonics
{
public:
	void run(SPHConfig *sphConfig)
	{
		SPHData h(sphConfig);
		int written = 0;
		for (int n = 0; n <= sphConfig->spec_n_max; n++)
			for (int m = 0; m <= std::min(sphConfig->spec_m_max, n); m++) {
				h.spec_update_lambda([&](int i_n, int i_m, std::complex<double> &o) {
					o = (i_n == n && i_m == m) ? 1.0 : 0.0;
				});
				char name[64];
				std::snprintf(name, sizeof(name), "SPH_n%i_m%i.csv", n, m);
				h.spat_write_file(name);
				written++;
			}
		if (written != sphConfig->spec_num_elems) {
			std::cerr << "P5: wrote " << written << " modes, expected " << sphConfig->spec_num_elems << std::endl;
			exit(1);
		}
		std::cout << "P5: wrote " << written << " spherical harmonics" << std::endl;
	}
};

#endif
