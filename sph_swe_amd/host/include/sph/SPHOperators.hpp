// SPHOperators over the C ABI (src/include/sph/SPHOperators.hpp): spectral band
// operators (diff_lon :30-54, diff_lat_mu :65-78, diff_lat_phi :86-92,
// spec_one_minus_mu_squared_diff_lat_mu :234-260,
// mu :268-297, mu2 :303-333, laplace :482-498) and the composite operators with
// their projection points (grad_lon :100-119, grad_lat :339-388, div_lon
// :397-402, div_lat :411-439, vort :507-513, div :523-529) and the
// Robert-function operators (robert_div_lon :129-145, robert_div_lat :155-172,
// robert_grad_lon :186-193, robert_grad_lat :203-210).
#ifndef SPHSWE_HOST_SPHOPERATORS_HPP
#define SPHSWE_HOST_SPHOPERATORS_HPP

#include "SPHData.hpp"

class SPHOperators
{
public:
	SPHData diff_lon(const SPHData &i) const { return spec(i, SPHSWE_OP_DIFF_LON); }
	SPHData spec_one_minus_mu_squared_diff_lat_mu(const SPHData &i) const { return spec(i, SPHSWE_OP_ONE_MINUS_MU2_DMU); }
	SPHData mu(const SPHData &i) const { return spec(i, SPHSWE_OP_MU); }
	SPHData mu2(const SPHData &i) const { return spec(i, SPHSWE_OP_MU2); }
	SPHData laplace(const SPHData &i) const { return spec(i, SPHSWE_OP_LAPLACE); }

	// d/dlambda / cos(phi): spectral i m, then division by sqrt(1 - mu^2) in physical space
	SPHData grad_lon(const SPHData &i) const { return metric(diff_lon(i), SPHSWE_SPAT_DIV_COSPHI); }
	// d/dphi: (1 - mu^2) d/dmu, then division by sqrt(1 - mu^2)
	SPHData grad_lat(const SPHData &i) const
	{
		return metric(spec_one_minus_mu_squared_diff_lat_mu(i), SPHSWE_SPAT_DIV_COSPHI);
	}
	SPHData div_lon(const SPHData &i) const { return grad_lon(i); }
	// (1 / cos phi) d(cos phi x)/dphi: times st, grad_lat, divided by st
	SPHData div_lat(const SPHData &i) const
	{
		SPHData o = metric(i, SPHSWE_SPAT_MUL_ST);
		return metric(grad_lat(o), SPHSWE_SPAT_DIV_ST);
	}
	// d/dmu: (1 - mu^2) d/dmu, then division by 1 - mu^2 = cos^2 phi (:65-78)
	SPHData diff_lat_mu(const SPHData &i) const
	{
		return metric(spec_one_minus_mu_squared_diff_lat_mu(i), SPHSWE_SPAT_DIV_ST2);
	}
	// d/dphi (:86-92)
	SPHData diff_lat_phi(const SPHData &i) const { return grad_lat(i); }
	// grad_lon(v) - grad_lat(u), without the metric term of the true vorticity
	SPHData vort(const SPHData &u, const SPHData &v) const { return grad_lon(v) - grad_lat(u); }
	SPHData div(const SPHData &u, const SPHData &v) const { return div_lon(u) + div_lat(v); }

	// Robert functions U = u cos phi, V = v cos phi: the divergences divide by
	// cos^2 phi in physical space (the cos-phi grid is the cogaussian st), the
	// gradients stay spectral
	SPHData robert_div_lon(const SPHData &i) const { return metric(diff_lon(i), SPHSWE_SPAT_DIV_ST2); }
	SPHData robert_div_lat(const SPHData &i) const
	{
		return metric(spec_one_minus_mu_squared_diff_lat_mu(i), SPHSWE_SPAT_DIV_ST2);
	}
	SPHData robert_grad_lon(const SPHData &i) const { return diff_lon(i); }
	SPHData robert_grad_lat(const SPHData &i) const { return spec_one_minus_mu_squared_diff_lat_mu(i); }

private:
	static SPHData spec(const SPHData &i, int op)
	{
		i.request_data_spectral();
		SPHData o(i.sphConfig);
		i.sphConfig->check(sphswe_spec_op(i.gpu(), op, i.d_spec, o.d_spec, 1, 1.0, 0.0, 0.0, 1.0), "SPHOperators");
		o.data_spec_valid = true;
		o.data_spat_valid = false;
		return o;
	}
	static SPHData metric(const SPHData &i, int op)
	{
		i.request_data_physical();
		SPHData o(i.sphConfig);
		i.sphConfig->check(sphswe_spat_op(i.gpu(), op, i.d_spat, nullptr, o.d_spat, 1.0, 1), "SPHOperators");
		o.data_spat_valid = true;
		o.data_spec_valid = false;
		return o;
	}
};

#endif
