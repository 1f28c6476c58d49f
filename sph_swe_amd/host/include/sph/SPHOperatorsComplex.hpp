// SPHOperatorsComplex over the C ABI (src/include/sph/SPHOperatorsComplex.hpp:
// diff_lon :27-60, diff_lat_mu :62-77, diff_lat_phi :87-92,
// spec_one_minus_mu_squared_diff_lat_mu :119-153, mu :155-187,
// mu2 :189-222, grad_lon :100-117, grad_lat :224-282, div_lon :284-298,
// div_lat :300-336 (division by cos(lat[j])), laplace :454-478, vort :480-494,
// div :513-528).
#ifndef SPHSWE_HOST_SPHOPERATORSCOMPLEX_HPP
#define SPHSWE_HOST_SPHOPERATORSCOMPLEX_HPP

#include "SPHDataComplex.hpp"

class SPHOperatorsComplex
{
public:
	SPHDataComplex diff_lon(const SPHDataComplex &i) const { return spec(i, SPHSWE_OP_DIFF_LON); }
	SPHDataComplex spec_one_minus_mu_squared_diff_lat_mu(const SPHDataComplex &i) const
	{
		return spec(i, SPHSWE_OP_ONE_MINUS_MU2_DMU);
	}
	SPHDataComplex mu(const SPHDataComplex &i) const { return spec(i, SPHSWE_OP_MU); }
	SPHDataComplex mu2(const SPHDataComplex &i) const { return spec(i, SPHSWE_OP_MU2); }
	SPHDataComplex laplace(const SPHDataComplex &i) const { return spec(i, SPHSWE_OP_LAPLACE); }

	SPHDataComplex grad_lon(const SPHDataComplex &i) const { return metric(diff_lon(i), SPHSWE_SPAT_DIV_COSPHI); }
	SPHDataComplex grad_lat(const SPHDataComplex &i) const
	{
		return metric(spec_one_minus_mu_squared_diff_lat_mu(i), SPHSWE_SPAT_DIV_COSPHI);
	}
	// d/dmu: (1 - mu^2) d/dmu, then division by 1 - mu^2
	SPHDataComplex diff_lat_mu(const SPHDataComplex &i) const
	{
		return metric(spec_one_minus_mu_squared_diff_lat_mu(i), SPHSWE_SPAT_DIV_ST2);
	}
	SPHDataComplex diff_lat_phi(const SPHDataComplex &i) const { return grad_lat(i); }
	SPHDataComplex div_lon(const SPHDataComplex &i) const { return grad_lon(i); }
	SPHDataComplex div_lat(const SPHDataComplex &i) const
	{
		SPHDataComplex o = metric(i, SPHSWE_SPAT_MUL_ST);
		return metric(grad_lat(o), SPHSWE_SPAT_DIV_COSLAT);
	}
	SPHDataComplex vort(const SPHDataComplex &u, const SPHDataComplex &v) const { return grad_lon(v) - grad_lat(u); }
	SPHDataComplex div(const SPHDataComplex &u, const SPHDataComplex &v) const { return div_lon(u) + div_lat(v); }

private:
	static SPHDataComplex spec(const SPHDataComplex &i, int op)
	{
		i.request_data_spectral();
		SPHDataComplex o(i.sphConfig);
		i.sphConfig->check(sphswe_spec_op_cplx(i.gpu(), op, i.d_spec, o.d_spec, 1, 1.0, 0.0, 0.0, 1.0, 0.0),
				   "SPHOperatorsComplex");
		o.data_spec_valid = true;
		o.data_spat_valid = false;
		return o;
	}
	static SPHDataComplex metric(const SPHDataComplex &i, int op)
	{
		i.request_data_physical();
		SPHDataComplex o(i.sphConfig);
		i.sphConfig->check(sphswe_spat_op_cplx(i.gpu(), op, i.d_spat, nullptr, o.d_spat, 1.0, 1), "SPHOperatorsComplex");
		o.data_spat_valid = true;
		o.data_spec_valid = false;
		return o;
	}
};

#endif
