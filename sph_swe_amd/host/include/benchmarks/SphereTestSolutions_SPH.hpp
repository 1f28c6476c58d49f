// One spherical harmonic as a test field of the self-test programs
// (src/include/benchmarks/SphereTestSolutions_SPH.hpp): f = Pbar_n^m(mu)
// cos(m lon) / sqrt(2 pi), Pbar normalised to unit L2 norm on [-1, 1] with the
// Condon-Shortley phase, and (1 - mu^2) df/dmu.  The reference lists the closed
// forms for n, m <= 3 (its n = m = 0 branch is unreachable: it tests n = 1
// twice); here any (n, m) is evaluated by the three-term recurrence and
//   (1 - mu^2) dP_n^m/dmu = (n + m) P_{n-1}^m - n mu P_n^m.
#ifndef SPHSWE_HOST_SPHERETESTSOLUTIONS_SPH_HPP
#define SPHSWE_HOST_SPHERETESTSOLUTIONS_SPH_HPP

#include <cmath>

class SphereTestSolutions_SPH
{
	int tn, tm;

	// unnormalised P_n^m(mu) with the Condon-Shortley phase, n >= m (0 for n < m)
	static double plm(int n, int m, double mu)
	{
		if (n < m) return 0.0;
		double pmm = 1.0;
		const double s = std::sqrt(1.0 - mu * mu);
		for (int k = 1; k <= m; k++) pmm *= -(2.0 * k - 1.0) * s;
		if (n == m) return pmm;
		double p0 = pmm, p1 = mu * (2.0 * m + 1.0) * pmm;
		for (int k = m + 2; k <= n; k++) {
			const double p2 = ((2.0 * k - 1.0) * mu * p1 - (k + m - 1.0) * p0) / (k - m);
			p0 = p1;
			p1 = p2;
		}
		return p1;
	}
	double norm() const
	{
		double r = (2.0 * tn + 1.0) / 2.0;
		for (int k = tn - tm + 1; k <= tn + tm; k++) r /= k;
		return std::sqrt(r) / std::sqrt(2.0 * M_PI);
	}

public:
	SphereTestSolutions_SPH(int i_n, int i_m) : tn(i_n), tm(i_m) {}

	void test_function__grid_gaussian(double lon, double mu, double &o_data) const
	{
		o_data = norm() * plm(tn, tm, mu) * std::cos(tm * lon);
	}
	// (1 - mu^2) d/dmu of the test function
	void correct_result_diff_mu__grid_gaussian(double lon, double mu, double &o_data) const
	{
		o_data = norm() * ((tn + tm) * plm(tn - 1, tm, mu) - tn * mu * plm(tn, tm, mu)) * std::cos(tm * lon);
	}
};

#endif
