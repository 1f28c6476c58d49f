// Host-side REXI setup (see sphswe_band.h).
#include "sphswe_band.h"
#include "sphswe_common.h"

#include <cmath>

namespace sphswe_host {

// Gaussian approximation of exp(-x^2/(4h^2))/sqrt(4 pi) by 2L+1 = 23 rational
// terms, L = 11: mean mu and weights a_l (Haut et al., the published table the
// reference uses at src/include/rexi/GaussianApproximation.hpp:51-81).
static const double kGaussMu = -4.315321510875024;
static const double kGaussA[23][2] = {
	{-1.0845749544592896e-7, 2.77075431662228e-8},
	{1.858753344202957e-8, -9.105375434750162e-7},
	{3.6743713227243024e-6, 7.073284346322969e-7},
	{-2.7990058083347696e-6, 0.0000112564827639346},
	{0.000014918577548849352, -0.0000316278486761932},
	{-0.0010751767283285608, -0.00047282220513073084},
	{0.003816465653840016, 0.017839810396560574},
	{0.12124105653274578, -0.12327042473830248},
	{-0.9774980792734348, -0.1877130220537587},
	{1.3432866123333178, 3.2034715228495942},
	{4.072408546157305, -6.123755543580666},
	{-9.442699917778205, 0.},
	{4.072408620272648, 6.123755841848161},
	{1.3432860877712938, -3.2034712658530275},
	{-0.9774985292598916, 0.18771238018072134},
	{0.1212417070363373, 0.12326987628935386},
	{0.0038169724770333343, -0.017839242222443888},
	{-0.0010756025812659208, 0.0004731874917343858},
	{0.000014713754789095218, 0.000031358475831136815},
	{-2.659323898804944e-6, -0.000011341571201752273},
	{3.6970377676364553e-6, -6.517457477594937e-7},
	{3.883933649142257e-9, 9.128496023863376e-7},
	{-1.0816457995911385e-7, -2.954309729192276e-8},
};

void rexi_coefficients(double h, int M, bool half, std::vector<cplx> &alpha, std::vector<cplx> &beta,
		       std::vector<cplx> *beta_im)
{
	const int L = 11;
	const int N = M + L;
	alpha.assign(2 * N + 1, cplx(0, 0));
	beta.assign(2 * N + 1, cplx(0, 0));
	if (beta_im) beta_im->assign(2 * N + 1, cplx(0, 0));
	// b_m = e^{h^2} e^{-i m h} (ExponentialApproximation.hpp:59-123); the weights of pole
	// n = l + m collect h Re(b_m) a_l and h Im(b_m) a_l
	for (int l = -L; l <= L; l++)
		for (int m = -M; m <= M; m++) {
			const int n = l + m;
			alpha[n + N] = h * (kGaussMu + cplx(0, n));
			const cplx bm = std::exp(h * h) * std::exp(-cplx(0, 1) * ((double)m * h));
			const cplx al(kGaussA[l + L][0], kGaussA[l + L][1]);
			beta[n + N] += bm.real() * h * al;
			if (beta_im) (*beta_im)[n + N] += bm.imag() * h * al;
		}
	if (half) {
		// the conjugate half of the poles contributes the complex conjugate:
		// keep n <= 0 and double all but the centre pole (REXI.hpp:114-130)
		alpha.resize(N + 1);
		beta.resize(N + 1);
		for (int i = 0; i < N; i++) beta[i] *= 2.0;
		if (beta_im) {
			beta_im->resize(N + 1);
			for (int i = 0; i < N; i++) (*beta_im)[i] *= 2.0;
		}
	}
}

int ncompact(int L, int n, int m)
{
	if (m < 0) {
		const int minc = L + m;
		return (((minc - 1) * minc) >> 1) + minc + (n + m);
	}
	return ((L * (L + 1)) >> 1) + ((m * (2 * L - m + 1)) >> 1) + n;
}

cplx *BandRows::row(int n, int m) { return data.data() + (size_t)ncompact(L, n, m) * nd; }

// rowElement_add (SPHMatrixComplex.hpp:181-206): the "n < m" tests use m, not |m|
static inline void row_add(BandRows &B, int n, int m, int rel, cplx v)
{
	if (n < m) return;
	const int nn = n + rel;
	if (nn < 0 || nn < m || nn > B.L) return;
	B.row(n, m)[rel + B.halo] += v;
}

void band_add(BandRows &B, int comp, cplx s, double r)
{
	const double ir2 = 1.0 / (r * r);
	const cplx I(0, 1);
	for (int m = -B.L; m <= B.L; m++)
		for (int n = std::abs(m); n <= B.L; n++) {
			const double dn = n, dm = m;
			switch (comp) {
			case 0: case 11: row_add(B, n, m, 0, s); break;
			case 1:
				row_add(B, n, m, -1, idR(dn - 1, dm) * s);
				row_add(B, n, m, +1, idS(dn + 1, dm) * s);
				break;
			case 2:
				row_add(B, n, m, -1, (-dn + 1.0) * idR(dn - 1, dm));
				row_add(B, n, m, +1, (dn + 2.0) * idS(dn + 1, dm));
				break;
			case 12:   // z2: mu^2
				row_add(B, n, m, -2, s * idA(dn - 2, dm));
				row_add(B, n, m, 0, s * idB(dn, dm));
				row_add(B, n, m, +2, s * idC(dn + 2, dm));
				break;
			case 13:   // z3: mu^4
				row_add(B, n, m, -4, s * (idA(dn - 2, dm) * idA(dn - 4.0, dm)));
				row_add(B, n, m, -2, s * (idA(dn - 2, dm) * idB(dn - 2, dm) + idB(dn, dm) * idA(dn - 2, dm)));
				row_add(B, n, m, 0, s * (idA(dn - 2, dm) * idC(dn, dm) + idB(dn, dm) * idB(dn, dm) + idC(dn + 2, dm) * idA(dn, dm)));
				row_add(B, n, m, +2, s * (idB(dn, dm) * idC(dn + 2, dm) + idC(dn + 2, dm) * idB(dn + 2, dm)));
				row_add(B, n, m, +4, s * (idC(dn + 2, dm) * idC(dn + 4, dm)));
				break;
			case 14:   // z4: grad_j(mu) grad_i(phi)
				row_add(B, n, m, 0, ir2 * s * (I * dm));
				break;
			case 15: { // z5: grad_j(mu) mu^2 grad_i(phi)
				const cplx fac = ir2 * s * (I * dm);
				row_add(B, n, m, -2, fac * idA(dn - 2, dm));
				row_add(B, n, m, 0, fac * idB(dn + 0, dm));
				row_add(B, n, m, +2, fac * idC(dn + 2, dm));
				break;
			}
			case 16:   // z6: grad_j(mu) mu grad_j(phi)
				row_add(B, n, m, -2, -ir2 * s * (idD(dn - 1, dm) * idR(dn - 2, dm) + (idE(dn, dm) - 3.0) * idA(dn - 2, dm)));
				row_add(B, n, m, 0, -ir2 * s * (1.0 + idD(dn - 1, dm) * idS(dn, dm) + (idE(dn, dm) - 3.0) * idB(dn, dm)));
				row_add(B, n, m, +2, -ir2 * s * (idE(dn, dm) - 3.0) * idC(dn + 2, dm));
				break;
			case 17:   // z7: laplace
				row_add(B, n, m, 0, -ir2 * s * dn * (dn + 1.0));
				break;
			case 18: { // z8: mu^2 laplace
				row_add(B, n, m, 0, ir2 * s * 2.0);
				const double fac = -ir2 * (dm * dm - dn * (dn + 1.0));
				row_add(B, n, m, -2, s * fac * idA(dn - 2, dm));
				row_add(B, n, m, 0, s * fac * idB(dn, dm));
				row_add(B, n, m, +2, s * fac * idC(dn + 2, dm));
				break;
			}
			// Robert-function components (SPHSolverComplex.hpp:241-370)
			case 19: { // z4robert: grad_j(mu) grad_i(phi) = (1 - mu^2) i m phi / r^2
				const cplx fac = ir2 * s * (I * dm);
				row_add(B, n, m, -2, -fac * idA(dn - 2, dm));
				row_add(B, n, m, 0, -fac * idB(dn, dm) + fac);
				row_add(B, n, m, +2, -fac * idC(dn + 2, dm));
				break;
			}
			case 20: { // z5robert: the five entries of (mu^2 - mu^4) i m / r^2, all on
				   // the -2 diagonal as the reference adds them (:305-309)
				const cplx fac = ir2 * s * (I * dm);
				row_add(B, n, m, -2, -fac * (idA(dn - 2, dm) * idA(dn - 4, dm)));
				row_add(B, n, m, -2, -fac * (idA(dn - 2, dm) * idB(dn - 2, dm) + idB(dn, dm) * idA(dn - 2, dm)) +
							     fac * idA(dn - 2, dm));
				row_add(B, n, m, -2, -fac * (idA(dn - 2, dm) * idC(dn, dm) + idB(dn, dm) * idB(dn, dm) +
							      idC(dn + 2, dm) * idA(dn, dm)) + fac * idB(dn, dm));
				row_add(B, n, m, -2, -fac * (idB(dn, dm) * idC(dn + 2, dm) + idC(dn + 2, dm) * idB(dn + 2, dm)) +
							     fac * idC(dn + 2, dm));
				row_add(B, n, m, -2, -fac * (idC(dn + 2, dm) * idC(dn + 4, dm)));
				break;
			}
			case 21: { // z6robert: grad_j(mu) mu grad_j(phi), times m / r^2
				const cplx s6 = ir2 * s * dm;
				const double gr = idG(dn - 1, dm) * idR(dn - 2, dm);
				const double gs = idG(dn - 1, dm) * idS(dn, dm) + idH(dn + 1, dm) * idR(dn, dm);
				const double hs = idH(dn + 1, dm) * idS(dn + 2, dm);
				row_add(B, n, m, -2, s6 * gr);
				row_add(B, n, m, 0, s6 * gs);
				row_add(B, n, m, +2, s6 * hs);
				row_add(B, n, m, -4, s6 * (-idA(dn - 4, dm) * gr));
				row_add(B, n, m, -2, s6 * (-idB(dn - 2, dm) * gr));
				row_add(B, n, m, 0, s6 * (-idC(dn, dm) * gr));
				row_add(B, n, m, -2, s6 * (-idA(dn - 2, dm) * gs));
				row_add(B, n, m, 0, s6 * (-idB(dn, dm) * gs));
				row_add(B, n, m, +2, s6 * (-idC(dn + 2, dm) * gs));
				row_add(B, n, m, 0, s6 * (-idA(dn, dm) * hs));
				row_add(B, n, m, +2, s6 * (-idB(dn + 2, dm) * hs));
				row_add(B, n, m, +4, s6 * (-idC(dn + 4, dm) * hs));
				break;
			}
			default: break;
			}
		}
}

int band_component_width(int comp)
{
	switch (comp) {
	case 0: case 11: case 14: case 17: return 0;
	case 1: case 2: return 1;
	case 12: case 15: case 16: case 18: case 19: case 20: return 2;
	case 13: case 21: return 4;
	default: return -1;
	}
}

int band_factor_block(const BandRows &B, int m, std::vector<cplx> &fac, std::vector<unsigned char> &piv)
{
	const int h = B.halo, nd = B.nd;
	const int kl = h, ku = h, kv = kl + ku;
	const int ldab = 2 * kl + ku + 1;
	const int N = B.L + 1 - std::abs(m);
	const cplx *A = B.data.data() + (size_t)ncompact(B.L, std::abs(m), m) * nd;
	std::vector<cplx> AB((size_t)ldab * N, cplx(0, 0));
#define ABX(i, j) AB[(size_t)(kv + (i) - (j)) + (size_t)(j) * ldab]
	// row-band -> column band (BandedMatrixSolver.hpp:344-368): entry i of row si
	// is A(si, si + i - h); entries outside the block are dropped
	for (int si = 0; si < N; si++)
		for (int i = 0; i < nd; i++) {
			const int sj = si + i - h;
			if (sj < 0 || sj >= N) continue;
			ABX(si, sj) = A[(size_t)si * nd + i];
		}
	int info = 0, ju = 0;
	std::vector<int> ipiv(N);
	for (int j = 0; j < N; j++) {
		const int km = std::min(kl, N - 1 - j);
		int jp = 0;
		double amax = -1;
		for (int i = 0; i <= km; i++) {
			const cplx v = ABX(j + i, j);
			const double a = std::fabs(v.real()) + std::fabs(v.imag());   // izamax uses |re|+|im|
			if (a > amax) { amax = a; jp = i; }
		}
		ipiv[j] = j + jp;
		if (ABX(j + jp, j) != cplx(0, 0)) {
			ju = std::max(ju, std::min(j + ku + jp, N - 1));
			if (jp != 0)
				for (int jj = j; jj <= ju; jj++) std::swap(ABX(j + jp, jj), ABX(j, jj));
			if (km > 0) {
				const cplx rp = 1.0 / ABX(j, j);
				for (int i = 1; i <= km; i++) ABX(j + i, j) *= rp;
				for (int jj = j + 1; jj <= ju; jj++) {
					const cplx t = ABX(j, jj);
					if (t != cplx(0, 0))
						for (int i = 1; i <= km; i++) ABX(j + i, jj) -= ABX(j + i, j) * t;
				}
			}
		} else if (info == 0) {
			info = j + 1;
		}
	}
	const int NE = 1 + kv + kl;
	fac.assign((size_t)N * NE, cplx(0, 0));
	piv.assign(N, 0);
	for (int j = 0; j < N; j++) {
		cplx *f = fac.data() + (size_t)j * NE;
		const cplx d = ABX(j, j);
		f[0] = (d != cplx(0, 0)) ? 1.0 / d : cplx(0, 0);
		for (int k = 1; k <= kv; k++) f[k] = (j - k >= 0) ? ABX(j - k, j) : cplx(0, 0);
		for (int k = 1; k <= kl; k++) f[kv + k] = (j + k < N) ? ABX(j + k, j) : cplx(0, 0);
		piv[j] = (unsigned char)(ipiv[j] - j);
	}
#undef ABX
	return info;
}

}  // namespace sphswe_host
