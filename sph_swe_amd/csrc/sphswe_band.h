// Host-side REXI setup: pole coefficients, per-m band assembly of the REXI
// operators and banded LU with partial pivoting (the algorithm of LAPACK's
// zgbtf2, used by zgbsv in the reference).
#pragma once

#include <complex>
#include <vector>

namespace sphswe_host {

typedef std::complex<double> cplx;

// REXI poles alpha_n and weights (REXI::setup, src/include/rexi/REXI.hpp:52-131):
// cos x ~ sum Re(beta_re_n / (ix + alpha_n)), sin x ~ sum Re(beta_im_n / (ix + alpha_n)).
// beta_im is filled when non-null.
void rexi_coefficients(double h, int M, bool half, std::vector<cplx> &alpha, std::vector<cplx> &beta_re,
		       std::vector<cplx> *beta_im = nullptr);

// Row-band storage of one SPHMatrixComplex: rows in the NCompact order
// (SPHConfig.hpp:136-156), num_diagonals = 2*halo+1 entries per row.
struct BandRows {
	int L, halo, nd;
	std::vector<cplx> data;
	BandRows(int L_, int halo_) : L(L_), halo(halo_), nd(2 * halo_ + 1), data((size_t)(L_ + 1) * (L_ + 1) * nd) {}
	cplx *row(int n, int m);
};

// component ids follow SPHSolverComplex: 0 scalar, 1 mu, 2 (1-mu^2) d/dmu,
// 11..18 rexi z1..z8 (SPHSolverComplex.hpp:88-420)
void band_add(BandRows &B, int comp, cplx scalar, double r);

// number of off-diagonals a component touches (0 for an unknown id)
int band_component_width(int comp);

// NCompact offset of (n = |m|, m): first row of the m block
int ncompact(int L, int n, int m);

// Factor the m-block of B (size N = L+1-|m|) with partial pivoting.  Output per
// row j (0-based in the block): inv_diag, U(j-1,j) .. U(j-kv,j), L(j+1,j) .. L(j+kl,j)
// (kv = 2*halo, kl = halo), and the pivot offset ipiv[j]-j in [0, kl].
// Returns 0, or j+1 if U(j,j) is exactly zero (zgbsv INFO).
int band_factor_block(const BandRows &B, int m, std::vector<cplx> &fac, std::vector<unsigned char> &piv);

}  // namespace sphswe_host
