// Banded complex solver: SPHSolverComplex<T> (src/include/sph/SPHSolverComplex.hpp:50-511)
// with its SPHMatrixComplex row-band storage and the per-m zgbsv solves of
// BandedMatrixSolver (BandedMatrixSolver.hpp:235-434).
//
// Assembly and the pivoted LU factorisation run on the host (they depend only
// on the matrix); the factors are uploaded once and every solve is one kernel
// with one thread per m (forward substitution with the row interchanges, then
// back substitution), reading and writing complex spectra in the l^2+l+m layout.
#include "sphswe_band.h"
#include "sphswe_common.h"

#include <cstdlib>

using sphswe_host::cplx;

struct sphswe_band {
	sphswe_ctx *ctx = nullptr;
	sphswe_host::BandRows rows;
	bool dirty = true;
	int halo = 0;
	double2 *d_fac = nullptr;           // [row][1 + 3 halo]
	unsigned char *d_piv = nullptr;     // [row]
	int *d_moff = nullptr;              // first row of each m block, index m + L
	explicit sphswe_band(int L, int h) : rows(L, h), halo(h) {}
};

namespace {

__device__ inline double2 bmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 bsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

template <int KL>
__global__ __launch_bounds__(64) void k_band_solve(const double2 *__restrict__ fac, const unsigned char *__restrict__ piv,
						    const int *__restrict__ moff, const double2 *__restrict__ rhs,
						    double2 *__restrict__ x, double2 *__restrict__ y, int L)
{
	constexpr int KV = 2 * KL, NE = 1 + KV + KL;
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t > 2 * L) return;
	const int m = t - L, am = abs(m);
	const int N = L + 1 - am;
	const size_t row0 = moff[m + L];
	auto at = [&](int j) -> double2 {
		if (j >= N) return make_double2(0, 0);
		const int n = am + j;
		return rhs[(size_t)n * n + n + m];
	};
	double2 w[KL + 1];
#pragma unroll
	for (int i = 0; i <= KL; i++) w[i] = at(i);
	for (int j = 0; j < N; j++) {
		const double2 *f = fac + (row0 + j) * NE;
		const int k = piv[row0 + j];
		// interchange rows j and j + k of the right-hand side, as pairwise
		// conditional swaps (a select chain reading w[k] becomes an indexed
		// access and moves the window to scratch memory)
#pragma unroll
		for (int i = 1; i <= KL; i++) {
			const bool s = (k == i);
			const double2 a = w[0], b = w[i];
			w[0] = make_double2(s ? b.x : a.x, s ? b.y : a.y);
			w[i] = make_double2(s ? a.x : b.x, s ? a.y : b.y);
		}
#pragma unroll
		for (int i = 1; i <= KL; i++) w[i] = bsub(w[i], bmul(f[KV + i], w[0]));
		y[row0 + j] = w[0];
#pragma unroll
		for (int i = 0; i < KL; i++) w[i] = w[i + 1];
		w[KL] = at(j + KL + 1);
	}
	double2 v[KV + 1];
#pragma unroll
	for (int i = 0; i <= KV; i++) v[i] = (N - 1 - i >= 0) ? y[row0 + N - 1 - i] : make_double2(0, 0);
	for (int j = N - 1; j >= 0; j--) {
		const double2 *f = fac + (row0 + j) * NE;
		const double2 xj = bmul(v[0], f[0]);
		const int n = am + j;
		x[(size_t)n * n + n + m] = xj;
#pragma unroll
		for (int i = 1; i <= KV; i++) v[i] = bsub(v[i], bmul(f[i], xj));
#pragma unroll
		for (int i = 0; i < KV; i++) v[i] = v[i + 1];
		v[KV] = (j - KV - 1 >= 0) ? y[row0 + j - KV - 1] : make_double2(0, 0);
	}
}

template <int KL>
void launch_band(sphswe_ctx *ctx, sphswe_band *b, const double2 *rhs, double2 *x, double2 *y)
{
	const int nth = 2 * ctx->L + 1;
	k_band_solve<KL><<<(nth + 63) / 64, 64, 0, ctx->stream>>>(b->d_fac, b->d_piv, b->d_moff, rhs, x, y, ctx->L);
}

int factorise(sphswe_ctx *ctx, sphswe_band *b)
{
	using namespace sphswe_host;
	const int L = ctx->L, h = b->halo, NE = 1 + 3 * h;
	const size_t rows = (size_t)(L + 1) * (L + 1);
	std::vector<double2> hf(rows * NE, make_double2(0, 0));
	std::vector<unsigned char> hp(rows, 0);
	std::vector<int> moff(2 * L + 1);
	std::vector<cplx> fac;
	std::vector<unsigned char> piv;
	for (int m = -L; m <= L; m++) {
		moff[m + L] = ncompact(L, std::abs(m), m);
		const int info = band_factor_block(b->rows, m, fac, piv);
		if (info)
			return sphswe_set_err(ctx, SPHSWE_ERR_SINGULAR, "band solve: zero pivot in the m = %d block (zgbsv INFO = %d)",
					      m, info);
		for (size_t j = 0; j < piv.size(); j++) {
			for (int e = 0; e < NE; e++)
				hf[(moff[m + L] + j) * NE + e] = make_double2(fac[j * NE + e].real(), fac[j * NE + e].imag());
			hp[moff[m + L] + j] = piv[j];
		}
	}
	if (!b->d_fac) {
		SPHSWE_HIP(ctx, hipMalloc((void **)&b->d_fac, sizeof(double2) * hf.size()));
		SPHSWE_HIP(ctx, hipMalloc((void **)&b->d_piv, hp.size()));
		SPHSWE_HIP(ctx, hipMalloc((void **)&b->d_moff, sizeof(int) * moff.size()));
	}
	// the stream may still read the previous factors
	SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
	SPHSWE_HIP(ctx, hipMemcpy(b->d_fac, hf.data(), sizeof(double2) * hf.size(), hipMemcpyHostToDevice));
	SPHSWE_HIP(ctx, hipMemcpy(b->d_piv, hp.data(), hp.size(), hipMemcpyHostToDevice));
	SPHSWE_HIP(ctx, hipMemcpy(b->d_moff, moff.data(), sizeof(int) * moff.size(), hipMemcpyHostToDevice));
	b->dirty = false;
	return SPHSWE_OK;
}

}  // namespace

extern "C" int sphswe_band_create(sphswe_ctx *ctx, int halo, sphswe_band **band)
{
	if (!ctx || !band) return SPHSWE_ERR_ARG;
	*band = nullptr;
	if (halo < 1 || halo > 4) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "band: halo %d outside 1..4", halo);
	sphswe_band *b = new sphswe_band(ctx->L, halo);
	b->ctx = ctx;
	ctx->bands.push_back(b);
	*band = b;
	return SPHSWE_OK;
}

// the context is being destroyed: the object keeps its own memory (freed by
// sphswe_band_destroy) and accepts no other call
void sphswe_band_detach(sphswe_band *b) { b->ctx = nullptr; }

extern "C" int sphswe_band_destroy(sphswe_band *band)
{
	if (!band) return SPHSWE_OK;
	if (band->ctx) {
		hipStreamSynchronize(band->ctx->stream);
		auto &v = band->ctx->bands;
		for (size_t i = 0; i < v.size(); i++)
			if (v[i] == band) { v.erase(v.begin() + i); break; }
	}
	if (band->d_fac) hipFree(band->d_fac);
	if (band->d_piv) hipFree(band->d_piv);
	if (band->d_moff) hipFree(band->d_moff);
	delete band;
	return SPHSWE_OK;
}

extern "C" int sphswe_band_add(sphswe_band *band, int component, double s_re, double s_im, double r)
{
	if (!band || !band->ctx) return SPHSWE_ERR_ARG;
	const int w = sphswe_host::band_component_width(component);
	if (w < 0) return sphswe_set_err(band->ctx, SPHSWE_ERR_ARG, "band: unknown component %d", component);
	// rowElement_add only asserts this (SPHMatrixComplex.hpp:200-204); an entry
	// outside the band would land in a neighbouring row
	if (w > band->halo)
		return sphswe_set_err(band->ctx, SPHSWE_ERR_ARG, "band: component %d needs halo %d, matrix has %d", component, w,
				      band->halo);
	if (!(r != 0.0)) return sphswe_set_err(band->ctx, SPHSWE_ERR_ARG, "band: radius must be non-zero");
	sphswe_host::band_add(band->rows, component, cplx(s_re, s_im), r);
	band->dirty = true;
	return SPHSWE_OK;
}

extern "C" int sphswe_band_solve(sphswe_ctx *ctx, sphswe_band *band, const double *d_rhs, double *d_x)
{
	if (!ctx || !band || !d_rhs || !d_x || band->ctx != ctx) return SPHSWE_ERR_ARG;
	if (band->dirty)
		if (int rc = factorise(ctx, band)) return rc;
	void *ys = nullptr;
	if (int rc = sphswe_scratch(ctx, sizeof(double2) * (size_t)ctx->ncplx, &ys)) return rc;
	double2 *y = (double2 *)ys;
	const double2 *rhs = (const double2 *)d_rhs;
	double2 *x = (double2 *)d_x;
	switch (band->halo) {
	case 1: launch_band<1>(ctx, band, rhs, x, y); break;
	case 2: launch_band<2>(ctx, band, rhs, x, y); break;
	case 3: launch_band<3>(ctx, band, rhs, x, y); break;
	default: launch_band<4>(ctx, band, rhs, x, y); break;
	}
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}
