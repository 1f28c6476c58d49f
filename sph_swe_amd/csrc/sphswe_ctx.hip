// Context, grid, recurrence tables, device memory pool, timers.
//
// Replaces the SHTNS plan set up by SPHConfig (src/include/sph/SPHConfig.hpp:
// 324-399) and the MemBlockAlloc pool (src/include/sweet/MemBlockAlloc.hpp:398-513).
#include "sphswe_common.h"

#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <map>

int sphswe_set_err(sphswe_ctx *ctx, int code, const char *fmt, ...)
{
	if (ctx) {
		char buf[1024];
		va_list ap;
		va_start(ap, fmt);
		vsnprintf(buf, sizeof(buf), fmt, ap);
		va_end(ap);
		ctx->err = buf;
	}
	return code;
}

static int max_prime_factor(int n)
{
	int mp = 1;
	for (int p = 2; n > 1; p++)
		while (n % p == 0) { mp = p; n /= p; }
	return mp;
}

extern "C" int sphswe_grid_auto(int lmax, int nl_order, int *nlat, int *nphi)
{
	if (lmax < 1 || nl_order < 1 || !nlat || !nphi) return SPHSWE_ERR_ARG;
	int nl = ((nl_order + 1) * lmax) / 2 + 1;
	nl += (nl & 1);
	int np = (nl_order + 1) * lmax + 1;
	while (max_prime_factor(np) > 7) np++;
	*nlat = nl;
	*nphi = np;
	return SPHSWE_OK;
}

// Gauss-Legendre nodes (Newton on P_n with the Tricomi initial guess), north first.
static void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
	x.assign(n, 0.0);
	w.assign(n, 0.0);
	const double pi = M_PI;
	for (int i = 0; i < (n + 1) / 2; i++) {
		double z = cos(pi * (4.0 * i + 3.0) / (4.0 * n + 2.0));
		double dp = 1.0;
		for (int it = 0; it < 50; it++) {
			double p0 = 1.0, p1 = z;
			for (int k = 2; k <= n; k++) {
				double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
				p0 = p1;
				p1 = p2;
			}
			if (n == 1) { p1 = z; p0 = 1.0; }
			dp = n * (p0 - z * p1) / (1.0 - z * z);
			double dz = p1 / dp;
			z -= dz;
			if (fabs(dz) < 1e-17) {
				double q0 = 1.0, q1 = z;
				for (int k = 2; k <= n; k++) {
					double q2 = ((2.0 * k - 1.0) * z * q1 - (k - 1.0) * q0) / k;
					q0 = q1;
					q1 = q2;
				}
				dp = n * (q0 - z * q1) / (1.0 - z * z);
				break;
			}
		}
		x[i] = z;
		x[n - 1 - i] = -z;
		w[i] = w[n - 1 - i] = 2.0 / (((1.0 - z) * (1.0 + z)) * dp * dp);
	}
	if (n & 1) x[n / 2] = 0.0;
}

static bool fft_plan(int n, FftPlan &p)
{
	p.n = n;
	p.npass = 0;
	int r = n;
	// radix 4 first (fewest passes), then 2, 3, 5, 7
	while (r % 4 == 0 && p.npass < SPHSWE_MAX_RADIX) { p.radix[p.npass++] = 4; r /= 4; }
	const int rs[4] = {2, 3, 5, 7};
	for (int k = 0; k < 4; k++)
		while (r % rs[k] == 0 && p.npass < SPHSWE_MAX_RADIX) { p.radix[p.npass++] = rs[k]; r /= rs[k]; }
	return r == 1;
}

// Seeds of the Legendre recurrence at the start s = m + seg*k of every degree
// segment and every northern ring: (Pbar_s, Pbar_{s+1}).  The recurrence runs
// from Pbar_m^m on values scaled by 2^-E (E an integer, renormalised by powers
// of two), so values far below the double range keep their scale on the way;
// the stored seeds are the true values, which below the double range round to
// subnormals or 0.  The kernels restart the recurrence from these values at
// every segment, so no chain runs further than seg degrees from a seed, and a
// seed that underflows only drops terms below 2^-981 (a segment of the
// recurrence grows by less than 2^41).
static void make_seeds(int L, int NJ, int seg, const std::vector<double> &ct, const std::vector<double> &st,
		       std::vector<int> &segoff, std::vector<double2> &seed)
{
	segoff.assign(L + 2, 0);
	for (int m = 0; m <= L; m++) segoff[m + 1] = segoff[m] + (L - m + seg) / seg;
	seed.assign((size_t)segoff[L + 1] * NJ, make_double2(0, 0));
	// running Pbar_m^m(mu_j) = xm[j] * 2^Em[j], built by the product
	// Pbar_m^m = -sqrt((2m+1)/(2m)) st Pbar_{m-1}^{m-1} (error ~ sqrt(m) eps)
	std::vector<double> xm(NJ, 1.0 / sqrt(4.0 * M_PI));
	std::vector<int> Em(NJ, 0);
	std::vector<double> av(L + 2), bv(L + 2);
	for (int m = 0; m <= L; m++) {
		for (int l = m + 2; l <= L; l++) {
			av[l] = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
			bv[l] = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
		}
		const int nseg = segoff[m + 1] - segoff[m];
		for (int j = 0; j < NJ; j++) {
			if (m > 0) {
				xm[j] *= -sqrt((2.0 * m + 1.0) / (2.0 * m)) * st[j];
				if (fabs(xm[j]) < 0x1p-200) {
					xm[j] = ldexp(xm[j], 200);
					Em[j] -= 200;
				}
			}
			// (x0, x1) = (P_l, P_l+1) * 2^-E
			int E = Em[j];
			double x0 = xm[j];
			double x1 = (m < L) ? sqrt(2.0 * m + 3.0) * ct[j] * x0 : 0.0;
			int l = m;
			for (int k = 0; k < nseg; k++) {
				const int s0 = m + k * seg;
				while (l < s0) {
					const int ln = l + 2;
					const double x2 = (ln <= L) ? av[ln] * (ct[j] * x1 - bv[ln] * x0) : 0.0;
					x0 = x1;
					x1 = x2;
					l++;
					if (fabs(x0) > 0x1p200 || fabs(x1) > 0x1p200) {
						x0 = ldexp(x0, -200);
						x1 = ldexp(x1, -200);
						E += 200;
					} else if (fabs(x0) < 0x1p-200 && fabs(x1) < 0x1p-200 && (x0 != 0.0 || x1 != 0.0)) {
						x0 = ldexp(x0, 200);
						x1 = ldexp(x1, 200);
						E -= 200;
					}
				}
				// true values (below the double range they round to subnormals or 0)
				const double y1 = (l + 1 <= L) ? x1 : 0.0;
				seed[(size_t)(segoff[m] + k) * NJ + j] = make_double2(ldexp(x0, E), ldexp(y1, E));
			}
		}
	}
}

template <typename T>
static int upload(sphswe_ctx *ctx, T **dst, const std::vector<T> &src)
{
	SPHSWE_HIP(ctx, hipMalloc((void **)dst, sizeof(T) * std::max<size_t>(1, src.size())));
	SPHSWE_HIP(ctx, hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
	return SPHSWE_OK;
}

extern "C" int sphswe_create(int lmax, int mmax, int nlat, int nphi, int device, sphswe_ctx **out)
{
	if (!out) return SPHSWE_ERR_ARG;
	*out = nullptr;
	if (lmax < 1 || mmax != lmax || nlat < 2 || (nlat & 1) || nphi < 2 * lmax + 1)
		return SPHSWE_ERR_ARG;
	sphswe_ctx *ctx = new sphswe_ctx();
	ctx->L = lmax;
	ctx->nlat = nlat;
	ctx->nphi = nphi;
	ctx->NJ = nlat / 2;
	ctx->nlm = (lmax + 1) * (lmax + 2) / 2;
	ctx->ncplx = (lmax + 1) * (lmax + 1);
	ctx->nspat = nlat * nphi;
	ctx->device = device;
	if (!fft_plan(nphi, ctx->plan)) {
		delete ctx;
		return SPHSWE_ERR_ARG;
	}
	int rc = SPHSWE_OK;
	auto fail = [&](int code) {
		*out = ctx;  // keep for last_error; caller destroys
		return code;
	};
	if (hipSetDevice(device) != hipSuccess) return fail(sphswe_set_err(ctx, SPHSWE_ERR_HIP, "hipSetDevice(%d) failed", device));
	if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
		return fail(sphswe_set_err(ctx, SPHSWE_ERR_HIP, "hipStreamCreate failed"));
	for (int i = 0; i < SPHSWE_MAX_TIMERS; i++)
		if (hipEventCreate(&ctx->ev[i]) != hipSuccess) return fail(sphswe_set_err(ctx, SPHSWE_ERR_HIP, "hipEventCreate failed"));

	// grid
	std::vector<double> x, w;
	gauss_legendre(nlat, x, w);
	ctx->ct = x;
	ctx->wg = w;
	ctx->st.resize(nlat);
	std::vector<double> sq(nlat), coslat(nlat);
	for (int j = 0; j < nlat; j++) {
		ctx->st[j] = sqrt((1.0 - x[j]) * (1.0 + x[j]));
		sq[j] = sqrt(1.0 - x[j] * x[j]);                 // SPHOperators.hpp:113
		coslat[j] = cos(M_PI * 0.5 - acos(x[j]));        // cos(lat[j]), SPHConfig.hpp:309-310
	}
	// recurrence coefficients, real spectral order
	std::vector<double2> coef(ctx->nlm);
	std::vector<int> mof(ctx->nlm);
	for (int m = 0; m <= lmax; m++)
		for (int l = m; l <= lmax; l++) {
			int k = lm_index(lmax, l, m);
			mof[k] = m;
			if (l >= m + 2) {
				double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
				double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
				coef[k] = make_double2(a, a * b);   // Pbar_l = (a mu) Pbar_{l-1} - (a b) Pbar_{l-2}
			} else {
				coef[k] = make_double2(0.0, 0.0);
			}
		}
	std::vector<int> lofc(ctx->ncplx);
	for (int l = 0; l <= lmax; l++)
		for (int m = -l; m <= l; m++) lofc[l * l + l + m] = l;
	std::vector<int> segoff;
	std::vector<double2> seed;
	// 32-degree segments (one MFMA analysis wavefront each)
	ctx->seg = 32;
	// seeds every 16 degrees (the fused analysis); those of the 32-degree
	// segments are every other one of them
	std::vector<int> segoff16;
	std::vector<double2> seed16;
	make_seeds(lmax, ctx->NJ, 16, ctx->ct, ctx->st, segoff16, seed16);
	segoff.assign(lmax + 2, 0);
	for (int m = 0; m <= lmax; m++) segoff[m + 1] = segoff[m] + (lmax - m + 32) / 32;
	seed.resize((size_t)segoff[lmax + 1] * ctx->NJ);
	for (int m = 0; m <= lmax; m++)
		for (int k = 0; k < segoff[m + 1] - segoff[m]; k++)
			std::copy(seed16.begin() + (size_t)(segoff16[m] + 2 * k) * ctx->NJ,
				  seed16.begin() + (size_t)(segoff16[m] + 2 * k + 1) * ctx->NJ,
				  seed.begin() + (size_t)(segoff[m] + k) * ctx->NJ);
	ctx->nseg_total = segoff[lmax + 1];
	std::vector<int> unit_m(ctx->nseg_total);
	for (int m = 0; m <= lmax; m++)
		for (int k = segoff[m]; k < segoff[m + 1]; k++) unit_m[k] = m;
	// 64-degree blocks of every order, (m, first degree), for the 4x4x4 analysis
	std::vector<int2> ablk;
	for (int m = 0; m <= lmax; m++)
		for (int d0 = m; d0 <= lmax; d0 += 64) ablk.push_back(make_int2(m, d0));
	ctx->nablk = (int)ablk.size();
	{
		int v = 0;
		if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess) ctx->max_lds = v;
	}
	std::vector<double2> tw(nphi);
	for (int k = 0; k < nphi; k++) {
		// long double argument reduction keeps every twiddle within 1 ulp
		long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)nphi;
		tw[k] = make_double2((double)cosl(ang), (double)sinl(ang));
	}

	if ((rc = upload(ctx, &ctx->d_ct, x))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_st, ctx->st))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_sq, sq))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_coslat, coslat))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_wg, w))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_coef, coef))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_seed, seed))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_seed16, seed16))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_segoff16, segoff16))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_segoff, segoff))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_unit_m, unit_m))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_ablk, ablk))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_mofidx, mof))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_lofcidx, lofc))) return fail(rc);
	if ((rc = upload(ctx, &ctx->d_tw, tw))) return fail(rc);
	{
		std::vector<float2> tw32(tw.size());
		for (size_t k = 0; k < tw.size(); k++) tw32[k] = make_float2((float)tw[k].x, (float)tw[k].y);
		if ((rc = upload(ctx, &ctx->d_tw32, tw32))) return fail(rc);
	}
	if (hipMalloc((void **)&ctx->d_flag, 64 * sizeof(double)) != hipSuccess)
		return fail(sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "hipMalloc flag"));
	if ((rc = sphswe_noalloc_init(ctx))) return fail(rc);
	*out = ctx;
	return SPHSWE_OK;
}

static void sphswe_pool_release(const sphswe_ctx *ctx);

extern "C" int sphswe_destroy(sphswe_ctx *ctx)
{
	if (!ctx) return SPHSWE_OK;
	hipSetDevice(ctx->device);
	if (ctx->stream) hipStreamSynchronize(ctx->stream);
	sphswe_swe_release(ctx);
	sphswe_leg_tables_free(ctx);
	sphswe_leg_bt_tables_free(ctx);
	// objects still alive on this context keep their own memory and are freed
	// by their destroy calls; they no longer refer to the context
	for (sphswe_rexi *r : ctx->rexis) sphswe_rexi_detach(r);
	for (sphswe_band *b : ctx->bands) sphswe_band_detach(b);
	ctx->rexis.clear();
	ctx->bands.clear();
	void *ptrs[] = {ctx->d_ct, ctx->d_st, ctx->d_sq, ctx->d_coslat, ctx->d_wg, ctx->d_coef,
			ctx->d_seed, ctx->d_segoff, ctx->d_seed16, ctx->d_segoff16, ctx->d_unit_m, ctx->d_ablk, ctx->d_mofidx, ctx->d_lofcidx, ctx->d_tw, ctx->d_tw32, ctx->d_scratch, ctx->d_escratch, ctx->d_flag};
	for (void *p : ptrs)
		if (p) hipFree(p);
	for (auto &kv : ctx->allocs) hipFree(kv.first);
	sphswe_pool_release(ctx);
	for (int i = 0; i < SPHSWE_MAX_TIMERS; i++)
		if (ctx->ev[i]) hipEventDestroy(ctx->ev[i]);
	if (ctx->stream) hipStreamDestroy(ctx->stream);
	delete ctx;
	return SPHSWE_OK;
}

extern "C" int sphswe_get_sizes(const sphswe_ctx *ctx, int *lmax, int *nlat, int *nphi, int *nlm, int *ncplx)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (lmax) *lmax = ctx->L;
	if (nlat) *nlat = ctx->nlat;
	if (nphi) *nphi = ctx->nphi;
	if (nlm) *nlm = ctx->nlm;
	if (ncplx) *ncplx = ctx->ncplx;
	return SPHSWE_OK;
}

extern "C" int sphswe_get_grid(const sphswe_ctx *ctx, double *ct, double *st, double *wg)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (ct) memcpy(ct, ctx->ct.data(), sizeof(double) * ctx->nlat);
	if (st) memcpy(st, ctx->st.data(), sizeof(double) * ctx->nlat);
	if (wg) memcpy(wg, ctx->wg.data(), sizeof(double) * ctx->nlat);
	return SPHSWE_OK;
}

extern "C" const char *sphswe_last_error(const sphswe_ctx *ctx)
{
	return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int sphswe_sync(sphswe_ctx *ctx)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return SPHSWE_OK;
}

extern "C" void *sphswe_stream(sphswe_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

static int grow(sphswe_ctx *ctx, void **buf, size_t *have, size_t bytes, void **p)
{
	if (bytes > *have) {
		// a captured step that points into the old buffer must be rebuilt
		ctx->buf_gen++;
		if (*buf) {
			SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
			SPHSWE_HIP(ctx, hipFree(*buf));
			*buf = nullptr;
		}
		size_t nb = std::max(bytes, *have * 3 / 2);
		if (hipMalloc(buf, nb) != hipSuccess) {
			*have = 0;
			return sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "scratch allocation of %zu bytes failed", nb);
		}
		*have = nb;
	}
	*p = *buf;
	return SPHSWE_OK;
}

int sphswe_scratch(sphswe_ctx *ctx, size_t bytes, void **p)
{
	return grow(ctx, &ctx->d_scratch, &ctx->scratch_bytes, bytes, p);
}

int sphswe_escratch(sphswe_ctx *ctx, size_t bytes, void **p)
{
	return grow(ctx, &ctx->d_escratch, &ctx->escratch_bytes, bytes, p);
}

// ---- device memory pool: freed blocks are cached by size and reused (the
// role of MemBlockAlloc); blocks are not zeroed, as in the reference.
namespace {
struct Pool {
	std::multimap<size_t, void *> free_blocks;
};
std::unordered_map<const sphswe_ctx *, Pool> g_pools;
}

// frees the blocks cached for a context and forgets it (sphswe_destroy, after
// the stream has drained), so that a context created later at the same host
// address starts with an empty pool
static void sphswe_pool_release(const sphswe_ctx *ctx)
{
	auto it = g_pools.find(ctx);
	if (it == g_pools.end()) return;
	for (auto &kv : it->second.free_blocks) (void)hipFree(kv.second);
	g_pools.erase(it);
}

extern "C" int sphswe_malloc(sphswe_ctx *ctx, size_t bytes, void **d_ptr)
{
	if (!ctx || !d_ptr) return SPHSWE_ERR_ARG;
	size_t sz = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
	Pool &pool = g_pools[ctx];
	auto it = pool.free_blocks.find(sz);
	if (it != pool.free_blocks.end()) {
		*d_ptr = it->second;
		pool.free_blocks.erase(it);
		ctx->allocs[*d_ptr] = sz;
		return SPHSWE_OK;
	}
	void *p = nullptr;
	if (hipMalloc(&p, sz) != hipSuccess) {
		// release cached blocks and retry once
		SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
		for (auto &kv : pool.free_blocks) hipFree(kv.second);
		pool.free_blocks.clear();
		if (hipMalloc(&p, sz) != hipSuccess)
			return sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "hipMalloc(%zu) failed", sz);
	}
	ctx->allocs[p] = sz;
	*d_ptr = p;
	return SPHSWE_OK;
}

extern "C" int sphswe_free(sphswe_ctx *ctx, void *d_ptr)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (!d_ptr) return SPHSWE_OK;
	auto it = ctx->allocs.find(d_ptr);
	if (it == ctx->allocs.end()) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "sphswe_free: unknown pointer");
	g_pools[ctx].free_blocks.emplace(it->second, d_ptr);
	ctx->allocs.erase(it);
	return SPHSWE_OK;
}

extern "C" int sphswe_memcpy_h2d(sphswe_ctx *ctx, void *d_dst, const void *h_src, size_t bytes)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	SPHSWE_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
	SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return SPHSWE_OK;
}

extern "C" int sphswe_memcpy_d2h(sphswe_ctx *ctx, void *h_dst, const void *d_src, size_t bytes)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	SPHSWE_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
	SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return SPHSWE_OK;
}

extern "C" int sphswe_memcpy_d2d(sphswe_ctx *ctx, void *d_dst, const void *d_src, size_t bytes)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	SPHSWE_HIP(ctx, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
	return SPHSWE_OK;
}

__global__ void k_fill_zero(uint64_t *p, size_t n8, unsigned char *tail, size_t ntail)
{
	size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
	size_t stride = (size_t)gridDim.x * blockDim.x;
	for (size_t k = i; k < n8; k += stride) p[k] = 0;
	if (i < ntail) tail[i] = 0;
}

extern "C" int sphswe_fill_zero(sphswe_ctx *ctx, void *d_ptr, size_t bytes)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (bytes == 0) return SPHSWE_OK;
	if (((uintptr_t)d_ptr & 7) != 0) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "fill_zero: unaligned pointer");
	size_t n8 = bytes / 8, ntail = bytes % 8;
	int blocks = (int)std::min<size_t>(4096, (n8 + 255) / 256 + 1);
	k_fill_zero<<<blocks, 256, 0, ctx->stream>>>((uint64_t *)d_ptr, n8, (unsigned char *)d_ptr + n8 * 8, ntail);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

extern "C" int sphswe_timer_mark(sphswe_ctx *ctx, int slot)
{
	if (!ctx || slot < 0 || slot >= SPHSWE_MAX_TIMERS) return SPHSWE_ERR_ARG;
	SPHSWE_HIP(ctx, hipEventRecord(ctx->ev[slot], ctx->stream));
	return SPHSWE_OK;
}

extern "C" int sphswe_timer_elapsed(sphswe_ctx *ctx, int s0, int s1, float *ms)
{
	if (!ctx || !ms || s0 < 0 || s1 < 0 || s0 >= SPHSWE_MAX_TIMERS || s1 >= SPHSWE_MAX_TIMERS) return SPHSWE_ERR_ARG;
	SPHSWE_HIP(ctx, hipEventSynchronize(ctx->ev[s1]));
	SPHSWE_HIP(ctx, hipEventElapsedTime(ms, ctx->ev[s0], ctx->ev[s1]));
	return SPHSWE_OK;
}
