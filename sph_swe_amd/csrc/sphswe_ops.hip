// SHT entry points, spectral operators (SPHOperators / SPHOperatorsComplex),
// physical pointwise operators (spat_update_lambda*, operator*) and reductions.
#include "sphswe_common.h"

#include <algorithm>

// ---- SHT pair ----------------------------------------------------------------

extern "C" int sphswe_sht_analysis(sphswe_ctx *ctx, const double *d_spat, double *d_spec, int nfields, int precision)
{
	if (!ctx || !d_spat || !d_spec || nfields < 0) return SPHSWE_ERR_ARG;
	if (nfields == 0) return SPHSWE_OK;
	void *four;
	if (int rc = sphswe_scratch(ctx, sizeof(double2) * (size_t)(ctx->L + 1) * ctx->nlat * nfields, &four)) return rc;
	if (int rc = sphswe_fft_r2c(ctx, d_spat, (double2 *)four, nfields)) return rc;
	return sphswe_leg_anal(ctx, (const double2 *)four, (double2 *)d_spec, nfields, precision);
}

extern "C" int sphswe_sht_synthesis(sphswe_ctx *ctx, const double *d_spec, double *d_spat, int nfields, int precision)
{
	if (!ctx || !d_spat || !d_spec || nfields < 0) return SPHSWE_ERR_ARG;
	if (nfields == 0) return SPHSWE_OK;
	void *four;
	if (int rc = sphswe_scratch(ctx, sizeof(double2) * (size_t)(ctx->L + 1) * ctx->nlat * nfields, &four)) return rc;
	if (int rc = sphswe_leg_synth(ctx, (const double2 *)d_spec, (double2 *)four, nfields, precision)) return rc;
	return sphswe_fft_c2r(ctx, (const double2 *)four, d_spat, nfields);
}

extern "C" int sphswe_sht_analysis_cplx(sphswe_ctx *ctx, const double *d_spat, double *d_spec, int nfields, int precision)
{
	if (!ctx || !d_spat || !d_spec || nfields < 0) return SPHSWE_ERR_ARG;
	if (nfields == 0) return SPHSWE_OK;
	void *four;
	if (int rc = sphswe_scratch(ctx, sizeof(double2) * (size_t)(2 * ctx->L + 1) * ctx->nlat * nfields, &four)) return rc;
	if (int rc = sphswe_fft_c2c_fwd(ctx, (const double2 *)d_spat, (double2 *)four, nfields)) return rc;
	return sphswe_leg_anal_cplx(ctx, (const double2 *)four, (double2 *)d_spec, nfields, precision);
}

extern "C" int sphswe_sht_synthesis_cplx(sphswe_ctx *ctx, const double *d_spec, double *d_spat, int nfields, int precision)
{
	if (!ctx || !d_spat || !d_spec || nfields < 0) return SPHSWE_ERR_ARG;
	if (nfields == 0) return SPHSWE_OK;
	void *four;
	if (int rc = sphswe_scratch(ctx, sizeof(double2) * (size_t)(2 * ctx->L + 1) * ctx->nlat * nfields, &four)) return rc;
	if (int rc = sphswe_leg_synth_cplx(ctx, (const double2 *)d_spec, (double2 *)four, nfields, precision)) return rc;
	return sphswe_fft_c2c_bwd(ctx, (const double2 *)four, (double2 *)d_spat, nfields);
}

// ---- spectral operators ------------------------------------------------------

namespace {

// value of op at (l, m) reading neighbours through get(l') (zero outside the triangle)
template <typename Get>
__device__ inline double2 spec_op_value(int op, int l, int m, int L, Get get, double2 s, double p0, double p1, double p2)
{
	const double dl = l, dm = m;
	double2 x = get(l), v;
	switch (op) {
	case SPHSWE_OP_IDENTITY: v = x; break;
	case SPHSWE_OP_DIFF_LON: v = make_double2(-dm * x.y, dm * x.x); break;
	case SPHSWE_OP_ONE_MINUS_MU2_DMU: {
		// ((1-n) R(n-1,m)) x_{n-1} + ((n+2) S(n+1,m)) x_{n+1}  (SPHOperators.hpp:248-252)
		double cm = (-dl + 1.0) * idR(dl - 1, dm), cp = (dl + 2.0) * idS(dl + 1, dm);
		double2 a = get(l - 1), b = get(l + 1);
		v = make_double2(cm * a.x + cp * b.x, cm * a.y + cp * b.y);
		break;
	}
	case SPHSWE_OP_MU: {
		double cm = idR(dl - 1, dm), cp = idS(dl + 1, dm);
		double2 a = get(l - 1), b = get(l + 1);
		v = make_double2(cm * a.x + cp * b.x, cm * a.y + cp * b.y);
		break;
	}
	case SPHSWE_OP_MU2: {
		double ca = idA(dl - 2, dm), cb = idB(dl, dm), cc = idC(dl + 2, dm);
		double2 a = get(l - 2), c = get(l + 2);
		v = make_double2(ca * a.x + cb * x.x + cc * c.x, ca * a.y + cb * x.y + cc * c.y);
		break;
	}
	case SPHSWE_OP_LAPLACE: {
		double f = -dl * (dl + 1.0);
		v = make_double2(x.x * f, x.y * f);
		break;
	}
	default: {  // HELMHOLTZ: x / (a + (-b n(n+1))), a = s, b = (p0 + i p2) / r^2, r = p1
		double r2 = p1 * p1;
		double br = p0 / r2, bi = p2 / r2;
		double n2 = dl * (dl + 1.0);
		double2 d = make_double2(s.x + (-br * n2), s.y + (-bi * n2));
		double den = d.x * d.x + d.y * d.y;
		return make_double2((x.x * d.x + x.y * d.y) / den, (x.y * d.x - x.x * d.y) / den);
	}
	}
	if (op == SPHSWE_OP_IDENTITY || s.x != 1.0 || s.y != 0.0) v = cmul(v, s);
	return v;
}

__global__ void k_spec_op(int op, const double2 *__restrict__ in, double2 *__restrict__ out, const int *__restrict__ mof,
			  int L, int nlm, long total, double2 s, double p0, double p1)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= total) return;
	int f = (int)(i / nlm);
	int idx = (int)(i - (long)f * nlm);
	int m = mof[idx];
	int base = lm_index(L, 0, m);
	int l = idx - base;
	const double2 *x = in + (long)f * nlm + base;
	auto get = [&](int ll) -> double2 {
		if (ll < m || ll > L) return make_double2(0, 0);
		return x[ll];
	};
	out[i] = spec_op_value(op, l, m, L, get, s, p0, p1, 0.0);
}

__global__ void k_spec_op_cplx(int op, const double2 *__restrict__ in, double2 *__restrict__ out, const int *__restrict__ lofc,
			       int L, int nc, long total, double2 s, double p0, double p1, double p2)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= total) return;
	int f = (int)(i / nc);
	int idx = (int)(i - (long)f * nc);
	int l = lofc[idx];
	int m = idx - l * l - l;
	const double2 *x = in + (long)f * nc;
	auto get = [&](int ll) -> double2 {
		if (ll < 0 || ll > L || abs(m) > ll) return make_double2(0, 0);
		return x[ll * ll + ll + m];
	};
	out[i] = spec_op_value(op, l, m, L, get, s, p0, p1, p2);
}

__global__ void k_axpby(long n, double2 a, const double2 *__restrict__ x, double2 b, const double2 *__restrict__ y,
			double2 *__restrict__ out)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= n) return;
	double2 v = cmul(a, x[i]);
	if (y) v = cadd(v, cmul(b, y[i]));
	out[i] = v;
}

__global__ void k_add_scalar(double2 *p, double2 v)
{
	p[0] = cadd(p[0], v);
}

inline int grid1(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

}  // namespace

extern "C" int sphswe_spec_op(sphswe_ctx *ctx, int op, const double *d_in, double *d_out, int nfields, double s_re,
			      double s_im, double p0, double p1)
{
	if (!ctx || !d_in || !d_out || nfields < 0 || op < 0 || op > SPHSWE_OP_HELMHOLTZ) return SPHSWE_ERR_ARG;
	if (d_in == d_out && op != SPHSWE_OP_IDENTITY && op != SPHSWE_OP_DIFF_LON && op != SPHSWE_OP_LAPLACE &&
	    op != SPHSWE_OP_HELMHOLTZ)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "spec_op: banded operators cannot run in place");
	long total = (long)ctx->nlm * nfields;
	if (total == 0) return SPHSWE_OK;
	k_spec_op<<<grid1(total), 256, 0, ctx->stream>>>(op, (const double2 *)d_in, (double2 *)d_out, ctx->d_mofidx, ctx->L,
							   ctx->nlm, total, make_double2(s_re, s_im), p0, p1);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

extern "C" int sphswe_spec_op_cplx(sphswe_ctx *ctx, int op, const double *d_in, double *d_out, int nfields, double s_re,
				   double s_im, double p0, double p1, double p2)
{
	if (!ctx || !d_in || !d_out || nfields < 0 || op < 0 || op > SPHSWE_OP_HELMHOLTZ) return SPHSWE_ERR_ARG;
	if (d_in == d_out && op != SPHSWE_OP_IDENTITY && op != SPHSWE_OP_DIFF_LON && op != SPHSWE_OP_LAPLACE &&
	    op != SPHSWE_OP_HELMHOLTZ)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "spec_op_cplx: banded operators cannot run in place");
	long total = (long)ctx->ncplx * nfields;
	if (total == 0) return SPHSWE_OK;
	k_spec_op_cplx<<<grid1(total), 256, 0, ctx->stream>>>(op, (const double2 *)d_in, (double2 *)d_out, ctx->d_lofcidx,
								ctx->L, ctx->ncplx, total, make_double2(s_re, s_im), p0, p1, p2);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

extern "C" int sphswe_spec_axpby(sphswe_ctx *ctx, size_t n, double a_re, double a_im, const double *d_x, double b_re,
				 double b_im, const double *d_y, double *d_out)
{
	if (!ctx || !d_x || !d_out) return SPHSWE_ERR_ARG;
	if (n == 0) return SPHSWE_OK;
	k_axpby<<<grid1((long)n), 256, 0, ctx->stream>>>((long)n, make_double2(a_re, a_im), (const double2 *)d_x,
							  make_double2(b_re, b_im), (const double2 *)d_y, (double2 *)d_out);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

namespace {
// real spectra between truncations Ls and Ld (m-major, m (2L - m + 1) / 2 + l):
// block y = order m of the destination, threads over its degrees; orders and
// degrees above Ls are zero
__global__ void k_copy_modes(const double2 *__restrict__ src, double2 *__restrict__ dst, int Ls, int Ld, long nlm_s,
			     long nlm_d, int nfields)
{
	const int m = blockIdx.y;
	const long ob_d = (long)m * (2 * Ld - m + 1) / 2, ob_s = (long)m * (2 * Ls - m + 1) / 2;
	for (int l = m + blockIdx.x * blockDim.x + threadIdx.x; l <= Ld; l += gridDim.x * blockDim.x)
		for (int f = 0; f < nfields; f++)
			dst[f * nlm_d + ob_d + l] = (m <= Ls && l <= Ls) ? src[f * nlm_s + ob_s + l] : make_double2(0, 0);
}
}  // namespace

extern "C" int sphswe_spec_copy_modes(sphswe_ctx *src_ctx, const double *d_src, sphswe_ctx *dst_ctx, double *d_dst,
				      int nfields)
{
	if (!src_ctx || !dst_ctx || !d_src || !d_dst || nfields < 0) return SPHSWE_ERR_ARG;
	if (src_ctx->device != dst_ctx->device)
		return sphswe_set_err(dst_ctx, SPHSWE_ERR_ARG, "spec_copy_modes: the contexts are on devices %d and %d",
				      src_ctx->device, dst_ctx->device);
	if (nfields == 0) return SPHSWE_OK;
	if (src_ctx != dst_ctx) {
		// the copy runs on the destination stream, after the work queued on the source stream
		hipEvent_t ev;
		SPHSWE_HIP(dst_ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		hipError_t e = hipEventRecord(ev, src_ctx->stream);
		if (e == hipSuccess) e = hipStreamWaitEvent(dst_ctx->stream, ev, 0);
		hipEventDestroy(ev);
		if (e != hipSuccess) return sphswe_set_err(dst_ctx, SPHSWE_ERR_HIP, "spec_copy_modes: %s", hipGetErrorString(e));
	}
	const int Ld = dst_ctx->L;
	dim3 grid((Ld + 1 + 255) / 256, Ld + 1);
	k_copy_modes<<<grid, 256, 0, dst_ctx->stream>>>((const double2 *)d_src, (double2 *)d_dst, src_ctx->L, Ld,
						       src_ctx->nlm, dst_ctx->nlm, nfields);
	SPHSWE_LAUNCH_CHECK(dst_ctx);
	return SPHSWE_OK;
}

extern "C" int sphswe_spec_add_scalar(sphswe_ctx *ctx, double *d_spec, double v_re, double v_im)
{
	if (!ctx || !d_spec) return SPHSWE_ERR_ARG;
	double s = std::sqrt(4.0 * M_PI);
	k_add_scalar<<<1, 1, 0, ctx->stream>>>((double2 *)d_spec, make_double2(v_re * s, v_im * s));
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

// ---- physical operators ------------------------------------------------------

namespace {

template <typename T>
__device__ inline T tmul(T a, double s);
template <>
__device__ inline double tmul(double a, double s) { return a * s; }
template <>
__device__ inline double2 tmul(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
template <typename T>
__device__ inline T tdiv(T a, double s);
template <>
__device__ inline double tdiv(double a, double s) { return a / s; }
template <>
__device__ inline double2 tdiv(double2 a, double s) { return make_double2(a.x / s, a.y / s); }
__device__ inline double tprod(double a, double b) { return b * a; }
__device__ inline double2 tprod(double2 a, double2 b) { return cmul(b, a); }

template <typename T>
__global__ void k_spat_op(int op, const T *__restrict__ a, const T *__restrict__ b, T *__restrict__ out, double s,
			  const double *__restrict__ ct, const double *__restrict__ st, const double *__restrict__ sq,
			  const double *__restrict__ coslat, int nlat, long total)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= total) return;
	int j = (int)(i % nlat);
	T x = a[i];
	switch (op) {
	case SPHSWE_SPAT_MUL: out[i] = tprod(x, b[i]); break;
	case SPHSWE_SPAT_DIV_COSPHI: out[i] = tdiv(x, sq[j]); break;
	case SPHSWE_SPAT_MUL_ST: out[i] = tmul(x, st[j]); break;
	case SPHSWE_SPAT_DIV_ST: out[i] = tdiv(x, st[j]); break;
	case SPHSWE_SPAT_MUL_MU: out[i] = tmul(x, ct[j]); break;
	case SPHSWE_SPAT_SCALE: out[i] = tmul(x, s); break;
	case SPHSWE_SPAT_DIV_ST2: out[i] = tdiv(x, st[j] * st[j]); break;
	default: out[i] = tdiv(x, coslat[j]); break;
	}
}

}  // namespace

extern "C" int sphswe_spat_op(sphswe_ctx *ctx, int op, const double *d_a, const double *d_b, double *d_out, double s,
			      int nfields)
{
	if (!ctx || !d_a || !d_out || nfields < 0 || op < 0 || op > SPHSWE_SPAT_DIV_ST2) return SPHSWE_ERR_ARG;
	if (op == SPHSWE_SPAT_MUL && !d_b) return SPHSWE_ERR_ARG;
	long total = (long)ctx->nspat * nfields;
	if (total == 0) return SPHSWE_OK;
	k_spat_op<double><<<grid1(total), 256, 0, ctx->stream>>>(op, d_a, d_b, d_out, s, ctx->d_ct, ctx->d_st, ctx->d_sq,
								  ctx->d_coslat, ctx->nlat, total);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

extern "C" int sphswe_spat_op_cplx(sphswe_ctx *ctx, int op, const double *d_a, const double *d_b, double *d_out,
				   double s, int nfields)
{
	if (!ctx || !d_a || !d_out || nfields < 0 || op < 0 || op > SPHSWE_SPAT_DIV_ST2) return SPHSWE_ERR_ARG;
	if (op == SPHSWE_SPAT_MUL && !d_b) return SPHSWE_ERR_ARG;
	long total = (long)ctx->nspat * nfields;
	if (total == 0) return SPHSWE_OK;
	k_spat_op<double2><<<grid1(total), 256, 0, ctx->stream>>>(op, (const double2 *)d_a, (const double2 *)d_b,
								   (double2 *)d_out, s, ctx->d_ct, ctx->d_st, ctx->d_sq,
								   ctx->d_coslat, ctx->nlat, total);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

namespace {

__global__ void k_real_to_cplx_f(const double *__restrict__ a, double2 *__restrict__ o, long n)
{
	const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i < n) o[i] = make_double2(a[i], 0.0);
}

__global__ void k_cplx_part(const double2 *__restrict__ a, double *__restrict__ o, int part, long n)
{
	const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i < n) o[i] = part ? a[i].y : a[i].x;
}

}  // namespace

extern "C" int sphswe_spat_real_to_cplx(sphswe_ctx *ctx, const double *d_real, double *d_cplx, int nfields)
{
	if (!ctx || !d_real || !d_cplx || nfields < 0) return SPHSWE_ERR_ARG;
	const long total = (long)ctx->nspat * nfields;
	if (total == 0) return SPHSWE_OK;
	k_real_to_cplx_f<<<grid1(total), 256, 0, ctx->stream>>>(d_real, (double2 *)d_cplx, total);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

extern "C" int sphswe_spat_cplx_part(sphswe_ctx *ctx, const double *d_cplx, double *d_real, int part, int nfields)
{
	if (!ctx || !d_real || !d_cplx || nfields < 0 || (part != 0 && part != 1)) return SPHSWE_ERR_ARG;
	const long total = (long)ctx->nspat * nfields;
	if (total == 0) return SPHSWE_OK;
	k_cplx_part<<<grid1(total), 256, 0, ctx->stream>>>((const double2 *)d_cplx, d_real, part, total);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

// ---- reductions --------------------------------------------------------------

namespace {

__global__ void k_nan(const double *__restrict__ x, long n, double *flag)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	long stride = (long)gridDim.x * blockDim.x;
	bool bad = false;
	for (long k = i; k < n; k += stride) bad |= !isfinite(x[k]);
	if (__any(bad) && (threadIdx.x & 63) == 0) flag[0] = 1.0;
}

__global__ void k_minmax(const double *__restrict__ x, long n, double *part)
{
	__shared__ double smn[256], smx[256], sab[256];
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	long stride = (long)gridDim.x * blockDim.x;
	double mn = INFINITY, mx = -INFINITY, ab = -1.0;
	for (long k = i; k < n; k += stride) {
		double v = x[k];
		mn = fmin(mn, v);
		mx = fmax(mx, v);
		ab = fmax(ab, fabs(v));
	}
	smn[threadIdx.x] = mn;
	smx[threadIdx.x] = mx;
	sab[threadIdx.x] = ab;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1) {
		if ((int)threadIdx.x < o) {
			smn[threadIdx.x] = fmin(smn[threadIdx.x], smn[threadIdx.x + o]);
			smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + o]);
			sab[threadIdx.x] = fmax(sab[threadIdx.x], sab[threadIdx.x + o]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		part[3 * blockIdx.x] = smn[0];
		part[3 * blockIdx.x + 1] = smx[0];
		part[3 * blockIdx.x + 2] = sab[0];
	}
}

}  // namespace

extern "C" int sphswe_any_nan(sphswe_ctx *ctx, const double *d_x, size_t n, int *h_flag)
{
	if (!ctx || !d_x || !h_flag) return SPHSWE_ERR_ARG;
	SPHSWE_HIP(ctx, hipMemsetAsync(ctx->d_flag, 0, sizeof(double), ctx->stream));
	int blocks = (int)std::min<long>(1024, grid1((long)n));
	if (blocks > 0) k_nan<<<blocks, 256, 0, ctx->stream>>>(d_x, (long)n, ctx->d_flag);
	SPHSWE_LAUNCH_CHECK(ctx);
	double f = 0;
	SPHSWE_HIP(ctx, hipMemcpyAsync(&f, ctx->d_flag, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
	SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
	*h_flag = f != 0.0;
	return SPHSWE_OK;
}

extern "C" int sphswe_reduce_minmax(sphswe_ctx *ctx, const double *d_x, size_t n, double *h_min, double *h_max,
				    double *h_absmax)
{
	if (!ctx || !d_x) return SPHSWE_ERR_ARG;
	const int blocks = 16;
	k_minmax<<<blocks, 256, 0, ctx->stream>>>(d_x, (long)n, ctx->d_flag);
	SPHSWE_LAUNCH_CHECK(ctx);
	double part[3 * blocks];
	SPHSWE_HIP(ctx, hipMemcpyAsync(part, ctx->d_flag, sizeof(part), hipMemcpyDeviceToHost, ctx->stream));
	SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
	double mn = INFINITY, mx = -INFINITY, ab = -1.0;
	for (int b = 0; b < blocks; b++) {
		mn = std::min(mn, part[3 * b]);
		mx = std::max(mx, part[3 * b + 1]);
		ab = std::max(ab, part[3 * b + 2]);
	}
	if (h_min) *h_min = mn;
	if (h_max) *h_max = mx;
	if (h_absmax) *h_absmax = ab;
	return SPHSWE_OK;
}
