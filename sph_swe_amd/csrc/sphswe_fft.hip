// Longitude FFTs of the SHT (the FFTW part of SHTNS).  One workgroup holds
// `nseq` complex sequences of length nphi in LDS and runs a mixed radix
// (4, 2, 3, 5, 7) Stockham autosort FFT on them, ping-ponging between two LDS
// buffers.  Real rings are packed two per complex sequence
// (z = f_a + i f_b), so a real r2c/c2r costs half a complex FFT.
// The physical layout is theta-contiguous ([i*nlat + j]); a workgroup reads
// 2*nseq (real) or nseq (complex) adjacent rings, and consecutive workgroups
// are gathered on one XCD so the 128-byte lines they share stay in its L2.
#include "sphswe_common.h"

namespace {

enum { MODE_R2C = 0, MODE_C2R = 1, MODE_C2C_FWD = 2, MODE_C2C_BWD = 3 };
constexpr int KB = 8;   // loads in flight per thread in the staging loops

struct FftArgs {
	int N;
	int npass;
	int radix[SPHSWE_MAX_RADIX];
};

__device__ inline double2 mul_i(double2 a, int sign)  // sign * i * a
{
	return sign > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <int R>
struct Dft;

template <>
struct Dft<2> {
	__device__ static void run(double2 *v, int) { double2 a = v[0], b = v[1]; v[0] = cadd(a, b); v[1] = csub(a, b); }
};

template <>
struct Dft<4> {
	__device__ static void run(double2 *v, int sign)
	{
		double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
		double2 a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
		double2 ia3 = mul_i(a3, sign);   // sign*i*a3
		v[0] = cadd(a0, a2);
		v[2] = csub(a0, a2);
		v[1] = cadd(a1, ia3);
		v[3] = csub(a1, ia3);
	}
};

// odd prime radix: direct O(R^2) with cos/sin(2 pi k / R) tables
template <int R>
struct OddTab;
template <>
struct OddTab<3> {
	static constexpr double c[3] = {1.0, -0.4999999999999998, -0.5000000000000004};
	static constexpr double s[3] = {0.0, 0.8660254037844387, -0.8660254037844384};
};
template <>
struct OddTab<5> {
	static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749473, -0.8090169943749476, 0.30901699437494723};
	static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924732, -0.587785252292473, -0.9510565162951536};
};
template <>
struct OddTab<7> {
	static constexpr double c[7] = {1.0, 0.6234898018587336, -0.22252093395631434, -0.900968867902419, -0.9009688679024191, -0.2225209339563146, 0.6234898018587334};
	static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.43388373911755823, -0.433883739117558, -0.9749279121818236, -0.7818314824680299};
};

template <int R>
struct DftOdd {
	__device__ static void run(double2 *v, int sign)
	{
		double2 out[R];
#pragma unroll
		for (int q = 0; q < R; q++) {
			double2 acc = v[0];
#pragma unroll
			for (int t = 1; t < R; t++) {
				const int e = (q * t) % R;
				const double wr = OddTab<R>::c[e], wi = sign * OddTab<R>::s[e];
				acc.x += v[t].x * wr - v[t].y * wi;
				acc.y += v[t].x * wi + v[t].y * wr;
			}
			out[q] = acc;
		}
#pragma unroll
		for (int q = 0; q < R; q++) v[q] = out[q];
	}
};
template <>
struct Dft<3> : DftOdd<3> {};
template <>
struct Dft<5> : DftOdd<5> {};
template <>
struct Dft<7> : DftOdd<7> {};

template <int R>
__device__ void stockham_pass(const double2 *__restrict__ src, double2 *__restrict__ dst, int N, int Ns, int nseq,
			      const double2 *__restrict__ tw, int sign)
{
	const int nb = N / R;
	const int total = nb * nseq;
	const int twstride = N / (Ns * R);
	for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
		int s = idx / nb;
		int b = idx - s * nb;
		const double2 *in = src + s * N;
		double2 *out = dst + s * N;
		int k = b % Ns;
		double2 v[R];
#pragma unroll
		for (int t = 0; t < R; t++) v[t] = in[b + t * nb];
		if (Ns > 1) {
#pragma unroll
			for (int t = 1; t < R; t++) {
				double2 w = tw[t * k * twstride];
				if (sign > 0) w.y = -w.y;
				v[t] = cmul(v[t], w);
			}
		}
		Dft<R>::run(v, sign);
		int o = (b / Ns) * Ns * R + k;
#pragma unroll
		for (int t = 0; t < R; t++) out[o + t * Ns] = v[t];
	}
}

// runs all passes; returns the buffer holding the result
__device__ double2 *fft_lds(double2 *a, double2 *b, const FftArgs &plan, int nseq, const double2 *__restrict__ tw, int sign)
{
	int Ns = 1;
	for (int p = 0; p < plan.npass; p++) {
		int R = plan.radix[p];
		switch (R) {
		case 2: stockham_pass<2>(a, b, plan.N, Ns, nseq, tw, sign); break;
		case 3: stockham_pass<3>(a, b, plan.N, Ns, nseq, tw, sign); break;
		case 4: stockham_pass<4>(a, b, plan.N, Ns, nseq, tw, sign); break;
		case 5: stockham_pass<5>(a, b, plan.N, Ns, nseq, tw, sign); break;
		default: stockham_pass<7>(a, b, plan.N, Ns, nseq, tw, sign); break;
		}
		__syncthreads();
		double2 *t = a;
		a = b;
		b = t;
		Ns *= R;
	}
	return a;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_fft(const void *__restrict__ in, void *__restrict__ out, FftArgs plan,
					     const double2 *__restrict__ tw, int nlat, int L, int nseq, int nblk,
					     int nblk_pad, double norm)
{
	extern __shared__ double2 sm[];
	const int lb = xcd_remap(blockIdx.x, nblk_pad);
	if (lb >= nblk) return;
	const int N = plan.N;
	const int f = blockIdx.y;
	const size_t nspat = (size_t)nlat * N;
	double2 *bufA = sm;
	double2 *bufB = sm + (size_t)nseq * N;
	const int tid = threadIdx.x, nt = blockDim.x;

	if (MODE == MODE_R2C) {
		const int RB = 2 * nseq;
		const int j0 = lb * RB;
		const double *src = (const double *)in + f * nspat;
		double *lds = (double *)bufA;
		for (int b = 0; b < N * RB; b += KB * nt) {
			double v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = min(b + u * nt + tid, N * RB - 1);
				int i = idx / RB, r = idx - i * RB;
				int j = min(j0 + r, nlat - 1);
				v[u] = src[(size_t)i * nlat + j];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = b + u * nt + tid;
				if (idx >= N * RB) break;
				int i = idx / RB, r = idx - i * RB;
				lds[2 * ((r >> 1) * N + i) + (r & 1)] = (j0 + r < nlat) ? v[u] : 0.0;
			}
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, -1);
		double2 *dst = (double2 *)out + (size_t)f * (L + 1) * nlat;
		for (int idx = tid; idx < (L + 1) * RB; idx += nt) {
			int m = idx / RB, r = idx - m * RB;
			int j = j0 + r;
			if (j >= nlat) continue;
			int s = r >> 1;
			double2 Z = res[s * N + m];
			double2 Zc = cconj(res[s * N + (m == 0 ? 0 : N - m)]);
			double2 v;
			if ((r & 1) == 0) v = make_double2(0.5 * (Z.x + Zc.x), 0.5 * (Z.y + Zc.y));
			else v = make_double2(0.5 * (Z.y - Zc.y), -0.5 * (Z.x - Zc.x));
			if (m == 0) v.y = 0.0;
			dst[(size_t)m * nlat + j] = make_double2(v.x * norm, v.y * norm);
		}
	} else if (MODE == MODE_C2R) {
		const int RB = 2 * nseq;
		const int j0 = lb * RB;
		const double2 *src = (const double2 *)in + (size_t)f * (L + 1) * nlat;
		// stage G[s][m][part] in bufB
		const int ng = (L + 1) * RB;
		for (int b = 0; b < ng; b += KB * nt) {
			double2 v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = min(b + u * nt + tid, ng - 1);
				int m = idx / RB, r = idx - m * RB;
				int j = min(j0 + r, nlat - 1);
				v[u] = src[(size_t)m * nlat + j];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = b + u * nt + tid;
				if (idx >= ng) break;
				int m = idx / RB, r = idx - m * RB;
				bufB[((r >> 1) * (L + 1) + m) * 2 + (r & 1)] = (j0 + r < nlat) ? v[u] : make_double2(0, 0);
			}
		}
		__syncthreads();
		for (int idx = tid; idx < nseq * N; idx += nt) {
			int s = idx / N, k = idx - s * N;
			double2 z = make_double2(0, 0);
			if (k == 0) {
				double2 A = bufB[(s * (L + 1)) * 2], B = bufB[(s * (L + 1)) * 2 + 1];
				z = make_double2(A.x, B.x);
			} else if (k <= L) {
				double2 A = bufB[(s * (L + 1) + k) * 2], B = bufB[(s * (L + 1) + k) * 2 + 1];
				z = make_double2(A.x - B.y, A.y + B.x);
			} else if (k >= N - L) {
				int m = N - k;
				double2 A = bufB[(s * (L + 1) + m) * 2], B = bufB[(s * (L + 1) + m) * 2 + 1];
				z = make_double2(A.x + B.y, -A.y + B.x);
			}
			bufA[idx] = z;
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, +1);
		double *dst = (double *)out + f * nspat;
		const double *rl = (const double *)res;
		for (int idx = tid; idx < N * RB; idx += nt) {
			int i = idx / RB, r = idx - i * RB;
			int j = j0 + r;
			if (j < nlat) dst[(size_t)i * nlat + j] = rl[2 * ((r >> 1) * N + i) + (r & 1)];
		}
	} else if (MODE == MODE_C2C_FWD) {
		const int j0 = lb * nseq;
		const double2 *src = (const double2 *)in + f * nspat;
		for (int b = 0; b < N * nseq; b += KB * nt) {
			double2 v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = min(b + u * nt + tid, N * nseq - 1);
				int i = idx / nseq, r = idx - i * nseq;
				v[u] = src[(size_t)i * nlat + min(j0 + r, nlat - 1)];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = b + u * nt + tid;
				if (idx >= N * nseq) break;
				int i = idx / nseq, r = idx - i * nseq;
				bufA[r * N + i] = (j0 + r < nlat) ? v[u] : make_double2(0, 0);
			}
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, -1);
		double2 *dst = (double2 *)out + (size_t)f * (2 * L + 1) * nlat;
		for (int idx = tid; idx < (2 * L + 1) * nseq; idx += nt) {
			int mi = idx / nseq, r = idx - mi * nseq;
			int j = j0 + r;
			if (j >= nlat) continue;
			int k = mi - L;
			if (k < 0) k += N;
			double2 z = res[r * N + k];
			dst[(size_t)mi * nlat + j] = make_double2(z.x * norm, z.y * norm);
		}
	} else {  // MODE_C2C_BWD
		const int j0 = lb * nseq;
		const double2 *src = (const double2 *)in + (size_t)f * (2 * L + 1) * nlat;
		for (int idx = tid; idx < nseq * N; idx += nt) bufA[idx] = make_double2(0, 0);
		__syncthreads();
		const int ng = (2 * L + 1) * nseq;
		for (int b = 0; b < ng; b += KB * nt) {
			double2 v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = min(b + u * nt + tid, ng - 1);
				int mi = idx / nseq, r = idx - mi * nseq;
				v[u] = src[(size_t)mi * nlat + min(j0 + r, nlat - 1)];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int idx = b + u * nt + tid;
				if (idx >= ng) break;
				int mi = idx / nseq, r = idx - mi * nseq;
				int k = mi - L;
				if (k < 0) k += N;
				bufA[r * N + k] = (j0 + r < nlat) ? v[u] : make_double2(0, 0);
			}
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, +1);
		double2 *dst = (double2 *)out + f * nspat;
		for (int idx = tid; idx < N * nseq; idx += nt) {
			int i = idx / nseq, r = idx - i * nseq;
			int j = j0 + r;
			if (j < nlat) dst[(size_t)i * nlat + j] = res[r * N + i];
		}
	}
}

int seqs_per_block(int N)
{
	int s = 2048 / N;
	if (s < 1) s = 1;
	if (s > 8) s = 8;
	return s;
}

template <int MODE>
int launch(sphswe_ctx *ctx, const void *in, void *out, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	FftArgs a;
	a.N = ctx->plan.n;
	a.npass = ctx->plan.npass;
	for (int i = 0; i < SPHSWE_MAX_RADIX; i++) a.radix[i] = ctx->plan.radix[i];
	int nseq = seqs_per_block(a.N);
	int rings = (MODE == MODE_R2C || MODE == MODE_C2R) ? 2 * nseq : nseq;
	int nblk = (ctx->nlat + rings - 1) / rings;
	int nblk_pad = (nblk + 7) / 8 * 8;
	// + 2*nseq slack: the c2r staging of G uses 2*(L+1) <= N+1 entries per sequence
	size_t lds = (2 * (size_t)nseq * a.N + 2 * nseq) * sizeof(double2);
	double norm = 2.0 * M_PI / a.N;
	dim3 grid(nblk_pad, nf);
	if (lds > 65536)
		SPHSWE_HIP(ctx, hipFuncSetAttribute((const void *)k_fft<MODE>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	k_fft<MODE><<<grid, 256, lds, ctx->stream>>>(in, out, a, ctx->d_tw, ctx->nlat, ctx->L, nseq, nblk, nblk_pad, norm);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

}  // namespace

int sphswe_fft_r2c(sphswe_ctx *ctx, const double *spat, double2 *four, int nf)
{
	return launch<MODE_R2C>(ctx, spat, four, nf);
}
int sphswe_fft_c2r(sphswe_ctx *ctx, const double2 *four, double *spat, int nf)
{
	return launch<MODE_C2R>(ctx, four, spat, nf);
}
int sphswe_fft_c2c_fwd(sphswe_ctx *ctx, const double2 *spat, double2 *four, int nf)
{
	return launch<MODE_C2C_FWD>(ctx, spat, four, nf);
}
int sphswe_fft_c2c_bwd(sphswe_ctx *ctx, const double2 *four, double2 *spat, int nf)
{
	return launch<MODE_C2C_BWD>(ctx, four, spat, nf);
}

extern "C" int sphswe_fft_analysis(sphswe_ctx *ctx, const double *d_spat, double *d_four, int nfields)
{
	if (!ctx || !d_spat || !d_four || nfields < 0) return SPHSWE_ERR_ARG;
	return sphswe_fft_r2c(ctx, d_spat, (double2 *)d_four, nfields);
}

extern "C" int sphswe_fft_synthesis(sphswe_ctx *ctx, const double *d_four, double *d_spat, int nfields)
{
	if (!ctx || !d_spat || !d_four || nfields < 0) return SPHSWE_ERR_ARG;
	return sphswe_fft_c2r(ctx, (const double2 *)d_four, d_spat, nfields);
}
