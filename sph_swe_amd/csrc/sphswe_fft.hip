// Longitude FFTs of the SHT (the FFTW part of SHTNS) on the theta-contiguous
// physical layout ([i*nlat + j]).  Real rings are packed two per complex
// sequence (z = f_a + i f_b), so a real r2c/c2r costs half a complex FFT.
//   - k_fft_wide: the real transforms at compile-time sizes (the grids of the
//     usual truncations), a tile of adjacent ring pairs per workgroup in one
//     in-place LDS buffer, first and last radix pass against HBM;
//   - k_fft: any size and the complex transforms, runtime mixed radix
//     (4, 2, 3, 5, 7) Stockham passes ping-ponging between two LDS buffers.
// Consecutive workgroups are gathered on one XCD so the 128-byte lines they
// share stay in its L2.
#include "sphswe_fft_blocks.h"

namespace {

enum { MODE_R2C = 0, MODE_C2R = 1, MODE_C2C_FWD = 2, MODE_C2C_BWD = 3 };
constexpr int KB = 8;   // loads in flight per thread in the staging loops

// n / d for 0 <= n < 2^20 without an integer division: the float quotient is
// within 1/8 of n / d, one correction step makes it exact
struct QDiv {
	int d;
	float inv;
};
__host__ inline QDiv qdiv_of(int d) { return QDiv{d, 1.0f / (float)d}; }
__device__ __forceinline__ int qdiv(int n, const QDiv &q, int &r)
{
	int t = (int)((float)n * q.inv);
	r = n - t * q.d;
	if (r < 0) {
		t--;
		r += q.d;
	} else if (r >= q.d) {
		t++;
		r -= q.d;
	}
	return t;
}

struct FftArgs {
	int N;
	int npass;
	int radix[SPHSWE_MAX_RADIX];
	QDiv nbq[SPHSWE_MAX_RADIX];   // N / radix of each pass
	QDiv nsq[SPHSWE_MAX_RADIX];   // product of the radices before each pass
	QDiv Nq, RBq, Lq;             // N, rings (or sequences) per workgroup, L + 1
};

// LDS element index of entry i of sequence s: one padding element every 8
// complex values spreads the strided accesses of the early passes over the banks
__device__ __forceinline__ int lidx(int s, int i, int NP) { return s * NP + i + (i >> 3); }
__host__ __device__ inline int padded_len(int N) { return N + (N >> 3) + 1; }

template <int R>
__device__ void stockham_pass(const double2 *__restrict__ src, double2 *__restrict__ dst, int N, int Ns, const QDiv &nbq,
			      const QDiv &nsq, int nseq, const double2 *__restrict__ tw, int sign)
{
	const int nb = nbq.d;
	const int total = nb * nseq;
	const int twstride = N / (Ns * R);
	const int NP = padded_len(N);
	for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
		int b;
		const int s = qdiv(idx, nbq, b);
		int k;
		const int jb = qdiv(b, nsq, k);
		double2 v[R];
#pragma unroll
		for (int t = 0; t < R; t++) v[t] = src[lidx(s, b + t * nb, NP)];
		if (Ns > 1) {
#pragma unroll
			for (int t = 1; t < R; t++) {
				double2 w = tw[t * k * twstride];
				if (sign > 0) w.y = -w.y;
				v[t] = cmul(v[t], w);
			}
		}
		Dft<R>::run(v, sign);
		const int o = jb * Ns * R + k;
#pragma unroll
		for (int t = 0; t < R; t++) dst[lidx(s, o + t * Ns, NP)] = v[t];
	}
}

// runs all passes; returns the buffer holding the result
__device__ double2 *fft_lds(double2 *a, double2 *b, const FftArgs &plan, int nseq, const double2 *__restrict__ tw, int sign)
{
	int Ns = 1;
	for (int p = 0; p < plan.npass; p++) {
		int R = plan.radix[p];
		switch (R) {
		case 2: stockham_pass<2>(a, b, plan.N, Ns, plan.nbq[p], plan.nsq[p], nseq, tw, sign); break;
		case 3: stockham_pass<3>(a, b, plan.N, Ns, plan.nbq[p], plan.nsq[p], nseq, tw, sign); break;
		case 4: stockham_pass<4>(a, b, plan.N, Ns, plan.nbq[p], plan.nsq[p], nseq, tw, sign); break;
		case 5: stockham_pass<5>(a, b, plan.N, Ns, plan.nbq[p], plan.nsq[p], nseq, tw, sign); break;
		default: stockham_pass<7>(a, b, plan.N, Ns, plan.nbq[p], plan.nsq[p], nseq, tw, sign); break;
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
	const int N = plan.N, NP = padded_len(N);
	const int f = blockIdx.y;
	const size_t nspat = (size_t)nlat * N;
	double2 *bufA = sm;
	double2 *bufB = sm + (size_t)nseq * NP;
	const int tid = threadIdx.x, nt = blockDim.x;
	const int RB = plan.RBq.d;   // rings (real modes) or sequences (complex modes) of this workgroup

	if (MODE == MODE_R2C) {
		const int j0 = lb * RB;
		const double *src = (const double *)in + f * nspat;
		double *lds = (double *)bufA;
		for (int b = 0; b < N * RB; b += KB * nt) {
			double v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int r;
				const int i = qdiv(min(b + u * nt + tid, N * RB - 1), plan.RBq, r);
				v[u] = src[(size_t)i * nlat + min(j0 + r, nlat - 1)];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				const int idx = b + u * nt + tid;
				if (idx >= N * RB) break;
				int r;
				const int i = qdiv(idx, plan.RBq, r);
				lds[2 * lidx(r >> 1, i, NP) + (r & 1)] = (j0 + r < nlat) ? v[u] : 0.0;
			}
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, -1);
		double2 *dst = (double2 *)out + (size_t)f * (L + 1) * nlat;
		for (int idx = tid; idx < (L + 1) * RB; idx += nt) {
			int r;
			const int m = qdiv(idx, plan.RBq, r);
			const int j = j0 + r;
			if (j >= nlat) continue;
			const int s = r >> 1;
			const double2 Z = res[lidx(s, m, NP)];
			const double2 Zc = cconj(res[lidx(s, m == 0 ? 0 : N - m, NP)]);
			double2 v;
			if ((r & 1) == 0) v = make_double2(0.5 * (Z.x + Zc.x), 0.5 * (Z.y + Zc.y));
			else v = make_double2(0.5 * (Z.y - Zc.y), -0.5 * (Z.x - Zc.x));
			if (m == 0) v.y = 0.0;
			dst[(size_t)m * nlat + j] = make_double2(v.x * norm, v.y * norm);
		}
	} else if (MODE == MODE_C2R) {
		const int j0 = lb * RB;
		const double2 *src = (const double2 *)in + (size_t)f * (L + 1) * nlat;
		// stage G[s][m][part] in bufB (2 (L + 1) <= N + 1 < NP entries per sequence)
		const int ng = (L + 1) * RB;
		for (int b = 0; b < ng; b += KB * nt) {
			double2 v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int r;
				const int m = qdiv(min(b + u * nt + tid, ng - 1), plan.RBq, r);
				v[u] = src[(size_t)m * nlat + min(j0 + r, nlat - 1)];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				const int idx = b + u * nt + tid;
				if (idx >= ng) break;
				int r;
				const int m = qdiv(idx, plan.RBq, r);
				bufB[((r >> 1) * (L + 1) + m) * 2 + (r & 1)] = (j0 + r < nlat) ? v[u] : make_double2(0, 0);
			}
		}
		__syncthreads();
		for (int idx = tid; idx < nseq * N; idx += nt) {
			int k;
			const int s = qdiv(idx, plan.Nq, k);
			double2 z = make_double2(0, 0);
			if (k == 0) {
				const double2 A = bufB[(s * (L + 1)) * 2], B = bufB[(s * (L + 1)) * 2 + 1];
				z = make_double2(A.x, B.x);
			} else if (k <= L) {
				const double2 A = bufB[(s * (L + 1) + k) * 2], B = bufB[(s * (L + 1) + k) * 2 + 1];
				z = make_double2(A.x - B.y, A.y + B.x);
			} else if (k >= N - L) {
				const int m = N - k;
				const double2 A = bufB[(s * (L + 1) + m) * 2], B = bufB[(s * (L + 1) + m) * 2 + 1];
				z = make_double2(A.x + B.y, -A.y + B.x);
			}
			bufA[lidx(s, k, NP)] = z;
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, +1);
		double *dst = (double *)out + f * nspat;
		const double *rl = (const double *)res;
		for (int idx = tid; idx < N * RB; idx += nt) {
			int r;
			const int i = qdiv(idx, plan.RBq, r);
			const int j = j0 + r;
			if (j < nlat) dst[(size_t)i * nlat + j] = rl[2 * lidx(r >> 1, i, NP) + (r & 1)];
		}
	} else if (MODE == MODE_C2C_FWD) {
		const int j0 = lb * RB;
		const double2 *src = (const double2 *)in + f * nspat;
		for (int b = 0; b < N * RB; b += KB * nt) {
			double2 v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int r;
				const int i = qdiv(min(b + u * nt + tid, N * RB - 1), plan.RBq, r);
				v[u] = src[(size_t)i * nlat + min(j0 + r, nlat - 1)];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				const int idx = b + u * nt + tid;
				if (idx >= N * RB) break;
				int r;
				const int i = qdiv(idx, plan.RBq, r);
				bufA[lidx(r, i, NP)] = (j0 + r < nlat) ? v[u] : make_double2(0, 0);
			}
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, -1);
		double2 *dst = (double2 *)out + (size_t)f * (2 * L + 1) * nlat;
		for (int idx = tid; idx < (2 * L + 1) * RB; idx += nt) {
			int r;
			const int mi = qdiv(idx, plan.RBq, r);
			const int j = j0 + r;
			if (j >= nlat) continue;
			int k = mi - L;
			if (k < 0) k += N;
			const double2 z = res[lidx(r, k, NP)];
			dst[(size_t)mi * nlat + j] = make_double2(z.x * norm, z.y * norm);
		}
	} else {  // MODE_C2C_BWD
		const int j0 = lb * RB;
		const double2 *src = (const double2 *)in + (size_t)f * (2 * L + 1) * nlat;
		for (int idx = tid; idx < nseq * NP; idx += nt) bufA[idx] = make_double2(0, 0);
		__syncthreads();
		const int ng = (2 * L + 1) * RB;
		for (int b = 0; b < ng; b += KB * nt) {
			double2 v[KB];
#pragma unroll
			for (int u = 0; u < KB; u++) {
				int r;
				const int mi = qdiv(min(b + u * nt + tid, ng - 1), plan.RBq, r);
				v[u] = src[(size_t)mi * nlat + min(j0 + r, nlat - 1)];
			}
#pragma unroll
			for (int u = 0; u < KB; u++) {
				const int idx = b + u * nt + tid;
				if (idx >= ng) break;
				int r;
				const int mi = qdiv(idx, plan.RBq, r);
				int k = mi - L;
				if (k < 0) k += N;
				bufA[lidx(r, k, NP)] = (j0 + r < nlat) ? v[u] : make_double2(0, 0);
			}
		}
		__syncthreads();
		double2 *res = fft_lds(bufA, bufB, plan, nseq, tw, +1);
		double2 *dst = (double2 *)out + f * nspat;
		for (int idx = tid; idx < N * RB; idx += nt) {
			int r;
			const int i = qdiv(idx, plan.RBq, r);
			const int j = j0 + r;
			if (j < nlat) dst[(size_t)i * nlat + j] = res[lidx(r, i, NP)];
		}
	}
}

// MODE_R2C: nf real fields [i * nlat + j] -> Fourier rows (L + 1, nlat);
// MODE_C2R: back.  nlat is even and N > 2 L (checked by the launcher).
// The first radix pass (Ns = 1) runs on values loaded from HBM and writes
// the LDS tile; the middle passes run in place; the last pass (Ns = N / R)
// reads the tile and writes HBM.  In the r2c direction a thread takes the
// last-pass butterflies k and Ns - k (0 and Ns / 2 together), whose outputs
// hold both Z[m] and Z[N - m] of every order m they produce, so the split
// into the two rings of a pair needs no further exchange.
template <int MODE, int N, int NSEQ, int NT>
__global__ __launch_bounds__(NT) void k_fft_wide(const void *__restrict__ in, void *__restrict__ out,
						 const double2 *__restrict__ tw, int nlat, int L, int nblk, int nblk_pad,
						 double norm)
{
	typedef WPad<N, NSEQ> W;
	typedef CPlan<N> PL;
	constexpr int RB = 2 * NSEQ;
	constexpr int NP = PL::npass();
	static_assert(NP >= 2, "two radix passes at least");
	constexpr int R0 = PL::radix(0), NB0 = N / R0;
	constexpr int RL = PL::radix(NP - 1), NSL = PL::ns(NP - 1);
	static_assert(NSL * RL == N && NSL % 2 == 0, "last pass");
	constexpr int SIGN = MODE == MODE_R2C ? -1 : 1;
	extern __shared__ double2 sm[];
	const int lb = xcd_remap(blockIdx.x, nblk_pad);
	if (lb >= nblk) return;
	const int f = blockIdx.y;
	const size_t nspat = (size_t)nlat * N;
	const int tid = threadIdx.x;
	const int j0 = lb * RB;
	const int h = nlat >> 1;   // ring pairs per longitude
	const int p0 = j0 >> 1;    // first ring pair of the tile

	// ---- first pass from HBM ----
	for (int it = tid; it < NSEQ * NB0; it += NT) {
		const int b = it / NSEQ, sq = it % NSEQ;
		const bool live = j0 + 2 * sq < nlat;
		double2 v[R0];
		if (MODE == MODE_R2C) {
			const double2 *src = (const double2 *)((const double *)in + f * nspat);
#pragma unroll
			for (int t = 0; t < R0; t++)
				v[t] = live ? src[(size_t)(b + t * NB0) * h + p0 + sq] : make_double2(0, 0);
		} else {
			// z[k] = A + iB (k <= L), conj(A) + i conj(B) of order N - k (k >= N - L)
			const double2 *src = (const double2 *)in + (size_t)f * (L + 1) * nlat;
			double2 A[R0], B[R0];
#pragma unroll
			for (int t = 0; t < R0; t++) {
				const int k = b + t * NB0;
				const int m = k <= L ? k : (k >= N - L ? N - k : -1);
				if (live && m >= 0) {
					const size_t o = (size_t)m * nlat + j0 + 2 * sq;
					A[t] = src[o];
					B[t] = src[o + 1];
				} else {
					A[t] = B[t] = make_double2(0, 0);
				}
			}
#pragma unroll
			for (int t = 0; t < R0; t++) {
				const int k = b + t * NB0;
				if (k == 0) v[t] = make_double2(A[t].x, B[t].x);
				else if (k <= L) v[t] = make_double2(A[t].x - B[t].y, A[t].y + B[t].x);
				else v[t] = make_double2(A[t].x + B[t].y, -A[t].y + B[t].x);
			}
		}
		Dft<R0>::run(v, SIGN);
#pragma unroll
		for (int t = 0; t < R0; t++) sm[W::at(sq, b * R0 + t)] = v[t];
	}
	__syncthreads();
	passes_ip_range<N, NSEQ, NT, SIGN, 1, NP - 1>(sm, tw);

	// ---- last pass to HBM ----
	if (MODE == MODE_R2C) {
		double2 *dst = (double2 *)out + (size_t)f * (L + 1) * nlat;
		constexpr int NH = NSL / 2;
		const double hn = 0.5 * norm;
		for (int it = tid; it < NSEQ * NH; it += NT) {
			const int pp = it / NSEQ, sq = it % NSEQ;
			const int j = j0 + 2 * sq;
			const int k1 = pp, k2 = pp == 0 ? NH : NSL - pp;
			double2 z1[RL], z2[RL];
#pragma unroll
			for (int t = 0; t < RL; t++) {
				z1[t] = sm[W::at(sq, k1 + t * NSL)];
				z2[t] = sm[W::at(sq, k2 + t * NSL)];
			}
#pragma unroll
			for (int t = 1; t < RL; t++) {
				z1[t] = cmul(z1[t], tw[t * k1]);
				z2[t] = cmul(z2[t], tw[t * k2]);
			}
			Dft<RL>::run(z1, SIGN);
			Dft<RL>::run(z2, SIGN);
			if (j >= nlat) continue;
			// orders m = k + t NSL of butterflies k1 and k2, with their partners
			auto emit = [&](int m, double2 Z, double2 P) __attribute__((always_inline)) {
				if (m > L) return;
				const double2 Zc = cconj(P);
				double2 a = make_double2(hn * (Z.x + Zc.x), hn * (Z.y + Zc.y));
				double2 c = make_double2(hn * (Z.y - Zc.y), -hn * (Z.x - Zc.x));
				if (m == 0) a.y = c.y = 0.0;
				dst[(size_t)m * nlat + j] = a;
				dst[(size_t)m * nlat + j + 1] = c;
			};
#pragma unroll
			for (int t = 0; t < RL; t++) {
				const double2 q1 = pp == 0 ? z1[(RL - t) % RL] : z2[RL - 1 - t];
				const double2 q2 = pp == 0 ? z2[RL - 1 - t] : z1[RL - 1 - t];
				emit(k1 + t * NSL, z1[t], q1);
				emit(k2 + t * NSL, z2[t], q2);
			}
		}
	} else {
		double2 *dst = (double2 *)((double *)out + f * nspat);
		for (int it = tid; it < NSEQ * NSL; it += NT) {
			const int k = it / NSEQ, sq = it % NSEQ;
			double2 v[RL];
#pragma unroll
			for (int t = 0; t < RL; t++) v[t] = sm[W::at(sq, k + t * NSL)];
#pragma unroll
			for (int t = 1; t < RL; t++) {
				double2 w = tw[t * k];
				w.y = -w.y;
				v[t] = cmul(v[t], w);
			}
			Dft<RL>::run(v, SIGN);
			if (j0 + 2 * sq < nlat) {
#pragma unroll
				for (int t = 0; t < RL; t++) dst[(size_t)(k + t * NSL) * h + p0 + sq] = v[t];
			}
		}
	}
}

// ---- FP32, ring-major internal fields ---------------------------------------
// The mixed-precision SWE stepper keeps its physical fields in FP32, ring-major
// (float[f][j][N], every ring one contiguous row), and its Fourier rows in FP32
// (float2[f][L + 1][nlat], the layout above).  A workgroup owns 2 NSEQ adjacent
// rings; sequence s is z_s[i] = f(j0 + 2s, i) + i f(j0 + 2s + 1, i).  The
// physical side is read and written a ring row at a time (consecutive lanes,
// consecutive longitudes); the Fourier side as NSEQ ring pairs of one order,
// 16 NSEQ contiguous bytes.  Passes, radices and the split into the two rings
// of a pair are those of k_fft_wide, in FP32 with FP32 twiddles.
template <int MODE, int N, int NSEQ, int NT>
__global__ __launch_bounds__(NT) void k_fft_rm32(const void *__restrict__ in, void *__restrict__ out,
						 const float2 *__restrict__ tw, int nlat, int L, int nblk, int nblk_pad,
						 float norm)
{
	typedef WPad<N, NSEQ> W;
	typedef CPlan<N> PL;
	constexpr int RB = 2 * NSEQ;
	constexpr int NP = PL::npass();
	static_assert(NP >= 2, "two radix passes at least");
	constexpr int R0 = PL::radix(0), NB0 = N / R0;
	constexpr int RL = PL::radix(NP - 1), NSL = PL::ns(NP - 1);
	static_assert(NSL * RL == N && NSL % 2 == 0, "last pass");
	constexpr int SIGN = MODE == MODE_R2C ? -1 : 1;
	extern __shared__ float2 smf[];
	const int lb = xcd_remap(blockIdx.x, nblk_pad);
	if (lb >= nblk) return;
	const int f = blockIdx.y;
	const size_t nspat = (size_t)nlat * N;
	const int tid = threadIdx.x;
	const int j0 = lb * RB;

	// ---- first pass from HBM ----
	for (int it = tid; it < NSEQ * NB0; it += NT) {
		const int b = it % NB0, sq = it / NB0;
		const int j = j0 + 2 * sq;
		const bool live = j < nlat;
		float2 v[R0];
		if (MODE == MODE_R2C) {
			const float *ra = (const float *)in + f * nspat + (size_t)(live ? j : 0) * N;
			const float *rb = ra + N;
#pragma unroll
			for (int t = 0; t < R0; t++)
				v[t] = live ? make_float2(ra[b + t * NB0], rb[b + t * NB0]) : make_float2(0.f, 0.f);
		} else {
			// z[k] = A + iB (k <= L), conj(A) + i conj(B) of order N - k (k >= N - L)
			const float2 *src = (const float2 *)in + (size_t)f * (L + 1) * nlat;
			float2 A[R0], B[R0];
#pragma unroll
			for (int t = 0; t < R0; t++) {
				const int k = b + t * NB0;
				const int m = k <= L ? k : (k >= N - L ? N - k : -1);
				if (live && m >= 0) {
					const float4 ab = *(const float4 *)(src + (size_t)m * nlat + j);
					A[t] = make_float2(ab.x, ab.y);
					B[t] = make_float2(ab.z, ab.w);
				} else {
					A[t] = B[t] = make_float2(0.f, 0.f);
				}
			}
#pragma unroll
			for (int t = 0; t < R0; t++) {
				const int k = b + t * NB0;
				if (k == 0) v[t] = make_float2(A[t].x, B[t].x);
				else if (k <= L) v[t] = make_float2(A[t].x - B[t].y, A[t].y + B[t].x);
				else v[t] = make_float2(A[t].x + B[t].y, -A[t].y + B[t].x);
			}
		}
		Dft<R0>::run(v, SIGN);
#pragma unroll
		for (int t = 0; t < R0; t++) smf[W::at(sq, b * R0 + t)] = v[t];
	}
	__syncthreads();
	passes_ip_range<N, NSEQ, NT, SIGN, 1, NP - 1>(smf, tw);

	// ---- last pass to HBM ----
	if (MODE == MODE_R2C) {
		float2 *dst = (float2 *)out + (size_t)f * (L + 1) * nlat;
		constexpr int NH = NSL / 2;
		const float hn = 0.5f * norm;
		for (int it = tid; it < NSEQ * NH; it += NT) {
			const int pp = it / NSEQ, sq = it % NSEQ;
			const int j = j0 + 2 * sq;
			const int k1 = pp, k2 = pp == 0 ? NH : NSL - pp;
			float2 z1[RL], z2[RL];
#pragma unroll
			for (int t = 0; t < RL; t++) {
				z1[t] = smf[W::at(sq, k1 + t * NSL)];
				z2[t] = smf[W::at(sq, k2 + t * NSL)];
			}
#pragma unroll
			for (int t = 1; t < RL; t++) {
				z1[t] = cmul(z1[t], tw[t * k1]);
				z2[t] = cmul(z2[t], tw[t * k2]);
			}
			Dft<RL>::run(z1, SIGN);
			Dft<RL>::run(z2, SIGN);
			if (j >= nlat) continue;
			// orders m = k + t NSL of butterflies k1 and k2, with their partners
			auto emit = [&](int m, float2 Z, float2 P) __attribute__((always_inline)) {
				if (m > L) return;
				const float2 Zc = cconj(P);
				float2 a = make_float2(hn * (Z.x + Zc.x), hn * (Z.y + Zc.y));
				float2 c = make_float2(hn * (Z.y - Zc.y), -hn * (Z.x - Zc.x));
				if (m == 0) a.y = c.y = 0.f;
				*(float4 *)(dst + (size_t)m * nlat + j) = make_float4(a.x, a.y, c.x, c.y);
			};
#pragma unroll
			for (int t = 0; t < RL; t++) {
				const float2 q1 = pp == 0 ? z1[(RL - t) % RL] : z2[RL - 1 - t];
				const float2 q2 = pp == 0 ? z2[RL - 1 - t] : z1[RL - 1 - t];
				emit(k1 + t * NSL, z1[t], q1);
				emit(k2 + t * NSL, z2[t], q2);
			}
		}
	} else {
		float *dst = (float *)out + f * nspat;
		for (int it = tid; it < NSEQ * NSL; it += NT) {
			const int k = it % NSL, sq = it / NSL;
			const int j = j0 + 2 * sq;
			float2 v[RL];
#pragma unroll
			for (int t = 0; t < RL; t++) v[t] = smf[W::at(sq, k + t * NSL)];
#pragma unroll
			for (int t = 1; t < RL; t++) {
				float2 w = tw[t * k];
				w.y = -w.y;
				v[t] = cmul(v[t], w);
			}
			Dft<RL>::run(v, SIGN);
			if (j < nlat) {
				float *ra = dst + (size_t)j * N, *rb = ra + N;
#pragma unroll
				for (int t = 0; t < RL; t++) {
					ra[k + t * NSL] = v[t].x;
					rb[k + t * NSL] = v[t].y;
				}
			}
		}
	}
}

template <int MODE, int N, int NSEQ>
int launch_rm32(sphswe_ctx *ctx, const void *in, void *out, int nf)
{
	constexpr int RB = 2 * NSEQ;
	constexpr int NT = 256;
	const int nblk = (ctx->nlat + RB - 1) / RB;
	const int nblk_pad = (nblk + 7) / 8 * 8;
	const size_t lds = (size_t)NSEQ * WPad<N, NSEQ>::len * sizeof(float2);
	static_assert(NSEQ * WPad<N, NSEQ>::len * sizeof(float2) <= 160 * 1024, "tile exceeds the LDS");
	if (lds > 65536)
		SPHSWE_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_rm32<MODE, N, NSEQ, NT>,
						    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	dim3 grid(nblk_pad, nf);
	k_fft_rm32<MODE, N, NSEQ, NT><<<grid, NT, lds, ctx->stream>>>(in, out, ctx->d_tw32, ctx->nlat, ctx->L, nblk,
								       nblk_pad, (float)(2.0 * M_PI / N));
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

// the FP32 ring-major transform for the grid sizes of the usual truncations,
// or an error for other sizes
template <int MODE>
int launch_rm32_fixed(sphswe_ctx *ctx, const void *in, void *out, int nf)
{
	if (ctx->nlat % 2 != 0 || 2 * ctx->L + 1 > ctx->plan.n)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "FP32 FFT: grid %d x %d", ctx->nlat, ctx->plan.n);
	if (!ctx->d_tw32) return sphswe_set_err(ctx, SPHSWE_ERR_INTERNAL, "FP32 FFT: no FP32 twiddles");
	switch (ctx->plan.n) {
	case 96: return launch_rm32<MODE, 96, 4>(ctx, in, out, nf);
	case 196: return launch_rm32<MODE, 196, 4>(ctx, in, out, nf);
	case 384: return launch_rm32<MODE, 384, 4>(ctx, in, out, nf);
	case 784: return launch_rm32<MODE, 784, 4>(ctx, in, out, nf);
	case 1536: return launch_rm32<MODE, 1536, 4>(ctx, in, out, nf);
	case 3072: return launch_rm32<MODE, 3072, 4>(ctx, in, out, nf);
	}
	return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "FP32 FFT: no compile-time transform of length %d", ctx->plan.n);
}

template <int MODE, int N, int NSEQ>
int launch_real_wide(sphswe_ctx *ctx, const void *in, void *out, int nf)
{
	constexpr int RB = 2 * NSEQ;
	constexpr int NT = NSEQ * N >= 6144 ? 512 : 256;
	const int nblk = (ctx->nlat + RB - 1) / RB;
	const int nblk_pad = (nblk + 7) / 8 * 8;
	const size_t lds = (size_t)NSEQ * WPad<N, NSEQ>::len * sizeof(double2);
	static_assert(NSEQ * WPad<N, NSEQ>::len * sizeof(double2) <= 160 * 1024, "tile exceeds the LDS");
	if (lds > 65536)
		SPHSWE_HIP(ctx, hipFuncSetAttribute((const void *)k_fft_wide<MODE, N, NSEQ, NT>,
						    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	dim3 grid(nblk_pad, nf);
	k_fft_wide<MODE, N, NSEQ, NT><<<grid, NT, lds, ctx->stream>>>(in, out, ctx->d_tw, ctx->nlat, ctx->L, nblk, nblk_pad,
								       2.0 * M_PI / N);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

// the compile-time kernel for the grid sizes of the usual truncations, or -1.
// nseq: complex sequences per workgroup (2 nseq rings); 0 takes the default,
// a value without an instantiation for this size falls back to it
template <int MODE>
int launch_real_fixed(sphswe_ctx *ctx, const void *in, void *out, int nf, int nseq)
{
	if (ctx->nlat % 2 != 0 || 2 * ctx->L + 1 > ctx->plan.n) return -1;
	switch (ctx->plan.n) {
	case 96:
		if (nseq == 16) return launch_real_wide<MODE, 96, 16>(ctx, in, out, nf);
		return launch_real_wide<MODE, 96, 8>(ctx, in, out, nf);
	case 196:
		if (nseq == 4) return launch_real_wide<MODE, 196, 4>(ctx, in, out, nf);
		return launch_real_wide<MODE, 196, 8>(ctx, in, out, nf);
	case 384:
		if (nseq == 2) return launch_real_wide<MODE, 384, 2>(ctx, in, out, nf);
		if (nseq == 8) return launch_real_wide<MODE, 384, 8>(ctx, in, out, nf);
		return launch_real_wide<MODE, 384, 4>(ctx, in, out, nf);
	case 784:
		if (nseq == 1) return launch_real_wide<MODE, 784, 1>(ctx, in, out, nf);
		if (nseq == 4) return launch_real_wide<MODE, 784, 4>(ctx, in, out, nf);
		return launch_real_wide<MODE, 784, 2>(ctx, in, out, nf);
	case 1536:
		if (nseq == 1) return launch_real_wide<MODE, 1536, 1>(ctx, in, out, nf);
		if (nseq == 4) return launch_real_wide<MODE, 1536, 4>(ctx, in, out, nf);
		return launch_real_wide<MODE, 1536, 2>(ctx, in, out, nf);
	case 3072:
		if (nseq == 1) return launch_real_wide<MODE, 3072, 1>(ctx, in, out, nf);
		return launch_real_wide<MODE, 3072, 2>(ctx, in, out, nf);
	}
	return -1;
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
	int Ns = 1;
	for (int i = 0; i < SPHSWE_MAX_RADIX; i++) {
		a.radix[i] = ctx->plan.radix[i];
		if (i < a.npass) {
			a.nbq[i] = qdiv_of(a.N / a.radix[i]);
			a.nsq[i] = qdiv_of(Ns);
			Ns *= a.radix[i];
		} else {
			a.nbq[i] = a.nsq[i] = qdiv_of(1);
		}
	}
	int nseq = seqs_per_block(a.N);
	int rings = (MODE == MODE_R2C || MODE == MODE_C2R) ? 2 * nseq : nseq;
	a.Nq = qdiv_of(a.N);
	a.RBq = qdiv_of(rings);
	a.Lq = qdiv_of(ctx->L + 1);
	int nblk = (ctx->nlat + rings - 1) / rings;
	int nblk_pad = (nblk + 7) / 8 * 8;
	// two padded buffers of nseq sequences (the c2r staging of G uses
	// 2 (L + 1) <= N + 1 < padded_len(N) entries per sequence)
	size_t lds = 2 * (size_t)nseq * padded_len(a.N) * sizeof(double2);
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
	if (nf <= 0) return SPHSWE_OK;
	const int rc = launch_real_fixed<MODE_R2C>(ctx, spat, four, nf, ctx->fft_nseq_r2c);
	return rc >= 0 ? rc : launch<MODE_R2C>(ctx, spat, four, nf);
}
int sphswe_fft_r2c32(sphswe_ctx *ctx, const float *spat, float2 *four, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	return launch_rm32_fixed<MODE_R2C>(ctx, spat, four, nf);
}
int sphswe_fft_c2r32(sphswe_ctx *ctx, const float2 *four, float *spat, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	return launch_rm32_fixed<MODE_C2R>(ctx, four, spat, nf);
}
// whether the FP32 transforms have a compile-time kernel for this grid
bool sphswe_fft32_available(const sphswe_ctx *ctx)
{
	if (ctx->nlat % 2 != 0 || 2 * ctx->L + 1 > ctx->plan.n) return false;
	switch (ctx->plan.n) {
	case 96: case 196: case 384: case 784: case 1536: case 3072: return true;
	}
	return false;
}
int sphswe_fft_c2r(sphswe_ctx *ctx, const double2 *four, double *spat, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	const int rc = launch_real_fixed<MODE_C2R>(ctx, four, spat, nf, ctx->fft_nseq_c2r);
	return rc >= 0 ? rc : launch<MODE_C2R>(ctx, four, spat, nf);
}
// the sequences-per-workgroup candidates of the compile-time real FFT of this
// grid (empty for other sizes), for the autotuner
int sphswe_fft_nseq_candidates(const sphswe_ctx *ctx, int *out, int max)
{
	static const int c96[] = {8, 16}, c196[] = {4, 8}, c384[] = {2, 4, 8}, c784[] = {1, 2, 4}, c1536[] = {1, 2, 4},
			 c3072[] = {1, 2};
	const int *c = nullptr;
	int n = 0;
	switch (ctx->plan.n) {
	case 96: c = c96; n = 2; break;
	case 196: c = c196; n = 2; break;
	case 384: c = c384; n = 3; break;
	case 784: c = c784; n = 3; break;
	case 1536: c = c1536; n = 3; break;
	case 3072: c = c3072; n = 2; break;
	}
	for (int i = 0; i < n && i < max; i++) out[i] = c[i];
	return n < max ? n : max;
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
