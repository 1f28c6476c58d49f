// Building blocks of the longitude FFTs, shared by the FFT kernels and the
// fused transform-and-pointwise kernel of the mixed SWE step: small DFTs
// written once for double2 and float2, the compile-time radix plan of a
// length, the padded LDS tile of NSEQ sequences, and one in-place Stockham
// pass of a tile.
#ifndef SPHSWE_FFT_BLOCKS_H
#define SPHSWE_FFT_BLOCKS_H

#include "sphswe_common.h"

namespace {

// the butterflies below are written once for double2 and float2; RealOf<C>
// is the scalar type, constants are rounded to it once
template <typename C>
struct RealOf;
template <>
struct RealOf<double2> {
	typedef double T;
};
template <>
struct RealOf<float2> {
	typedef float T;
};
template <typename C>
__device__ __forceinline__ C cmk(typename RealOf<C>::T re, typename RealOf<C>::T im)
{
	C z;
	z.x = re;
	z.y = im;
	return z;
}

template <typename C>
__device__ inline C mul_i(C a, int sign)  // sign * i * a
{
	return sign > 0 ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x);
}

template <int R>
struct Dft;

template <>
struct Dft<2> {
	template <typename C>
	__device__ static void run(C *v, int)
	{
		C a = v[0], b = v[1];
		v[0] = cadd(a, b);
		v[1] = csub(a, b);
	}
};

template <>
struct Dft<4> {
	template <typename C>
	__device__ static void run(C *v, int sign)
	{
		C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
		C a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
		C ia3 = mul_i(a3, sign);   // sign*i*a3
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
	template <typename C>
	__device__ static void run(C *v, int sign)
	{
		typedef typename RealOf<C>::T Rl;
		C out[R];
#pragma unroll
		for (int q = 0; q < R; q++) {
			C acc = v[0];
#pragma unroll
			for (int t = 1; t < R; t++) {
				const int e = (q * t) % R;
				const Rl wr = (Rl)OddTab<R>::c[e], wi = (Rl)(sign * OddTab<R>::s[e]);
				acc.x += v[t].x * wr - v[t].y * wi;
				acc.y += v[t].x * wi + v[t].y * wr;
			}
			out[q] = acc;
		}
#pragma unroll
		for (int q = 0; q < R; q++) v[q] = out[q];
	}
};
// radix 3: X0 = x0 + (x1 + x2), X1,2 = x0 - (x1 + x2) / 2 +- sign i sqrt(3)/2 (x1 - x2)
template <>
struct Dft<3> {
	template <typename C>
	__device__ static void run(C *v, int sign)
	{
		typedef typename RealOf<C>::T Rl;
		constexpr Rl s3 = (Rl)0.86602540378443864676, hf = (Rl)0.5;
		const C t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
		const C m1 = cmk<C>(v[0].x - hf * t1.x, v[0].y - hf * t1.y);
		const C m2 = cmk<C>(-sign * s3 * t2.y, sign * s3 * t2.x);
		v[0] = cadd(v[0], t1);
		v[1] = cadd(m1, m2);
		v[2] = csub(m1, m2);
	}
};
template <>
struct Dft<5> : DftOdd<5> {};
template <>
struct Dft<7> : DftOdd<7> {};

// ----------------------------------------------------------------------------
// real transforms at compile-time sizes: radix-16/8 passes, constant strides
// ----------------------------------------------------------------------------

// DFT of 8: radix 2 then two DFTs of 4, w8^n = exp(sign i 2 pi n / 8)
template <>
struct Dft<8> {
	template <typename C>
	__device__ static void run(C *v, int sign)
	{
		typedef typename RealOf<C>::T Rl;
		constexpr Rl h = (Rl)0.70710678118654752440;
		C a[4], b[4];
#pragma unroll
		for (int n = 0; n < 4; n++) {
			a[n] = cadd(v[n], v[n + 4]);
			b[n] = csub(v[n], v[n + 4]);
		}
		b[1] = cmul(b[1], cmk<C>(h, sign * h));
		b[2] = mul_i(b[2], sign);
		b[3] = cmul(b[3], cmk<C>(-h, sign * h));
		Dft<4>::run(a, sign);
		Dft<4>::run(b, sign);
#pragma unroll
		for (int k = 0; k < 4; k++) {
			v[2 * k] = a[k];
			v[2 * k + 1] = b[k];
		}
	}
};

// DFT of 16 as 4 x 4: DFTs of 4 over n1 of x[4 n1 + n2], twiddles
// w16^(n2 k1), DFTs of 4 over n2; X[k1 + 4 k2]
template <>
struct Dft<16> {
	template <typename C>
	__device__ static void run(C *v, int sign)
	{
		typedef typename RealOf<C>::T Rl;
		constexpr Rl c1 = (Rl)0.92387953251128675613, s1 = (Rl)0.38268343236508977173, h = (Rl)0.70710678118654752440;
		// w16^e for e = n2 k1 in {0, 1, 2, 3, 4, 6, 9}: (cos, sin) of 2 pi e / 16
		C y[4][4];
#pragma unroll
		for (int n2 = 0; n2 < 4; n2++) {
			C t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
			Dft<4>::run(t, sign);
#pragma unroll
			for (int k1 = 0; k1 < 4; k1++) y[n2][k1] = t[k1];
		}
		const C w1 = cmk<C>(c1, sign * s1), w2 = cmk<C>(h, sign * h), w3 = cmk<C>(s1, sign * c1);
		const C w6 = cmk<C>(-h, sign * h), w9 = cmk<C>(-c1, -sign * s1);
		y[1][1] = cmul(y[1][1], w1);
		y[1][2] = cmul(y[1][2], w2);
		y[1][3] = cmul(y[1][3], w3);
		y[2][1] = cmul(y[2][1], w2);
		y[2][2] = mul_i(y[2][2], sign);
		y[2][3] = cmul(y[2][3], w6);
		y[3][1] = cmul(y[3][1], w3);
		y[3][2] = cmul(y[3][2], w6);
		y[3][3] = cmul(y[3][3], w9);
#pragma unroll
		for (int k1 = 0; k1 < 4; k1++) {
			C t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
			Dft<4>::run(t, sign);
#pragma unroll
			for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = t[k2];
		}
	}
};

// DFT of R = A B as B-point DFTs over n1 of x[A n1 + n2], twiddles
// w_R^(n2 k1), A-point DFTs over n2; X[k1 + B k2].  R = 6 (A = 2) and R = 12
// (A = 4) with B = 3; the twiddles are multiples of 30 degrees.
template <int A>
struct DftX3 {
	template <typename C>
	__device__ static void run(C *v, int sign)
	{
		typedef typename RealOf<C>::T Rl;
		constexpr int R = 3 * A;
		constexpr double c30 = 0.86602540378443864676;
		// (cos, sin) of 30 e degrees, e = 0..6
		constexpr double cw[7] = {1.0, c30, 0.5, 0.0, -0.5, -c30, -1.0};
		constexpr double sw[7] = {0.0, 0.5, c30, 1.0, c30, 0.5, 0.0};
		C y[A][3];
#pragma unroll
		for (int n2 = 0; n2 < A; n2++) {
			C t[3] = {v[n2], v[A + n2], v[2 * A + n2]};
			Dft<3>::run(t, sign);
#pragma unroll
			for (int k1 = 0; k1 < 3; k1++) {
				const int e = (n2 * k1) * (12 / R);   // in units of 30 degrees
				y[n2][k1] = e == 0 ? t[k1] : cmul(t[k1], cmk<C>((Rl)cw[e], (Rl)(sign * sw[e])));
			}
		}
#pragma unroll
		for (int k1 = 0; k1 < 3; k1++) {
			C t[A];
#pragma unroll
			for (int n2 = 0; n2 < A; n2++) t[n2] = y[n2][k1];
			Dft<A>::run(t, sign);
#pragma unroll
			for (int k2 = 0; k2 < A; k2++) v[k1 + 3 * k2] = t[k2];
		}
	}
};
template <>
struct Dft<6> : DftX3<2> {};
template <>
struct Dft<12> : DftX3<4> {};

// radix of the next pass for a remaining length n: 16 while it divides, a
// last 12 or 6 in one pass, then 8, 4, 2, 3, 5, 7
// RULE 0: 16 while it divides, a last 12 or 6 in one pass, then 8, 4, 2, 3, 5, 7
// (few passes of large radix: the tile kernels, whose threads each take many
// butterflies); RULE 1: 8 while it divides, then 4, 2, 3, 5, 7 (more, smaller
// butterflies per pass: a workgroup of 256 threads on one sequence keeps three
// quarters of its threads busy at N = 1536 = 8 8 8 3, against three eighths
// with 16 16 6)
template <int RULE>
constexpr int first_radix(int n)
{
	if (RULE == 1) return n % 8 == 0 ? 8 : n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : 7;
	return n % 16 == 0 ? 16 : n == 12 ? 12 : n == 6 ? 6 : n % 8 == 0 ? 8 : n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : n % 5 == 0 ? 5 : 7;
}
template <int N, int RULE = 0>
struct CPlan {
	static constexpr int npass()
	{
		int n = N, c = 0;
		while (n > 1) {
			n /= first_radix<RULE>(n);
			c++;
		}
		return c;
	}
	static constexpr int radix(int p)
	{
		int n = N;
		for (int i = 0; i < p; i++) n /= first_radix<RULE>(n);
		return first_radix<RULE>(n);
	}
	static constexpr int ns(int p)
	{
		int s = 1, n = N;
		for (int i = 0; i < p; i++) {
			const int r = first_radix<RULE>(n);
			s *= r;
			n /= r;
		}
		return s;
	}
};

// ---- reference layout, wide ring tiles ----------------------------------------
// A workgroup owns 2 NSEQ adjacent rings; ring pair s is the complex sequence
// z_s[i] = f(i, j0 + 2s) + i f(i, j0 + 2s + 1), which is one 16-byte load or
// store at [i * nlat + j0 + 2s], so a row of the tile is 32 NSEQ contiguous
// bytes.  The sequences live in one LDS buffer; every pass reads its inputs
// into registers, waits at a barrier and writes its outputs in place, so the
// tile is twice as wide as with two buffers.  The sequence stride is padded to
// 8 / NSEQ (mod 8) elements beyond the every-16 padding, so the 8 lanes of a
// 128-byte LDS access that hold one row's ring pairs (and the next rows) fall
// on distinct banks.
template <int N, int NSEQ>
struct WPad {
	static constexpr int base = N + N / 16;
	static constexpr int want = NSEQ >= 8 ? 1 : 8 / NSEQ;
	static constexpr int len = base + ((want - base % 8) + 8) % 8;
	__device__ static __forceinline__ int at(int s, int i) { return s * len + i + (i >> 4); }
};

template <int N, int NSEQ, int NT, int R, int Ns, int SIGN, typename C>
__device__ __forceinline__ void pass_ip(C *__restrict__ buf, const C *__restrict__ tw)
{
	constexpr int NB = N / R, TOT = NSEQ * NB, TWS = N / (Ns * R);
	constexpr int IT = (TOT + NT - 1) / NT;
	C v[IT][R];
#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int idx = (int)threadIdx.x + it * NT;
		if (TOT % NT != 0 && idx >= TOT) break;
		const int sq = idx / NB, b = idx % NB;
		const int k = b % Ns;
#pragma unroll
		for (int t = 0; t < R; t++) v[it][t] = buf[WPad<N, NSEQ>::at(sq, b + t * NB)];
		if (Ns > 1) {
#pragma unroll
			for (int t = 1; t < R; t++) {
				C w = tw[t * k * TWS];
				if (SIGN > 0) w.y = -w.y;
				v[it][t] = cmul(v[it][t], w);
			}
		}
		Dft<R>::run(v[it], SIGN);
	}
	__syncthreads();
#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int idx = (int)threadIdx.x + it * NT;
		if (TOT % NT != 0 && idx >= TOT) break;
		const int sq = idx / NB, b = idx % NB;
		const int k = b % Ns, jb = b / Ns;
		const int o = jb * Ns * R + k;
#pragma unroll
		for (int t = 0; t < R; t++) buf[WPad<N, NSEQ>::at(sq, o + t * Ns)] = v[it][t];
	}
	__syncthreads();
}

// the same pass from one LDS buffer into another: one barrier per pass, twice
// the LDS of the in-place form
template <int N, int NSEQ, int NT, int R, int Ns, int SIGN, typename C>
__device__ __forceinline__ void pass_pp(const C *__restrict__ src, C *__restrict__ dst, const C *__restrict__ tw)
{
	constexpr int NB = N / R, TOT = NSEQ * NB, TWS = N / (Ns * R);
	constexpr int IT = (TOT + NT - 1) / NT;
#pragma unroll
	for (int it = 0; it < IT; it++) {
		const int idx = (int)threadIdx.x + it * NT;
		if (TOT % NT != 0 && idx >= TOT) break;
		const int sq = idx / NB, b = idx % NB;
		const int k = b % Ns, jb = b / Ns;
		C v[R];
#pragma unroll
		for (int t = 0; t < R; t++) v[t] = src[WPad<N, NSEQ>::at(sq, b + t * NB)];
		if (Ns > 1) {
#pragma unroll
			for (int t = 1; t < R; t++) {
				C w = tw[t * k * TWS];
				if (SIGN > 0) w.y = -w.y;
				v[t] = cmul(v[t], w);
			}
		}
		Dft<R>::run(v, SIGN);
		const int o = jb * Ns * R + k;
#pragma unroll
		for (int t = 0; t < R; t++) dst[WPad<N, NSEQ>::at(sq, o + t * Ns)] = v[t];
	}
	__syncthreads();
}

// passes P .. PEND - 1 alternating between buffers a and b, starting from a;
// the result is in a when PEND - P is even, in b otherwise
template <int N, int NSEQ, int NT, int SIGN, int P, int PEND, typename C, int RULE = 0>
__device__ __forceinline__ void passes_pp_range(C *a, C *b, const C *__restrict__ tw)
{
	if constexpr (P < PEND) {
		pass_pp<N, NSEQ, NT, CPlan<N, RULE>::radix(P), CPlan<N, RULE>::ns(P), SIGN>(a, b, tw);
		passes_pp_range<N, NSEQ, NT, SIGN, P + 1, PEND, C, RULE>(b, a, tw);
	}
}

template <int N, int NSEQ, int NT, int SIGN, int P, typename C>
__device__ __forceinline__ void passes_ip(C *buf, const C *__restrict__ tw)
{
	if constexpr (P < CPlan<N>::npass()) {
		pass_ip<N, NSEQ, NT, CPlan<N>::radix(P), CPlan<N>::ns(P), SIGN>(buf, tw);
		passes_ip<N, NSEQ, NT, SIGN, P + 1>(buf, tw);
	}
}

template <int N, int NSEQ, int NT, int SIGN, int P, int PEND, typename C, int RULE = 0>
__device__ __forceinline__ void passes_ip_range(C *buf, const C *__restrict__ tw)
{
	if constexpr (P < PEND) {
		pass_ip<N, NSEQ, NT, CPlan<N, RULE>::radix(P), CPlan<N, RULE>::ns(P), SIGN>(buf, tw);
		passes_ip_range<N, NSEQ, NT, SIGN, P + 1, PEND, C, RULE>(buf, tw);
	}
}

}  // namespace

#endif
