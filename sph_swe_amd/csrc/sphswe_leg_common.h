// Pieces shared by the FP64 and the mixed-precision Legendre kernels: the
// geometry handed to the kernels, the slot addressing of the spectral and
// Fourier layouts, and one step of the FP64 associated-Legendre recurrence.
#ifndef SPHSWE_LEG_COMMON_H
#define SPHSWE_LEG_COMMON_H

#include "sphswe_common.h"

#include <algorithm>

namespace {

struct Geo {
	const double2 *coef;
	const double2 *seed;    // (Pbar_s0, Pbar_s0+1) at each degree-segment start s0
	const int *segoff;
	const double *ct;
	const double *wg;
	const int *moff;        // pole layout: first row of each m block, index m + L
	int L, nlat, NJ, seg;
	int P;                  // pole layout: number of poles (row stride)
};

// slot addressing ------------------------------------------------------------
// LAY 0 real:            q(l) = spec[f*nlm + m(2L-m+1)/2 + l], Fourier row f*(L+1)*nlat + m*nlat
// LAY 1 complex:         q(l) = spec[f*ncplx + l*l + l + m],   Fourier row f*(2L+1)*nlat + (m+L)*nlat
// LAY 2 complex, poles:  q(l) = spec[(moff[m+L] + l - |m|)*P + p] (pole fastest), Fourier row as LAY 1
enum { LAY_REAL = 0, LAY_CPLX = 1, LAY_POLE = 2 };

template <int LAY>
struct Slot {
	long qb;      // spectral base
	long qs;      // spectral stride per degree (LAY_POLE)
	long fb;      // Fourier row base
	double sgn;   // (-1)^m for negative complex m
	bool valid;
	__device__ inline long q(int l, int am) const
	{
		return LAY == LAY_REAL ? qb + l : (LAY == LAY_CPLX ? qb + (long)l * l + l : qb + (long)(l - am) * qs);
	}
};

template <int LAY, int NS>
__device__ inline void make_slots(Slot<LAY> (&sl)[NS], int am, int group, int nfields, const Geo &g)
{
	const int L = g.L, nlat = g.nlat;
	const long nlm = (long)(L + 1) * (L + 2) / 2, ncplx = (long)(L + 1) * (L + 1);
#pragma unroll
	for (int s = 0; s < NS; s++) {
		if (LAY == LAY_REAL) {
			int f = group * NS + s;
			sl[s].valid = f < nfields;
			if (f >= nfields) f = nfields - 1;
			sl[s].qb = f * nlm + lm_index(L, 0, am);
			sl[s].qs = 1;
			sl[s].fb = (long)f * (L + 1) * nlat + (long)am * nlat;
			sl[s].sgn = 1.0;
		} else {
			int f = group * (NS / 2) + (s >> 1);
			int m = (s & 1) ? -am : am;
			sl[s].valid = f < nfields && !((s & 1) && am == 0);
			if (f >= nfields) f = nfields - 1;
			if (LAY == LAY_CPLX) {
				sl[s].qb = f * ncplx + m;
				sl[s].qs = 1;
			} else {
				sl[s].qb = (long)g.moff[m + L] * g.P + f;
				sl[s].qs = g.P;
			}
			sl[s].fb = (long)f * (2 * L + 1) * nlat + (long)(m + L) * nlat;
			sl[s].sgn = ((s & 1) && (am & 1)) ? -1.0 : 1.0;
		}
	}
}

// address of q(l) of slot ss of order am (what make_slots + Slot::q give),
// computed from ss without indexing a slot array
template <int LAY, int NS>
__device__ inline long slot_qaddr(int ss, int am, int l, int group, int nfields, const Geo &g)
{
	const int L = g.L;
	if (LAY == LAY_REAL) {
		int f = group * NS + ss;
		if (f >= nfields) f = nfields - 1;
		return (long)f * ((long)(L + 1) * (L + 2) / 2) + lm_index(L, l, am);
	}
	int f = group * (NS / 2) + (ss >> 1);
	if (f >= nfields) f = nfields - 1;
	const int m = (ss & 1) ? -am : am;
	if (LAY == LAY_CPLX) return (long)f * ((long)(L + 1) * (L + 1)) + (long)l * l + l + m;
	return ((long)g.moff[m + L] + (l - am)) * g.P + f;
}

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));

enum { PM_F64 = SPHSWE_PREC_FP64, PM_F32 = SPHSWE_PREC_FP32, PM_BF16 = SPHSWE_PREC_BF16, PM_BF16X3 = SPHSWE_PREC_BF16X3 };

template <int PM>
struct Terms {
	static constexpr int n = PM == PM_BF16 ? 2 : 3;
};

// bf16 of a finite x, rounded to nearest even
__device__ inline short bf16_rn(float x)
{
	const unsigned u = __float_as_uint(x);
	return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf16_f(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

// split four values into NT bf16 terms: term k is the bf16 of what the
// previous terms leave (each difference is exact in FP32)
template <int NT>
__device__ inline void split4(const float (&v)[4], s4v (&t)[NT])
{
#pragma unroll
	for (int e = 0; e < 4; e++) {
		float r = v[e];
#pragma unroll
		for (int k = 0; k < NT; k++) {
			const short h = bf16_rn(r);
			t[k][e] = h;
			r -= bf16_f(h);
		}
	}
}


// the column of an MFMA lane: slot = n / 2, re/im = n % 2; its validity, the
// sign (-1)^m of a negative-m complex slot, spectral address and Fourier row
// base, all from arithmetic on the slot index
template <int LAY, int NS>
struct Column {
	bool valid;
	double sgn;
	int f, m;
};

template <int LAY, int NS>
__device__ inline Column<LAY, NS> column_of(int slot, int am, int group, int nfields)
{
	Column<LAY, NS> c;
	if (LAY == LAY_REAL) {
		c.f = group * NS + slot;
		c.m = am;
		c.valid = slot < NS && c.f < nfields;
		c.sgn = 1.0;
	} else {
		c.f = group * (NS / 2) + (slot >> 1);
		c.m = (slot & 1) ? -am : am;
		c.valid = slot < NS && c.f < nfields && !((slot & 1) && am == 0);
		c.sgn = ((slot & 1) && (am & 1)) ? -1.0 : 1.0;
	}
	if (c.f >= nfields) c.f = nfields - 1;
	return c;
}

template <int LAY, int NS>
__device__ inline long column_q(const Column<LAY, NS> &c, int am, int l, const Geo &g)
{
	const int L = g.L;
	if (LAY == LAY_REAL) return (long)c.f * ((long)(L + 1) * (L + 2) / 2) + lm_index(L, l, am);
	if (LAY == LAY_CPLX) return (long)c.f * ((long)(L + 1) * (L + 1)) + (long)l * l + l + c.m;
	return ((long)g.moff[c.m + L] + (l - am)) * g.P + c.f;
}

template <int LAY, int NS>
__device__ inline long column_fb(const Column<LAY, NS> &c, const Geo &g)
{
	const int L = g.L, nlat = g.nlat;
	if (LAY == LAY_REAL) return (long)c.f * (L + 1) * nlat + (long)c.m * nlat;
	return (long)c.f * (2 * L + 1) * nlat + (long)(c.m + L) * nlat;
}

// one recurrence step: Pbar_l = (a mu) Pbar_{l-1} - (a b) Pbar_{l-2}
__device__ inline double rec_step(double2 c, double mu, double p1, double p0)
{
	return fma(c.x * mu, p1, -(c.y * p0));
}


// degrees 2..SEG-1 of a segment from (p0, p1) = degrees 0 and 1, with the
// coefficients cf[t] in LDS; store(t, value) writes the tile.  In batches of 8:
// 8 coefficient reads, 8 steps in registers, 8 stores (a coefficient read that
// follows a tile store would otherwise wait its full LDS latency every step).
template <int SEG, typename F>
__device__ __forceinline__ void rec_tile(const double2 *cf, double mu, double p0, double p1, F store)
{
#pragma unroll
	for (int tb = 2; tb < SEG; tb += 8) {
		double2 c[8];
		double v[8];
#pragma unroll
		for (int k = 0; k < 8; k++)
			if (tb + k < SEG) c[k] = cf[tb + k];
#pragma unroll
		for (int k = 0; k < 8; k++)
			if (tb + k < SEG) {
				v[k] = rec_step(c[k], mu, p1, p0);
				p0 = p1;
				p1 = v[k];
			}
#pragma unroll
		for (int k = 0; k < 8; k++)
			if (tb + k < SEG) store(tb + k, v[k]);
	}
}

// W wavefronts of a workgroup share a pair of orders with nseg0 and nseg1
// degree segments: the number given to the first order that minimises the
// largest share
template <int W>
__host__ __device__ inline int synth_split(int nseg0, int nseg1)
{
	if (nseg1 == 0) return W;
	int best = 1, cost = 1 << 30;
	for (int k = 1; k < W; k++) {
		const int a = (nseg0 + k - 1) / k, b = (nseg1 + W - k - 1) / (W - k);
		const int c = a > b ? a : b;
		if (c < cost) { cost = c; best = k; }
	}
	return best;
}

// a barrier and LDS fence within one wavefront
__device__ inline void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

Geo geo_of(sphswe_ctx *ctx)
{
	Geo g;
	g.coef = ctx->d_coef;
	g.seed = ctx->d_seed;
	g.segoff = ctx->d_segoff;
	g.seg = ctx->seg;
	g.ct = ctx->d_ct;
	g.wg = ctx->d_wg;
	g.moff = nullptr;
	g.L = ctx->L;
	g.nlat = ctx->nlat;
	g.NJ = ctx->NJ;
	g.P = 0;
	return g;
}

// ---- analysis operands ------------------------------------------------------
// E[ring][column] = w_j (F_N + F_S) (even parity of l - m) and w_j (F_N - F_S)
// (odd), for the 16 MFMA columns of a slot group, written once per (order,
// 64-ring chunk) in the order in which the analysis lanes consume them, so
// that every 32-degree segment of the order reuses them with coalesced loads.
// Rings past NJ get weight 0.  Lane n + 16 kq (column n) holds QN values q,
// stored at [q][lane]; the columns of absent fields are not written (a stale
// value there reaches only its own MFMA output column, which is not stored).
//   FP64 / FP32: q = 16 parity + s, s = 0..15: ring 4 s + kq;
//   bf16 (NT terms): q = (4 parity + s) NT + term, s = 0..3: four values for
//                rings 16 s + 4 kq + e (e = 0..3).
template <int PM>
__host__ __device__ constexpr int eop_qn()
{
	return (PM == PM_F64 || PM == PM_F32) ? 32 : 8 * Terms<PM>::n;
}
template <int PM>
__device__ __forceinline__ int eop_at(int n, int q, int kq)
{
	return q * 64 + kq * 16 + n;
}

template <int PM>
__host__ __device__ inline size_t eop_bytes_per_chunk()
{
	// both parities
	return PM == PM_F64 ? 2 * 16 * 64 * 8 : (PM == PM_F32 ? 2 * 16 * 64 * 4 : (size_t)2 * 4 * Terms<PM>::n * 64 * 8);
}

template <int LAY, int NS, int PM>
__global__ __launch_bounds__(256) void k_leg_eprep(const double2 *__restrict__ four, void *__restrict__ E, Geo g,
						   int group0, int nfields)
{
	const int nch = (g.NJ + 63) / 64;
	const long tid = blockIdx.x * 256L + threadIdx.x;
	if (tid >= (long)(g.L + 1) * nch * 64) return;
	const int lane = (int)(tid & 63);
	const long mc = tid >> 6;                 // am * nch + c
	const int c = (int)(mc % nch), am = (int)(mc / nch);
	const int group = group0 + (int)blockIdx.y;
	const int n = lane & 15, kq = lane >> 4;
	const Column<LAY, NS> col = column_of<LAY, NS>(n >> 1, am, group, nfields);
	const double *prow = (const double *)four + 2 * column_fb<LAY, NS>(col, g) + (n & 1);
	const double cmask = col.valid ? 1.0 : 0.0;
	const int NJ = g.NJ, nlat = g.nlat;
	double be[16], bo[16];   // computed for every lane, stored for present columns
#pragma unroll
	for (int s = 0; s < 16; s++) {
		const int rr = (PM == PM_F64 || PM == PM_F32) ? 4 * s + kq : 16 * (s >> 2) + 4 * kq + (s & 3);
		const int rj = c * 64 + rr;
		const int rc = min(rj, NJ - 1);
		const double fn = prow[2 * rc], fs = prow[2 * (nlat - 1 - rc)];
		const double w = rj < NJ ? g.wg[rc] * cmask : 0.0;
		be[s] = w * (fn + fs);
		bo[s] = w * (fn - fs);
	}
	const size_t base = (size_t)blockIdx.y * (g.L + 1) * nch * eop_bytes_per_chunk<PM>() + mc * eop_bytes_per_chunk<PM>();
	char *eb = (char *)E + base;
	if (!col.valid) return;
	if (PM == PM_F64) {
		double *e = (double *)eb;
#pragma unroll
		for (int s = 0; s < 16; s++) {
			e[eop_at<PM>(n, s, kq)] = be[s];
			e[eop_at<PM>(n, 16 + s, kq)] = bo[s];
		}
	} else if (PM == PM_F32) {
		float *e = (float *)eb;
#pragma unroll
		for (int s = 0; s < 16; s++) {
			e[eop_at<PM>(n, s, kq)] = (float)be[s];
			e[eop_at<PM>(n, 16 + s, kq)] = (float)bo[s];
		}
	} else {
		constexpr int NT = Terms<PM>::n;
		s4v *e = (s4v *)eb;
#pragma unroll
		for (int p = 0; p < 2; p++)
#pragma unroll
			for (int s = 0; s < 4; s++) {
				float v4[4];
#pragma unroll
				for (int k = 0; k < 4; k++) v4[k] = (float)(p ? bo[4 * s + k] : be[4 * s + k]);
				s4v t[NT];
				split4<NT>(v4, t);
#pragma unroll
				for (int k = 0; k < NT; k++) e[eop_at<PM>(n, (p * 4 + s) * NT + k, kq)] = t[k];
			}
	}
}

// the operand buffer for as many slot groups as fit in ANALYSIS_MAX_EBYTES (at
// least one), then for each batch: operands, analysis.  anal(E, group0, ngroups).
constexpr size_t ANALYSIS_MAX_EBYTES = (size_t)256 << 20;
template <int LAY, int NS, int PM, typename F>
int analysis_batches(sphswe_ctx *ctx, const Geo &g, const double2 *four, int nf, F anal)
{
	const int per = LAY == LAY_REAL ? NS : NS / 2;
	const int groups = (nf + per - 1) / per;
	const int nch = (g.NJ + 63) / 64;
	const size_t gbytes = (size_t)(g.L + 1) * nch * eop_bytes_per_chunk<PM>();
	int gb = (int)std::max<size_t>(1, std::min<size_t>(groups, ANALYSIS_MAX_EBYTES / gbytes));
	void *E = nullptr;
	if (ctx->ext_escratch) {
		// a caller-provided buffer: as many groups per batch as it holds
		const size_t fit = ctx->ext_escratch_bytes / gbytes;
		if (fit < 1)
			return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "analysis operands need %zu bytes, the scratch holds %zu",
					      gbytes, ctx->ext_escratch_bytes);
		gb = (int)std::min<size_t>(groups, fit);
		E = ctx->ext_escratch;
	} else if (int rc = sphswe_escratch(ctx, gbytes * gb, &E)) {
		return rc;
	}
	for (int g0 = 0; g0 < groups; g0 += gb) {
		const int ng = std::min(gb, groups - g0);
		const long nthr = (long)(g.L + 1) * nch * 64;
		dim3 grid((unsigned)((nthr + 255) / 256), ng);
		k_leg_eprep<LAY, NS, PM><<<grid, 256, 0, ctx->stream>>>(four, E, g, g0, nf);
		SPHSWE_LAUNCH_CHECK(ctx);
		if (int rc = anal(E, g0, ng)) return rc;
	}
	return SPHSWE_OK;
}

}  // namespace

#endif
