// Pieces shared by the FP64 and the mixed-precision Legendre kernels: the
// geometry handed to the kernels, the slot addressing of the spectral and
// Fourier layouts, and one step of the FP64 associated-Legendre recurrence.
#ifndef SPHSWE_LEG_COMMON_H
#define SPHSWE_LEG_COMMON_H

#include "sphswe_common.h"

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

// one recurrence step: Pbar_l = (a mu) Pbar_{l-1} - (a b) Pbar_{l-2}
__device__ inline double rec_step(double2 c, double mu, double p1, double p0)
{
	return fma(c.x * mu, p1, -(c.y * p0));
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

}  // namespace

#endif
