// Associated-Legendre half of the SHT (the Legendre part of SHTNS' spat_to_SH /
// SH_to_spat and their complex variants).
//
// Pbar_l^m(mu_j) is never stored: every lane runs the FP64 three-term
// recurrence Pbar_l = a_lm (mu Pbar_{l-1} - b_lm Pbar_{l-2}) for its own
// latitude, started from per-(m, j) seeds computed at setup (so tiny values
// near the poles never underflow).  Only the northern rings are visited; the
// southern ones follow from Pbar_l^m(-mu) = (-1)^(l-m) Pbar_l^m(mu), so every
// pass splits into an even and an odd partial sum.
//
// Synthesis: one lane per (m, j); the spectral coefficient of degree l is
// uniform over the wavefront (scalar loads), so a step costs 3 FP64 ops of
// recurrence plus 2 FMAs per (field, re/im).
// Analysis: the sum over j is a reduction, so a wavefront writes a 16-degree
// tile of Pbar (64 rings) to LDS, then re-reads it with lanes mapped to
// (degree, quarter of the rings), accumulates in registers, and reduces the
// quarters with cross-lane shuffles and the wavefronts through LDS.
//
// "Slots": a real field is one slot; a complex field is two (+m and -m, which
// share Pbar up to the factor (-1)^m).
#include "sphswe_common.h"

namespace {

__device__ inline int wave_min(int v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
	return v;
}
__device__ inline int wave_max(int v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
	return v;
}

struct Geo {
	const double2 *coef;
	const int *l0;
	const double2 *seed;
	const double *ct;
	const double *wg;
	int L, nlat, NJ;
};

// slot addressing ------------------------------------------------------------
// real:    spectral q(l) = spec[f*nlm + m(2L-m+1)/2 + l], Fourier row f*(L+1)*nlat + m*nlat
// complex: spectral q(l) = spec[f*ncplx + l*l + l + m],   Fourier row f*(2L+1)*nlat + (m+L)*nlat
template <bool CPLX>
struct Slot {
	long qb;      // spectral base
	long fb;      // Fourier row base
	double sgn;   // (-1)^m for negative complex m
	bool valid;
	__device__ static inline long off(int l) { return CPLX ? (long)l * l + l : (long)l; }
};

template <bool CPLX, int NS>
__device__ inline void make_slots(Slot<CPLX> (&sl)[NS], int am, int group, int nfields, int L, int nlat)
{
	const long nlm = (long)(L + 1) * (L + 2) / 2, ncplx = (long)(L + 1) * (L + 1);
#pragma unroll
	for (int s = 0; s < NS; s++) {
		if (!CPLX) {
			int f = group * NS + s;
			sl[s].valid = f < nfields;
			if (f >= nfields) f = nfields - 1;
			sl[s].qb = f * nlm + lm_index(L, 0, am);
			sl[s].fb = (long)f * (L + 1) * nlat + (long)am * nlat;
			sl[s].sgn = 1.0;
		} else {
			int f = group * (NS / 2) + (s >> 1);
			int m = (s & 1) ? -am : am;
			sl[s].valid = f < nfields && !((s & 1) && am == 0);
			if (f >= nfields) f = nfields - 1;
			sl[s].qb = f * ncplx + m;
			sl[s].fb = (long)f * (2 * L + 1) * nlat + (long)(m + L) * nlat;
			sl[s].sgn = ((s & 1) && (am & 1)) ? -1.0 : 1.0;
		}
	}
}

// ----------------------------------------------------------------------------
// synthesis
// ----------------------------------------------------------------------------

template <bool CPLX, int NS>
__global__ __launch_bounds__(64) void k_leg_synth(const double2 *__restrict__ spec, double2 *__restrict__ four, Geo g,
						  int nfields)
{
	const int am = blockIdx.y;
	const int L = g.L, NJ = g.NJ;
	const int j = blockIdx.x * 64 + threadIdx.x;
	const int jc = j < NJ ? j : NJ - 1;
	const double mu = g.ct[jc];
	const int ls = g.l0[(long)am * NJ + jc];
	const double2 sd = g.seed[(long)am * NJ + jc];
	const int cb = lm_index(L, 0, am);
	Slot<CPLX> sl[NS];
	make_slots<CPLX, NS>(sl, am, blockIdx.z, nfields, L, g.nlat);

	double eR[NS], eI[NS], oR[NS], oI[NS];
#pragma unroll
	for (int s = 0; s < NS; s++) eR[s] = eI[s] = oR[s] = oI[s] = 0.0;

	const int lmn = wave_min(ls), lmx = wave_max(ls);
	double p0 = 0.0, p1 = 0.0;
	int l = lmn;
	// phase 1: degrees where some lane injects its seeds
	const int l1 = min(lmx + 1, L);
	for (; l <= l1; ++l) {
		const double2 c = g.coef[cb + l];
		const double rec = c.x * (mu * p1 - c.y * p0);
		const double P = (l == ls) ? sd.x : ((l == ls + 1) ? sd.y : rec);
		p0 = p1;
		p1 = P;
		const bool odd = (l - am) & 1;
#pragma unroll
		for (int s = 0; s < NS; s++) {
			const double2 q = spec[sl[s].qb + Slot<CPLX>::off(l)];
			if (odd) { oR[s] = fma(q.x, P, oR[s]); oI[s] = fma(q.y, P, oI[s]); }
			else { eR[s] = fma(q.x, P, eR[s]); eI[s] = fma(q.y, P, eI[s]); }
		}
	}
	// phase 2: plain recurrence, unrolled by parity
#define STEP(ACR, ACI)                                                           \
	{                                                                        \
		const double2 c = g.coef[cb + l];                                \
		const double P = c.x * (mu * p1 - c.y * p0);                     \
		p0 = p1;                                                         \
		p1 = P;                                                          \
		_Pragma("unroll") for (int s = 0; s < NS; s++)                   \
		{                                                                \
			const double2 q = spec[sl[s].qb + Slot<CPLX>::off(l)];   \
			ACR[s] = fma(q.x, P, ACR[s]);                            \
			ACI[s] = fma(q.y, P, ACI[s]);                            \
		}                                                                \
		++l;                                                             \
	}
	if (l <= L && ((l - am) & 1)) STEP(oR, oI)
	for (; l + 1 <= L;) {
		STEP(eR, eI)
		STEP(oR, oI)
	}
	if (l <= L) STEP(eR, eI)
#undef STEP

	if (j < NJ) {
		const int js = g.nlat - 1 - j;
#pragma unroll
		for (int s = 0; s < NS; s++) {
			if (!sl[s].valid) continue;
			double2 n = make_double2(sl[s].sgn * (eR[s] + oR[s]), sl[s].sgn * (eI[s] + oI[s]));
			double2 so = make_double2(sl[s].sgn * (eR[s] - oR[s]), sl[s].sgn * (eI[s] - oI[s]));
			if (!CPLX && am == 0) { n.y = 0.0; so.y = 0.0; }  // Im q_l0 is ignored
			four[sl[s].fb + j] = n;
			four[sl[s].fb + js] = so;
		}
	}
}

// ----------------------------------------------------------------------------
// analysis
// ----------------------------------------------------------------------------

constexpr int LT = 16;      // degrees per tile
constexpr int MAXW = 4;     // wavefronts per workgroup
constexpr int MAXCH = 4;    // 64-ring chunks per wavefront

template <bool CPLX, int NS>
__global__ __launch_bounds__(256) void k_leg_anal(const double2 *__restrict__ four, double2 *__restrict__ spec, Geo g,
						  int nfields, int nw, int nch)
{
	__shared__ double Ps[MAXW][LT][65];
	__shared__ double2 Es[MAXW][2][64][NS];
	__shared__ double2 Qp[MAXW][LT][NS];

	const int am = blockIdx.x;
	const int L = g.L, NJ = g.NJ, nlat = g.nlat;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int cb = lm_index(L, 0, am);
	Slot<CPLX> sl[NS];
	make_slots<CPLX, NS>(sl, am, blockIdx.y, nfields, L, nlat);

	// per-chunk recurrence state
	double p0[MAXCH], p1[MAXCH], mu[MAXCH], w[MAXCH];
	double2 sd[MAXCH];
	int ls[MAXCH], lmn[MAXCH], lmx[MAXCH];
	bool val[MAXCH];
#pragma unroll
	for (int k = 0; k < MAXCH; k++) {
		int j = (wave + k * nw) * 64 + lane;
		val[k] = (k < nch) && (j < NJ);
		int jc = j < NJ ? j : NJ - 1;
		mu[k] = g.ct[jc];
		w[k] = g.wg[jc];
		ls[k] = val[k] ? g.l0[(long)am * NJ + jc] : L + 1;
		sd[k] = g.seed[(long)am * NJ + jc];
		lmn[k] = wave_min(ls[k]);
		lmx[k] = wave_max(val[k] ? ls[k] : -1);
		p0[k] = p1[k] = 0.0;
	}

	const int lq = lane & (LT - 1), part = lane >> 4;
	for (int lt0 = am; lt0 <= L; lt0 += LT) {
		double2 acc[NS];
#pragma unroll
		for (int s = 0; s < NS; s++) acc[s] = make_double2(0.0, 0.0);

#pragma unroll
		for (int k = 0; k < MAXCH; k++) {
			if (k >= nch) break;
			const int j = (wave + k * nw) * 64 + lane;
			// even / odd combinations of the two hemispheres, Gauss weight folded in
#pragma unroll
			for (int s = 0; s < NS; s++) {
				double2 e = make_double2(0, 0), o = make_double2(0, 0);
				if (val[k]) {
					double2 fn = four[sl[s].fb + j];
					double2 fs = four[sl[s].fb + (nlat - 1 - j)];
					e = make_double2(w[k] * (fn.x + fs.x), w[k] * (fn.y + fs.y));
					o = make_double2(w[k] * (fn.x - fs.x), w[k] * (fn.y - fs.y));
				}
				Es[wave][0][lane][s] = e;
				Es[wave][1][lane][s] = o;
			}
			// Pbar tile for degrees lt0 .. lt0+LT-1
#pragma unroll 4
			for (int t = 0; t < LT; t++) {
				const int l = lt0 + t;
				double P = 0.0;
				if (l <= L && l >= lmn[k]) {
					const double2 c = g.coef[cb + l];
					const double rec = c.x * (mu[k] * p1[k] - c.y * p0[k]);
					if (l <= lmx[k] + 1) P = (l == ls[k]) ? sd[k].x : ((l == ls[k] + 1) ? sd[k].y : rec);
					else P = rec;
					p0[k] = p1[k];
					p1[k] = P;
				}
				Ps[wave][t][lane] = val[k] ? P : 0.0;
			}
			__syncthreads();
			const int par = (lt0 + lq - am) & 1;
#pragma unroll 4
			for (int t = 0; t < 16; t++) {
				const int jl = part * 16 + t;
				const double P = Ps[wave][lq][jl];
#pragma unroll
				for (int s = 0; s < NS; s++) {
					const double2 e = Es[wave][par][jl][s];
					acc[s].x = fma(P, e.x, acc[s].x);
					acc[s].y = fma(P, e.y, acc[s].y);
				}
			}
			__syncthreads();
		}
		// reduce the four ring quarters
#pragma unroll
		for (int s = 0; s < NS; s++) {
			acc[s].x += __shfl_xor(acc[s].x, 16);
			acc[s].y += __shfl_xor(acc[s].y, 16);
			acc[s].x += __shfl_xor(acc[s].x, 32);
			acc[s].y += __shfl_xor(acc[s].y, 32);
		}
		if (part == 0) {
#pragma unroll
			for (int s = 0; s < NS; s++) Qp[wave][lq][s] = acc[s];
		}
		__syncthreads();
		if (threadIdx.x < LT * NS) {
			const int q = threadIdx.x & (LT - 1), s = threadIdx.x / LT;
			const int l = lt0 + q;
			double2 v = make_double2(0, 0);
			for (int ww = 0; ww < nw; ww++) {
				double2 x = Qp[ww][q][s];
				v.x += x.x;
				v.y += x.y;
			}
			// the slot loop is static, select it here
#pragma unroll
			for (int ss = 0; ss < NS; ss++)
				if (ss == s && l <= L && sl[ss].valid)
					spec[sl[ss].qb + Slot<CPLX>::off(l)] = make_double2(sl[ss].sgn * v.x, sl[ss].sgn * v.y);
		}
		__syncthreads();
	}
}

void wave_layout(int NJ, int &nw, int &nch)
{
	int nchunks = (NJ + 63) / 64;
	nw = nchunks < MAXW ? nchunks : MAXW;
	nch = (nchunks + nw - 1) / nw;
	nw = (nchunks + nch - 1) / nch;
}

Geo geo_of(sphswe_ctx *ctx)
{
	Geo g;
	g.coef = ctx->d_coef;
	g.l0 = ctx->d_l0;
	g.seed = ctx->d_seed;
	g.ct = ctx->d_ct;
	g.wg = ctx->d_wg;
	g.L = ctx->L;
	g.nlat = ctx->nlat;
	g.NJ = ctx->NJ;
	return g;
}

template <bool CPLX, int NS>
int launch_synth(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf)
{
	const int per = CPLX ? NS / 2 : NS;
	dim3 grid((ctx->NJ + 63) / 64, ctx->L + 1, (nf + per - 1) / per);
	k_leg_synth<CPLX, NS><<<grid, 64, 0, ctx->stream>>>(spec, four, geo_of(ctx), nf);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

template <bool CPLX, int NS>
int launch_anal(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf)
{
	const int per = CPLX ? NS / 2 : NS;
	int nw, nch;
	wave_layout(ctx->NJ, nw, nch);
	if (nch > MAXCH) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "nlat too large for the analysis kernel");
	dim3 grid(ctx->L + 1, (nf + per - 1) / per);
	k_leg_anal<CPLX, NS><<<grid, 64 * nw, 0, ctx->stream>>>(four, spec, geo_of(ctx), nf, nw, nch);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

int check_prec(sphswe_ctx *ctx, int prec)
{
	if (prec != SPHSWE_PREC_FP64)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "precision %d is not available for this transform", prec);
	return SPHSWE_OK;
}

}  // namespace

int sphswe_leg_synth(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	const long nlm = ctx->nlm;
	const long nfour = (long)(ctx->L + 1) * ctx->nlat;
	int rc = SPHSWE_OK;
	int f = 0;
	// groups of 8 slots amortise the recurrence best
	if (nf - f >= 8) {
		int g = (nf - f) / 8;
		if ((rc = launch_synth<false, 8>(ctx, spec + f * nlm, four + f * nfour, g * 8))) return rc;
		f += g * 8;
	}
	int r = nf - f;
	if (r >= 4) { if ((rc = launch_synth<false, 4>(ctx, spec + f * nlm, four + f * nfour, 4))) return rc; f += 4; r -= 4; }
	if (r == 3) rc = launch_synth<false, 3>(ctx, spec + f * nlm, four + f * nfour, 3);
	else if (r == 2) rc = launch_synth<false, 2>(ctx, spec + f * nlm, four + f * nfour, 2);
	else if (r == 1) rc = launch_synth<false, 1>(ctx, spec + f * nlm, four + f * nfour, 1);
	return rc;
}

int sphswe_leg_anal(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	const long nlm = ctx->nlm;
	const long nfour = (long)(ctx->L + 1) * ctx->nlat;
	int rc = SPHSWE_OK;
	int f = 0;
	if (nf - f >= 3) {
		int g = (nf - f) / 3;
		if ((rc = launch_anal<false, 3>(ctx, four, spec, g * 3))) return rc;
		f += g * 3;
	}
	int r = nf - f;
	if (r == 2) rc = launch_anal<false, 2>(ctx, four + f * nfour, spec + f * nlm, 2);
	else if (r == 1) rc = launch_anal<false, 1>(ctx, four + f * nfour, spec + f * nlm, 1);
	return rc;
}

int sphswe_leg_synth_cplx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	const long nc = ctx->ncplx;
	const long nfour = (long)(2 * ctx->L + 1) * ctx->nlat;
	int rc = SPHSWE_OK;
	int f = 0;
	if (nf >= 4) {
		int g = nf / 4;
		if ((rc = launch_synth<true, 8>(ctx, spec, four, g * 4))) return rc;
		f = g * 4;
	}
	int r = nf - f;
	if (r >= 2) { if ((rc = launch_synth<true, 4>(ctx, spec + f * nc, four + f * nfour, 2))) return rc; f += 2; r -= 2; }
	if (r == 1) rc = launch_synth<true, 2>(ctx, spec + f * nc, four + f * nfour, 1);
	return rc;
}

int sphswe_leg_anal_cplx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	return launch_anal<true, 2>(ctx, four, spec, nf);
}

extern "C" int sphswe_legendre_analysis(sphswe_ctx *ctx, const double *d_four, double *d_spec, int nfields, int precision)
{
	if (!ctx || !d_four || !d_spec || nfields < 0) return SPHSWE_ERR_ARG;
	return sphswe_leg_anal(ctx, (const double2 *)d_four, (double2 *)d_spec, nfields, precision);
}

extern "C" int sphswe_legendre_synthesis(sphswe_ctx *ctx, const double *d_spec, double *d_four, int nfields, int precision)
{
	if (!ctx || !d_four || !d_spec || nfields < 0) return SPHSWE_ERR_ARG;
	return sphswe_leg_synth(ctx, (const double2 *)d_spec, (double2 *)d_four, nfields, precision);
}
