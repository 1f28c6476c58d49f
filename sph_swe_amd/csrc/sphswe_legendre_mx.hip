// Mixed-precision Legendre transforms.  The associated-Legendre values come
// from the same FP64 recurrence as in the FP64 path (seeds every 32 degrees);
// only the contraction runs in lower precision, on the matrix cores, with FP32
// accumulation:
//   PM_F32     v_mfma_f32_16x16x4f32, operands rounded to FP32;
//   PM_BF16    v_mfma_f32_16x16x16bf16_1k, each operand split into
//              hi + lo bf16 (16 bits), products hi*hi + hi*lo + lo*hi;
//   PM_BF16X3  the same with hi + mid + lo (all 24 bits of the FP32 value),
//              products of total order <= 2 (six).
//
// Synthesis: a workgroup takes 64 rings and the pair of orders (m, L - m); its
// four wavefronts split the 32-degree segments of the two orders.  Per segment
// a wavefront runs the recurrence with lane = ring, writes the Pbar tile
// (degrees x rings, split by parity of l - m) to its LDS region, and adds
//   D[ring][column] += Pbar^T[ring][degree] * Q[degree][column]
// for four 16-ring tiles and both parities; the columns are (slot, re/im), 16
// of them.  The partial sums of the wavefronts of an order are reduced through
// LDS in a fixed order.
// Analysis: one wavefront per (m, 32-degree segment), looping over 64-ring
// chunks, as in the FP64 analysis:
//   D[degree][column] += Pbar[degree][ring] * E[ring][column]
// with E the hemisphere-combined Fourier rows times the Gauss weight, rounded
// (FP32) or split (bf16) once per (m, chunk) by the operand pass.
//
// Operand layouts (wave64), checked on the device: v_mfma_f32_16x16x4f32
// A[i][k] in lane i + 16k, B[k][n] in lane n + 16k; v_mfma_f32_16x16x16bf16_1k
// A[i][k] in lane i + 16 (k / 4) element k % 4, B[k][n] in lane n + 16 (k / 4)
// element k % 4; both D[i][n] in lane n + 16 (i / 4) element i % 4.
#include "sphswe_leg_common.h"

#include <type_traits>

namespace {

// acc += A * B over the split terms of total order below NT
template <int NT>
__device__ inline f4v mfma_split(const s4v (&a)[NT], const s4v (&b)[NT], f4v acc)
{
	if (NT == 3) {
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], b[1], acc, 0, 0, 0);
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[2], acc, 0, 0, 0);
		acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[2], b[0], acc, 0, 0, 0);
	}
	acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[1], acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], b[0], acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[0], acc, 0, 0, 0);
	return acc;
}

constexpr int MX_SEG = 32;      // degrees per segment (16 of each parity)
constexpr int MX_WAVES = 4;     // wavefronts per workgroup
constexpr int MX_NS = 8;        // slots per group: 16 MFMA columns

// ----------------------------------------------------------------------------
// synthesis
// ----------------------------------------------------------------------------

// row stride (elements) of the Pbar tile: conflict-free A-operand reads
// (FP32: 16 mod 32 for one value per lane, bf16: 12 mod 32 for four degrees
// per lane; FP64: 16 consecutive doubles per lane group in any case)
template <int PM>
struct SxStride {
	static constexpr int v = PM == PM_F64 ? 65 : (PM == PM_F32 ? 80 : 76);
};
// the tile and accumulator types of a mode: FP64 for PM_F64
// (v_mfma_f64_16x16x4f64, D[i][n] in lane n + 16 (i % 4), element i / 4),
// FP32 otherwise
template <int PM>
using TileT = std::conditional_t<PM == PM_F64, double, float>;
template <int PM>
using AccT = std::conditional_t<PM == PM_F64, d4v, f4v>;

template <int LAY, int NS, int PM>
__global__ __launch_bounds__(64 * MX_WAVES) void k_leg_synth_mx(const double2 *__restrict__ spec, double2 *__restrict__ four,
								Geo g, int nfields, const double *__restrict__ oscale)
{
	constexpr int PS = SxStride<PM>::v;
	typedef TileT<PM> T;
	__shared__ T Pt[MX_WAVES][2][16][PS];
	__shared__ double2 Cfs[MX_WAVES][MX_SEG];
	const int L = g.L, NJ = g.NJ;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int ma = blockIdx.y, mb = L - (int)blockIdx.y;
	const bool two = mb > ma;
	const int nrow0 = L - ma + 1, nrow1 = two ? L - mb + 1 : 0;
	const int nseg0 = (nrow0 + MX_SEG - 1) / MX_SEG, nseg1 = (nrow1 + MX_SEG - 1) / MX_SEG;
	const int kA = synth_split<MX_WAVES>(nseg0, nseg1);
	const int part = __builtin_amdgcn_readfirstlane(wave >= kA);
	const int am = __builtin_amdgcn_readfirstlane(part ? mb : ma);
	const int nw = part ? MX_WAVES - kA : kA, wi = part ? wave - kA : wave, ns = part ? nseg1 : nseg0;
	const int sa = wi * ns / nw, sb = (wi + 1) * ns / nw;
	const int so = g.segoff[am];

	const int j = blockIdx.x * 64 + lane;
	const int jc = j < NJ ? j : NJ - 1;
	const double mu = g.ct[jc];
	const int n = lane & 15, kq = lane >> 4;
	const Column<LAY, NS> col = column_of<LAY, NS>(n >> 1, am, blockIdx.z, nfields);
	const int comp = n & 1;
	const double *sd_ = (const double *)spec;

	AccT<PM> acc[2][4];
#pragma unroll
	for (int p = 0; p < 2; p++)
#pragma unroll
		for (int t = 0; t < 4; t++) acc[p][t] = AccT<PM>{0, 0, 0, 0};
	T (*P)[16][PS] = Pt[wave];
	double2 *cf = Cfs[wave];

	for (int k = sa; k < sb; k++) {
		const int s0 = am + k * MX_SEG, se = min(s0 + MX_SEG - 1, L);
		if (lane < MX_SEG) cf[lane] = g.coef[lm_index(L, min(s0 + lane, se), am)];
		// B operands: Q[degree][column]; degrees past se are 0
		T bq[2][4];
		s4v bs[2][Terms<PM>::n];
		if (PM == PM_F32 || PM == PM_F64) {
#pragma unroll
			for (int p = 0; p < 2; p++)
#pragma unroll
				for (int s = 0; s < 4; s++) {
					const int l = s0 + 2 * (4 * s + kq) + p;
					const int lc = min(l, se);
					const double v = sd_[2 * column_q<LAY, NS>(col, am, lc, g) + comp];
					bq[p][s] = (col.valid && l <= se) ? (T)(col.sgn * v) : (T)0;
				}
		} else {
#pragma unroll
			for (int p = 0; p < 2; p++) {
				float v4[4];
#pragma unroll
				for (int e = 0; e < 4; e++) {
					const int l = s0 + 2 * (4 * kq + e) + p;
					const int lc = min(l, se);
					const double v = sd_[2 * column_q<LAY, NS>(col, am, lc, g) + comp];
					v4[e] = (col.valid && l <= se) ? (float)(col.sgn * v) : 0.0f;
				}
				split4<Terms<PM>::n>(v4, bs[p]);
			}
		}
		wave_sync();
		// Pbar tile, lane = ring; rows past se hold finite values that meet Q = 0
		{
			const double2 sdv = g.seed[((long)so + k) * NJ + jc];
			double p0 = sdv.x, p1 = sdv.y;
			P[0][0][lane] = (T)p0;
			P[1][0][lane] = (T)p1;
			rec_tile<MX_SEG>(cf, mu, p0, p1, [&](int t, double v) { P[t & 1][t >> 1][lane] = (T)v; });
		}
		wave_sync();
		if constexpr (PM == PM_F64) {
#pragma unroll
			for (int p = 0; p < 2; p++)
#pragma unroll
				for (int s = 0; s < 4; s++)
#pragma unroll
					for (int t = 0; t < 4; t++)
						acc[p][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(P[p][4 * s + kq][16 * t + n], bq[p][s],
												 acc[p][t], 0, 0, 0);
		} else if constexpr (PM == PM_F32) {
#pragma unroll
			for (int p = 0; p < 2; p++)
#pragma unroll
				for (int s = 0; s < 4; s++)
#pragma unroll
					for (int t = 0; t < 4; t++)
						acc[p][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(P[p][4 * s + kq][16 * t + n], bq[p][s],
												 acc[p][t], 0, 0, 0);
		} else {
#pragma unroll
			for (int p = 0; p < 2; p++)
#pragma unroll
				for (int t = 0; t < 4; t++) {
					float a4[4];
#pragma unroll
					for (int e = 0; e < 4; e++) a4[e] = P[p][4 * kq + e][16 * t + n];
					s4v as[Terms<PM>::n];
					split4<Terms<PM>::n>(a4, as);
					acc[p][t] = mfma_split<Terms<PM>::n>(as, bs[p], acc[p][t]);
				}
		}
		wave_sync();
	}

	// reduce the wavefronts of each order (LDS reused), leaders write
	__syncthreads();
	T *red = &Pt[0][0][0][0];
	const bool leader = wi == 0;
	if (!leader) {
		T *rw = red + (size_t)(wave - 1) * 32 * 64;
#pragma unroll
		for (int p = 0; p < 2; p++)
#pragma unroll
			for (int t = 0; t < 4; t++)
#pragma unroll
				for (int r = 0; r < 4; r++) rw[((p * 4 + t) * 4 + r) * 64 + lane] = acc[p][t][r];
	}
	__syncthreads();
	if (!leader) return;
	for (int w = wave + 1; w < wave + nw; w++) {
		const T *rw = red + (size_t)(w - 1) * 32 * 64;
#pragma unroll
		for (int p = 0; p < 2; p++)
#pragma unroll
			for (int t = 0; t < 4; t++)
#pragma unroll
				for (int r = 0; r < 4; r++) acc[p][t][r] += rw[((p * 4 + t) * 4 + r) * 64 + lane];
	}
	if (!col.valid) return;
	const bool zero_im = LAY == LAY_REAL && am == 0 && comp == 1;   // Im q_l0 is ignored
	double *fo = (double *)four;
	const long fb = column_fb<LAY, NS>(col, g);
#pragma unroll
	for (int t = 0; t < 4; t++)
#pragma unroll
		for (int r = 0; r < 4; r++) {
			// ring of element r: D row i = 4 kq + r (FP32 layouts), kq + 4 r (FP64)
			const int jj = blockIdx.x * 64 + 16 * t + (PM == PM_F64 ? kq + 4 * r : 4 * kq + r);
			if (jj >= NJ) continue;
			const double sc = zero_im ? 0.0 : (oscale ? oscale[jj] : 1.0);
			const double E = acc[0][t][r], O = acc[1][t][r];
			fo[2 * (fb + jj) + comp] = sc * (E + O);
			fo[2 * (fb + g.nlat - 1 - jj) + comp] = sc * (E - O);
		}
}

// ----------------------------------------------------------------------------
// analysis
// ----------------------------------------------------------------------------

// row stride (floats) of the Pbar tile: 2 mod 32 (one value per lane), a
// multiple of 4 for the four-ring reads of the bf16 layout
template <int PM>
struct AxStride {
	static constexpr int v = PM == PM_F32 ? 66 : 68;
};

template <int LAY, int NS, int PM>
__global__ __launch_bounds__(64 * MX_WAVES) void k_leg_anal_mx(const void *__restrict__ E, double2 *__restrict__ spec,
							       Geo g, const int *__restrict__ unit_m, int nunits, int group0,
							       int nfields)
{
	constexpr int PS = AxStride<PM>::v;
	__shared__ __align__(16) float Pt[MX_WAVES][2][16][PS];
	__shared__ double2 Cfs[MX_WAVES][MX_SEG];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int u = blockIdx.x * MX_WAVES + wave;
	if (u >= nunits) return;   // whole wavefront; no workgroup barrier below
	const int am = unit_m[u];
	const int L = g.L, NJ = g.NJ;
	const int so = g.segoff[am];
	const int s0 = am + (u - so) * MX_SEG, se = min(s0 + MX_SEG - 1, L);
	const int nch = (NJ + 63) / 64;

	double2 *cf = Cfs[wave];
	if (lane < MX_SEG) cf[lane] = g.coef[lm_index(L, min(s0 + lane, se), am)];
	wave_sync();

	const int n = lane & 15, kq = lane >> 4;
	const Column<LAY, NS> col = column_of<LAY, NS>(n >> 1, am, group0 + (int)blockIdx.y, nfields);
	const int comp = n & 1;
	const char *Em = (const char *)E + ((size_t)blockIdx.y * (L + 1) + am) * nch * eop_bytes_per_chunk<PM>();

	f4v acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
	float (*P)[16][PS] = Pt[wave];
	for (int c = 0; c < nch; c++) {
		const char *Ec = Em + (size_t)c * eop_bytes_per_chunk<PM>();
		// B operands, in operand order (see k_leg_eprep)
		float be[16], bo[16];
		constexpr int NT = Terms<PM>::n;
		s4v bs[2][4][NT];
		if (PM == PM_F32) {
			const float *e = (const float *)Ec;
#pragma unroll
			for (int s = 0; s < 16; s++) {
				be[s] = col.valid ? e[eop_at<PM>(n, s, kq)] : 0.0f;
				bo[s] = col.valid ? e[eop_at<PM>(n, 16 + s, kq)] : 0.0f;
			}
		} else {
			const s4v *e = (const s4v *)Ec;
#pragma unroll
			for (int p = 0; p < 2; p++)
#pragma unroll
				for (int s = 0; s < 4; s++)
#pragma unroll
					for (int k = 0; k < NT; k++)
						bs[p][s][k] = col.valid ? e[eop_at<PM>(n, (p * 4 + s) * NT + k, kq)] : s4v{0, 0, 0, 0};
		}
		// Pbar tile: lane = ring
		{
			const int j = c * 64 + lane;
			const int jc = j < NJ ? j : NJ - 1;
			const double mu = g.ct[jc];
			const double2 sdv = g.seed[(long)u * NJ + jc];
			double p0 = sdv.x, p1 = sdv.y;
			P[0][0][lane] = (float)p0;
			P[1][0][lane] = (float)p1;
			rec_tile<MX_SEG>(cf, mu, p0, p1, [&](int t, double v) { P[t & 1][t >> 1][lane] = (float)v; });
		}
		wave_sync();
		if (PM == PM_F32) {
#pragma unroll
			for (int s = 0; s < 16; s++) {
				acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(P[0][n][4 * s + kq], be[s], acc0, 0, 0, 0);
				acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(P[1][n][4 * s + kq], bo[s], acc1, 0, 0, 0);
			}
		} else {
#pragma unroll
			for (int s = 0; s < 4; s++) {
				float a4[4];
				s4v as[NT];
#pragma unroll
				for (int e = 0; e < 4; e++) a4[e] = P[0][n][16 * s + 4 * kq + e];
				split4<NT>(a4, as);
				acc0 = mfma_split<NT>(as, bs[0][s], acc0);
#pragma unroll
				for (int e = 0; e < 4; e++) a4[e] = P[1][n][16 * s + 4 * kq + e];
				split4<NT>(a4, as);
				acc1 = mfma_split<NT>(as, bs[1][s], acc1);
			}
		}
		wave_sync();
	}
	// D[i][n] in lane n + 16 (i / 4), element i % 4; row i of parity p is degree s0 + 2i + p
	if (!col.valid) return;
	double *sdo = (double *)spec;
#pragma unroll
	for (int r = 0; r < 4; r++) {
		const int i = 4 * kq + r;
#pragma unroll
		for (int p = 0; p < 2; p++) {
			const int l = s0 + 2 * i + p;
			if (l <= se) sdo[2 * column_q<LAY, NS>(col, am, l, g) + comp] = col.sgn * (double)(p ? acc1[r] : acc0[r]);
		}
	}
}

template <int LAY, int PM>
int launch_synth_mx(sphswe_ctx *ctx, const Geo &g, const double2 *spec, double2 *four, int nf, const double *oscale)
{
	const int per = LAY == LAY_REAL ? MX_NS : MX_NS / 2;
	dim3 grid((ctx->NJ + 63) / 64, ctx->L / 2 + 1, (nf + per - 1) / per);
	k_leg_synth_mx<LAY, MX_NS, PM><<<grid, 64 * MX_WAVES, 0, ctx->stream>>>(spec, four, g, nf, oscale);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

template <int LAY, int PM>
int launch_anal_mx(sphswe_ctx *ctx, const Geo &g, const double2 *four, double2 *spec, int nf)
{
	const int nunits = ctx->nseg_total;
	return analysis_batches<LAY, MX_NS, PM>(ctx, g, four, nf, [&](void *E, int g0, int ng) -> int {
		dim3 grid((nunits + MX_WAVES - 1) / MX_WAVES, ng);
		k_leg_anal_mx<LAY, MX_NS, PM><<<grid, 64 * MX_WAVES, 0, ctx->stream>>>(E, spec, g, ctx->d_unit_m, nunits, g0, nf);
		SPHSWE_LAUNCH_CHECK(ctx);
		return SPHSWE_OK;
	});
}

template <int LAY>
int synth_mx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec)
{
	const Geo g = geo_of(ctx);
	switch (prec) {
	case PM_F32: return launch_synth_mx<LAY, PM_F32>(ctx, g, spec, four, nf, nullptr);
	case PM_BF16: return launch_synth_mx<LAY, PM_BF16>(ctx, g, spec, four, nf, nullptr);
	case PM_BF16X3: return launch_synth_mx<LAY, PM_BF16X3>(ctx, g, spec, four, nf, nullptr);
	}
	return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "unknown precision %d", prec);
}

template <int LAY>
int anal_mx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec)
{
	const Geo g = geo_of(ctx);
	switch (prec) {
	case PM_F32: return launch_anal_mx<LAY, PM_F32>(ctx, g, four, spec, nf);
	case PM_BF16: return launch_anal_mx<LAY, PM_BF16>(ctx, g, four, spec, nf);
	case PM_BF16X3: return launch_anal_mx<LAY, PM_BF16X3>(ctx, g, four, spec, nf);
	}
	return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "unknown precision %d", prec);
}

}  // namespace

// pole layout (REXI): nf columns of row stride P, m-block offsets moff, on
// the matrix cores in any precision (FP64: v_mfma_f64_16x16x4f64); synthesis
// output times oscale[j]
int sphswe_leg_synth_pole_mx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int P, const int *d_moff,
			     const double *oscale, int prec)
{
	Geo g = geo_of(ctx);
	g.moff = d_moff;
	g.P = P;
	switch (prec) {
	case PM_F64: return launch_synth_mx<LAY_POLE, PM_F64>(ctx, g, spec, four, nf, oscale);
	case PM_F32: return launch_synth_mx<LAY_POLE, PM_F32>(ctx, g, spec, four, nf, oscale);
	case PM_BF16: return launch_synth_mx<LAY_POLE, PM_BF16>(ctx, g, spec, four, nf, oscale);
	case PM_BF16X3: return launch_synth_mx<LAY_POLE, PM_BF16X3>(ctx, g, spec, four, nf, oscale);
	}
	return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "unknown precision %d", prec);
}

int sphswe_leg_anal_pole_mx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int P, const int *d_moff,
			    int prec)
{
	Geo g = geo_of(ctx);
	g.moff = d_moff;
	g.P = P;
	switch (prec) {
	case PM_F32: return launch_anal_mx<LAY_POLE, PM_F32>(ctx, g, four, spec, nf);
	case PM_BF16: return launch_anal_mx<LAY_POLE, PM_BF16>(ctx, g, four, spec, nf);
	case PM_BF16X3: return launch_anal_mx<LAY_POLE, PM_BF16X3>(ctx, g, four, spec, nf);
	}
	return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "unknown precision %d", prec);
}

// FP64 real synthesis on the FP64 matrix cores (chosen by the autotuner for
// many fields)
int sphswe_leg_synth_f64_mfma(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf)
{
	return launch_synth_mx<LAY_REAL, PM_F64>(ctx, geo_of(ctx), spec, four, nf, nullptr);
}

int sphswe_leg_synth_mx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec)
{
	return synth_mx<LAY_REAL>(ctx, spec, four, nf, prec);
}

int sphswe_leg_anal_mx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec)
{
	return anal_mx<LAY_REAL>(ctx, four, spec, nf, prec);
}

int sphswe_leg_synth_cplx_mx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec)
{
	return synth_mx<LAY_CPLX>(ctx, spec, four, nf, prec);
}

int sphswe_leg_anal_cplx_mx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec)
{
	return anal_mx<LAY_CPLX>(ctx, four, spec, nf, prec);
}
