// FP32 Legendre transforms of many real fields from stored tables.
//
// The FP32 contraction mode rounds every associated-Legendre value to FP32
// before the product.  Those rounded values do not depend on the fields, so
// for many fields they are computed once per context (the same FP64
// recurrence from the same seeds as the recurrence kernels) and stored:
//   TS(m, p)[k][j]  ring j fastest   (synthesis A operand: Pbar^T[ring][degree])
//   TA(m, p)[j][k]  degree k fastest (analysis A operand: Pbar[degree][ring])
// with degree l = m + p + 2k of parity p = (l - m) % 2 and j one of the NJ
// northern rings; both tables hold nlm * NJ floats (201 MB each at T511).
// A transform of nf fields is then one batched GEMM over the orders on the
// FP32 matrix cores (v_mfma_f32_16x16x4f32):
//   synthesis  E_p[ring][col] = sum_k TS(m,p)[k][ring] Q[m + p + 2k][col],
//              F_N = E_0 + E_1, F_S = E_0 - E_1;
//   analysis   Q[m + p + 2k][col] = sum_j TA(m,p)[j][k] w_j (F_N +- F_S)[j][col];
// the 2 nf real columns are the (field, re / im) pairs.  A workgroup takes an
// order, TAB_RT 16-row tiles (rings or degrees) of both parities and TAB_CT
// 16-column tiles; the table rows of one k step are contiguous over the tile.
// Operand layouts: A[i][k] in lane i + 16 k, B[k][n] in lane n + 16 k,
// D[i][n] in lane n + 16 (i / 4), element i % 4.
#include "sphswe_leg_common.h"

namespace {

constexpr int TAB_RT = 4;   // 16-row tiles per workgroup
constexpr int TAB_CT = 2;   // 16-column tiles per workgroup
constexpr int TAB_UN = 2;   // k steps whose loads are issued together
constexpr int TAB_SEG = 32; // degrees per recurrence segment (seeds every 32 degrees)

__host__ __device__ inline int tab_np(int L, int m, int p) { return (L + 1 - m + 1 - p) / 2; }

// one thread per (order, northern ring): the FP64 recurrence of the order from
// the stored segment seeds, rounded into both tables
__global__ __launch_bounds__(64) void k_tab_gen(const double2 *__restrict__ coef, const double2 *__restrict__ seed,
						const int *__restrict__ segoff, const double *__restrict__ ct,
						const long *__restrict__ off, float *__restrict__ TS, float *__restrict__ TA,
						int L, int NJ)
{
	const int m = blockIdx.y;
	const int j = blockIdx.x * 64 + threadIdx.x;
	if (j >= NJ) return;
	const double mu = ct[j];
	const int np0 = tab_np(L, m, 0), np1 = tab_np(L, m, 1);
	auto put = [&](int l, double v) {
		const int d = l - m, p = d & 1, k = d >> 1;
		const long o = off[2 * m + p];
		TS[o + (long)k * NJ + j] = (float)v;
		TA[o + (long)j * (p ? np1 : np0) + k] = (float)v;
	};
	const int nseg = segoff[m + 1] - segoff[m];
	for (int s = 0; s < nseg; s++) {
		const int s0 = m + TAB_SEG * s, se = min(s0 + TAB_SEG - 1, L);
		const double2 sd = seed[(long)(segoff[m] + s) * NJ + j];
		double p0 = sd.x, p1 = sd.y;
		put(s0, p0);
		if (s0 + 1 <= se) put(s0 + 1, p1);
		for (int l = s0 + 2; l <= se; l++) {
			const double v = rec_step(coef[lm_index(L, l, m)], mu, p1, p0);
			put(l, v);
			p0 = p1;
			p1 = v;
		}
	}
}

// A workgroup of TAB_W wavefronts takes an order, TAB_RT 16-row tiles and
// TAB_CT 16-column tiles; its wavefronts split the contraction (k steps
// interleaved) and their partial sums are added through LDS in a fixed order.
constexpr int TAB_W = 4;

template <int NP>
__device__ __forceinline__ void tab_reduce(f4v (&acc)[NP][TAB_RT][TAB_CT], float *red, int wave, int lane)
{
	constexpr int NV = NP * TAB_RT * TAB_CT * 4;
	if (wave > 0) {
		float *rw = red + (size_t)(wave - 1) * NV * 64;
#pragma unroll
		for (int p = 0; p < NP; p++)
#pragma unroll
			for (int t = 0; t < TAB_RT; t++)
#pragma unroll
				for (int c = 0; c < TAB_CT; c++)
#pragma unroll
					for (int e = 0; e < 4; e++) rw[(((p * TAB_RT + t) * TAB_CT + c) * 4 + e) * 64 + lane] = acc[p][t][c][e];
	}
	__syncthreads();
	if (wave == 0) {
		for (int w = 1; w < TAB_W; w++) {
			const float *rw = red + (size_t)(w - 1) * NV * 64;
#pragma unroll
			for (int p = 0; p < NP; p++)
#pragma unroll
				for (int t = 0; t < TAB_RT; t++)
#pragma unroll
					for (int c = 0; c < TAB_CT; c++)
#pragma unroll
						for (int e = 0; e < 4; e++)
							acc[p][t][c][e] += rw[(((p * TAB_RT + t) * TAB_CT + c) * 4 + e) * 64 + lane];
		}
	}
}

// synthesis: tile (m, first ring j0), TAB_CT * 16 columns per blockIdx.y
__global__ __launch_bounds__(64 * TAB_W) void k_tab_synth(const float *__restrict__ TS, const long *__restrict__ off,
							  const int2 *__restrict__ tiles, const double2 *__restrict__ spec,
							  double2 *__restrict__ four, int L, int nlat, int NJ, int nf)
{
	__shared__ float red[(TAB_W - 1) * 2 * TAB_RT * TAB_CT * 4 * 64];
	const int2 tl = tiles[blockIdx.x];
	const int m = tl.x, j0 = tl.y;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
	const long nlm = (long)(L + 1) * (L + 2) / 2;
	const int nt = min(TAB_RT, (NJ - j0 + 15) / 16);
	const double *q[TAB_CT];
	bool cv[TAB_CT];
#pragma unroll
	for (int c = 0; c < TAB_CT; c++) {
		const int col = (blockIdx.y * TAB_CT + c) * 16 + i;   // real column: field col / 2, component col % 2
		cv[c] = col < 2 * nf;
		const int f = cv[c] ? col >> 1 : 0;
		q[c] = (const double *)spec + 2 * ((long)f * nlm + lm_index(L, m, m)) + (col & 1);   // q[2 d] = Q_{m+d}
	}
	f4v acc[2][TAB_RT][TAB_CT];
#pragma unroll
	for (int p = 0; p < 2; p++)
#pragma unroll
		for (int t = 0; t < TAB_RT; t++)
#pragma unroll
			for (int c = 0; c < TAB_CT; c++) acc[p][t][c] = f4v{0, 0, 0, 0};
#pragma unroll
	for (int p = 0; p < 2; p++) {
		const int np = tab_np(L, m, p);
		const float *T = TS + off[2 * m + p] + j0 + i;
		for (int kk = 4 * wave; kk < np; kk += 4 * TAB_W * TAB_UN) {
			float b[TAB_UN][TAB_CT], a[TAB_UN][TAB_RT];
#pragma unroll
			for (int u = 0; u < TAB_UN; u++) {
				const int k = kk + 4 * TAB_W * u + kq;
				const bool kv = k < np;
#pragma unroll
				for (int c = 0; c < TAB_CT; c++) b[u][c] = (kv && cv[c]) ? (float)q[c][2 * (p + 2 * k)] : 0.0f;
				const float *Tk = T + (long)(kv ? k : 0) * NJ;
#pragma unroll
				for (int t = 0; t < TAB_RT; t++)
					a[u][t] = (t < nt && kv && j0 + 16 * t + i < NJ) ? Tk[16 * t] : 0.0f;
			}
#pragma unroll
			for (int u = 0; u < TAB_UN; u++)
#pragma unroll
				for (int t = 0; t < TAB_RT; t++)
#pragma unroll
					for (int c = 0; c < TAB_CT; c++)
						acc[p][t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u][t], b[u][c], acc[p][t][c], 0, 0, 0);
		}
	}
	tab_reduce<2>(acc, red, wave, lane);
	if (wave != 0) return;
#pragma unroll
	for (int c = 0; c < TAB_CT; c++) {
		const int col = (blockIdx.y * TAB_CT + c) * 16 + i;
		if (!cv[c]) continue;
		const int f = col >> 1, comp = col & 1;
		const bool zero_im = m == 0 && comp == 1;   // Im q_l0 is ignored
		double *fo = (double *)four + 2 * (((long)f * (L + 1) + m) * nlat) + comp;
#pragma unroll
		for (int t = 0; t < TAB_RT; t++) {
			if (t >= nt) break;
#pragma unroll
			for (int e = 0; e < 4; e++) {
				const int j = j0 + 16 * t + 4 * kq + e;
				if (j >= NJ) continue;
				const double E = acc[0][t][c][e], O = acc[1][t][c][e];
				fo[2 * (long)j] = zero_im ? 0.0 : E + O;
				fo[2 * (long)(nlat - 1 - j)] = zero_im ? 0.0 : E - O;
			}
		}
	}
}

// analysis: tile (m, first degree row r0 of both parities), TAB_CT * 16 columns per blockIdx.y
__global__ __launch_bounds__(64 * TAB_W) void k_tab_anal(const float *__restrict__ TA, const long *__restrict__ off,
							 const int2 *__restrict__ tiles, const double2 *__restrict__ four,
							 const double *__restrict__ wg, double2 *__restrict__ spec, int L,
							 int nlat, int NJ, int nf)
{
	__shared__ float red[(TAB_W - 1) * 2 * TAB_RT * TAB_CT * 4 * 64];
	const int2 tl = tiles[blockIdx.x];
	const int m = tl.x, r0 = tl.y;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, i = lane & 15, kq = lane >> 4;
	const int np0 = tab_np(L, m, 0), np1 = tab_np(L, m, 1);
	const int nt0 = min(TAB_RT, (np0 - r0 + 15) / 16), nt1 = min(TAB_RT, max(0, (np1 - r0 + 15) / 16));
	const double *fr[TAB_CT];
	bool cv[TAB_CT];
#pragma unroll
	for (int c = 0; c < TAB_CT; c++) {
		const int col = (blockIdx.y * TAB_CT + c) * 16 + i;
		cv[c] = col < 2 * nf;
		const int f = cv[c] ? col >> 1 : 0;
		fr[c] = (const double *)four + 2 * (((long)f * (L + 1) + m) * nlat) + (col & 1);
	}
	const float *T0 = TA + off[2 * m] + r0 + i, *T1 = TA + off[2 * m + 1] + r0 + i;
	f4v acc[2][TAB_RT][TAB_CT];
#pragma unroll
	for (int p = 0; p < 2; p++)
#pragma unroll
		for (int t = 0; t < TAB_RT; t++)
#pragma unroll
			for (int c = 0; c < TAB_CT; c++) acc[p][t][c] = f4v{0, 0, 0, 0};
	for (int kk = 4 * wave; kk < NJ; kk += 4 * TAB_W * TAB_UN) {
		float b0[TAB_UN][TAB_CT], b1[TAB_UN][TAB_CT], a0[TAB_UN][TAB_RT], a1[TAB_UN][TAB_RT];
#pragma unroll
		for (int u = 0; u < TAB_UN; u++) {
			const int j = kk + 4 * TAB_W * u + kq;
			const bool jv = j < NJ;
			const int jc = jv ? j : 0;
			const double w = jv ? wg[jc] : 0.0;
#pragma unroll
			for (int c = 0; c < TAB_CT; c++) {
				const double fn = (jv && cv[c]) ? fr[c][2 * (long)jc] : 0.0;
				const double fs = (jv && cv[c]) ? fr[c][2 * (long)(nlat - 1 - jc)] : 0.0;
				b0[u][c] = (float)(w * (fn + fs));
				b1[u][c] = (float)(w * (fn - fs));
			}
#pragma unroll
			for (int t = 0; t < TAB_RT; t++) {
				a0[u][t] = (t < nt0 && jv && r0 + 16 * t + i < np0) ? T0[(long)jc * np0 + 16 * t] : 0.0f;
				a1[u][t] = (t < nt1 && jv && r0 + 16 * t + i < np1) ? T1[(long)jc * np1 + 16 * t] : 0.0f;
			}
		}
#pragma unroll
		for (int u = 0; u < TAB_UN; u++)
#pragma unroll
			for (int t = 0; t < TAB_RT; t++)
#pragma unroll
				for (int c = 0; c < TAB_CT; c++) {
					acc[0][t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[u][t], b0[u][c], acc[0][t][c], 0, 0, 0);
					acc[1][t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[u][t], b1[u][c], acc[1][t][c], 0, 0, 0);
				}
	}
	tab_reduce<2>(acc, red, wave, lane);
	if (wave != 0) return;
	const long nlm = (long)(L + 1) * (L + 2) / 2;
#pragma unroll
	for (int c = 0; c < TAB_CT; c++) {
		const int col = (blockIdx.y * TAB_CT + c) * 16 + i;
		if (!cv[c]) continue;
		double *qo = (double *)spec + 2 * ((long)(col >> 1) * nlm + lm_index(L, m, m)) + (col & 1);
#pragma unroll
		for (int p = 0; p < 2; p++) {
			const int np = p ? np1 : np0, nt = p ? nt1 : nt0;
#pragma unroll
			for (int t = 0; t < TAB_RT; t++) {
				if (t >= nt) break;
#pragma unroll
				for (int e = 0; e < 4; e++) {
					const int row = r0 + 16 * t + 4 * kq + e;
					if (row < np) qo[2 * (long)(p + 2 * row)] = (double)acc[p][t][c][e];
				}
			}
		}
	}
}

}  // namespace

// the tables and the tile lists, built on first use (allocates; not inside a
// stream capture)
int sphswe_leg_tables(sphswe_ctx *ctx)
{
	if (ctx->d_tab_s) return SPHSWE_OK;
	const int L = ctx->L, NJ = ctx->NJ;
	if (ctx->seg != TAB_SEG) return sphswe_set_err(ctx, SPHSWE_ERR_INTERNAL, "Legendre tables: seed spacing %d", ctx->seg);
	std::vector<long> off(2 * (L + 1));
	long tot = 0;
	for (int m = 0; m <= L; m++)
		for (int p = 0; p < 2; p++) {
			off[2 * m + p] = tot;
			tot += (long)tab_np(L, m, p) * NJ;
		}
	std::vector<int2> ts, ta;
	for (int m = 0; m <= L; m++) {
		for (int j0 = 0; j0 < NJ; j0 += 16 * TAB_RT) ts.push_back(make_int2(m, j0));
		for (int r0 = 0; r0 < tab_np(L, m, 0); r0 += 16 * TAB_RT) ta.push_back(make_int2(m, r0));
	}
	float *s = nullptr, *a = nullptr;
	long *doff = nullptr;
	int2 *dts = nullptr, *dta = nullptr;
	auto fail = [&](const char *what) {
		for (void *p : {(void *)s, (void *)a, (void *)doff, (void *)dts, (void *)dta})
			if (p) hipFree(p);
		return sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "Legendre tables: %s", what);
	};
	if (hipMalloc((void **)&s, sizeof(float) * tot) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&a, sizeof(float) * tot) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&doff, sizeof(long) * off.size()) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&dts, sizeof(int2) * ts.size()) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&dta, sizeof(int2) * ta.size()) != hipSuccess) return fail("allocation");
	if (hipMemcpy(doff, off.data(), sizeof(long) * off.size(), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dts, ts.data(), sizeof(int2) * ts.size(), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dta, ta.data(), sizeof(int2) * ta.size(), hipMemcpyHostToDevice) != hipSuccess)
		return fail("upload");
	k_tab_gen<<<dim3((NJ + 63) / 64, L + 1), 64, 0, ctx->stream>>>(ctx->d_coef, ctx->d_seed, ctx->d_segoff, ctx->d_ct,
								       doff, s, a, L, NJ);
	if (hipGetLastError() != hipSuccess) return fail("generation");
	ctx->d_tab_s = s;
	ctx->d_tab_a = a;
	ctx->d_tab_off = doff;
	ctx->d_tab_tiles_s = dts;
	ctx->d_tab_tiles_a = dta;
	ctx->tab_ntiles_s = (int)ts.size();
	ctx->tab_ntiles_a = (int)ta.size();
	return SPHSWE_OK;
}

void sphswe_leg_tables_free(sphswe_ctx *ctx)
{
	ctx->buf_gen++;   // a captured FP32 step may read the tables
	for (void *p : {(void *)ctx->d_tab_s, (void *)ctx->d_tab_a, (void *)ctx->d_tab_off, (void *)ctx->d_tab_tiles_s,
			(void *)ctx->d_tab_tiles_a})
		if (p) hipFree(p);
	ctx->d_tab_s = ctx->d_tab_a = nullptr;
	ctx->d_tab_off = nullptr;
	ctx->d_tab_tiles_s = ctx->d_tab_tiles_a = nullptr;
}

int sphswe_leg_synth_tab(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf)
{
	if (int rc = sphswe_leg_tables(ctx)) return rc;
	dim3 grid(ctx->tab_ntiles_s, (2 * nf + 16 * TAB_CT - 1) / (16 * TAB_CT));
	k_tab_synth<<<grid, 64 * TAB_W, 0, ctx->stream>>>(ctx->d_tab_s, ctx->d_tab_off, ctx->d_tab_tiles_s, spec, four, ctx->L,
						  ctx->nlat, ctx->NJ, nf);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

int sphswe_leg_anal_tab(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf)
{
	if (int rc = sphswe_leg_tables(ctx)) return rc;
	dim3 grid(ctx->tab_ntiles_a, (2 * nf + 16 * TAB_CT - 1) / (16 * TAB_CT));
	k_tab_anal<<<grid, 64 * TAB_W, 0, ctx->stream>>>(ctx->d_tab_a, ctx->d_tab_off, ctx->d_tab_tiles_a, four, ctx->d_wg, spec,
						 ctx->L, ctx->nlat, ctx->NJ, nf);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}
