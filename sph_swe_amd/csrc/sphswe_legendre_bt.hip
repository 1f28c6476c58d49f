// Real Legendre transforms of many fields from stored bf16 hi/lo tables on the
// double-K bf16 matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation.
//
// Every associated-Legendre value Pbar_l^m(mu_j) of the northern rings is
// computed once per context by the FP64 recurrence (the same segment seeds as
// the FP64 kernels) and stored as two bf16 terms, hi = bf16(P) and
// lo = bf16(P - hi), so that hi + lo carries 16 significant bits.  The field
// operands are split the same way when they are staged, and a product is
//   hi_a hi_b + hi_a lo_b + lo_a hi_b
// (the lo * lo term is below the representation error of the split).
//
// Per order m and parity p = (l - m) mod 2, with degree l = m + p + 2k,
// k = 0..np-1, and the NJ northern rings j:
//   synthesis  E_p[j][c] = sum_k Pbar(m, p, k, j) Q[l][c],
//              F_N = E_0 + E_1, F_S = E_0 - E_1 (Pbar_l(-mu) = (-1)^(l-m) Pbar_l(mu));
//   analysis   Q[l][c] = sum_j Pbar(m, p, k, j) w_j (F_N + (-1)^p F_S)[j][c];
// the columns c are (field, re / im) pairs.  A workgroup takes the pair of
// orders (m, L - m), whose degree counts add up to L + 2, so that every
// workgroup does the same work, and a group of 16 CT columns.
//
// Tables, in MFMA operand order.  A fragment is one 16 x 32 A operand: 64
// lanes x 8 bf16 of hi terms, then the same of lo terms (2 KB); lane
// i + 16 g holds row i and the 8 consecutive contraction indices 8 g .. 8 g + 7.
//   TS (synthesis, rows = rings, contraction = degrees): fragment
//      soff[2m + p] + rt * nkb(m, p) + kb for ring tile rt (16 rings) and degree
//      block kb (32 degrees k), nkb = ceil(np / 32);
//   TA (analysis, rows = degrees, contraction = rings): fragment
//      aoff[2m + p] + dt * NKR + kb for degree tile dt (16 degrees k) and ring
//      block kb (32 rings), NKR = ceil(NJ / 32).
// Entries past np or NJ are zero.  At T511: 227 MB (TS) + 214 MB (TA).
// Operand layouts of v_mfma_f32_16x16x32_bf16 (checked on the device):
//   A[i][k] in lane i + 16 (k / 8), element k % 8; B[k][n] in lane n + 16 (k / 8),
//   element k % 8; D[i][n] in lane n + 16 (i / 4), element i % 4.
#include "sphswe_leg_common.h"

namespace {

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef short s8v __attribute__((ext_vector_type(8)));

constexpr int BT_CTS = 5;   // synthesis: most column fragments per workgroup
constexpr int BT_CTA = 2;   // analysis: most column fragments per workgroup
constexpr int BT_SEG = 32;  // degrees per recurrence segment (seed spacing)

__host__ __device__ inline int bt_np(int L, int m, int p) { return (L + 1 - m + 1 - p) / 2; }

__device__ __forceinline__ f4v mfma_hl(uint4 ah, uint4 al, uint4 bh, uint4 bl, f4v acc)
{
	const bf8v a0 = __builtin_bit_cast(bf8v, ah), a1 = __builtin_bit_cast(bf8v, al);
	const bf8v b0 = __builtin_bit_cast(bf8v, bh), b1 = __builtin_bit_cast(bf8v, bl);
	acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
	return acc;
}

// hi and lo bf16 terms of eight values, packed as 16-byte lane operands; the
// conversions round to nearest even (v_cvt_pk_bf16_f32, two values each)
__device__ __forceinline__ void split8(const float (&v)[8], uint4 &hi, uint4 &lo)
{
	bf8v h, l;
#pragma unroll
	for (int e = 0; e < 8; e++) {
		h[e] = (__bf16)v[e];
		l[e] = (__bf16)(v[e] - (float)h[e]);
	}
	hi = __builtin_bit_cast(uint4, h);
	lo = __builtin_bit_cast(uint4, l);
}

// one thread per (order, northern ring): the FP64 recurrence from the stored
// 32-degree segment seeds, split into both tables
__global__ __launch_bounds__(64) void k_bt_gen(const double2 *__restrict__ coef, const double2 *__restrict__ seed,
					       const int *__restrict__ segoff, const double *__restrict__ ct,
					       const int *__restrict__ soff, const int *__restrict__ aoff,
					       short *__restrict__ TS, short *__restrict__ TA, int L, int NJ)
{
	const int m = blockIdx.y;
	const int j = blockIdx.x * 64 + threadIdx.x;
	if (j >= NJ) return;
	const double mu = ct[j];
	const int NKR = (NJ + 31) >> 5;
	const int nkb0 = (bt_np(L, m, 0) + 31) >> 5, nkb1 = (bt_np(L, m, 1) + 31) >> 5;
	auto put = [&](int l, double v) {
		const int d = l - m, p = d & 1, k = d >> 1;
		const short hi = bf16_rn((float)v);
		const short lo = bf16_rn((float)(v - (double)bf16_f(hi)));
		{
			const long fr = soff[2 * m + p] + (long)(j >> 4) * (p ? nkb1 : nkb0) + (k >> 5);
			const int ln = (j & 15) + 16 * ((k & 31) >> 3);
			const long o = fr * 1024 + ln * 8 + (k & 7);
			TS[o] = hi;
			TS[o + 512] = lo;
		}
		{
			const long fr = aoff[2 * m + p] + (long)(k >> 4) * NKR + (j >> 5);
			const int ln = (k & 15) + 16 * ((j & 31) >> 3);
			const long o = fr * 1024 + ln * 8 + (j & 7);
			TA[o] = hi;
			TA[o + 512] = lo;
		}
	};
	const int nseg = segoff[m + 1] - segoff[m];
	for (int s = 0; s < nseg; s++) {
		const int s0 = m + BT_SEG * s, se = min(s0 + BT_SEG - 1, L);
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

// the value of lane n ^ 1 (the other lane of a (re, im) column pair): a DPP
// quad permute [1, 0, 3, 2], without a trip through LDS
__device__ __forceinline__ float xor1(float v)
{
	return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xb1, 0xf, 0xf, false));
}

template <typename FT>
__device__ __forceinline__ FT mk2(float a, float b);
template <>
__device__ __forceinline__ float2 mk2<float2>(float a, float b) { return make_float2(a, b); }
template <>
__device__ __forceinline__ double2 mk2<double2>(float a, float b) { return make_double2(a, b); }
// two consecutive complex values (an even index of a row): one 16-byte store in FP32
template <typename FT>
__device__ __forceinline__ void st2(FT *p, float a, float b, float c, float d);
template <>
__device__ __forceinline__ void st2<float2>(float2 *p, float a, float b, float c, float d)
{
	*(float4 *)p = make_float4(a, b, c, d);
}
template <>
__device__ __forceinline__ void st2<double2>(double2 *p, float a, float b, float c, float d)
{
	p[0] = make_double2(a, b);
	p[1] = make_double2(c, d);
}

// ---- column operands ---------------------------------------------------------
// Both contractions take their column operands pre-split into bf16 hi/lo in
// MFMA B-operand order, in 2 KB fragments of 16 columns (8 fields x re / im):
// 64 lanes x 8 bf16 of hi terms, then of lo terms; lane n + 16 g holds column n
// and the 8 contraction indices 8 g .. 8 g + 7 of a 32-long block.
//   SB (synthesis): fragment (kboff[2m + p] + kb) * NCT + ct, degrees
//      k = 32 kb + 8 g + e of parity p, fields 8 ct .. 8 ct + 7;
//   G  (analysis): fragment ((m * NKR + kb) * NCT + ct) * 2 + p, rings
//      j = 32 kb + 8 g + e, value w_j (F_N + (-1)^p F_S);
// NCT = ceil(nf / 8).  Columns of absent fields are left unwritten: a column
// reaches only its own MFMA output column, which is not stored.

// Synthesis operands, optionally with a spectral operator applied first: the
// recipe RCP maps output field fo = member * FO + q to a source field
// member * FS + s and an operator (identity, i m x, or D' x = (1 - mu^2) d/dmu x
// as the band (1 - l) R(l - 1, m) x_{l-1} + (l + 2) S(l + 1, m) x_{l+1},
// SPHOperators.hpp:234-260):
//   BT_RCP_COPY  FO = FS = 1, the field itself;
//   BT_RCP_SWE9  FO = 9, FS = 3: (h, u, v) -> h, u, v, i m h, D'h, i m u, D'u,
//                i m v, D'v (level 0 of the SWE tendency);
//   BT_RCP_SWE2  FO = 2, FS = 4: i m of source 0, D' of source 1 (level 2);
//   BT_RCP_SWE6  FO = 6, FS = 3: (h, u, v) -> h, u, v, D'h, D'u, D'v (level 0
//                when the physical stage forms i m x from the rows of x: the
//                synthesis is linear in each order, so the rows of i m x are
//                i m times the rows of x).
enum { BT_RCP_COPY = 0, BT_RCP_SWE9 = 1, BT_RCP_SWE2 = 2, BT_RCP_SWE6 = 3 };
enum { BT_OP_ID = 0, BT_OP_IM = 1, BT_OP_DMU = 2 };
template <int RCP>
struct Recipe {
	static constexpr int FO = RCP == BT_RCP_SWE9 ? 9 : RCP == BT_RCP_SWE6 ? 6 : RCP == BT_RCP_SWE2 ? 2 : 1;
	static constexpr int FS = RCP == BT_RCP_SWE9 || RCP == BT_RCP_SWE6 ? 3 : RCP == BT_RCP_SWE2 ? 4 : 1;
	static constexpr int FU = RCP == BT_RCP_SWE2 ? 2 : FS;   // the sources used: the first FU of FS
	__device__ static void of(int q, int &s, int &op)
	{
		if (RCP == BT_RCP_SWE9) {
			s = q < 3 ? q : (q - 3) >> 1;
			op = q < 3 ? BT_OP_ID : ((q - 3) & 1) ? BT_OP_DMU : BT_OP_IM;
		} else if (RCP == BT_RCP_SWE6) {
			s = q < 3 ? q : q - 3;
			op = q < 3 ? BT_OP_ID : BT_OP_DMU;
		} else if (RCP == BT_RCP_SWE2) {
			s = q;
			op = q ? BT_OP_DMU : BT_OP_IM;
		} else {
			s = 0;
			op = BT_OP_ID;
		}
	}
};

// A workgroup takes one (order m, 64-degree block) unit and MB members: it
// copies the source fields' degrees d0 - 1 .. d0 + 64 of the unit, and the D'
// band of its degrees, into LDS (consecutive lanes, consecutive degrees), then
// eight threads per output field apply the operator from there: lane s of a
// wavefront takes field 8 w + (s >> 1) & 7, the real (s & 1 = 0) or imaginary
// part, and the 16 degrees 16 (s >> 4) .. + 15 of the unit, and writes 16-byte
// slot s of a fragment's hi and lo halves, so that every store of a wavefront
// covers 1 KB contiguously.  The register budget keeps WPE wavefronts per SIMD
// (three workgroups of the level-0 recipe per CU).
template <int RCP>
struct SprepShape {
	static constexpr int MB = RCP == BT_RCP_SWE9 || RCP == BT_RCP_SWE6 ? 8 : 16;   // members per workgroup
	static constexpr int NSRC = MB * Recipe<RCP>::FU;        // source fields staged
	static constexpr int NOUT = MB * Recipe<RCP>::FO;        // output fields written
	static constexpr int NT = 8 * NOUT;
	static constexpr int DW = 66;                             // staged degrees per field
	static constexpr int WPE = RCP == BT_RCP_SWE9 || RCP == BT_RCP_SWE6 ? 7 : 4;
	static_assert(NOUT % 8 == 0 && NT <= 1024, "fragment-aligned output fields");
};

template <int RCP>
__global__ __launch_bounds__(SprepShape<RCP>::NT) __attribute__((amdgpu_waves_per_eu(SprepShape<RCP>::WPE)))
void k_bt_sprep(const double2 *__restrict__ spec, uint4 *__restrict__ SB, const int2 *__restrict__ mkb,
		const int *__restrict__ kboff, const double2 *__restrict__ dmu, int L, int nmem, int NCT)
{
	typedef Recipe<RCP> RC;
	typedef SprepShape<RCP> SH;
	__shared__ double2 xs[SH::NSRC][SH::DW];
	__shared__ double2 cs[64];   // (cm, cp) of D' for the unit's degrees
	const int2 e = mkb[blockIdx.x];
	const int m = e.x, kb = e.y;
	const int mem0 = blockIdx.y * SH::MB;
	const long nlm = (long)(L + 1) * (L + 2) / 2;
	const long lm0 = lm_index(L, m, m);
	const int dmax = L - m;
	const int db = 64 * kb - 1;   // l - m of staged degree 0
	// all loads in flight before the first LDS store
	constexpr int NIT = (SH::NSRC * SH::DW + SH::NT - 1) / SH::NT;
	double2 v[NIT];
#pragma unroll
	for (int it = 0; it < NIT; it++) {
		const int i = threadIdx.x + it * SH::NT;
		const int f = i / SH::DW, dd = i - f * SH::DW;
		const int d = db + dd;
		const int ms = f / RC::FU, fs = f - ms * RC::FU;   // member, source of the member
		const bool ok = i < SH::NSRC * SH::DW && mem0 + ms < nmem && d >= 0 && d <= dmax;
		v[it] = ok ? spec[(long)((mem0 + ms) * RC::FS + fs) * nlm + lm0 + d] : make_double2(0.0, 0.0);
	}
	double2 cv = make_double2(0.0, 0.0);
	if (threadIdx.x < 64 && 64 * kb + (int)threadIdx.x <= dmax) cv = dmu[lm0 + 64 * kb + threadIdx.x];
#pragma unroll
	for (int it = 0; it < NIT; it++) {
		const int i = threadIdx.x + it * SH::NT;
		if (i < SH::NSRC * SH::DW) (&xs[0][0])[i] = v[it];
	}
	if (threadIdx.x < 64) cs[threadIdx.x] = cv;
	__syncthreads();
	const int t = threadIdx.x, sl = t & 63;
	const int fl = 8 * (t >> 6) + ((sl >> 1) & 7), ri = sl & 1, g = sl >> 4;
	const int fo = mem0 * RC::FO + fl;
	if (fo >= nmem * RC::FO) return;
	const int ml = fl / RC::FO, qd = fl - ml * RC::FO;
	int sf, op;
	RC::of(qd, sf, op);
	// x[i] = degree 16 g + i of the unit (x[-1] and x[16] are the neighbours)
	const double2 *x = xs[ml * RC::FU + sf] + 1 + 16 * g;
	const double2 *c = cs + 16 * g;
	float q[2][8];
	if (op == BT_OP_DMU) {
#pragma unroll
		for (int i = 0; i < 16; i++) {
			const double2 cc = c[i], a = x[i - 1], b = x[i + 1];
			q[i & 1][i >> 1] = (float)(ri ? cc.x * a.y + cc.y * b.y : cc.x * a.x + cc.y * b.x);
		}
	} else {
		const double dm = op == BT_OP_IM ? m : 0.0;
#pragma unroll
		for (int i = 0; i < 16; i++) {
			const double2 w = x[i];   // zero past the order's last degree
			q[i & 1][i >> 1] = (float)(op == BT_OP_IM ? (ri ? dm * w.x : -dm * w.y) : (ri ? w.y : w.x));
		}
	}
#pragma unroll
	for (int p = 0; p < 2; p++) {
		if (kb >= (bt_np(L, m, p) + 31) >> 5) continue;
		uint4 h, l;
		uint4 *o = SB + ((long)(kboff[2 * m + p] + kb) * NCT + (fo >> 3)) * 128 + sl;
		split8(q[p], h, l);
		o[0] = h;
		o[64] = l;
	}
}

// A workgroup takes (order m, ring block kb) and the 8 fields of one column
// fragment: one thread per (field, ring) reads both hemispheres' Fourier
// values (consecutive lanes, consecutive rings) and forms w_j (F_N +- F_S);
// one thread per (field, 8-ring group, quantity) then gathers its 8 values
// through LDS and writes their hi and lo lanes.
template <typename FT>
__global__ __launch_bounds__(256) void k_bt_gprep(const FT *__restrict__ four, const double *__restrict__ wg,
						  uint4 *__restrict__ G, int L, int nlat, int NJ, int nf, int NCT)
{
	__shared__ float gv[8][4][33];   // [field][er, ei, or, oi][ring]
	const int NKR = (NJ + 31) >> 5;
	const int mkb = blockIdx.x;
	const int m = mkb / NKR;
	const int ct = blockIdx.y;
	const int tid = threadIdx.x;
	{
		const int fl = tid >> 5, r = tid & 31;
		const int f = ct * 8 + fl;
		const int j = (mkb - m * NKR) * 32 + r;
		float er = 0.f, ei = 0.f, orr = 0.f, oi = 0.f;
		if (f < nf && j < NJ) {
			const FT *row = four + ((long)f * (L + 1) + m) * nlat;
			const FT a = row[j], b = row[nlat - 1 - j];
			const double w = wg[j];
			er = (float)(w * ((double)a.x + (double)b.x));
			ei = (float)(w * ((double)a.y + (double)b.y));
			orr = (float)(w * ((double)a.x - (double)b.x));
			oi = (float)(w * ((double)a.y - (double)b.y));
		}
		gv[fl][0][r] = er;
		gv[fl][1][r] = ei;
		gv[fl][2][r] = orr;
		gv[fl][3][r] = oi;
	}
	__syncthreads();
	if (tid >= 128) return;
	const int fl = tid >> 4, g = (tid >> 2) & 3, qn = tid & 3;   // quantity 2 p + re/im
	const int f = ct * 8 + fl;
	if (f >= nf) return;
	float v[8];
#pragma unroll
	for (int i = 0; i < 8; i++) v[i] = gv[fl][qn][8 * g + i];
	uint4 h, l;
	split8(v, h, l);
	const int p = qn >> 1, ri = qn & 1;
	uint4 *o = G + (((long)mkb * NCT + ct) * 2 + p) * 128 + 2 * fl + ri + 16 * g;
	o[0] = h;
	o[64] = l;
}

// blocks of one order pair (all its ring and column groups) on one XCD, in
// consecutive dispatch order, so that the table fragments read by the column
// groups and the column fragments read by the ring groups come from its L2
__device__ __forceinline__ int bt_block(int nblk)
{
	const int per = (nblk + 7) >> 3;
	return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// ---- synthesis ----------------------------------------------------------------
// Workgroup (pair q, ring group rg of 2 RT ring tiles = 32 RT rings, column
// group cg of CT fragments) of four wavefronts: wavefront (parity p, ring half
// h) takes parity p of ring tiles 2 RT rg + RT h + r, r < RT, for all CT
// column fragments (4 RT CT accumulator registers).  The 32-degree blocks of both
// orders of the pair form one sequence of steps.  Per step the workgroup copies
// the column fragments of both parities into one of two LDS buffers while the
// wavefronts contract the other: each wavefront loads its RT table fragments
// straight from HBM (each is used by this wavefront only) one step ahead and
// issues 3 MFMAs per (ring tile, column fragment), 3 RT CT per step.  At the end
// of an order the odd-parity wavefronts hand their sums to the even-parity
// ones through the (drained) LDS buffers, which form F_N = E_0 + E_1 and
// F_S = E_0 - E_1 and write both hemispheres.
//
// EM = 1 (the divergence level of the SWE tendency, fields (P0, P1) per
// member): instead of Fourier rows, every wavefront emits the analysis operand
// of the per-ring combination O = -(P0 / sq + P1 / (sq st)) scale for its own
// parity.  The combination has coefficients constant along a ring, so it
// commutes with the longitude transforms: the analysis Fourier rows of O are
// 2 pi times the combination of the synthesised rows (nphi > 2 L, no products),
// and the analysis operand w_j (F_N + (-1)^p F_S) is 4 pi w_j times the
// combination of the parity-p sums E_p.  It is written pre-split in the
// analysis fragment order (one field per member, NCT1 = ceil(nmem / 8)).
template <int CT, typename FT, int DA = 2, int RT = 4, int WPE = 2, int EM = 0>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
void k_bt_synth(const uint4 *__restrict__ TS, const int *__restrict__ soff, const int *__restrict__ kboff,
		const uint4 *__restrict__ SB, FT *__restrict__ four, int L, int nlat, int NJ, int nf, int NCT, int nrg,
		int ncg, uint4 *__restrict__ G1, const double *__restrict__ sq, const double *__restrict__ st,
		const double *__restrict__ wg, float scale)
{
	static_assert(EM == 0 || RT == 2, "one 32-ring block per wavefront");
	__shared__ uint4 Bs[2][2][CT][128];   // [buf][parity][ct][hi/lo, lane]
	constexpr int NSTG = CT;              // 16-byte copies per thread per step (256 CT in all)
	const int nblk = ((L + 2) / 2) * nrg * ncg;
	const int lb = bt_block(nblk);
	if (lb >= nblk) return;
	const int cg = lb % ncg, rg = (lb / ncg) % nrg, q = lb / (ncg * nrg);
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const int wp = wave & 1, wh = wave >> 1;
	const int NJt = (NJ + 15) >> 4;
	const int rt0 = rg * 2 * RT + wh * RT;
	// the two orders of the pair (L even: the middle order once); per-order
	// values are selected by the order index, never held in indexed arrays
	// (which would live in scratch memory)
	struct Ord {
		int m, nk0, nk1, kbo0, kbo1, nkw;
		long fa;
	};
	auto ord_of = [&](int m) {
		Ord r;
		r.m = m;
		r.nk0 = (bt_np(L, m, 0) + 31) >> 5;
		r.nk1 = (bt_np(L, m, 1) + 31) >> 5;
		r.kbo0 = kboff[2 * m];
		r.kbo1 = kboff[2 * m + 1];
		r.nkw = wp ? r.nk1 : r.nk0;
		r.fa = soff[2 * m + wp];
		return r;
	};
	const Ord O0 = ord_of(q), O1 = ord_of(L - q);
	const int nord = L - q > q ? 2 : 1;
	const int T0 = O0.nk0, T = T0 + (nord == 2 ? O1.nk0 : 0);
	auto sload = [&](int t, uint4 (&v)[NSTG]) {
		const bool o = t >= T0;
		const int kb = o ? t - T0 : t;
		const int nk0 = o ? O1.nk0 : O0.nk0, nk1 = o ? O1.nk1 : O0.nk1;
		const int kbo0 = o ? O1.kbo0 : O0.kbo0, kbo1 = o ? O1.kbo1 : O0.kbo1;
#pragma unroll
		for (int i = 0; i < NSTG; i++) {
			const int idx = tid + i * 256;
			const int p = idx / (128 * CT), r = idx % (128 * CT), ct = r >> 7, w = r & 127;
			const int ctg = cg * CT + ct;
			const bool ok = kb < (p ? nk1 : nk0) && ctg < NCT;
			v[i] = ok ? SB[((long)((p ? kbo1 : kbo0) + kb) * NCT + ctg) * 128 + w] : uint4{0, 0, 0, 0};
		}
	};
	auto sstore = [&](int buf, const uint4 (&v)[NSTG]) {
#pragma unroll
		for (int i = 0; i < NSTG; i++) (&Bs[buf][0][0][0])[tid + i * 256] = v[i];
	};
	auto aload = [&](int t, uint4 (&a)[RT][2]) {
		const bool o = t >= T0;
		const int kb = o ? t - T0 : t;
		const int nkw = o ? O1.nkw : O0.nkw;
		const long fa = o ? O1.fa : O0.fa;
		if (kb >= nkw) return;
#pragma unroll
		for (int r = 0; r < RT; r++) {
			if (rt0 + r >= NJt) continue;
			const long fr = fa + (long)(rt0 + r) * nkw + kb;
			a[r][0] = TS[fr * 128 + lane];
			a[r][1] = TS[fr * 128 + 64 + lane];
		}
	};
	f4v acc[RT][CT];
#pragma unroll
	for (int r = 0; r < RT; r++)
#pragma unroll
		for (int c = 0; c < CT; c++) acc[r][c] = f4v{0.f, 0.f, 0.f, 0.f};
	// end of order m: odd-parity sums through LDS, then F_N, F_S from the
	// even-parity wavefronts; lane n + 16 g holds rings 4 g + e of column n: the
	// even lane of a (re, im) pair writes rings 4 g, 4 g + 1, the odd lane
	// rings 4 g + 2, 4 g + 3
	// EM = 1: the analysis operand of this wavefront's parity and 32-ring block
	auto emit_g1 = [&](int m) {
		// per wavefront: [column][ring] of one fragment half, then the ring
		// factors (a, b, weight) of its 32 rings
		float *T = (float *)&Bs[0][0][0][0] + wave * (8 * 33 + 96);
		float *Fc = T + 8 * 33;
		const int n = lane & 15, g = lane >> 4, ri = n & 1;
		const int kb1 = rt0 >> 1, NKR = (NJ + 31) >> 5, nmem = nf >> 1, NCT1 = (nmem + 7) >> 3;
		__syncthreads();
		if (lane < 32) {
			const int j = min(32 * kb1 + lane, NJ - 1);
			const double ic = 1.0 / sq[j];
			Fc[lane] = (float)(-scale * ic);
			Fc[32 + lane] = (float)(-scale * ic / st[j]);
			Fc[64 + lane] = 32 * kb1 + lane < NJ ? (float)(4.0 * M_PI * wg[j]) : 0.f;
		}
		wave_sync();
#pragma unroll
		for (int c = 0; c < CT; c++) {
			const int ctg = cg * CT + c;   // columns 16 ctg + n: member 4 ctg + n / 4, field (n >> 1) & 1
			if (ctg >= NCT) break;
#pragma unroll
			for (int r = 0; r < RT; r++)
#pragma unroll
				for (int e = 0; e < 4; e++) {
					const float x = acc[r][c][e];
					const float y = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4e, 0xf, 0xf, false));
					const int jj = 16 * r + 4 * g + e;
					float v = (x * Fc[jj] + y * Fc[32 + jj]) * Fc[64 + jj];
					if (m == 0 && ri) v = 0.f;
					if (!(n & 2)) T[(2 * (n >> 2) + ri) * 33 + jj] = v;
				}
			wave_sync();
			if (lane < 32) {
				const int n1 = lane & 7, g1 = lane >> 3;
				float v8[8];
#pragma unroll
				for (int i = 0; i < 8; i++) v8[i] = T[n1 * 33 + 8 * g1 + i];
				uint4 h, l;
				split8(v8, h, l);
				const int mem1 = 4 * ctg + (n1 >> 1);
				if (mem1 < nmem && kb1 < NKR) {   // ring groups rounded up past the last block write nothing
					uint4 *o = G1 + (((long)(m * NKR + kb1) * NCT1 + (mem1 >> 3)) * 2 + wp) * 128 +
						   2 * (mem1 & 7) + (n1 & 1) + 16 * g1;
					o[0] = h;
					o[64] = l;
				}
			}
			wave_sync();
		}
#pragma unroll
		for (int r = 0; r < RT; r++)
#pragma unroll
			for (int c = 0; c < CT; c++) acc[r][c] = f4v{0.f, 0.f, 0.f, 0.f};
		__syncthreads();
	};
	auto emit = [&](int m) {
		if constexpr (EM == 1) {
			emit_g1(m);
			return;
		}
		f4v *X = (f4v *)&Bs[0][0][0][0];   // [h][r][c][lane]: 2 * RT * CT * 64 * 16 bytes, within the LDS of Bs
		__syncthreads();
		if (wp) {
#pragma unroll
			for (int r = 0; r < RT; r++)
#pragma unroll
				for (int c = 0; c < CT; c++) X[((wh * RT + r) * CT + c) * 64 + lane] = acc[r][c];
		}
		__syncthreads();
		if (!wp) {
			const int n = lane & 15, g = lane >> 4, odd = n & 1;
#pragma unroll
			for (int r = 0; r < RT; r++) {
				const int rt = rt0 + r;
#pragma unroll
				for (int c = 0; c < CT; c++) {
					const f4v e1 = X[((wh * RT + r) * CT + c) * 64 + lane];
					float fn[4], fs[4], pn[4], ps[4];
#pragma unroll
					for (int e = 0; e < 4; e++) {
						fn[e] = acc[r][c][e] + e1[e];
						fs[e] = acc[r][c][e] - e1[e];
					}
#pragma unroll
					for (int e = 0; e < 4; e++) {
						pn[e] = xor1(fn[e]);
						ps[e] = xor1(fs[e]);
					}
					const int f = (cg * 16 * CT + 16 * c + n) >> 1;
					if (f >= nf || rt >= NJt) continue;
					FT *row = four + ((long)f * (L + 1) + m) * nlat;
					// rings j0 and j0 + 1 (elements 2 odd, 2 odd + 1, selected
					// without indexing by a lane value); both hemispheres' pairs
					// are adjacent in the row
					const int j0 = rt * 16 + 4 * g + 2 * odd;
					float nr[2], ni[2], sr[2], si[2];
#pragma unroll
					for (int h = 0; h < 2; h++) {
						nr[h] = odd ? pn[2 + h] : fn[h];
						ni[h] = odd ? fn[2 + h] : pn[h];
						sr[h] = odd ? ps[2 + h] : fs[h];
						si[h] = odd ? fs[2 + h] : ps[h];
					}
					if (j0 + 1 < NJ) {
						st2<FT>(row + j0, nr[0], ni[0], nr[1], ni[1]);
						st2<FT>(row + nlat - 2 - j0, sr[1], si[1], sr[0], si[0]);
					} else if (j0 < NJ) {
						row[j0] = mk2<FT>(nr[0], ni[0]);
						row[nlat - 1 - j0] = mk2<FT>(sr[0], si[0]);
					}
				}
			}
		}
#pragma unroll
		for (int r = 0; r < RT; r++)
#pragma unroll
			for (int c = 0; c < CT; c++) acc[r][c] = f4v{0.f, 0.f, 0.f, 0.f};
		__syncthreads();
	};
	// table fragments DA - 1 steps ahead in a ring of DA register sets (DA = 1:
	// loaded at the start of their step, fewest registers)
	uint4 qa[DA][RT][2];
	uint4 sv[NSTG];
	sload(0, sv);
#pragma unroll
	for (int u = 0; u + 1 < DA; u++)
		if (u < T) aload(u, qa[u]);
	sstore(0, sv);
	__syncthreads();
	auto step = [&](int t, uint4 (&a)[RT][2], uint4 (&na)[RT][2]) {
		const bool o = t >= T0;
		const int kb = o ? t - T0 : t;
		const int buf = t & 1;
		const bool more = t + 1 < T;
		if (DA == 1) aload(t, a);
		if (more) sload(t + 1, sv);
		if (DA > 1 && t + DA - 1 < T) aload(t + DA - 1, na);
		if (kb < (o ? O1.nkw : O0.nkw)) {
#pragma unroll
			for (int c = 0; c < CT; c++) {
				const uint4 bh = Bs[buf][wp][c][lane], bl = Bs[buf][wp][c][64 + lane];
#pragma unroll
				for (int r = 0; r < RT; r++) acc[r][c] = mfma_hl(a[r][0], a[r][1], bh, bl, acc[r][c]);
			}
		}
		if (t == T0 - 1 || t == T - 1) emit(o ? O1.m : O0.m);
		if (more) sstore(buf ^ 1, sv);
		__syncthreads();
	};
	for (int t0 = 0; t0 < T; t0 += DA) {
#pragma unroll
		for (int u = 0; u < DA; u++) {
			if (t0 + u >= T) break;
			step(t0 + u, qa[u], qa[(u + DA - 1) % DA]);
		}
	}
}

// ---- analysis -----------------------------------------------------------------
// Workgroup (pair q, column group cg, share hs of SPL) of AW wavefronts.  The
// degree tiles of the pair, ordered (order, parity, tile), are dealt out in
// runs of RA consecutive tiles per wavefront, AW RA tiles per share; the
// contraction runs over 32-ring blocks.  Per block the workgroup copies the
// column fragments of the orders and parities its tiles touch into one of two
// LDS buffers; a wavefront reloads them only where its (order, parity)
// changes.  Table fragments and staged column fragments are loaded D blocks
// ahead through register queues.
template <int CT, int RA, int AW, int D>
__global__ __launch_bounds__(64 * AW) void k_bt_anal(const uint4 *__restrict__ TA, const int *__restrict__ aoff,
						     const uint4 *__restrict__ G, double2 *__restrict__ spec, int L, int NJ,
						     int nf, int NCT, int ncg, int SPL)
{
	__shared__ uint4 Bs[2][2][2][CT][128];   // [buf][order][parity][ct][hi/lo, lane]
	constexpr int NSTG = (512 * CT + 64 * AW - 1) / (64 * AW);
	const int nblk = ((L + 2) / 2) * ncg * SPL;
	const int lb = bt_block(nblk);
	if (lb >= nblk) return;
	const int hs = lb % SPL, cg = (lb / SPL) % ncg, q = lb / (SPL * ncg);
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	const int NKR = (NJ + 31) >> 5;
	const long nlm = (long)(L + 1) * (L + 2) / 2;
	const int m1 = L - q;
	const int no = (m1 > q) ? 2 : 1;
	const int c0 = (bt_np(L, q, 0) + 15) >> 4;
	const int c1 = c0 + ((bt_np(L, q, 1) + 15) >> 4);
	const int c2 = c1 + (no == 2 ? (bt_np(L, m1, 0) + 15) >> 4 : 0);
	const int c3 = c2 + (no == 2 ? (bt_np(L, m1, 1) + 15) >> 4 : 0);
	// this wavefront's tiles: (order, parity) group and tile of tiles RA w + s
	int tg[RA], tdt[RA], tfr[RA];
	bool tv[RA];
#pragma unroll
	for (int s = 0; s < RA; s++) {
		const int t = (hs * AW + wave) * RA + s;
		const int grp = (t >= c0) + (t >= c1) + (t >= c2);
		const int base = grp == 0 ? 0 : grp == 1 ? c0 : grp == 2 ? c1 : c2;
		tv[s] = t < c3;
		tg[s] = tv[s] ? grp : 0;
		tdt[s] = tv[s] ? t - base : 0;
		const int m = (tg[s] >> 1) ? m1 : q;
		tfr[s] = tv[s] ? aoff[2 * m + (tg[s] & 1)] + tdt[s] * NKR : 0;
	}
	// the (order, parity) groups 2 o + p that this workgroup's tiles touch:
	// only their column fragments are staged
	const int tw0 = hs * AW * RA, tw1 = min(tw0 + AW * RA, c3) - 1;
	const int glo = (tw0 >= c0) + (tw0 >= c1) + (tw0 >= c2);
	const int ghi = (tw1 >= c0) + (tw1 >= c1) + (tw1 >= c2);
	auto sload = [&](int kb, uint4 (&v)[NSTG]) {
#pragma unroll
		for (int i = 0; i < NSTG; i++) {
			const int idx = tid + i * 64 * AW;
			// idx = ((o * CT + ct) * 2 + p) * 128 + w in G order; LDS order [o][p][ct]
			const int o = idx / (256 * CT), r = idx % (256 * CT), ct = r >> 8;
			const int ctg = cg * CT + ct;
			const int gi = 2 * o + ((r >> 7) & 1);
			const bool ok = idx < 512 * CT && o < no && ctg < NCT && gi >= glo && gi <= ghi;
			const int m = o ? m1 : q;
			v[i] = ok ? G[((long)(m * NKR + kb) * NCT + ctg) * 256 + (r & 255)] : uint4{0, 0, 0, 0};
		}
	};
	auto sstore = [&](int buf, const uint4 (&v)[NSTG]) {
#pragma unroll
		for (int i = 0; i < NSTG; i++) {
			const int idx = tid + i * 64 * AW;
			if (idx >= 512 * CT) continue;
			const int o = idx / (256 * CT), r = idx % (256 * CT), ct = r >> 8, p = (r >> 7) & 1, w = r & 127;
			Bs[buf][o][p][ct][w] = v[i];
		}
	};
	auto aload = [&](int kb, uint4 (&a)[RA][2]) {
#pragma unroll
		for (int s = 0; s < RA; s++) {
			if (!tv[s]) continue;
			const long fr = tfr[s] + kb;
			a[s][0] = TA[fr * 128 + lane];
			a[s][1] = TA[fr * 128 + 64 + lane];
		}
	};
	f4v acc[RA][CT];
#pragma unroll
	for (int s = 0; s < RA; s++)
#pragma unroll
		for (int c = 0; c < CT; c++) acc[s][c] = f4v{0.f, 0.f, 0.f, 0.f};
	uint4 qa[D][RA][2], qs[D][NSTG];
#pragma unroll
	for (int u = 0; u < D; u++)
		if (u < NKR) {
			aload(u, qa[u]);
			sload(u, qs[u]);
		}
	sstore(0, qs[0]);
	__syncthreads();
	for (int k0 = 0; k0 < NKR; k0 += D) {
#pragma unroll
		for (int u = 0; u < D; u++) {
			const int kb = k0 + u;
			if (kb >= NKR) break;
			const int buf = kb & 1;
			uint4 bh[CT], bl[CT];
#pragma unroll
			for (int s = 0; s < RA; s++) {
				if (!tv[s]) continue;
				if (s == 0 || tg[s] != tg[s - 1]) {
					const int o = tg[s] >> 1, p = tg[s] & 1;
#pragma unroll
					for (int c = 0; c < CT; c++) {
						bh[c] = Bs[buf][o][p][c][lane];
						bl[c] = Bs[buf][o][p][c][64 + lane];
					}
				}
#pragma unroll
				for (int c = 0; c < CT; c++) acc[s][c] = mfma_hl(qa[u][s][0], qa[u][s][1], bh[c], bl[c], acc[s][c]);
			}
			if (kb + 1 < NKR) sstore(buf ^ 1, qs[(u + 1) % D]);
			if (kb + D < NKR) {
				aload(kb + D, qa[u]);
				sload(kb + D, qs[u]);
			}
			__syncthreads();
		}
	}
	// lane n + 16 g holds degrees 4 g + e of column n: the even lane of a
	// (re, im) pair writes e = 0, 1, the odd lane e = 2, 3
	const int n = lane & 15, g = lane >> 4, odd = n & 1;
#pragma unroll
	for (int s = 0; s < RA; s++) {
		if (!tv[s]) continue;
		const int p = tg[s] & 1, m = (tg[s] >> 1) ? m1 : q;
		const int np = bt_np(L, m, p);
#pragma unroll
		for (int c = 0; c < CT; c++) {
			float v[4], pv[4];
#pragma unroll
			for (int e = 0; e < 4; e++) v[e] = acc[s][c][e];
#pragma unroll
			for (int e = 0; e < 4; e++) pv[e] = xor1(v[e]);
			const int f = (cg * 16 * CT + 16 * c + n) >> 1;
			if (f >= nf) continue;
			double2 *qo = spec + (long)f * nlm + lm_index(L, m, m) + p;
#pragma unroll
			for (int h = 0; h < 2; h++) {
				const int k = 16 * tdt[s] + 4 * g + 2 * odd + h;
				if (k >= np) continue;
				// element 2 odd + h, selected without indexing by a lane value
				qo[2 * k] = odd ? make_double2(pv[2 + h], v[2 + h]) : make_double2(v[h], pv[h]);
			}
		}
	}
}

// groups of column fragments: the fewest groups of at most maxct fragments,
// with the fragments spread evenly over them
int bt_ct(int NCT, int maxct)
{
	const int ng = (NCT + maxct - 1) / maxct;
	return (NCT + ng - 1) / ng;
}

constexpr int BT_DS = 3;   // synthesis: steps of look-ahead of the loads
constexpr int BT_DA = 3;   // analysis

// two ring tiles per wavefront (64 rings per workgroup) at three wavefronts
// per SIMD: 233 against 244 us (four tiles, two per SIMD) for 144 fields at T511
template <int CT, typename FT, int EM = 0>
int launch_bt_synth(sphswe_ctx *ctx, const uint4 *SB, FT *four, int nf, int NCT, uint4 *G1 = nullptr,
		    float scale = 0.f)
{
	constexpr int RT = 2;
	const int NJt = (ctx->NJ + 15) / 16;
	const int nrg = (NJt + 2 * RT - 1) / (2 * RT), ncg = (NCT + CT - 1) / CT;
	const int nblk = ((ctx->L + 2) / 2) * nrg * ncg;
	k_bt_synth<CT, FT, 2, RT, 3, EM><<<(nblk + 7) / 8 * 8, 256, 0, ctx->stream>>>(
		ctx->d_bt_s, ctx->d_bt_soff, ctx->d_bt_kboff, SB, four, ctx->L, ctx->nlat, ctx->NJ, nf, NCT, nrg, ncg, G1,
		ctx->d_sq, ctx->d_st, ctx->d_wg, scale);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

template <int CT, int RA, int AW, int D = BT_DA>
int launch_bt_anal(sphswe_ctx *ctx, const uint4 *G, double2 *spec, int nf, int NCT, int spl = 1)
{
	const int ncg = (NCT + CT - 1) / CT;
	const int nblk = ((ctx->L + 2) / 2) * ncg * spl;
	k_bt_anal<CT, RA, AW, D><<<(nblk + 7) / 8 * 8, 64 * AW, 0, ctx->stream>>>(ctx->d_bt_a, ctx->d_bt_aoff, G, spec,
										   ctx->L, ctx->NJ, nf, NCT, ncg, spl);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

}  // namespace

// bytes of the pre-split column operands of nf fields
size_t sphswe_bt_sb_bytes(const sphswe_ctx *ctx, int nf) { return (size_t)ctx->bt_nkb * ((nf + 7) / 8) * 2048; }
size_t sphswe_bt_g_bytes(const sphswe_ctx *ctx, int nf)
{
	return (size_t)(ctx->L + 1) * ((ctx->NJ + 31) / 32) * ((nf + 7) / 8) * 4096;
}

template <int RCP>
int bt_sprep(sphswe_ctx *ctx, const double2 *spec, void *SB, int nmem)
{
	if (nmem <= 0) return SPHSWE_OK;
	typedef SprepShape<RCP> SH;
	const int nfo = nmem * Recipe<RCP>::FO;
	dim3 grid(ctx->bt_nmkb, (nmem + SH::MB - 1) / SH::MB);
	k_bt_sprep<RCP><<<grid, SH::NT, 0, ctx->stream>>>(spec, (uint4 *)SB, ctx->d_bt_mkb, ctx->d_bt_kboff, ctx->d_bt_dmu,
							  ctx->L, nmem, (nfo + 7) / 8);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}
int sphswe_bt_sprep(sphswe_ctx *ctx, const double2 *spec, void *SB, int nf)
{
	return bt_sprep<BT_RCP_COPY>(ctx, spec, SB, nf);
}
int sphswe_bt_sprep_swe9(sphswe_ctx *ctx, const double2 *X, void *SB, int nmem)
{
	return bt_sprep<BT_RCP_SWE9>(ctx, X, SB, nmem);
}
int sphswe_bt_sprep_swe6(sphswe_ctx *ctx, const double2 *X, void *SB, int nmem)
{
	return bt_sprep<BT_RCP_SWE6>(ctx, X, SB, nmem);
}
int sphswe_bt_sprep_swe2(sphswe_ctx *ctx, const double2 *S4, void *SB, int nmem)
{
	return bt_sprep<BT_RCP_SWE2>(ctx, S4, SB, nmem);
}

template <typename FT>
int bt_gprep(sphswe_ctx *ctx, const FT *four, void *G, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	const int NCT = (nf + 7) / 8;
	dim3 grid((ctx->L + 1) * ((ctx->NJ + 31) / 32), NCT);
	k_bt_gprep<FT><<<grid, 256, 0, ctx->stream>>>(four, ctx->d_wg, (uint4 *)G, ctx->L, ctx->nlat, ctx->NJ, nf, NCT);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}
int sphswe_bt_gprep(sphswe_ctx *ctx, const double2 *four, void *G, int nf) { return bt_gprep(ctx, four, G, nf); }
int sphswe_bt_gprep32(sphswe_ctx *ctx, const float2 *four, void *G, int nf) { return bt_gprep(ctx, four, G, nf); }

template <typename FT>
int bt_synth_split(sphswe_ctx *ctx, const void *SB, FT *four, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	const int NCT = (nf + 7) / 8;
	const uint4 *sb = (const uint4 *)SB;
	switch (bt_ct(NCT, BT_CTS)) {
	case 1: return launch_bt_synth<1, FT>(ctx, sb, four, nf, NCT);
	case 2: return launch_bt_synth<2, FT>(ctx, sb, four, nf, NCT);
	case 3: return launch_bt_synth<3, FT>(ctx, sb, four, nf, NCT);
	case 4: return launch_bt_synth<4, FT>(ctx, sb, four, nf, NCT);
	default: return launch_bt_synth<5, FT>(ctx, sb, four, nf, NCT);
	}
}
int sphswe_bt_synth(sphswe_ctx *ctx, const void *SB, double2 *four, int nf) { return bt_synth_split(ctx, SB, four, nf); }
int sphswe_bt_synth_div(sphswe_ctx *ctx, const void *SB, void *G1, int nmem, float scale)
{
	if (nmem <= 0) return SPHSWE_OK;
	if (ctx->nlat % 2 != 0 || 2 * ctx->L + 1 > ctx->plan.n)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "divergence level: needs an even ring count and nphi > 2 L");
	const int NCT = (2 * nmem + 7) / 8;
	const uint4 *sb = (const uint4 *)SB;
	uint4 *g = (uint4 *)G1;
	switch (bt_ct(NCT, BT_CTS)) {
	case 1: return launch_bt_synth<1, float2, 1>(ctx, sb, nullptr, 2 * nmem, NCT, g, scale);
	case 2: return launch_bt_synth<2, float2, 1>(ctx, sb, nullptr, 2 * nmem, NCT, g, scale);
	case 3: return launch_bt_synth<3, float2, 1>(ctx, sb, nullptr, 2 * nmem, NCT, g, scale);
	case 4: return launch_bt_synth<4, float2, 1>(ctx, sb, nullptr, 2 * nmem, NCT, g, scale);
	default: return launch_bt_synth<5, float2, 1>(ctx, sb, nullptr, 2 * nmem, NCT, g, scale);
	}
}
int sphswe_bt_synth32(sphswe_ctx *ctx, const void *SB, float2 *four, int nf) { return bt_synth_split(ctx, SB, four, nf); }

int sphswe_bt_anal(sphswe_ctx *ctx, const void *Gv, double2 *spec, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	const int NCT = (nf + 7) / 8;
	const uint4 *G = (const uint4 *)Gv;
	// (tiles per wavefront, wavefronts): every degree tile of the fullest order
	// pair in one workgroup for up to 32 tiles
	const int tm = ctx->bt_tmax;
	const int ct = bt_ct(NCT, BT_CTA);
	if (tm <= 16) return ct == 1 ? launch_bt_anal<1, 1, 16>(ctx, G, spec, nf, NCT) : launch_bt_anal<2, 1, 16>(ctx, G, spec, nf, NCT);
	if (tm <= 32) return ct == 1 ? launch_bt_anal<1, 2, 16>(ctx, G, spec, nf, NCT) : launch_bt_anal<2, 2, 16>(ctx, G, spec, nf, NCT);
	// more tiles: workgroups of four wavefronts with three tiles each, the
	// pair's tiles split over ceil(tm / 12) workgroups (more, smaller
	// workgroups in flight: 154 against 185 us for the 64-field analysis at T511)
	const int spl = (tm + 11) / 12;
	return ct == 1 ? launch_bt_anal<1, 3, 4, 2>(ctx, G, spec, nf, NCT, spl) : launch_bt_anal<2, 3, 4, 2>(ctx, G, spec, nf, NCT, spl);
}

// the tables, built on first use (allocates; not inside a stream capture)
int sphswe_leg_bt_tables(sphswe_ctx *ctx)
{
	if (ctx->d_bt_s) return SPHSWE_OK;
	const int L = ctx->L, NJ = ctx->NJ;
	if (ctx->seg != BT_SEG)
		return sphswe_set_err(ctx, SPHSWE_ERR_INTERNAL, "bf16 Legendre tables: seed spacing %d", ctx->seg);
	const int NJt = (NJ + 15) / 16, NKR = (NJ + 31) / 32;
	std::vector<int> so(2 * (L + 1)), ao(2 * (L + 1));
	long ns = 0, na = 0;
	for (int m = 0; m <= L; m++)
		for (int p = 0; p < 2; p++) {
			const int np = bt_np(L, m, p);
			so[2 * m + p] = (int)ns;
			ao[2 * m + p] = (int)na;
			ns += (long)NJt * ((np + 31) / 32);
			na += (long)((np + 15) / 16) * NKR;
		}
	if (ns > 0x7fffffffL || na > 0x7fffffffL)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "bf16 Legendre tables: too many fragments");
	int tmax = 0;
	for (int q = 0; q < (L + 2) / 2; q++) {
		int t = 0;
		for (int p = 0; p < 2; p++) {
			t += (bt_np(L, q, p) + 15) / 16;
			if (L - q > q) t += (bt_np(L, L - q, p) + 15) / 16;
		}
		tmax = std::max(tmax, t);
	}
	// the (m, p, kb) blocks of the synthesis column operands, in fragment order
	std::vector<int> kbo(2 * (L + 1));
	std::vector<int2> kbl;
	for (int m = 0; m <= L; m++)
		for (int p = 0; p < 2; p++) {
			kbo[2 * m + p] = (int)kbl.size();
			for (int kb = 0; kb < (bt_np(L, m, p) + 31) / 32; kb++) kbl.push_back(make_int2(2 * m + p, kb));
		}
	// the band of D' = (1 - mu^2) d/dmu per degree: (1 - l) R(l - 1, m) on x_{l-1}
	// and (l + 2) S(l + 1, m) on x_{l+1} (SPHOperators.hpp:234-260)
	std::vector<double2> dmu((size_t)(L + 1) * (L + 2) / 2);
	for (int m = 0; m <= L; m++)
		for (int l = m; l <= L; l++)
			dmu[lm_index(L, l, m)] = make_double2((-(double)l + 1.0) * idR(l - 1, m), ((double)l + 2.0) * idS(l + 1, m));
	// the (m, 64-degree block) units of the operand split
	std::vector<int2> mkb;
	for (int m = 0; m <= L; m++)
		for (int kb = 0; kb < (bt_np(L, m, 0) + 31) / 32; kb++) mkb.push_back(make_int2(m, kb));
	uint4 *s = nullptr, *a = nullptr;
	int *dso = nullptr, *dao = nullptr, *dkbo = nullptr;
	int2 *dkbl = nullptr, *dmkb = nullptr;
	double2 *ddmu = nullptr;
	auto fail = [&](const char *what) {
		for (void *p : {(void *)s, (void *)a, (void *)dso, (void *)dao, (void *)dkbo, (void *)dkbl, (void *)dmkb, (void *)ddmu})
			if (p) hipFree(p);
		return sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "bf16 Legendre tables: %s", what);
	};
	if (hipMalloc((void **)&s, (size_t)ns * 2048) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&a, (size_t)na * 2048) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&dso, sizeof(int) * so.size()) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&dao, sizeof(int) * ao.size()) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&dkbo, sizeof(int) * kbo.size()) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&dkbl, sizeof(int2) * kbl.size()) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&dmkb, sizeof(int2) * mkb.size()) != hipSuccess) return fail("allocation");
	if (hipMalloc((void **)&ddmu, sizeof(double2) * dmu.size()) != hipSuccess) return fail("allocation");
	if (hipMemcpy(dso, so.data(), sizeof(int) * so.size(), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dao, ao.data(), sizeof(int) * ao.size(), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dkbo, kbo.data(), sizeof(int) * kbo.size(), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dkbl, kbl.data(), sizeof(int2) * kbl.size(), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dmkb, mkb.data(), sizeof(int2) * mkb.size(), hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(ddmu, dmu.data(), sizeof(double2) * dmu.size(), hipMemcpyHostToDevice) != hipSuccess)
		return fail("upload");
	if (hipMemsetAsync(s, 0, (size_t)ns * 2048, ctx->stream) != hipSuccess ||
	    hipMemsetAsync(a, 0, (size_t)na * 2048, ctx->stream) != hipSuccess)
		return fail("clear");
	k_bt_gen<<<dim3((NJ + 63) / 64, L + 1), 64, 0, ctx->stream>>>(ctx->d_coef, ctx->d_seed, ctx->d_segoff, ctx->d_ct, dso,
								      dao, (short *)s, (short *)a, L, NJ);
	if (hipGetLastError() != hipSuccess) return fail("generation");
	ctx->d_bt_s = s;
	ctx->d_bt_a = a;
	ctx->d_bt_soff = dso;
	ctx->d_bt_aoff = dao;
	ctx->d_bt_kboff = dkbo;
	ctx->d_bt_kbl = dkbl;
	ctx->bt_nkb = (int)kbl.size();
	ctx->d_bt_mkb = dmkb;
	ctx->bt_nmkb = (int)mkb.size();
	ctx->d_bt_dmu = ddmu;
	ctx->bt_tmax = tmax;
	ctx->bt_bytes = (size_t)(ns + na) * 2048;
	return SPHSWE_OK;
}

void sphswe_leg_bt_tables_free(sphswe_ctx *ctx)
{
	ctx->buf_gen++;   // a captured step may read the tables
	for (void *p : {(void *)ctx->d_bt_s, (void *)ctx->d_bt_a, (void *)ctx->d_bt_soff, (void *)ctx->d_bt_aoff,
			(void *)ctx->d_bt_kboff, (void *)ctx->d_bt_kbl, (void *)ctx->d_bt_mkb, (void *)ctx->d_bt_dmu})
		if (p) hipFree(p);
	ctx->d_bt_mkb = nullptr;
	ctx->d_bt_dmu = nullptr;
	ctx->bt_nmkb = 0;
	ctx->d_bt_s = ctx->d_bt_a = nullptr;
	ctx->d_bt_soff = ctx->d_bt_aoff = ctx->d_bt_kboff = nullptr;
	ctx->d_bt_kbl = nullptr;
	ctx->bt_bytes = 0;
	ctx->bt_nkb = 0;
}

// the transforms of the FP64 entry points: Fourier rows in FP64, the column
// operands in the context's operand scratch
int sphswe_leg_synth_bt(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	if (int rc = sphswe_leg_bt_tables(ctx)) return rc;
	void *SB;
	if (int rc = sphswe_escratch(ctx, sphswe_bt_sb_bytes(ctx, nf), &SB)) return rc;
	if (int rc = sphswe_bt_sprep(ctx, spec, SB, nf)) return rc;
	return sphswe_bt_synth(ctx, SB, four, nf);
}
int sphswe_leg_anal_bt(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf)
{
	if (nf <= 0) return SPHSWE_OK;
	if (int rc = sphswe_leg_bt_tables(ctx)) return rc;
	void *G;
	if (int rc = sphswe_escratch(ctx, sphswe_bt_g_bytes(ctx, nf), &G)) return rc;
	if (int rc = sphswe_bt_gprep(ctx, four, G, nf)) return rc;
	return sphswe_bt_anal(ctx, G, spec, nf);
}
