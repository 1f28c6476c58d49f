// Associated-Legendre half of the SHT (the Legendre part of SHTNS' spat_to_SH /
// SH_to_spat and their complex variants).
//
// Pbar_l^m(mu_j) is never stored: every lane runs the FP64 three-term
// recurrence Pbar_l = a_lm (mu Pbar_{l-1} - b_lm Pbar_{l-2}) for its own
// latitude, restarted every 32 degrees from seeds computed at setup (values
// that underflow there stay below 2^-981 over the segment).  Only the
// northern rings are visited; the southern ones follow from
// Pbar_l^m(-mu) = (-1)^(l-m) Pbar_l^m(mu), so every pass splits into an even
// and an odd partial sum.
//
// Synthesis: lanes are rings; the spectral coefficient of degree l is uniform
// over the wavefront (scalar loads), so a step costs 3 FP64 ops of recurrence
// plus 2 FMAs per (field, re/im).
// Analysis: the sum over rings is a contraction, done on the FP64 matrix
// cores: a wavefront writes a tile of Pbar (32 degrees x 64 rings) to LDS and
// multiplies it into the hemisphere-combined, weighted Fourier rows.
//
// "Slots": a real field is one slot; a complex field is two (+m and -m, which
// share Pbar up to the factor (-1)^m).
#include "sphswe_leg_common.h"

namespace {

// ----------------------------------------------------------------------------
// synthesis
// ----------------------------------------------------------------------------

// Synthesis: a workgroup handles the pair of orders (m, L - m) (together L + 2
// degrees, so every workgroup has the same work) for 128 rings, two per lane
// (two independent recurrence chains per lane).  Its wavefronts take
// contiguous groups of 32-degree segments of one order each and restart the
// recurrence from the stored seeds at every segment.  Recurrence coefficients
// and spectral values are uniform over a wavefront (scalar loads); the partial
// sums of the wavefronts of an order are reduced through LDS in a fixed order
// (deterministic).
// W wavefronts per workgroup, RPL rings per lane (64 RPL rings per
// workgroup); the partial sums of the wavefronts of one order are added in a
// fixed pairwise tree through LDS, at most W/2 wavefronts' partials at a time.
template <int NS, int W, int RPL>
__host__ __device__ inline size_t synth_lds_bytes()
{
	return sizeof(double) * (size_t)(W / 2) * NS * 4 * 64 * RPL;
}


// partial sums of degrees [s0, se] (s0 at a segment start) of order am over
// this lane's rings.  Every segment restarts the recurrence from its stored
// seeds; a seed that underflows only drops terms below 2^-981 (the recurrence
// grows by less than 2^41 over a segment).
template <int LAY, int NS, int SY_RPL>
__device__ __forceinline__ void synth_range(const double2 *__restrict__ spec, const Slot<LAY> (&sl)[NS], int am, int s0,
					    int se, int so, const Geo &g, const int (&jcr)[SY_RPL], const double (&mu)[SY_RPL],
					    double (&eR)[SY_RPL][NS], double (&eI)[SY_RPL][NS], double (&oR)[SY_RPL][NS],
					    double (&oI)[SY_RPL][NS])
{
	const int L = g.L, NJ = g.NJ, seg = g.seg;
	for (int sb = s0; sb <= se; sb += seg) {
		const int segend = min(sb + seg - 1, se);
		const long k = (long)so + (sb - am) / seg;
		double p0[SY_RPL], p1[SY_RPL];
#pragma unroll
		for (int r = 0; r < SY_RPL; r++) {
			const double2 sd = g.seed[k * NJ + jcr[r]];
			p0[r] = sd.x;
			p1[r] = sd.y;
		}
		// degrees sb and sb + 1 are the seeds (sb - am is even; the seed of a
		// degree past L is 0)
		{
			double2 qe[NS], qo[NS];
#pragma unroll
			for (int s = 0; s < NS; s++) {
				qe[s] = spec[sl[s].q(sb, am)];
				qo[s] = sb + 1 <= segend ? spec[sl[s].q(sb + 1, am)] : make_double2(0.0, 0.0);
			}
#pragma unroll
			for (int r = 0; r < SY_RPL; r++)
#pragma unroll
				for (int s = 0; s < NS; s++) {
					eR[r][s] = fma(qe[s].x, p0[r], eR[r][s]);
					eI[r][s] = fma(qe[s].y, p0[r], eI[r][s]);
					oR[r][s] = fma(qo[s].x, p1[r], oR[r][s]);
					oI[r][s] = fma(qo[s].y, p1[r], oI[r][s]);
				}
		}
		// pairs of degrees; coefficients and spectral values are uniform over
		// the wavefront and come through scalar loads, issued one pair ahead
		// so that their latency overlaps the arithmetic of the current pair
		// (the last pair of the segment reloads itself: every load stays
		// within [sb, segend])
		int l = sb + 2;
		if (l + 1 <= segend) {
			double2 ce = g.coef[lm_index(L, l, am)], co = g.coef[lm_index(L, l + 1, am)];
			double2 qe[NS], qo[NS];
#pragma unroll
			for (int s = 0; s < NS; s++) { qe[s] = spec[sl[s].q(l, am)]; qo[s] = spec[sl[s].q(l + 1, am)]; }
			for (;;) {
				const bool more = l + 3 <= segend;
				const int ln = more ? l + 2 : l;
				const double2 ce_n = g.coef[lm_index(L, ln, am)], co_n = g.coef[lm_index(L, ln + 1, am)];
				double2 qe_n[NS], qo_n[NS];
#pragma unroll
				for (int s = 0; s < NS; s++) { qe_n[s] = spec[sl[s].q(ln, am)]; qo_n[s] = spec[sl[s].q(ln + 1, am)]; }
#pragma unroll
				for (int r = 0; r < SY_RPL; r++) {
					const double Pe = rec_step(ce, mu[r], p1[r], p0[r]);
					const double Po = rec_step(co, mu[r], Pe, p1[r]);
					p0[r] = Pe;
					p1[r] = Po;
#pragma unroll
					for (int s = 0; s < NS; s++) {
						eR[r][s] = fma(qe[s].x, Pe, eR[r][s]);
						eI[r][s] = fma(qe[s].y, Pe, eI[r][s]);
						oR[r][s] = fma(qo[s].x, Po, oR[r][s]);
						oI[r][s] = fma(qo[s].y, Po, oI[r][s]);
					}
				}
				l += 2;
				if (!more) break;
				ce = ce_n;
				co = co_n;
#pragma unroll
				for (int s = 0; s < NS; s++) {
					qe[s] = qe_n[s];
					qo[s] = qo_n[s];
				}
			}
		}
		if (l == segend) {
			// a last single degree (even parity)
			const double2 c = g.coef[lm_index(L, l, am)];
			double2 q[NS];
#pragma unroll
			for (int s = 0; s < NS; s++) q[s] = spec[sl[s].q(l, am)];
#pragma unroll
			for (int r = 0; r < SY_RPL; r++) {
				const double P = rec_step(c, mu[r], p1[r], p0[r]);
#pragma unroll
				for (int s = 0; s < NS; s++) {
					eR[r][s] = fma(q[s].x, P, eR[r][s]);
					eI[r][s] = fma(q[s].y, P, eI[r][s]);
				}
			}
		}
	}
}

template <int LAY, int NS, int SY_WAVES, int SY_RPL>
__global__ __launch_bounds__(64 * SY_WAVES) void k_leg_synth(const double2 *__restrict__ spec,
							     double2 *__restrict__ four, Geo g, int nfields,
							     const double *__restrict__ oscale)
{
	constexpr int SY_RW = 64 * SY_RPL;
	extern __shared__ __align__(16) unsigned char lds_raw[];
	double *red = (double *)lds_raw;     // [W / 2][NS*4][SY_RW]
	const int L = g.L, NJ = g.NJ, seg = g.seg;
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int ma = blockIdx.y, mb = L - (int)blockIdx.y;
	const bool two = mb > ma;            // the middle order of an even L stands alone
	const int nrow0 = L - ma + 1, nrow1 = two ? L - mb + 1 : 0;
	const int nseg0 = (nrow0 + seg - 1) / seg, nseg1 = (nrow1 + seg - 1) / seg;
	const int kA = synth_split<SY_WAVES>(nseg0, nseg1);
	// this wavefront's order and segment range [sa, sb)
	const int part = __builtin_amdgcn_readfirstlane(wave >= kA);
	const int am = __builtin_amdgcn_readfirstlane(part ? mb : ma);
	const int nw = part ? SY_WAVES - kA : kA, wi = part ? wave - kA : wave, ns = part ? nseg1 : nseg0;
	const int sa = wi * ns / nw, sb = (wi + 1) * ns / nw;

	int jr[SY_RPL], jcr[SY_RPL];
	double mu[SY_RPL];
#pragma unroll
	for (int r = 0; r < SY_RPL; r++) {
		jr[r] = blockIdx.x * SY_RW + r * 64 + lane;
		jcr[r] = jr[r] < NJ ? jr[r] : NJ - 1;
		mu[r] = g.ct[jcr[r]];
	}
	Slot<LAY> sl[NS];
	make_slots<LAY, NS>(sl, am, blockIdx.z, nfields, g);

	double eR[SY_RPL][NS], eI[SY_RPL][NS], oR[SY_RPL][NS], oI[SY_RPL][NS];
#pragma unroll
	for (int r = 0; r < SY_RPL; r++)
#pragma unroll
		for (int s = 0; s < NS; s++) eR[r][s] = eI[r][s] = oR[r][s] = oI[r][s] = 0.0;
	if (sb > sa)
		synth_range<LAY, NS, SY_RPL>(spec, sl, am, __builtin_amdgcn_readfirstlane(am + sa * seg),
					     __builtin_amdgcn_readfirstlane(min(am + sb * seg - 1, L)), g.segoff[am], g, jcr,
					     mu, eR, eI, oR, oI);
	// pairwise tree over the nw wavefronts of this order: in each round the
	// upper half [h, n) writes its partials, the lower half adds them; every
	// wavefront takes part in the same number of rounds (barriers)
	constexpr int REG = NS * 4 * SY_RW;
	const int roff = part ? kA / 2 : 0;   // first LDS region of this order
	int n = nw;
	for (int span = max(kA, SY_WAVES - kA); span > 1; span = (span + 1) / 2) {
		const int h = (n + 1) / 2;
		if (n > 1 && wi >= h && wi < n) {
			double *rw = red + (size_t)(roff + wi - h) * REG;
#pragma unroll
			for (int r = 0; r < SY_RPL; r++)
#pragma unroll
				for (int s = 0; s < NS; s++) {
					const int o = r * 64 + lane;
					rw[(4 * s + 0) * SY_RW + o] = eR[r][s];
					rw[(4 * s + 1) * SY_RW + o] = eI[r][s];
					rw[(4 * s + 2) * SY_RW + o] = oR[r][s];
					rw[(4 * s + 3) * SY_RW + o] = oI[r][s];
				}
		}
		__syncthreads();
		if (n > 1 && wi < n - h) {
			const double *rw = red + (size_t)(roff + wi) * REG;
#pragma unroll
			for (int r = 0; r < SY_RPL; r++)
#pragma unroll
				for (int s = 0; s < NS; s++) {
					const int o = r * 64 + lane;
					eR[r][s] += rw[(4 * s + 0) * SY_RW + o];
					eI[r][s] += rw[(4 * s + 1) * SY_RW + o];
					oR[r][s] += rw[(4 * s + 2) * SY_RW + o];
					oI[r][s] += rw[(4 * s + 3) * SY_RW + o];
				}
		}
		__syncthreads();
		if (n > 1) n = h;
	}
	if (wi != 0) return;
#pragma unroll
	for (int r = 0; r < SY_RPL; r++) {
		const int j = jr[r];
		if (j >= NJ) continue;
		const int js = g.nlat - 1 - j;
#pragma unroll
		for (int s = 0; s < NS; s++) {
			if (!sl[s].valid) continue;
			const double sc = oscale ? sl[s].sgn * oscale[j] : sl[s].sgn;   // oscale symmetric in j
			double2 nn = make_double2(sc * (eR[r][s] + oR[r][s]), sc * (eI[r][s] + oI[r][s]));
			double2 ss2 = make_double2(sc * (eR[r][s] - oR[r][s]), sc * (eI[r][s] - oI[r][s]));
			if (!(LAY != LAY_REAL) && am == 0) { nn.y = 0.0; ss2.y = 0.0; }  // Im q_l0 is ignored
			four[sl[s].fb + j] = nn;
			four[sl[s].fb + js] = ss2;
		}
	}
}

// ----------------------------------------------------------------------------
// analysis (FP64 MFMA)
// ----------------------------------------------------------------------------

// One wavefront per (m, 32-degree segment) covers every ring, so no reduction
// crosses wavefronts.  For each 64-ring chunk the wavefront
//   - loads its B operands, the weighted hemisphere combinations E of the
//     chunk's rings for its column (written once per (m, chunk) by the
//     operand pass; columns of absent fields read as 0);
//   - runs the recurrence with lane = ring over the 32 degrees of its segment
//     and writes the Pbar tile to its own LDS region, split by parity of l-m;
//   - contracts over rings with v_mfma_f64_16x16x4f64:
//       Q[16 rows of one parity][16 columns] += Pbar[rows][4 rings] * E[4 rings][cols]
//     where the columns are (slot, re/im).
// Operand layouts of v_mfma_f64_16x16x4f64 (wave64): A[i][k] in lane i + 16k,
// B[k][n] in lane n + 16k, D[i][n] in lane n + 16 (i % 4), element i / 4.
constexpr int AN_SEG = 32;        // degrees per wavefront (16 of each parity)
constexpr int AN_WAVES = 4;       // independent wavefronts per workgroup
constexpr int AN_PS = 65;         // LDS row stride (doubles) of the Pbar tile
constexpr int AN_NS = 8;          // slots per group: 16 MFMA columns

template <int LAY, int NS>
__global__ __launch_bounds__(64 * AN_WAVES) void k_leg_anal(const double *__restrict__ E, double2 *__restrict__ spec, Geo g,
							    const int *__restrict__ unit_m, int nunits, int group0,
							    int nfields)
{
	__shared__ double Pt[AN_WAVES][2][16][AN_PS];
	__shared__ double2 Cfs[AN_WAVES][AN_SEG];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int u = blockIdx.x * AN_WAVES + wave;
	if (u >= nunits) return;   // whole wavefront; no workgroup barrier below
	const int am = unit_m[u];
	const int L = g.L, NJ = g.NJ;
	const int so = g.segoff[am];
	const int s0 = am + (u - so) * AN_SEG, se = min(s0 + AN_SEG - 1, L);
	const int nch = (NJ + 63) / 64;
	const int n = lane & 15, kq = lane >> 4;
	const Column<LAY, NS> col = column_of<LAY, NS>(n >> 1, am, group0 + (int)blockIdx.y, nfields);
	const int comp = n & 1;

	// recurrence coefficients of the segment, read back as LDS broadcasts
	double2 *cf = Cfs[wave];
	if (lane < AN_SEG) cf[lane] = g.coef[lm_index(L, min(s0 + lane, se), am)];
	wave_sync();

	const double *Em = E + ((size_t)blockIdx.y * (L + 1) + am) * nch * (2 * 16 * 64);
	d4v acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
	double (*P)[16][AN_PS] = Pt[wave];
	// operands of chunk c: E values (B) and the Pbar tile (lane = ring; rings
	// past NJ give finite values that meet weight 0; rows past se, of a short
	// last segment, are never stored)
	auto load_e = [&](int c, double (&be)[16], double (&bo)[16]) __attribute__((always_inline)) {
		const double *Ec = Em + (size_t)c * (2 * 16 * 64);
		// unconditional loads (the buffer spans all 16 columns): the columns of
		// absent fields hold stale values, which reach only their own output
		// column of the MFMA, and that column is never stored
#pragma unroll
		for (int s = 0; s < 16; s++) {
			be[s] = Ec[eop_at<PM_F64>(n, s, kq)];
			bo[s] = Ec[eop_at<PM_F64>(n, 16 + s, kq)];
		}
	};
	// seed and mu of this lane's ring in chunk c (chunks past the last reload
	// the last); they are loaded two chunks ahead, so that the recurrence never
	// waits for the E loads issued after them (loads complete in order)
	auto ring_of = [&](int c, double2 &sd, double &mu) __attribute__((always_inline)) {
		const int j = min(c, nch - 1) * 64 + lane;
		const int jc = j < NJ ? j : NJ - 1;
		mu = g.ct[jc];
		sd = g.seed[(long)u * NJ + jc];
	};
	auto make_tile = [&](double2 sd, double mu) __attribute__((always_inline)) {
		P[0][0][lane] = sd.x;
		P[1][0][lane] = sd.y;
		rec_tile<AN_SEG>(cf, mu, sd.x, sd.y, [&](int t, double v) { P[t & 1][t >> 1][lane] = v; });
	};
	// software pipeline: the tile of chunk c is read into registers, its MFMAs
	// are issued, and while the matrix core works through them the recurrence
	// of chunk c + 1 rewrites the tile and its E values are loaded
	double be[16], bo[16];
	double2 sd, sdn;
	double mu, mun;
	ring_of(0, sd, mu);
	ring_of(1, sdn, mun);
	load_e(0, be, bo);
	make_tile(sd, mu);
	wave_sync();
	for (int c = 0; c < nch; c++) {
		double a0[16], a1[16];
#pragma unroll
		for (int s = 0; s < 16; s++) {
			a0[s] = P[0][n][4 * s + kq];
			a1[s] = P[1][n][4 * s + kq];
		}
		wave_sync();   // every lane has its A operands: the tile is free
		// contraction over the 64 rings of the chunk, 4 per MFMA
#pragma unroll
		for (int s = 0; s < 16; s++) {
			acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[s], be[s], acc0, 0, 0, 0);
			acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[s], bo[s], acc1, 0, 0, 0);
		}
		if (c + 1 < nch) {
			sd = sdn;
			mu = mun;
			ring_of(c + 2, sdn, mun);
			load_e(c + 1, be, bo);
			make_tile(sd, mu);
			wave_sync();
		}
	}
	// D[i][n] in lane n + 16 (i % 4), element i / 4; row i of parity p is degree s0 + 2i + p
	if (!col.valid) return;
	double *sdo = (double *)spec;
#pragma unroll
	for (int r = 0; r < 4; r++) {
		const int i = kq + 4 * r;
#pragma unroll
		for (int p = 0; p < 2; p++) {
			const int l = s0 + 2 * i + p;
			if (l <= se) sdo[2 * column_q<LAY, NS>(col, am, l, g) + comp] = col.sgn * (p ? acc1[r] : acc0[r]);
		}
	}
}

// ----------------------------------------------------------------------------
// analysis on the 4x4x4 FP64 matrix instruction, no operand pass
// ----------------------------------------------------------------------------

// v_mfma_f64_4x4x4f64 computes four independent 4x4x4 products per wavefront
// (it sustains a higher FP64 rate than the 16x16x4 form).  Lane layout
// (checked on the device): lane = 16 k + 4 blk + x, block blk = (lane >> 2) & 3,
//   A_blk[i][k] in lane 16 k + 4 blk + i,
//   B_blk[k][n] in lane 16 k + 4 blk + n,
//   D_blk[i][n] in lane 16 i + 4 blk + n.
// A workgroup of AW wavefronts takes AW * 32 degrees of one order m (a block of
// the table ablk) and a group of up to four real fields.  Wavefront w takes the
// 32 degrees H = d0 + 32 w ...: lanes 0-31 run the recurrence for the first 16
// degrees, lanes 32-63 for the next 16 (each from the seeds of its 16-degree
// segment), both over the 32 rings of a chunk (lane & 31).  The tile
// T[ring][row] holds Pbar with rows 16 h + (even degrees 0-7 | odd 8-15) of
// half h.  Per 4-ring k-step t and row set s (= half), blocks blk are
// (parity blk >> 1, degree quad blk & 1), and column group cg (two slots) is
// one instruction:
//   D[s][cg] += Pbar[4 degrees][4 rings] * E[4 rings][4 columns],
// E[ring][col] = w_j (F_N + F_S) for col < 8, w_j (F_N - F_S) for col >= 8,
// col = 8 parity + 2 slot + re/im, staged per chunk by the workgroup from the
// Fourier rows (coalesced) into one of two LDS buffers.
constexpr int A4_PS = 33;   // LDS row stride (doubles) of a tile row (one ring)
constexpr int A4_ES = 17;   // LDS row stride of a staged E row

template <int AW, int NSL>
__global__ __launch_bounds__(64 * AW) void k_leg_anal4(const double2 *__restrict__ four, double2 *__restrict__ spec,
							Geo g, const double2 *__restrict__ seed16,
							const int *__restrict__ segoff16, const int2 *__restrict__ ablk,
							int nblk, int nblk_pad, int group, int nfields)
{
	constexpr int NCG = NSL / 2;            // column groups of four columns (two slots)
	constexpr int ES = 4 * NSL + 1;         // LDS row stride of a staged E row (padded)
	constexpr int NST = (NSL * 32 + 64 * AW - 1) / (64 * AW);   // staging pairs per thread
	__shared__ double T[AW][32][A4_PS];
	__shared__ double Eb[2][32][ES];
	__shared__ double2 Cf[AW][32];
	const int lb = xcd_remap(blockIdx.x, nblk_pad);
	if (lb >= nblk) return;   // whole workgroup
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int2 bk = ablk[lb];
	const int am = bk.x;
	const int L = g.L, NJ = g.NJ, nlat = g.nlat;
	const int H = bk.y + 32 * wave;             // first degree of this wavefront
	const bool active = H <= L;                 // wavefront-uniform
	const int rh = lane >> 5, rr = lane & 31;   // recurrence: half, ring of the chunk
	const int h0 = H + 16 * rh;                 // first degree of this lane's half
	const bool ract = h0 <= L;
	const int useg = segoff16[am] + (min(h0, L) - am) / 16;
	const int nch = (NJ + 31) / 32;
	// MFMA lane roles
	const int mk = lane >> 4, mblk = (lane >> 2) & 3, mx = lane & 3;
	const int f0 = group * NSL;

	// staging: pair = (slot, ring) of a chunk, NST pairs per thread
	struct Stage {
		double2 fn[NST], fs[NST];
		double w[NST];
	};
	auto stage_load = [&](int c) __attribute__((always_inline)) {
		Stage t;
#pragma unroll
		for (int k = 0; k < NST; k++) {
			const int pr = threadIdx.x + k * 64 * AW;
			const int sl = pr >> 5, r = pr & 31;
			const int j = c * 32 + r, jc = min(j, NJ - 1);
			const bool ok = pr < NSL * 32 && f0 + sl < nfields;
			const double2 *Fs = four + ((long)(ok ? f0 + sl : 0) * (L + 1) + am) * nlat;
			t.w[k] = (ok && j < NJ) ? g.wg[jc] : 0.0;
			t.fn[k] = ok ? Fs[jc] : make_double2(0.0, 0.0);
			t.fs[k] = ok ? Fs[nlat - 1 - jc] : make_double2(0.0, 0.0);
		}
		return t;
	};
	auto stage_store = [&](int buf, const Stage &t) __attribute__((always_inline)) {
#pragma unroll
		for (int k = 0; k < NST; k++) {
			const int pr = threadIdx.x + k * 64 * AW;
			if (pr >= NSL * 32) continue;
			const int sl = pr >> 5, r = pr & 31;
			double *e = Eb[buf][r];
			e[2 * sl] = t.w[k] * (t.fn[k].x + t.fs[k].x);
			e[2 * sl + 1] = t.w[k] * (t.fn[k].y + t.fs[k].y);
			e[2 * NSL + 2 * sl] = t.w[k] * (t.fn[k].x - t.fs[k].x);
			e[2 * NSL + 2 * sl + 1] = t.w[k] * (t.fn[k].y - t.fs[k].y);
		}
	};
	double2 *cf = Cf[wave];
	if (lane < 32) cf[lane] = g.coef[lm_index(L, min(H + lane, L), am)];
	double (*Tw)[A4_PS] = T[wave];
	// seed and mu of this lane's ring in chunk c, loaded one chunk ahead
	struct RingIn {
		double2 sd;
		double mu;
	};
	auto ring_in = [&](int c) __attribute__((always_inline)) {
		const int jc = min(c * 32 + rr, NJ - 1);
		RingIn r;
		r.mu = g.ct[jc];
		r.sd = (active && ract) ? seed16[(long)useg * NJ + jc] : make_double2(0.0, 0.0);
		return r;
	};
	// the tile rows of this lane's half: degree h0 + t is row 16 rh + (t odd ? 8 + t/2 : t/2)
	auto make_tile = [&](const RingIn &ri) __attribute__((always_inline)) {
		double *row = Tw[rr] + 16 * rh;
		double p0 = ri.sd.x, p1 = ri.sd.y;
		row[0] = p0;
		row[8] = p1;
#pragma unroll
		for (int tb = 2; tb < 16; tb += 7) {
			double2 cc[7];
			double v[7];
#pragma unroll
			for (int k = 0; k < 7; k++) cc[k] = cf[16 * rh + tb + k];
#pragma unroll
			for (int k = 0; k < 7; k++) {
				v[k] = rec_step(cc[k], ri.mu, p1, p0);
				p0 = p1;
				p1 = v[k];
			}
#pragma unroll
			for (int k = 0; k < 7; k++) {
				const int t = tb + k;
				row[(t & 1) ? 8 + (t >> 1) : (t >> 1)] = v[k];
			}
		}
	};
	double acc[2][NCG];
#pragma unroll
	for (int s = 0; s < 2; s++)
#pragma unroll
		for (int cg = 0; cg < NCG; cg++) acc[s][cg] = 0.0;

	// global loads run two chunks ahead of their use (HBM latency is thousands
	// of cycles); the loop is unrolled by two so that the register sets
	// alternate without copies (a copy would wait for the load in flight)
	{
		const Stage t0 = stage_load(0);
		stage_store(0, t0);
	}
	const RingIn r0 = ring_in(0);
	Stage sA = stage_load(min(1, nch - 1)), sB = stage_load(min(2, nch - 1));
	RingIn rA = ring_in(min(1, nch - 1)), rB = ring_in(min(2, nch - 1));
	wave_sync();   // the coefficients
	if (active) make_tile(r0);
	__syncthreads();
	// chunk c: contract the tile and staged operands of c, make the tile of
	// c + 1 from r_use, stage c + 1 from s_use, then load c + 3 into them
	auto chunk = [&](int c, Stage &s_use, RingIn &r_use) __attribute__((always_inline)) {
		const bool more = c + 1 < nch;
		if (active) {
			double a[8][2], b[8][NCG];
#pragma unroll
			for (int t = 0; t < 8; t++) {
				const double *tr = Tw[4 * t + mk];
				a[t][0] = tr[4 * mblk + mx];
				a[t][1] = tr[16 + 4 * mblk + mx];
				const double *er = Eb[c & 1][4 * t + mk] + 2 * NSL * (mblk >> 1) + mx;
#pragma unroll
				for (int cg = 0; cg < NCG; cg++) b[t][cg] = er[4 * cg];
			}
			wave_sync();   // the tile is free
#pragma unroll
			for (int t = 0; t < 8; t++)
#pragma unroll
				for (int s = 0; s < 2; s++)
#pragma unroll
					for (int cg = 0; cg < NCG; cg++)
						acc[s][cg] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[t][s], b[t][cg], acc[s][cg], 0, 0, 0);
			if (more) make_tile(r_use);
		}
		if (more) {
			stage_store((c + 1) & 1, s_use);
			__syncthreads();
		}
		s_use = stage_load(min(c + 3, nch - 1));
		r_use = ring_in(min(c + 3, nch - 1));
	};
	for (int c = 0; c < nch; c += 2) {
		chunk(c, sA, rA);
		if (c + 1 < nch) chunk(c + 1, sB, rB);
	}
	if (!active) return;
	// D[i][n] of block mblk in lane 16 i + 4 mblk + n: i = mk, n = mx; row
	// 16 s + 4 mblk + i -> degree H + 16 s + 2 (4 (mblk & 1) + i) + (mblk >> 1)
	const int par = mblk >> 1, q = 4 * (mblk & 1) + mk;
	const long nlm = (long)(L + 1) * (L + 2) / 2;
#pragma unroll
	for (int cg = 0; cg < NCG; cg++) {
		const int f = f0 + 2 * cg + (mx >> 1);
		if (f >= nfields) continue;
		double *sdo = (double *)spec + 2 * (f * nlm + lm_index(L, 0, am)) + (mx & 1);
#pragma unroll
		for (int s = 0; s < 2; s++) {
			const int l = H + 16 * s + 2 * q + par;
			if (l <= L) sdo[2 * (long)l] = acc[s][cg];
		}
	}
}

// groups of NSL fields, workgroups of AW wavefronts (32 AW degrees; the block
// table has 64-degree blocks, so AW = 2)
template <int AW, int NSL>
int launch_anal4(sphswe_ctx *ctx, const Geo &g, const double2 *four, double2 *spec, int nf)
{
	static_assert(AW == 2, "block table of 64 degrees");
	const int nblk = ctx->nablk, nblk_pad = (nblk + 7) & ~7;
	const int2 *tab = ctx->d_ablk;
	for (int gr = 0; gr < (nf + NSL - 1) / NSL; gr++) {
		k_leg_anal4<AW, NSL><<<nblk_pad, 64 * AW, 0, ctx->stream>>>(four, spec, g, ctx->d_seed16, ctx->d_segoff16,
									     tab, nblk, nblk_pad, gr, nf);
		SPHSWE_LAUNCH_CHECK(ctx);
	}
	return SPHSWE_OK;
}

template <int LAY, int NS, int W, int RPL>
int launch_synth_wr(sphswe_ctx *ctx, const Geo &g, const double2 *spec, double2 *four, int nf, const double *oscale)
{
	const int per = (LAY != LAY_REAL) ? NS / 2 : NS;
	const size_t lds = synth_lds_bytes<NS, W, RPL>();
	if ((long)lds > ctx->max_lds) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "synthesis needs %zu bytes of LDS", lds);
	if (lds > 65536)
		SPHSWE_HIP(ctx, hipFuncSetAttribute((const void *)k_leg_synth<LAY, NS, W, RPL>,
						    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	// one workgroup per (64 RPL rings, pair of orders (m, L - m), group of slots)
	dim3 grid((ctx->NJ + 64 * RPL - 1) / (64 * RPL), ctx->L / 2 + 1, (nf + per - 1) / per);
	k_leg_synth<LAY, NS, W, RPL><<<grid, 64 * W, lds, ctx->stream>>>(spec, four, g, nf, oscale);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

// the workgroup shape chosen for the context: 8 wavefronts x 3 rings per lane
// (the default, 0) or 4 x 2 (1)
template <int LAY, int NS>
int launch_synth(sphswe_ctx *ctx, const Geo &g, const double2 *spec, double2 *four, int nf, const double *oscale)
{
	if (ctx->synth_shape == 1) return launch_synth_wr<LAY, NS, 4, 2>(ctx, g, spec, four, nf, oscale);
	return launch_synth_wr<LAY, NS, 8, 3>(ctx, g, spec, four, nf, oscale);
}

template <int LAY>
int launch_anal(sphswe_ctx *ctx, const Geo &g, const double2 *four, double2 *spec, int nf)
{
	const int nunits = ctx->nseg_total;
	return analysis_batches<LAY, AN_NS, PM_F64>(ctx, g, four, nf, [&](void *E, int g0, int ng) -> int {
		dim3 grid((nunits + AN_WAVES - 1) / AN_WAVES, ng);
		k_leg_anal<LAY, AN_NS><<<grid, 64 * AN_WAVES, 0, ctx->stream>>>((const double *)E, spec, g, ctx->d_unit_m,
										     nunits, g0, nf);
		SPHSWE_LAUNCH_CHECK(ctx);
		return SPHSWE_OK;
	});
}

// real transforms take every mode, the complex and pole-layout ones FP64 .. BF16X3
int check_prec(sphswe_ctx *ctx, int prec, bool real = false)
{
	const int top = real ? SPHSWE_PREC_MIXED : SPHSWE_PREC_BF16X3;
	if (prec < SPHSWE_PREC_FP64 || prec > top)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "precision %d is not available for this transform", prec);
	return SPHSWE_OK;
}

}  // namespace

int sphswe_leg_synth(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec, true)) return rc;
	if (prec == SPHSWE_PREC_MIXED) return sphswe_leg_synth_bt(ctx, spec, four, nf);
	if (prec == SPHSWE_PREC_FP32 && ctx->tab_min_synth > 0 && nf >= ctx->tab_min_synth)
		return sphswe_leg_synth_tab(ctx, spec, four, nf);
	if (prec != SPHSWE_PREC_FP64) return sphswe_leg_synth_mx(ctx, spec, four, nf, prec);
	if (ctx->synth_mfma_min > 0 && nf >= ctx->synth_mfma_min) return sphswe_leg_synth_f64_mfma(ctx, spec, four, nf);
	const long nlm = ctx->nlm;
	const long nfour = (long)(ctx->L + 1) * ctx->nlat;
	// groups of 4 and 3 fields, no group of one where avoidable: a group
	// reruns the recurrence, so a lone field costs nearly as much as four
	int n4 = nf / 4, n3 = 0, n2 = 0, n1 = 0;
	if (ctx->synth_shape == 0) {
		// the 8 x 3 shape: groups of three (its four-field form is slower)
		n4 = 0;
		n3 = nf / 3;
		n2 = nf % 3 == 2;
		n1 = nf % 3 == 1;
	} else {
		switch (nf % 4) {
		case 1:
			if (nf >= 9) { n4 -= 2; n3 = 3; } else if (nf == 5) { n4 = 0; n3 = 1; n2 = 1; } else n1 = 1;
			break;
		case 2:
			if (nf >= 6) { n4 -= 1; n3 = 2; } else n2 = 1;
			break;
		case 3: n3 = 1; break;
		}
	}
	const Geo g = geo_of(ctx);
	int rc = SPHSWE_OK;
	int f = 0;
	if (n4 && (rc = launch_synth<LAY_REAL, 4>(ctx, g, spec + f * nlm, four + f * nfour, 4 * n4, nullptr))) return rc;
	f += 4 * n4;
	if (n3 && (rc = launch_synth<LAY_REAL, 3>(ctx, g, spec + f * nlm, four + f * nfour, 3 * n3, nullptr))) return rc;
	f += 3 * n3;
	if (n2 && (rc = launch_synth<LAY_REAL, 2>(ctx, g, spec + f * nlm, four + f * nfour, 2 * n2, nullptr))) return rc;
	f += 2 * n2;
	if (n1) rc = launch_synth<LAY_REAL, 1>(ctx, g, spec + f * nlm, four + f * nfour, 1, nullptr);
	return rc;
}

int sphswe_leg_anal(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec, true)) return rc;
	if (prec == SPHSWE_PREC_MIXED) return sphswe_leg_anal_bt(ctx, four, spec, nf);
	if (prec == SPHSWE_PREC_FP32 && ctx->tab_min_anal > 0 && nf >= ctx->tab_min_anal && !ctx->ext_escratch)
		return sphswe_leg_anal_tab(ctx, four, spec, nf);
	if (prec != SPHSWE_PREC_FP64) return sphswe_leg_anal_mx(ctx, four, spec, nf, prec);
	// up to four fields on the 4x4x4 kernel (default; it runs one group of four
	// in about two thirds of the 16x16x4 kernel's time), more fields in groups of
	// eight on the 16x16x4 kernel (which wins once its columns are full)
	if (ctx->anal_variant == 0 && nf <= 4) return launch_anal4<2, 4>(ctx, geo_of(ctx), four, spec, nf);
	return launch_anal<LAY_REAL>(ctx, geo_of(ctx), four, spec, nf);
}

int sphswe_leg_synth_cplx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	if (prec != SPHSWE_PREC_FP64) return sphswe_leg_synth_cplx_mx(ctx, spec, four, nf, prec);
	const long nc = ctx->ncplx;
	const long nfour = (long)(2 * ctx->L + 1) * ctx->nlat;
	int rc = SPHSWE_OK;
	int f = 0;
	if (nf >= 2) {
		int g = nf / 2;
		if ((rc = launch_synth<LAY_CPLX, 4>(ctx, geo_of(ctx), spec, four, g * 2, nullptr))) return rc;
		f = g * 2;
	}
	int r = nf - f;
	if (r == 1) rc = launch_synth<LAY_CPLX, 2>(ctx, geo_of(ctx), spec + f * nc, four + f * nfour, 1, nullptr);
	return rc;
}

int sphswe_leg_anal_cplx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	if (prec != SPHSWE_PREC_FP64) return sphswe_leg_anal_cplx_mx(ctx, four, spec, nf, prec);
	return launch_anal<LAY_CPLX>(ctx, geo_of(ctx), four, spec, nf);
}

// pole layout (REXI): nfields columns, row stride P, m-block offsets moff;
// synthesis output multiplied by oscale[j] when given.  Hundreds of columns:
// the synthesis contraction runs on the matrix cores in every precision, the
// FP64 analysis on the FP64 analysis kernel.
int sphswe_leg_synth_pole(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int P, const int *d_moff,
			  const double *oscale, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	return sphswe_leg_synth_pole_mx(ctx, spec, four, nf, P, d_moff, oscale, prec);
}

int sphswe_leg_anal_pole(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int P, const int *d_moff, int prec)
{
	if (int rc = check_prec(ctx, prec)) return rc;
	if (prec != SPHSWE_PREC_FP64) return sphswe_leg_anal_pole_mx(ctx, four, spec, nf, P, d_moff, prec);
	Geo g = geo_of(ctx);
	g.moff = d_moff;
	g.P = P;
	return launch_anal<LAY_POLE>(ctx, g, four, spec, nf);
}

size_t sphswe_leg_anal_group_bytes(const sphswe_ctx *ctx, int prec)
{
	const int nch = (ctx->NJ + 63) / 64;
	size_t per_chunk = 0;
	switch (prec) {
	case SPHSWE_PREC_FP64: per_chunk = eop_bytes_per_chunk<PM_F64>(); break;
	case SPHSWE_PREC_FP32: per_chunk = eop_bytes_per_chunk<PM_F32>(); break;
	case SPHSWE_PREC_BF16: per_chunk = eop_bytes_per_chunk<PM_BF16>(); break;
	default: per_chunk = eop_bytes_per_chunk<PM_BF16X3>(); break;
	}
	return (size_t)(ctx->L + 1) * nch * per_chunk;
}

// sizes the analysis operand buffer for nf real fields in a precision, so that
// later analyses allocate nothing (needed before a stream capture)
int sphswe_leg_anal_reserve(sphswe_ctx *ctx, int nf, int prec)
{
	if (int rc = check_prec(ctx, prec, true)) return rc;
	// the table form needs no operand buffer, only its tables
	if (prec == SPHSWE_PREC_MIXED) return sphswe_leg_bt_tables(ctx);
	const int nch = (ctx->NJ + 63) / 64;
	size_t per_chunk = 0;
	switch (prec) {
	case SPHSWE_PREC_FP64: per_chunk = eop_bytes_per_chunk<PM_F64>(); break;
	case SPHSWE_PREC_FP32: per_chunk = eop_bytes_per_chunk<PM_F32>(); break;
	case SPHSWE_PREC_BF16: per_chunk = eop_bytes_per_chunk<PM_BF16>(); break;
	default: per_chunk = eop_bytes_per_chunk<PM_BF16X3>(); break;
	}
	const size_t gbytes = (size_t)(ctx->L + 1) * nch * per_chunk;
	const int groups = (nf + 7) / 8;
	const size_t gb = std::max<size_t>(1, std::min<size_t>(groups, ANALYSIS_MAX_EBYTES / gbytes));
	void *E;
	return sphswe_escratch(ctx, gb * gbytes, &E);
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
