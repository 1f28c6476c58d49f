// Fused nonlinear SWE tendency + classic RK4 over nmembers independent states.
//
// Follows AppSWE::p_run_euler_timestep_update (src/AppSWE.hpp:574-659) and
// TimesteppingRK::run_rk_timestep order 4 (src/include/sweet/TimesteppingRK.hpp:
// 310-392) with the same projection points: every field the reference analyses
// (products, metric-divided gradients) is analysed here too, every operator is
// applied in the same space.  The difference is batching: the reference runs
// about 30 single-field transforms per tendency through a lazy graph; here the
// fields of one projection level go through one batched transform:
//
//   level 0  synthesis of 9 fields:  h, u, v, im*h, D'h, im*u, D'u, im*v, D'v
//            (D' = (1-mu^2) d/dmu, SPHOperators.hpp:234-260)
//   pointwise h*u, h*v*st (div_lat, SPHOperators.hpp:411-439) and the u, v
//            tendencies (grad_lon/grad_lat divide by sqrt(1-mu^2), f = 2 Omega mu);
//            nonlinear = 0 the linear equations (h0 for h in the fluxes, no
//            advection), nonlinear = 2 the difference of the two (flux of
//            h - h0 and advection only: the part a split scheme leaves to RK4);
//            linear with Robert functions (AppSWE.hpp:606-620): the h
//            tendency -(h0/a)(im U + D'V)/cos^2 phi is complete here, the u, v
//            ones have no metric division, and levels 2-3 are skipped
//   level 1  analysis of 4 fields, im* and D' on the two flux fields
//   level 2  synthesis of 2 fields, divide by sqrt(1-mu^2) (and st)
//   level 3  analysis of 1 field (the h tendency)
//   spectral viscosity nu/a^2 laplace and the RK4 combinations.
//
// Sums that the reference forms in spectral space after separate analyses
// (e.g. div_lon(hu) + div_lat(hv)) are formed here in physical space before
// one analysis; the analysis is linear, so the two agree to round-off.
#include "sphswe_common.h"
#include "sphswe_fft_blocks.h"

#include <cstring>
#include <type_traits>

namespace {

struct SweWs {
	size_t bytes = 0;
	void *base = nullptr;
};
std::unordered_map<const sphswe_ctx *, SweWs> g_ws;

// spectral prep: 9 fields per member from (h, u, v)
__global__ void k_prep9(const double2 *__restrict__ X, double2 *__restrict__ S9, const int *__restrict__ mof, int L,
			int nlm, int nmem)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= (long)nlm * nmem) return;
	int mem = (int)(i / nlm);
	int idx = (int)(i - (long)mem * nlm);
	int m = mof[idx];
	int base = lm_index(L, 0, m);
	int l = idx - base;
	const double dl = l, dm = m;
	const double cm = (-dl + 1.0) * idR(dl - 1, dm), cp = (dl + 2.0) * idS(dl + 1, dm);
	const double2 *x = X + (long)mem * 3 * nlm;
	double2 *o = S9 + (long)mem * 9 * nlm;
#pragma unroll
	for (int f = 0; f < 3; f++) {
		const double2 *xf = x + (long)f * nlm + base;
		double2 v = xf[l];
		double2 a = (l - 1 >= m) ? xf[l - 1] : make_double2(0, 0);
		double2 b = (l + 1 <= L) ? xf[l + 1] : make_double2(0, 0);
		o[(long)f * nlm + idx] = v;
		o[(long)(3 + 2 * f) * nlm + idx] = make_double2(-dm * v.y, dm * v.x);
		o[(long)(4 + 2 * f) * nlm + idx] = make_double2(cm * a.x + cp * b.x, cm * a.y + cp * b.y);
	}
}

// pointwise level 0 -> 4 fields per member
__global__ void k_phys1(const double *__restrict__ P9, double *__restrict__ P4, const double *__restrict__ ct,
			const double *__restrict__ st, const double *__restrict__ sq, int nlat, long nspat,
			int nmem, double ga, double two_omega, double ia, int nonlinear, double h0)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= nspat * nmem) return;
	int mem = (int)(i / nspat);
	long k = i - (long)mem * nspat;
	int j = (int)(k % nlat);   // ring of point k ([i * nlat + j])
	const double *p = P9 + (long)mem * 9 * nspat + k;
	double H = p[0], U = p[nspat], V = p[2 * nspat];
	double Hl = p[3 * nspat], Hm = p[4 * nspat];
	double Ul = p[5 * nspat], Um = p[6 * nspat], Vl = p[7 * nspat], Vm = p[8 * nspat];
	const double c = sq[j], mu = ct[j];
	double *o = P4 + (long)mem * 4 * nspat + k;
	if (nonlinear == 1) {
		o[0] = U * H;                         // h*u
		o[nspat] = (V * H) * st[j];           // (h*v)*cos(phi)
		double tu = -ga * (Hl / c) + two_omega * (V * mu) - ((Ul / c) * U + (Um / c) * V) * ia;
		double tv = -ga * (Hm / c) - two_omega * (U * mu) - ((Vl / c) * U + (Vm / c) * V) * ia;
		o[2 * nspat] = tu;
		o[3 * nspat] = tv;
	} else if (nonlinear == 3) {
		// linear, Robert functions: robert_div_lon(U) + robert_div_lat(V) divide
		// i m U and D'V by cos^2 phi (= st^2), robert_grad_lon/lat(h) = i m h, D'h
		o[0] = -((Ul + Vm) / (st[j] * st[j])) * h0;   // h0 carries 1/a here
		o[nspat] = 0.0;
		o[2 * nspat] = -ga * Hl + two_omega * (V * mu);
		o[3 * nspat] = -ga * Hm - two_omega * (U * mu);
	} else if (nonlinear == 2) {
		// the full tendency minus its linear part: flux of h - h0, advection
		const double Hp = H - h0;
		o[0] = U * Hp;
		o[nspat] = (V * Hp) * st[j];
		o[2 * nspat] = -((Ul / c) * U + (Um / c) * V) * ia;
		o[3 * nspat] = -((Vl / c) * U + (Vm / c) * V) * ia;
	} else {
		o[0] = U;
		o[nspat] = V * st[j];
		o[2 * nspat] = -ga * (Hl / c) + two_omega * (V * mu);
		o[3 * nspat] = -ga * (Hm / c) - two_omega * (U * mu);
	}
}

// spectral level 1: im * A(hu), D' A(hv st)
__global__ void k_prep2(const double2 *__restrict__ S4, double2 *__restrict__ S2, const int *__restrict__ mof, int L,
			int nlm, int nmem)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= (long)nlm * nmem) return;
	int mem = (int)(i / nlm);
	int idx = (int)(i - (long)mem * nlm);
	int m = mof[idx];
	int base = lm_index(L, 0, m);
	int l = idx - base;
	const double dl = l, dm = m;
	const double2 *a1 = S4 + (long)mem * 4 * nlm;
	const double2 *a2 = a1 + nlm + base;
	double2 v = a1[idx];
	double2 *o = S2 + (long)mem * 2 * nlm;
	o[idx] = make_double2(-dm * v.y, dm * v.x);
	const double cm = (-dl + 1.0) * idR(dl - 1, dm), cp = (dl + 2.0) * idS(dl + 1, dm);
	double2 a = (l - 1 >= m) ? a2[l - 1] : make_double2(0, 0);
	double2 b = (l + 1 <= L) ? a2[l + 1] : make_double2(0, 0);
	o[nlm + idx] = make_double2(cm * a.x + cp * b.x, cm * a.y + cp * b.y);
}

// pointwise level 2: h tendency in physical space
__global__ void k_phys2(const double *__restrict__ P2, double *__restrict__ P1, const double *__restrict__ st,
			const double *__restrict__ sq, int nlat, long nspat, int nmem, double scale)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= nspat * nmem) return;
	int mem = (int)(i / nspat);
	long k = i - (long)mem * nspat;
	int j = (int)(k % nlat);
	const double *p = P2 + (long)mem * 2 * nspat + k;
	double c = sq[j];
	double dlon = p[0] / c;
	double dlat = (p[nspat] / c) / st[j];
	P1[(long)mem * nspat + k] = -(dlon + dlat) * scale;
}

// ---- the mixed form (SPHSWE_PREC_MIXED): FP32 ring-major physical fields ----
// The same pointwise terms as k_phys1 in FP32 on fields laid out
// [member][field][ring j][longitude], so that ring j of point k is k / nphi.
__global__ void k_phys1_mx(const float *__restrict__ P9, float *__restrict__ P4, const double *__restrict__ ct,
			   const double *__restrict__ st, const double *__restrict__ sq, int nphi, long nspat, int nmem,
			   float ga, float two_omega, float ia, int nonlinear, float h0)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= nspat * nmem) return;
	int mem = (int)(i / nspat);
	long k = i - (long)mem * nspat;
	int j = (int)(k / nphi);
	const float *p = P9 + (long)mem * 9 * nspat + k;
	float H = p[0], U = p[nspat], V = p[2 * nspat];
	float Hl = p[3 * nspat], Hm = p[4 * nspat];
	float Ul = p[5 * nspat], Um = p[6 * nspat], Vl = p[7 * nspat], Vm = p[8 * nspat];
	// the metric divisions as products with reciprocals rounded once from FP64
	const float ic = (float)(1.0 / sq[j]), mu = (float)ct[j], sj = (float)st[j];
	float *o = P4 + (long)mem * 4 * nspat + k;
	if (nonlinear == 1) {
		o[0] = U * H;
		o[nspat] = (V * H) * sj;
		o[2 * nspat] = -ga * (Hl * ic) + two_omega * (V * mu) - ((Ul * ic) * U + (Um * ic) * V) * ia;
		o[3 * nspat] = -ga * (Hm * ic) - two_omega * (U * mu) - ((Vl * ic) * U + (Vm * ic) * V) * ia;
	} else if (nonlinear == 3) {
		o[0] = -((Ul + Vm) * (float)(1.0 / (st[j] * st[j]))) * h0;
		o[nspat] = 0.f;
		o[2 * nspat] = -ga * Hl + two_omega * (V * mu);
		o[3 * nspat] = -ga * Hm - two_omega * (U * mu);
	} else if (nonlinear == 2) {
		const float Hp = H - h0;
		o[0] = U * Hp;
		o[nspat] = (V * Hp) * sj;
		o[2 * nspat] = -((Ul * ic) * U + (Um * ic) * V) * ia;
		o[3 * nspat] = -((Vl * ic) * U + (Vm * ic) * V) * ia;
	} else {
		o[0] = U;
		o[nspat] = V * sj;
		o[2 * nspat] = -ga * (Hl * ic) + two_omega * (V * mu);
		o[3 * nspat] = -ga * (Hm * ic) - two_omega * (U * mu);
	}
}

// The mixed form's physical stage without physical fields in HBM: one
// workgroup per (member, ring pair j, j + 1) runs the inverse FFTs of the
// 9 input fields (from 6 row sets), which leave each thread the longitudes
// k + NT w + t NSL of its last-pass butterflies in registers, forms the
// pointwise terms there (those of k_phys1_mx), and runs the forward FFTs of
// the 4 outputs from LDS to the output Fourier rows.  (The divergence level
// needs no physical stage in this form: its pointwise step is a per-ring
// linear combination, applied to the synthesis sums by the table synthesis.)  The two rings are the real and imaginary parts of one
// complex sequence.  Fields go through the passes FI (inputs) or FO (outputs)
// at a time as the sequences of one tile; one at a time measured fastest (at
// T511 three inputs per round need twice the registers and run 1.9x slower),
// and the next field's Fourier rows are loaded while the current one is
// transformed.  Passes, radices and normalisation are those of the FFT kernels
// with radices of 8 first (at N = 1536, 8 8 8 3 keeps 192 of the 256 threads
// busy in the first passes, 16 16 6 only 96: 527 against 563 us for 16
// members); lengths whose last pass has fewer butterflies than a wavefront
// take the separate kernels.
template <int N, int NTMAX = 256>
struct FusedPlan {
	static constexpr int RULE = 1;
	typedef CPlan<N, RULE> PL;
	static constexpr int NP = PL::npass();
	static constexpr int R0 = PL::radix(0);
	static constexpr int RL = PL::radix(NP - 1), NSL = PL::ns(NP - 1);
	static constexpr int NT = NSL < NTMAX ? NSL : NTMAX;   // threads per workgroup
	static constexpr int NW = NSL / NT;                 // last-pass butterflies per thread and field
	static constexpr int NV = NW * RL;                  // longitudes held per thread
	static constexpr bool ok = NP >= 2 && NSL % 64 == 0 && NSL % NT == 0 && NV <= 8;
};

// The input holds 6 row sets per member (h, u, v, D'h, D'u, D'v); the rows of
// the i m fields are formed from those of h, u, v as they are loaded (the
// synthesis is linear in each order: rows of i m x = i m times rows of x).
template <int MODE, int N>
__global__ __launch_bounds__(FusedPlan<N>::NT) void k_phys_fft_mx(
	const float2 *__restrict__ Fin, float2 *__restrict__ Fout, const float2 *__restrict__ twg,
	const double *__restrict__ ct, const double *__restrict__ st, const double *__restrict__ sq, int nlat, int L,
	int nblk, int nblk_pad, float ga, float two_omega, float ia, float h0)
{
	constexpr int mode = MODE;
	typedef FusedPlan<N> FP;
	constexpr int NIN = 9, NOUT = 4, NSRC = 6;
	// the input row set of field q, and whether it is multiplied by i m
	auto src_of = [](int q) constexpr { return q < 3 ? q : (q & 1) ? (q - 3) >> 1 : 3 + ((q - 4) >> 1); };
	auto im_of = [](int q) constexpr { return q >= 3 && (q & 1); };
	// LINF (modes 1, 2): the terms of O2 and O3 that are linear with per-ring
	// coefficients (pressure gradient and Coriolis) are added to the output rows
	// in Fourier space, where they are 2 pi times the same combination of the
	// input rows (no products, so the transform pair is exact there); i m h and
	// D'h are then not transformed: 7 of the 9 inverse FFTs.  At N = 384 the
	// prefetched rows would cost a wave per SIMD (64-thread workgroups), so the
	// short rings keep all 9.
	constexpr bool LINF = (MODE == 1 || MODE == 2) && N >= 1024;
	constexpr int NQ = LINF ? 7 : NIN;
	auto field_at = [](int qi) constexpr { return !LINF ? qi : qi < 3 ? qi : qi + 2; };
	constexpr int FI = 1, FO = 1;   // fields per round, in and out
	typedef WPad<N, FI> WI;
	typedef WPad<N, FO> WO;
	constexpr int NT = FP::NT, NP = FP::NP, RL = FP::RL, NSL = FP::NSL, NW = FP::NW, NV = FP::NV;
	constexpr int SMN = FI * WI::len > FO * WO::len ? FI * WI::len : FO * WO::len;
	// two buffers: the passes alternate between them (one barrier each); the
	// results of NP - 1 passes from sm are in smr
	__shared__ float2 sm[SMN], sm2[SMN];
	float2 *const smr = (NP - 1) % 2 ? sm2 : sm;
	// the twiddles in LDS: read there, they are not hoisted out of the
	// unrolled field loop and kept in registers
	__shared__ float2 tws[N];
	const int lb = xcd_remap(blockIdx.x, nblk_pad);
	if (lb >= nblk) return;
	const int npair = nlat >> 1;
	const int mem = lb / npair, j = 2 * (lb - mem * npair);
	const int tid = threadIdx.x;
	for (int k = tid; k < N; k += NT) tws[k] = twg[k];
	const long frow = (long)(L + 1) * nlat;   // one field's Fourier rows

	// Each input field's values (ring j in x, ring j + 1 in y, at this thread's
	// longitudes) are folded into the outputs O as the field arrives, so that
	// only U, V and the outputs stay live.  In field order H, U, V, i m H,
	// D'H, i m U, D'U, i m V, D'V:
	//   mode 1  O0 = U H, O1 = V H st, O2 = -ga Hl/c + 2 Omega V mu - (Ul/c U + Um/c V)/a,
	//           O3 = -ga Hm/c - 2 Omega U mu - (Vl/c U + Vm/c V)/a;
	//   mode 0  O0 = U, O1 = V st, O2, O3 without the advection terms;
	//   mode 2  O0 = U (H - h0), O1 = V (H - h0) st, O2, O3 the advection terms only;
	//   mode 3  O0 = -(Ul + Vm) h0 / st^2, O1 = 0, O2 = -ga Hl + 2 Omega V mu,
	//           O3 = -ga Hm - 2 Omega U mu.
	// per ring: 1 / sqrt(1 - mu^2), mu, cos phi and the reciprocal of cos^2 phi,
	// rounded once from FP64 (the terms multiply by them)
	float ric[2], rmu[2], rst[2], ris2[2];
#pragma unroll
	for (int r = 0; r < 2; r++) {
		ric[r] = (float)(1.0 / sq[j + r]);
		rmu[r] = (float)ct[j + r];
		rst[r] = (float)st[j + r];
		ris2[r] = (float)(1.0 / (st[j + r] * st[j + r]));
	}
	float Hs[NV][2], Us[NV][2], Vs[NV][2], O[NOUT][NV][2];
	auto fold = [&](int q, int e, int r, float x) __attribute__((always_inline)) {
		const float ic = ric[r], mu = rmu[r], sj = rst[r];
		switch (q) {
		case 0: Hs[e][r] = mode == 2 ? x - h0 : x; break;
		case 1:
			Us[e][r] = x;
			O[0][e][r] = (mode == 1 || mode == 2) ? x * Hs[e][r] : x;
			break;
		case 2:
			Vs[e][r] = x;
			O[1][e][r] = mode == 3 ? 0.f : (mode == 0 ? x * sj : (x * Hs[e][r]) * sj);
			O[2][e][r] = (mode == 2 || LINF) ? 0.f : two_omega * (x * mu);
			O[3][e][r] = (mode == 2 || LINF) ? 0.f : -(two_omega * (Us[e][r] * mu));
			break;
		case 3:
			if (mode != 2) O[2][e][r] = -ga * (mode == 3 ? x : x * ic) + O[2][e][r];
			break;
		case 4:
			if (mode != 2) O[3][e][r] = -ga * (mode == 3 ? x : x * ic) + O[3][e][r];
			break;
		case 5:
			if (mode == 3) O[0][e][r] = x;
			else if (mode != 0) O[2][e][r] -= ((x * ic) * Us[e][r]) * ia;
			break;
		case 6:
			if (mode == 1 || mode == 2) O[2][e][r] -= ((x * ic) * Vs[e][r]) * ia;
			break;
		case 7:
			if (mode == 1 || mode == 2) O[3][e][r] -= ((x * ic) * Us[e][r]) * ia;
			break;
		case 8:
			if (mode == 3) O[0][e][r] = -((O[0][e][r] + x) * ris2[r]) * h0;
			else if (mode != 0) O[3][e][r] -= ((x * ic) * Vs[e][r]) * ia;
			break;
		}
	};

	// The Fourier rows of the next round are loaded one round ahead: thread
	// tid holds orders m = tid + NT u of each field, both rings (16 bytes); the
	// sequence z[m] = A + iB, z[N - m] = conj(A) + i conj(B) (0 between L and
	// N - L) is written to LDS in natural order and every pass runs from there.
	constexpr int NMU = (N / 2 + NT) / NT;   // orders per thread: m <= L < N / 2
	float4 ab[FI][NMU];
	auto fetch = [&](int q0) __attribute__((always_inline)) {
#pragma unroll
		for (int f = 0; f < FI; f++) {
			const float2 *src = Fin + (long)(mem * NSRC + src_of(field_at(q0 + f))) * frow;
#pragma unroll
			for (int u = 0; u < NMU; u++) {
				const int m = tid + u * NT;
				ab[f][u] = (q0 + f < NQ && m <= L) ? *(const float4 *)(src + (size_t)m * nlat + j)
								   : make_float4(0.f, 0.f, 0.f, 0.f);
			}
		}
	};
	fetch(0);
	// unrolled: every register array is indexed by compile-time numbers only
#pragma unroll
	for (int q0 = 0; q0 < NQ; q0 += FI) {
#pragma unroll
		for (int f = 0; f < FI; f++) {
#pragma unroll
			for (int u = 0; u < NMU; u++) {
				const int m = tid + u * NT;
				float4 v = ab[f][u];
				if (im_of(field_at(q0 + f))) {   // i m A, i m B
					const float fm = (float)m;
					v = make_float4(-fm * v.y, fm * v.x, -fm * v.w, fm * v.z);
				}
				if (m == 0) {
					sm[WI::at(f, 0)] = make_float2(v.x, v.z);
				} else if (m <= L) {
					sm[WI::at(f, m)] = make_float2(v.x - v.w, v.y + v.z);
					sm[WI::at(f, N - m)] = make_float2(v.x + v.w, -v.y + v.z);
				}
			}
			for (int k = L + 1 + tid; k < N - L; k += NT) sm[WI::at(f, k)] = make_float2(0.f, 0.f);
		}
		if (q0 + FI < NQ) fetch(q0 + FI);
		__syncthreads();
		passes_pp_range<N, FI, NT, +1, 0, NP - 1, float2, FP::RULE>(sm, sm2, tws);
		// last pass into registers, folded into the outputs
#pragma unroll
		for (int f = 0; f < FI; f++) {
			if (q0 + f >= NQ) break;
#pragma unroll
			for (int w = 0; w < NW; w++) {
				const int k = tid + w * NT;
				float2 v[RL];
#pragma unroll
				for (int t = 0; t < RL; t++) v[t] = smr[WI::at(f, k + t * NSL)];
#pragma unroll
				for (int t = 1; t < RL; t++) {
					float2 tw = tws[t * k];
					tw.y = -tw.y;
					v[t] = cmul(v[t], tw);
				}
				Dft<RL>::run(v, +1);
#pragma unroll
				for (int t = 0; t < RL; t++) {
					fold(field_at(q0 + f), w * RL + t, 0, v[t].x);
					fold(field_at(q0 + f), w * RL + t, 1, v[t].y);
				}
			}
		}
		__syncthreads();
	}

	// the outputs, FO per round: natural order into LDS, forward passes, last
	// pass to the Fourier rows (butterflies k and NSL - k give both Z[m] and
	// Z[N - m])
	const float hn = (float)(M_PI / N);   // half the analysis normalisation 2 pi / N
	// LINF: 2 pi times the per-ring coefficients of the linear terms
	float lga[2], lom[2];
#pragma unroll
	for (int r = 0; r < 2; r++) {
		lga[r] = (float)(-2.0 * M_PI * ga / sq[j + r]);
		lom[r] = (float)(2.0 * M_PI * two_omega * ct[j + r]);
	}
	constexpr int NH = NSL / 2;
#pragma unroll
	for (int o0 = 0; o0 < NOUT; o0 += FO) {
#pragma unroll
	for (int o = 0; o < FO; o++)
#pragma unroll
		for (int w = 0; w < NW; w++)
#pragma unroll
			for (int t = 0; t < RL; t++)
				sm[WO::at(o, tid + w * NT + t * NSL)] =
					make_float2(O[o0 + o][w * RL + t][0], O[o0 + o][w * RL + t][1]);
	// LINF: the input rows of the linear terms of this output, loaded before its
	// passes (thread tid takes butterflies tid and NSL - tid, orders k + t NSL)
	float4 lp[2][RL], ls[2][RL];
	if (LINF && MODE == 1 && (o0 == 2 || o0 == 3)) {
		static_assert(FO * NH <= NT, "at most one last-pass butterfly pair per thread");
		const float2 *rp = Fin + (long)(mem * NSRC + (o0 == 2 ? 0 : 3)) * frow + j;
		const float2 *rs = Fin + (long)(mem * NSRC + (o0 == 2 ? 2 : 1)) * frow + j;
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int t = 0; t < RL; t++) {
				const int k = h ? (tid == 0 ? NH : NSL - tid) : tid;
				const int m = k + t * NSL;
				const bool ok = tid < FO * NH && m <= L;
				lp[h][t] = ok ? *(const float4 *)(rp + (size_t)m * nlat) : make_float4(0.f, 0.f, 0.f, 0.f);
				ls[h][t] = ok ? *(const float4 *)(rs + (size_t)m * nlat) : make_float4(0.f, 0.f, 0.f, 0.f);
			}
	}
	__syncthreads();
	passes_pp_range<N, FO, NT, -1, 0, NP - 1, float2, FP::RULE>(sm, sm2, tws);
	for (int it = tid; it < FO * NH; it += NT) {
		const int o = it / NH, pp = it - o * NH;
		float2 *dst = Fout + (long)(mem * NOUT + o0 + o) * frow;
		const int k1 = pp, k2 = pp == 0 ? NH : NSL - pp;
		float2 z1[RL], z2[RL];
#pragma unroll
		for (int t = 0; t < RL; t++) {
			z1[t] = smr[WO::at(o, k1 + t * NSL)];
			z2[t] = smr[WO::at(o, k2 + t * NSL)];
		}
#pragma unroll
		for (int t = 1; t < RL; t++) {
			z1[t] = cmul(z1[t], tws[t * k1]);
			z2[t] = cmul(z2[t], tws[t * k2]);
		}
		Dft<RL>::run(z1, -1);
		Dft<RL>::run(z2, -1);
		auto emit = [&](int m, float2 Z, float2 Q, const float4 &P, const float4 &S) __attribute__((always_inline)) {
			if (m > L) return;
			const float2 Zc = cconj(Q);
			float2 a = make_float2(hn * (Z.x + Zc.x), hn * (Z.y + Zc.y));
			float2 c = make_float2(hn * (Z.y - Zc.y), -hn * (Z.x - Zc.x));
			if (LINF && MODE == 1 && (o0 == 2 || o0 == 3)) {
				// O2 += 2 pi (-ga / c i m H + 2 Omega mu V), O3 += 2 pi (-ga / c D'H - 2 Omega mu U)
				// from the rows P, S of ring j (x, y) and j + 1 (z, w)
				const float fm = o0 == 2 ? (float)m : 1.f;
				const float sg = o0 == 2 ? 1.f : -1.f;
				// o0 == 2: i m (x + i y) = (-m y, m x); o0 == 3: D'H itself
				const float p0x = o0 == 2 ? -fm * P.y : P.x, p0y = o0 == 2 ? fm * P.x : P.y;
				const float p1x = o0 == 2 ? -fm * P.w : P.z, p1y = o0 == 2 ? fm * P.z : P.w;
				a.x += lga[0] * p0x + sg * lom[0] * S.x;
				a.y += lga[0] * p0y + sg * lom[0] * S.y;
				c.x += lga[1] * p1x + sg * lom[1] * S.z;
				c.y += lga[1] * p1y + sg * lom[1] * S.w;
			}
			if (m == 0) a.y = c.y = 0.f;
			*(float4 *)(dst + (size_t)m * nlat + j) = make_float4(a.x, a.y, c.x, c.y);
		};
#pragma unroll
		for (int t = 0; t < RL; t++) {
			const float2 q1 = pp == 0 ? z1[(RL - t) % RL] : z2[RL - 1 - t];
			const float2 q2 = pp == 0 ? z2[RL - 1 - t] : z1[RL - 1 - t];
			emit(k1 + t * NSL, z1[t], q1, lp[0][t], ls[0][t]);
			emit(k2 + t * NSL, z2[t], q2, lp[1][t], ls[1][t]);
		}
	}
	__syncthreads();
	}
}

// tendencies k = (T_h, T_u, T_v) + nu/a^2 laplace(Y) and the RK4 bookkeeping
//   stage < 3: acc (=|+=) b_s k ; Ynext = X + dt*(a_{s+1} k)
//   stage 3:   X += dt*(acc + b_3 k)
__global__ void k_stage(const double2 *__restrict__ TH, int th_stride, const double2 *__restrict__ S4,
			const double2 *__restrict__ Y,
			double2 *__restrict__ X, double2 *__restrict__ ACC, double2 *__restrict__ YN,
			const int *__restrict__ mof, int L, int nlm, int nmem, double visc, int stage, double bs,
			double anext, double dt)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= (long)nlm * nmem) return;
	int mem = (int)(i / nlm);
	int idx = (int)(i - (long)mem * nlm);
	int m = mof[idx];
	int l = idx - lm_index(L, 0, m);
	const double lap = -(double)l * ((double)l + 1.0);
	const double2 *y = Y + (long)mem * 3 * nlm;
	double2 *x = X + (long)mem * 3 * nlm;
	double2 *acc = ACC + (long)mem * 3 * nlm;
	double2 *yn = YN + (long)mem * 3 * nlm;
	const double2 *s4 = S4 + (long)mem * 4 * nlm;
	double2 t[3];
	t[0] = TH[(long)mem * th_stride * nlm + idx];
	t[1] = s4[2 * nlm + idx];
	t[2] = s4[3 * nlm + idx];
#pragma unroll
	for (int f = 0; f < 3; f++) {
		double2 k = t[f];
		if (visc != 0.0) {
			double2 yv = y[(long)f * nlm + idx];
			k.x += (yv.x * lap) * visc;
			k.y += (yv.y * lap) * visc;
		}
		long o = (long)f * nlm + idx;
		double2 bk = make_double2(bs * k.x, bs * k.y);
		if (stage == 0) {
			acc[o] = bk;
		} else if (stage < 3) {
			acc[o] = cadd(acc[o], bk);
		}
		if (stage < 3) {
			double2 xv = x[o];
			double2 ak = make_double2(anext * k.x, anext * k.y);
			yn[o] = make_double2(xv.x + dt * ak.x, xv.y + dt * ak.y);
		} else {
			double2 s = cadd(acc[o], bk);
			double2 xv = x[o];
			x[o] = make_double2(xv.x + dt * s.x, xv.y + dt * s.y);
		}
	}
}

inline int g1(long n) { return (int)((n + 255) / 256); }


// device buffers of one RK4 step over nmem members
struct SweBufs {
	double2 *Ya, *ACC, *Yb, *S9, *S4, *S2, *TH, *F;
	double *P9, *P4, *P2, *P1;
	// the mixed form: pre-split synthesis and analysis operands, FP32 Fourier
	// rows and FP32 ring-major physical fields
	void *SB, *G;
	float2 *F32;
	float *Q9, *Q4;
};

// one captured step per context, rebuilt when its arguments change
struct SweGraph {
	hipGraphExec_t exec = nullptr;
	sphswe_swe_params prm;
	double *state = nullptr;
	int nmem = 0;
	unsigned long gen = 0;   // ctx->buf_gen at capture
};
std::unordered_map<const sphswe_ctx *, SweGraph> g_graph;

int swe_buffers(sphswe_ctx *ctx, int nmem, int prec, SweBufs &b)
{
	const long nlm = ctx->nlm, nspat = ctx->nspat;
	const long nfour = (long)(ctx->L + 1) * ctx->nlat;
	const bool mixed = prec == SPHSWE_PREC_MIXED;
	b = SweBufs{};
	size_t bytes;
	size_t sb_b = 0, g_b = 0;
	if (mixed) {
		if (!sphswe_fft32_available(ctx))
			return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "swe_rk4: the mixed form has no FP32 FFT for %d longitudes",
					      ctx->nphi);
		if (int rc = sphswe_leg_bt_tables(ctx)) return rc;
		if (ctx->nlat % 2 != 0 || 2 * ctx->L + 1 > ctx->nphi)
			return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "swe_rk4: the mixed form needs an even ring count and nphi > 2 L");
		// Y, ACC, YN (3 nlm each), S4 (4), TH (1) spectral in FP64; operands of
		// 9 and 4 fields; FP32 Fourier rows of 9 + 4 fields; FP32 fields 9 + 4
		// (only for grids without a fused physical stage)
		sb_b = (sphswe_bt_sb_bytes(ctx, 9 * nmem) + 255) & ~(size_t)255;
		g_b = (sphswe_bt_g_bytes(ctx, 4 * nmem) + 255) & ~(size_t)255;
		bytes = (size_t)nlm * nmem * (3 + 3 + 3 + 4 + 1) * 16 + sb_b + g_b + (size_t)nfour * nmem * 13 * 8 +
			(size_t)nspat * nmem * (9 + 4) * 4 + 1024;
	} else {
		// workspace: Y (3 nlm), ACC (3), YN (3), S9 (9), S4 (4), S2 (2), TH (1) spectral;
		// P9 (9), P4 (4), P2 (2), P1 (1) physical; Fourier for 9 fields
		const size_t spec_e = (size_t)nlm * nmem * (3 + 3 + 3 + 9 + 4 + 2 + 1);
		const size_t spat_e = (size_t)nspat * nmem * (9 + 4 + 2 + 1);
		const size_t four_e = (size_t)nfour * nmem * 9;
		bytes = spec_e * 16 + spat_e * 8 + four_e * 16 + 1024;
	}
	SweWs &ws = g_ws[ctx];
	if (ws.bytes < bytes) {
		if (ws.base) {
			SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
			// a captured step points into the old workspace: drop it
			auto gi = g_graph.find(ctx);
			if (gi != g_graph.end() && gi->second.exec) {
				SPHSWE_HIP(ctx, hipGraphExecDestroy(gi->second.exec));
				gi->second.exec = nullptr;
			}
			SPHSWE_HIP(ctx, hipFree(ws.base));
			ws.base = nullptr;
			ws.bytes = 0;
		}
		if (hipMalloc(&ws.base, bytes) != hipSuccess)
			return sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "swe workspace of %zu bytes", bytes);
		ws.bytes = bytes;
	}
	if (mixed) {
		char *c = (char *)ws.base;
		auto take = [&](size_t n) {
			char *r = c;
			c += (n + 255) & ~(size_t)255;
			return (void *)r;
		};
		b.Ya = (double2 *)take((size_t)3 * nlm * nmem * 16);
		b.ACC = (double2 *)take((size_t)3 * nlm * nmem * 16);
		b.Yb = (double2 *)take((size_t)3 * nlm * nmem * 16);
		b.S4 = (double2 *)take((size_t)4 * nlm * nmem * 16);
		b.TH = (double2 *)take((size_t)nlm * nmem * 16);
		b.SB = take(sb_b);
		b.G = take(g_b);
		b.F32 = (float2 *)take((size_t)13 * nfour * nmem * 8);   // 9 synthesised + 4 analysed fields
		b.Q9 = (float *)take((size_t)9 * nspat * nmem * 4);
		b.Q4 = (float *)take((size_t)4 * nspat * nmem * 4);
		return SPHSWE_OK;
	}
	double2 *sp = (double2 *)ws.base;
	b.Ya = sp; sp += 3 * nlm * nmem;
	b.ACC = sp; sp += 3 * nlm * nmem;
	b.Yb = sp; sp += 3 * nlm * nmem;
	b.S9 = sp; sp += 9 * nlm * nmem;
	b.S4 = sp; sp += 4 * nlm * nmem;
	b.S2 = sp; sp += 2 * nlm * nmem;
	b.TH = sp; sp += 1 * nlm * nmem;
	b.F = sp; sp += 9 * nfour * nmem;
	double *dp = (double *)sp;
	b.P9 = dp; dp += 9 * nspat * nmem;
	b.P4 = dp; dp += 4 * nspat * nmem;
	b.P2 = dp; dp += 2 * nspat * nmem;
	b.P1 = dp;
	return SPHSWE_OK;
}

// one classic RK4 step of nmem members, X updated in place; launches only
// (no allocation, no synchronisation), so that it can be captured
int swe_step(sphswe_ctx *ctx, const sphswe_swe_params *p, double2 *X, int nm, const SweBufs &b)
{
	const long nlm = ctx->nlm, nspat = ctx->nspat;
	const double a = p->earth_radius, g = p->gravitation;
	const double ga = g / a, ia = 1.0 / a, two_omega = 2.0 * p->coriolis_omega;
	const double visc = p->viscosity / (a * a);
	const double dt = p->dt;
	const double bcoef[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
	const double anext[3] = {0.5, 0.5, 1.0};
	const int prec = p->precision;   // of the Legendre contractions; everything else is FP64
	// 3: the linear equations with Robert functions (the reference reads its flag
	// in the linear branch only)
	const int mode = (p->nonlinear == 0 && p->robert_functions) ? 3 : p->nonlinear;
	const int L = ctx->L, nlat = ctx->nlat;
	hipStream_t st = ctx->stream;
	int rc;
	// the physical buffers keep the reference's layout, whose wide-tile FFTs
	// run at about twice the rate of a ring-major two-buffer kernel
	auto c2r = [&](const double2 *f, double *ph, int n) { return sphswe_fft_c2r(ctx, f, ph, n); };
	auto r2c = [&](const double *ph, double2 *f, int n) { return sphswe_fft_r2c(ctx, ph, f, n); };
	const double2 *Y = X;
	double2 *YN = b.Ya;
	for (int s = 0; s < 4; s++) {
		k_prep9<<<g1(nlm * nm), 256, 0, st>>>(Y, b.S9, ctx->d_mofidx, L, (int)nlm, nm);
		if ((rc = sphswe_leg_synth(ctx, b.S9, b.F, 9 * nm, prec))) return rc;
		if ((rc = c2r(b.F, b.P9, 9 * nm))) return rc;
		k_phys1<<<g1(nspat * nm), 256, 0, st>>>(b.P9, b.P4, ctx->d_ct, ctx->d_st, ctx->d_sq, nlat, nspat, nm, ga,
							 two_omega, ia, mode, mode == 3 ? p->h0 / a : p->h0);
		if ((rc = r2c(b.P4, b.F, 4 * nm))) return rc;
		if ((rc = sphswe_leg_anal(ctx, b.F, b.S4, 4 * nm, prec))) return rc;
		// the h tendency: field 0 of each member's four (Robert functions: complete
		// after level 1), else levels 2-3
		const double2 *TH = b.S4;
		int th_stride = 4;
		if (mode != 3) {
			k_prep2<<<g1(nlm * nm), 256, 0, st>>>(b.S4, b.S2, ctx->d_mofidx, L, (int)nlm, nm);
			if ((rc = sphswe_leg_synth(ctx, b.S2, b.F, 2 * nm, prec))) return rc;
			if ((rc = c2r(b.F, b.P2, 2 * nm))) return rc;
			k_phys2<<<g1(nspat * nm), 256, 0, st>>>(b.P2, b.P1, ctx->d_st, ctx->d_sq, nlat, nspat, nm,
								 p->nonlinear ? ia : p->h0 / a);
			if ((rc = r2c(b.P1, b.F, nm))) return rc;
			if ((rc = sphswe_leg_anal(ctx, b.F, b.TH, nm, prec))) return rc;
			TH = b.TH;
			th_stride = 1;
		}
		k_stage<<<g1(nlm * nm), 256, 0, st>>>(TH, th_stride, b.S4, Y, X, b.ACC, YN, ctx->d_mofidx, L, (int)nlm, nm,
						       visc, s, bcoef[s], s < 3 ? anext[s] : 0.0, dt);
		SPHSWE_LAUNCH_CHECK(ctx);
		if (s < 3) {
			Y = YN;
			YN = (YN == b.Ya) ? b.Yb : b.Ya;
		}
	}
	return SPHSWE_OK;
}

// whether the context's grid has a fused physical stage
bool phys_fft_fused(const sphswe_ctx *ctx)
{
	return ctx->nlat % 2 == 0 && 2 * ctx->L + 1 <= ctx->plan.n && (ctx->plan.n == 384 || ctx->plan.n == 1536);
}

// the fused physical stage for the context's longitude count (the input holds
// h, u, v, D'h, D'u, D'v per member), or 1 when there is no fused kernel for it
int phys_fft_mx(sphswe_ctx *ctx, const float2 *Fin, float2 *Fout, int nm, float ga, float two_omega, float ia,
		int mode, float h0)
{
	const int nblk = nm * (ctx->nlat / 2);
	const int nblk_pad = (nblk + 7) / 8 * 8;
	auto by_n = [&](auto md) -> int {
		constexpr int M = decltype(md)::value;
		auto go = [&](auto kern, int nt) -> int {
			kern<<<nblk_pad, nt, 0, ctx->stream>>>(Fin, Fout, ctx->d_tw32, ctx->d_ct, ctx->d_st, ctx->d_sq,
							       ctx->nlat, ctx->L, nblk, nblk_pad, ga, two_omega, ia, h0);
			SPHSWE_LAUNCH_CHECK(ctx);
			return SPHSWE_OK;
		};
		switch (ctx->plan.n) {
		case 384: return go(k_phys_fft_mx<M, 384>, FusedPlan<384>::NT);
		case 1536: return go(k_phys_fft_mx<M, 1536>, FusedPlan<1536>::NT);
		}
		return 1;
	};
	if (!phys_fft_fused(ctx)) return 1;
	switch (mode) {
	case 0: return by_n(std::integral_constant<int, 0>{});
	case 1: return by_n(std::integral_constant<int, 1>{});
	case 2: return by_n(std::integral_constant<int, 2>{});
	default: return by_n(std::integral_constant<int, 3>{});
	}
}
static_assert(FusedPlan<384>::ok && FusedPlan<1536>::ok, "fused plans");

// one RK4 step in the mixed form: the Legendre contractions on the bf16 matrix
// cores from the stored tables (operands split into bf16 hi + lo), Fourier rows,
// longitude FFTs and pointwise terms in FP32, the state and the stage
// combinations in FP64.  The projection levels are those of swe_step; the
// divergence level's pointwise step, a combination with coefficients constant
// along each ring, goes straight from the synthesis sums to the analysis
// operand (equal to the transform round trip up to rounding, since nphi > 2 L).
int swe_step_mixed(sphswe_ctx *ctx, const sphswe_swe_params *p, double2 *X, int nm, const SweBufs &b)
{
	const long nlm = ctx->nlm, nspat = ctx->nspat;
	const double a = p->earth_radius, g = p->gravitation;
	const double ga = g / a, ia = 1.0 / a, two_omega = 2.0 * p->coriolis_omega;
	const double visc = p->viscosity / (a * a);
	const double dt = p->dt;
	const double bcoef[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
	const double anext[3] = {0.5, 0.5, 1.0};
	const int mode = (p->nonlinear == 0 && p->robert_functions) ? 3 : p->nonlinear;
	const int L = ctx->L, nphi = ctx->nphi;
	const long nfour = (long)(L + 1) * ctx->nlat;
	hipStream_t st = ctx->stream;
	int rc;
	const double2 *Y = X;
	double2 *YN = b.Ya;
	// with the fused physical stage only h, u, v and their D' are synthesised:
	// that kernel forms the rows of i m x from those of x
	const bool six = phys_fft_fused(ctx);
	for (int s = 0; s < 4; s++) {
		if ((rc = six ? sphswe_bt_sprep_swe6(ctx, Y, b.SB, nm) : sphswe_bt_sprep_swe9(ctx, Y, b.SB, nm))) return rc;
		if ((rc = sphswe_bt_synth32(ctx, b.SB, b.F32, (six ? 6 : 9) * nm))) return rc;
		const float h0f = (float)(mode == 3 ? p->h0 / a : p->h0);
		float2 *F4 = b.F32 + (long)9 * nfour * nm;   // output rows after the (at most 9) synthesised fields
		rc = six ? phys_fft_mx(ctx, b.F32, F4, nm, (float)ga, (float)two_omega, (float)ia, mode, h0f) : 1;
		if (rc == 1) {
			if ((rc = sphswe_fft_c2r32(ctx, b.F32, b.Q9, 9 * nm))) return rc;
			k_phys1_mx<<<g1(nspat * nm), 256, 0, st>>>(b.Q9, b.Q4, ctx->d_ct, ctx->d_st, ctx->d_sq, nphi, nspat,
								   nm, (float)ga, (float)two_omega, (float)ia, mode, h0f);
			if ((rc = sphswe_fft_r2c32(ctx, b.Q4, F4, 4 * nm))) return rc;
		} else if (rc) {
			return rc;
		}
		if ((rc = sphswe_bt_gprep32(ctx, F4, b.G, 4 * nm))) return rc;
		if ((rc = sphswe_bt_anal(ctx, b.G, b.S4, 4 * nm))) return rc;
		const double2 *TH = b.S4;
		int th_stride = 4;
		if (mode != 3) {
			// the divergence level: its pointwise step is a per-ring linear
			// combination, applied to the synthesis sums in place of a
			// transform round trip
			if ((rc = sphswe_bt_sprep_swe2(ctx, b.S4, b.SB, nm))) return rc;
			const float sc = (float)(p->nonlinear ? ia : p->h0 / a);
			if ((rc = sphswe_bt_synth_div(ctx, b.SB, b.G, nm, sc))) return rc;
			if ((rc = sphswe_bt_anal(ctx, b.G, b.TH, nm))) return rc;
			TH = b.TH;
			th_stride = 1;
		}
		k_stage<<<g1(nlm * nm), 256, 0, st>>>(TH, th_stride, b.S4, Y, X, b.ACC, YN, ctx->d_mofidx, L, (int)nlm, nm,
						       visc, s, bcoef[s], s < 3 ? anext[s] : 0.0, dt);
		SPHSWE_LAUNCH_CHECK(ctx);
		if (s < 3) {
			Y = YN;
			YN = (YN == b.Ya) ? b.Yb : b.Ya;
		}
	}
	return SPHSWE_OK;
}

int check_params(sphswe_ctx *ctx, const sphswe_swe_params *p)
{
	if (p->precision < SPHSWE_PREC_FP64 || p->precision > SPHSWE_PREC_MIXED)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "swe_rk4: precision %d not available", p->precision);
	if (p->nonlinear < 0 || p->nonlinear > 2)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "swe_rk4: nonlinear must be 0, 1 or 2");
	if (p->robert_functions != 0 && p->robert_functions != 1)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "swe_rk4: robert_functions must be 0 or 1");
	return SPHSWE_OK;
}

}  // namespace

// frees the workspace and the captured step of a context (called by
// sphswe_destroy after the stream has drained)
void sphswe_swe_release(sphswe_ctx *ctx)
{
	auto g = g_graph.find(ctx);
	if (g != g_graph.end()) {
		if (g->second.exec) (void)hipGraphExecDestroy(g->second.exec);
		g_graph.erase(g);
	}
	auto w = g_ws.find(ctx);
	if (w != g_ws.end()) {
		if (w->second.base) (void)hipFree(w->second.base);
		g_ws.erase(w);
	}
}

extern "C" int sphswe_swe_rk4(sphswe_ctx *ctx, const sphswe_swe_params *p, double *d_state, int nmem, int nsteps)
{
	if (!ctx || !p || !d_state || nmem < 1 || nsteps < 0) return SPHSWE_ERR_ARG;
	if (int rc = check_params(ctx, p)) return rc;
	SweBufs b;
	if (int rc = swe_buffers(ctx, nmem, p->precision, b)) return rc;
	for (int n = 0; n < nsteps; n++)
		if (int rc = p->precision == SPHSWE_PREC_MIXED ? swe_step_mixed(ctx, p, (double2 *)d_state, nmem, b)
							       : swe_step(ctx, p, (double2 *)d_state, nmem, b))
			return rc;
	return SPHSWE_OK;
}

extern "C" int sphswe_swe_rk4_graph(sphswe_ctx *ctx, const sphswe_swe_params *p, double *d_state, int nmem, int nsteps)
{
	if (!ctx || !p || !d_state || nmem < 1 || nsteps < 0) return SPHSWE_ERR_ARG;
	if (int rc = check_params(ctx, p)) return rc;
	SweGraph &G = g_graph[ctx];
	// every buffer the step uses exists before the check below, so that a
	// regrowth it causes is seen by the generation test
	SweBufs b;
	if (int rc = swe_buffers(ctx, nmem, p->precision, b)) return rc;
	if (int rc = sphswe_leg_anal_reserve(ctx, 4 * nmem, p->precision)) return rc;
	if (p->precision == SPHSWE_PREC_FP32 && (ctx->tab_min_synth > 0 || ctx->tab_min_anal > 0))
		if (int rc = sphswe_leg_tables(ctx)) return rc;
	if (!G.exec || G.state != d_state || G.nmem != nmem || G.gen != ctx->buf_gen ||
	    memcmp(&G.prm, p, sizeof *p) != 0) {
		if (G.exec) {
			SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
			SPHSWE_HIP(ctx, hipGraphExecDestroy(G.exec));
			G.exec = nullptr;
		}
		SPHSWE_HIP(ctx, hipStreamSynchronize(ctx->stream));
		hipGraph_t graph;
		SPHSWE_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
		const int rc = p->precision == SPHSWE_PREC_MIXED ? swe_step_mixed(ctx, p, (double2 *)d_state, nmem, b)
								   : swe_step(ctx, p, (double2 *)d_state, nmem, b);
		hipGraph_t captured = nullptr;
		const hipError_t ec = hipStreamEndCapture(ctx->stream, &captured);
		if (rc) {
			if (captured) hipGraphDestroy(captured);
			return rc;
		}
		SPHSWE_HIP(ctx, ec);
		graph = captured;
		const hipError_t ei = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
		hipGraphDestroy(graph);
		SPHSWE_HIP(ctx, ei);
		G.prm = *p;
		G.state = d_state;
		G.nmem = nmem;
		G.gen = ctx->buf_gen;
	}
	for (int n = 0; n < nsteps; n++) SPHSWE_HIP(ctx, hipGraphLaunch(G.exec, ctx->stream));
	return SPHSWE_OK;
}
