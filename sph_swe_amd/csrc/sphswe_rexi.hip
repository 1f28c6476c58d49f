// Linear SWE time step by REXI (AppSWE.hpp:336-565 with SWERexiSPH::solve,
// src/SWERexiSPH.hpp:137-244), batched over all poles.
//
// Per time step:
//  1. pole-independent fields, following the reference's projection graph with
//     the GPU complex SHT: div0, eta0 (div/vort of u0, v0 with their metric
//     divisions), phi0 in complex form, the Coriolis forcing pieces
//     W1..W3 = A(g * S(u0 | mu2 u0 | mu v0)) with g = S((2 Omega / a) A(grad_lat mu)),
//     and their mu / mu2 images;
//  2. one kernel over (m, pole): assemble the phi right-hand side
//        rhs = a^2 foo + (2W)^2 mu2(foo) + (Phi/a) Fc     (SWERexiSPH.hpp:176-195)
//     from those fields and the pole scalars, solve the 9-diagonal system with
//     the cached pivoted LU factors, emit i m phi and D' phi for the gradients,
//     and reduce beta*phi over the poles of the wavefront;
//  3. grad_lon / grad_lat of phi = A(S(.) / sqrt(1-mu^2)) per m, done without
//     FFT: Legendre synthesis to (m, ring) space, scaling by 2 pi / sqrt(1-mu^2),
//     Legendre analysis (the FFT pair would be an identity on one m up to the
//     2 pi of its normalisation);
//  4. one kernel over (m, pole): a = u0 + G_lon/r, b = v0 + G_lat/r,
//     rhs_a = alpha a - 2W mu(b), rhs_b = 2W mu(a) + alpha b, the two
//     5-diagonal solves, reduction of beta*u, beta*v;
//  5. sum of the pole partials (FP64, fixed order), real part through the
//     complex/real SHT pair, h = phi/g, implicit viscosity.
//
// Layout of per-pole data ("pole layout"): row = moff[m+L] + (n - |m|) in the
// reference's NCompact order (SPHConfig.hpp:136-156), pole fastest, so the
// threads of a wavefront (64 poles of one m) read consecutive addresses.
#include "sphswe_band.h"
#include "sphswe_common.h"

#include <algorithm>
#include <cstring>
#include <thread>
#include <type_traits>

using sphswe_host::cplx;

struct sphswe_rexi {
	sphswe_ctx *ctx = nullptr;
	sphswe_rexi_params prm;
	int P = 0;
	int coriolis = 0;
	std::vector<cplx> alpha0, beta0;   // as returned by REXI::setup
	double2 *d_alpha = nullptr, *d_beta = nullptr;   // divided by dt (SWERexiSPH.hpp:82-83)
	int *d_moff = nullptr;
	void *d_fphi = nullptr, *d_fvel = nullptr;   // parity-compact LU factors, double2 or float2
	int fac32 = 0;                     // factors stored in FP32
	int robert = 0;                    // SWERexiSPHRobert: Robert-function velocities
	int refine = 0;                    // one sweep of FP64 iterative refinement per solve
	double2 *d_kphi = nullptr;         // refinement: pole-independent pieces of the phi bands [rows][4][5]
	double2 *d_kvel = nullptr;         // refinement: W^2 mu^2 band of the velocity systems [rows][3]
	unsigned char *d_pphi = nullptr, *d_pvel = nullptr;
	double *d_oscale = nullptr;        // 2 pi / sqrt(1 - mu^2)
	// work
	double2 *d_X = nullptr, *d_G = nullptr, *d_Y = nullptr;
	// the gradient operator per |m| and parity of n - |m| (pole independent):
	// symmetric n_p x n_p, FP64 and rounded to FP32
	double *d_gop = nullptr;
	float *d_gop32 = nullptr;
	long *d_goff = nullptr;            // offset of the (|m|, parity) matrix, index 2|m| + parity
	int4 *d_gtiles = nullptr;          // (m, parity, first row, -) of each output block of the apply kernel
	int ngtiles = 0;
	double2 *d_part = nullptr, *d_basis = nullptr;
	double2 *d_cphys = nullptr, *d_cspec = nullptr;
	double2 *d_in = nullptr;           // phi = g h, u, v of a step (3 real spectra)
	int ngroups = 0;
	size_t rows = 0;
};

namespace {

constexpr int NBASIS = 12;   // B1..B10, u0c, v0c
enum { B_DIV = 0, B_MUETA, B_PHI, B_MU2PHI, B_MU2DIV, B_MU2MUETA, B_MU2MU2PHI, B_W1, B_W2, B_W3, B_U0, B_V0 };

__device__ inline double2 cm(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cs(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ inline double2 ca(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 cb(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cinv(double2 a)
{
	const double d = a.x * a.x + a.y * a.y;
	return make_double2(a.x / d, -a.y / d);
}

__device__ inline double2 wave_sum(double2 v)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1) {
		v.x += __shfl_xor(v.x, o);
		v.y += __shfl_xor(v.y, o);
	}
	return v;
}

// R independent wavefront sums, interleaved so that their cross-lane
// exchanges wait together
template <int R>
__device__ inline void wave_sum_n(double2 (&v)[R])
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
#pragma unroll
		for (int r = 0; r < R; r++) {
			v[r].x += __shfl_xor(v[r].x, o);
			v[r].y += __shfl_xor(v[r].y, o);
		}
}

// RD: rows of a band sweep whose loads are issued together. The sweeps are
// sequential in the row, so without this every row waits a full memory
// latency. 8 rows: the phi sweeps at T256 / 200 poles 570 -> 502 us against 4,
// more rows add nothing. The refining forms hold 256 VGPRs with 8 rows (one
// wave per SIMD, two with 4): they take 8 rows up to T319 and 4 above, where
// the occupancy is worth more (refined step at T383 150 -> 154 steps/s, at
// T256 342 against 333 with 8 rows; T1023: 36 -> 31 ms per phi + vel pair)
constexpr int sweep_rows(bool refine, int L) { return refine && L >= 320 ? 4 : 8; }

// m of workgroup row y: 0, -1, 1, -2, 2, ... (longest systems first)
__device__ inline int m_of(int y) { return (y & 1) ? -((y + 1) >> 1) : (y >> 1); }

// swap window[0] <-> window[k] for a runtime even k in [0, K], as pairwise
// conditional swaps on registers (a select chain reading w[k] is turned into an
// indexed access by the compiler, which moves the window to scratch memory).
// The REXI bands couple degrees of equal parity only, so pivots are even.
template <int K>
__device__ inline void pivot_swap(double2 (&w)[K + 1], int k)
{
#pragma unroll
	for (int i = 2; i <= K; i += 2) {
		const bool s = (k == i);
		const double2 a = w[0], b = w[i];
		w[0] = make_double2(s ? b.x : a.x, s ? b.y : a.y);
		w[i] = make_double2(s ? a.x : b.x, s ? a.y : b.y);
	}
}

// LU factors of the REXI bands, parity-compact (only the even offsets are
// nonzero): per row j of an m block, for halo h, entry 0 = 1 / U(j,j), entries
// 1..h = U(j-2e, j), entries h+1..h+h/2 = L(j+2e, j), e = 1, 2, ...; stored in
// FP64 (double2) or rounded to FP32 (float2), the arithmetic is FP64
constexpr int rexi_nec(int halo) { return 1 + halo + halo / 2; }
__device__ inline double2 ldf(const double2 *p) { return *p; }
__device__ inline double2 ldf(const float2 *p)
{
	const float2 v = *p;
	return make_double2(v.x, v.y);
}
template <typename FT>
__device__ inline FT stf(double2 v);
template <>
__device__ inline double2 stf<double2>(double2 v) { return v; }
template <>
__device__ inline float2 stf<float2>(double2 v) { return make_float2((float)v.x, (float)v.y); }

// LU factorisation with partial pivoting of the REXI bands on the device, one
// thread per (pole, m, parity): the zgbtf2 steps that zgbsv_ runs on the band
// (BandedMatrixSolver.hpp:416-427; pivot = the largest |re| + |im|, interchanges
// applied to the current and later columns only) on the compact index of one
// parity class, rows j = 2c + par of the m block, which couple only among
// themselves.  The band is A(alpha) = alpha^D I + sum_k s_k(alpha) K_k with the
// pole-independent pieces K_k at the even offsets of every row:
//   phi (KC = 2, D = 4): K0, K2, K1, K-1 with s = 1, alpha^2, alpha, 1 / alpha;
//   winds (KC = 1, D = 2): K0 with s = 1.
// The window holds KC + 1 rows from the pivot column over 2 KC + 1 columns.
// Writes the parity-compact factors (rexi_nec(2 KC) entries per row: 1 / U(j,j),
// U(j-2e, j), L(j+2e, j)) and the even pivot offsets; *info = the smallest
// 1-based row (within its block) whose pivot is zero, as zgbtf2's INFO.
template <int KC, typename FT>
__global__ __launch_bounds__(64) void k_rexi_factor(const double2 *__restrict__ alpha, const int *__restrict__ moff,
						    const double2 *__restrict__ K, FT *__restrict__ fac,
						    unsigned char *__restrict__ piv, int *__restrict__ info, int L, int P)
{
	constexpr int NK = KC == 2 ? 4 : 1;   // pieces
	constexpr int WD = 2 * KC + 1;        // entries per row of a piece (offsets -2KC .. 2KC, even)
	constexpr int NC = 1 + 2 * KC + KC;   // = rexi_nec(2 KC)
	const int pole = blockIdx.x * 64 + threadIdx.x;
	if (pole >= P) return;
	const int par = blockIdx.y & 1;
	const int m = m_of(blockIdx.y >> 1);
	const int N = L + 1 - abs(m);
	const int Np = (N - par + 1) >> 1;
	if (Np <= 0) return;
	const size_t r0 = moff[m + L];
	const double2 al = alpha[pole];
	double2 sc[NK];
	double2 diag;
	if (KC == 2) {
		const double2 a2 = cm(al, al);
		sc[0] = make_double2(1, 0);
		sc[NK > 1 ? 1 : 0] = a2;
		sc[NK > 2 ? 2 : 0] = al;
		sc[NK > 3 ? 3 : 0] = cinv(al);
		diag = cm(a2, a2);
	} else {
		sc[0] = make_double2(1, 0);
		diag = cm(al, al);
	}
	// entry (compact row cr, compact offset e) of A(alpha), zero outside the block
	auto entry = [&](int cr, int e) __attribute__((always_inline)) -> double2 {
		const int cc = cr + e;
		if (cr >= Np || cc < 0 || cc >= Np) return make_double2(0, 0);
		const size_t q = r0 + 2 * cr + par;
		double2 v = e == 0 ? diag : make_double2(0, 0);
#pragma unroll
		for (int k = 0; k < NK; k++) v = ca(v, cm(sc[k], K[(q * NK + k) * WD + KC + e]));
		return v;
	};
	auto slot = [&](int cr, int e) __attribute__((always_inline)) -> FT * {
		return fac + ((r0 + 2 * cr + par) * NC + e) * (size_t)P + pole;
	};
	double2 w[KC + 1][2 * KC + 1];
#pragma unroll
	for (int i = 0; i <= KC; i++)
#pragma unroll
		for (int d = 0; d <= 2 * KC; d++) w[i][d] = (d - i >= -KC && d - i <= KC) ? entry(i, d - i) : make_double2(0, 0);
	for (int c = 0; c < Np; c++) {
		const int km = min(KC, Np - 1 - c);
		int ip = 0;
		double amax = -1.0;
#pragma unroll
		for (int i = 0; i <= KC; i++) {
			const double a = fabs(w[i][0].x) + fabs(w[i][0].y);
			if (i <= km && a > amax) {
				amax = a;
				ip = i;
			}
		}
		// interchange rows 0 and ip of the window (selects on registers)
#pragma unroll
		for (int i = 1; i <= KC; i++) {
			const bool sw = ip == i;
#pragma unroll
			for (int d = 0; d <= 2 * KC; d++) {
				const double2 a = w[0][d], b = w[i][d];
				w[0][d] = make_double2(sw ? b.x : a.x, sw ? b.y : a.y);
				w[i][d] = make_double2(sw ? a.x : b.x, sw ? a.y : b.y);
			}
		}
		const bool nz = w[0][0].x != 0.0 || w[0][0].y != 0.0;
		if (!nz) atomicMin(info, 2 * c + par + 1);
		const double2 rp = nz ? cinv(w[0][0]) : make_double2(0, 0);
		*slot(c, 0) = stf<FT>(rp);
		// U(c, c + e) belongs to column c + e: entry e of that row
#pragma unroll
		for (int e = 1; e <= 2 * KC; e++)
			if (c + e < Np) *slot(c + e, e) = stf<FT>(w[0][e]);
#pragma unroll
		for (int i = 1; i <= KC; i++) {
			const double2 l = (i <= km && nz) ? cm(w[i][0], rp) : make_double2(0, 0);
			*slot(c, 2 * KC + i) = stf<FT>(l);
			if (i <= km && nz)
#pragma unroll
				for (int d = 1; d <= 2 * KC; d++) w[i][d] = cb(w[i][d], cm(l, w[0][d]));
		}
		piv[(r0 + 2 * c + par) * (size_t)P + pole] = (unsigned char)(2 * ip);
		// next pivot column: shift the window, the row c + 1 + KC enters
#pragma unroll
		for (int i = 0; i < KC; i++) {
#pragma unroll
			for (int d = 0; d < 2 * KC; d++) w[i][d] = w[i + 1][d + 1];
			w[i][2 * KC] = make_double2(0, 0);
		}
#pragma unroll
		for (int d = 0; d <= 2 * KC; d++) w[KC][d] = entry(c + 1 + KC, d - KC);
	}
}

// without rotation the phi system is diagonal: 1 / (alpha^2 + (Phi / r^2) n (n + 1))
// (SWERexiSPH.hpp:230-231), entry 0 of every row
template <typename FT>
__global__ void k_rexi_diag(const double2 *__restrict__ alpha, const int *__restrict__ moff, FT *__restrict__ fac,
			    int L, int P, int NC, double g_over_r2)
{
	const int m = m_of(blockIdx.y);
	const int N = L + 1 - abs(m);
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (long)N * P) return;
	const int j = (int)(t / P), pole = (int)(t % P);
	const int n = abs(m) + j;
	const double2 al = alpha[pole];
	double2 d = cm(al, al);
	d.x += g_over_r2 * (double)n * ((double)n + 1.0);
	fac[((size_t)(moff[m + L] + j) * NC) * P + pole] = stf<FT>(cinv(d));
}

// right-hand side of the phi system of one pole at degree n, order m
// (SWERexiSPH.hpp:162-195, with the pole-independent fields in the basis B)
__device__ inline double2 phi_rhs(const double2 *__restrict__ B, long ncplx, int coriolis, double G, double W, int m, int n,
				  double2 al)
{
	const long k = (long)n * n + n + m;
	const double2 b1 = B[B_DIV * ncplx + k], b3 = B[B_PHI * ncplx + k];
	if (!coriolis) {
		// rhs = Phi div0 + alpha phi0 (SWERexiSPH.hpp:230)
		return ca(cs(b1, G), cm(al, b3));
	}
	const double W2 = W * W;
	const double2 ial = cinv(al), al2 = cm(al, al);
	const double2 b2 = B[B_MUETA * ncplx + k], b4 = B[B_MU2PHI * ncplx + k];
	const double2 b5 = B[B_MU2DIV * ncplx + k], b6 = B[B_MU2MUETA * ncplx + k];
	const double2 b7 = B[B_MU2MU2PHI * ncplx + k];
	const double2 w1 = B[B_W1 * ncplx + k], w2 = B[B_W2 * ncplx + k];
	const double2 w3 = B[B_W3 * ncplx + k];
	const double2 wa = cs(ial, W), w2a = cs(ial, W2);
	// foo = G (div0 - (W/alpha) mu(eta0)) + (alpha phi0 + (W^2/alpha) mu2(phi0))
	const double2 foo = ca(cs(cb(b1, cm(wa, b2)), G), ca(cm(al, b3), cm(w2a, b4)));
	const double2 mfoo = ca(cs(cb(b5, cm(wa, b6)), G), ca(cm(al, b4), cm(w2a, b7)));
	// Fc = -alpha^2 W1 + W^2 W2 + 2 alpha W W3 (the W's carry the factor W/a)
	const double2 fc = ca(cb(cs(w2, W2), cm(al2, w1)), cm(cs(al, 2.0 * W), w3));
	return ca(ca(cm(al2, foo), cs(mfoo, W2)), cm(cs(ial, G), fc));
}

// the phi right-hand sides of every (row, pole) into the pole layout (the
// forward sweep then streams them like the factors)
__global__ __launch_bounds__(256) void k_rexi_phi_rhs(const double2 *__restrict__ alpha, const int *__restrict__ moff,
						      const double2 *__restrict__ B, double2 *__restrict__ out, int L, int P,
						      long rs, long ncplx, double G, double W)
{
	const int m = (int)blockIdx.y - L, am = abs(m), N = L + 1 - am;
	const long t = blockIdx.x * 256L + threadIdx.x;
	if (t >= (long)N * P) return;
	const int j = (int)(t / P), p = (int)(t % P);
	out[((long)moff[m + L] + j) * rs + p] = phi_rhs(B, ncplx, 1, G, W, m, am + j, alpha[p]);
}

template <typename FT>
struct PhiArgs {
	const double2 *alpha, *beta;
	const int *moff;
	const FT *fac;
	const unsigned char *piv;
	const double2 *basis;     // NBASIS complex spectral fields, standard layout
	double2 *Y;               // forward-substitution scratch, pole layout
	double2 *X;               // [rows][2P]: i m phi | D' phi
	double2 *part;            // [group][3][rows]: pole partial sums
	int L, P, ncplx;
	double two_omega, avg_geo;
	int coriolis;
	size_t rows;
	// iterative refinement: right-hand sides, then residuals, in R (row stride
	// 2P, first half); the pole-independent pieces of A per row and even
	// offset -4..4: K[row][kind][5], kinds K0, K2 (times alpha^2), K1 (alpha),
	// Km1 (1/alpha); A = alpha^4 I + K0 + alpha^2 K2 + alpha K1 + Km1 / alpha
	double2 *R;
	const double2 *K;
};

// phi: assemble the right-hand side, solve, emit gradient inputs, reduce
// beta*phi.  REF: the solve with the stored (FP32) factors is refined once:
// x0 (kept in the first half of X), r = b - A x0 in FP64, the correction
// solve, x = x0 + dx.
template <int KL, typename FT, bool REF, int RD>
__global__ __launch_bounds__(64) void k_rexi_phi(PhiArgs<FT> a)
{
	constexpr int KV = 2 * KL, NE = rexi_nec(KL);
	const int m = m_of(blockIdx.y);
	const int am = abs(m);
	const int L = a.L, P = a.P;
	const int N = L + 1 - am;
	const int p = blockIdx.x * 64 + threadIdx.x;
	const bool pv = p < P;
	const int pc = pv ? p : P - 1;
	const size_t row0 = a.moff[m + L];
	const double2 al = a.alpha[pc], be = a.beta[pc];
	const double G = a.avg_geo, W = a.two_omega;
	const double2 *B = a.basis;

	auto rhs_at = [&](int j) __attribute__((always_inline)) -> double2 {
		return phi_rhs(B, a.ncplx, a.coriolis, G, W, m, am + j, al);
	};

	double2 phi[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};   // phi(j), phi(j+1), phi(j+2)

	if (!a.coriolis) {
		// diagonal Helmholtz: phi = rhs / (alpha^2 + (Phi / r^2) n (n+1)) is done by
		// the caller through factors holding 1 / diag; rows independent
		for (int j = 0; j < N; j++) {
			const double2 x = cm(rhs_at(j), ldf(a.fac + (row0 + j) * (size_t)NE * P + pc));
			if (pv) a.X[(row0 + j) * 2 * P + p] = make_double2(-m * x.y, m * x.x);
			a.Y[(row0 + j) * P + pc] = x;
			double2 bx = wave_sum(pv ? cm(be, x) : make_double2(0, 0));
			if (threadIdx.x == 0) a.part[((size_t)blockIdx.x * 3 + 0) * a.rows + row0 + j] = bx;
		}
		// D' phi from the stored phi
		for (int j = 0; j < N; j++) {
			const int n = am + j;
			const double dn = n, dm = m;
			const double c1 = (-dn + 1.0) * idR(dn - 1, dm), c2 = (dn + 2.0) * idS(dn + 1, dm);
			const double2 xm = (j > 0) ? a.Y[(row0 + j - 1) * P + pc] : make_double2(0, 0);
			const double2 xp = (j + 1 < N) ? a.Y[(row0 + j + 1) * P + pc] : make_double2(0, 0);
			if (pv) a.X[(row0 + j) * 2 * P + P + p] = ca(cs(xm, c1), cs(xp, c2));
		}
		return;
	}

	// forward substitution with row interchanges (L of the pivoted LU): the
	// right-hand sides of R (row stride rs) into Y.  R may be Y itself: row j
	// is stored after the rows ahead of it have been loaded.
	auto fwd = [&](const double2 *Rr, size_t rs) __attribute__((always_inline)) {
		double2 w[KL + 1];
#pragma unroll
		for (int i = 0; i <= KL; i++) w[i] = (i < N) ? Rr[(row0 + i) * rs + pc] : make_double2(0, 0);
		for (int jb = 0; jb < N; jb += RD) {
			double2 lf[RD][KL / 2], rn[RD];
			int kp[RD];
#pragma unroll
			for (int d = 0; d < RD; d++) {
				const int j = min(jb + d, N - 1);
				const FT *f = a.fac + (row0 + j) * (size_t)NE * P + pc;
#pragma unroll
				for (int e = 0; e < KL / 2; e++) lf[d][e] = ldf(f + (size_t)(KL + 1 + e) * P);
				kp[d] = a.piv[(row0 + j) * P + pc];
				const int jn = jb + d + KL + 1;
				rn[d] = (jn < N) ? Rr[(row0 + jn) * rs + pc] : make_double2(0, 0);
			}
#pragma unroll
			for (int d = 0; d < RD; d++) {
				const int j = jb + d;
				if (j < N) {
					pivot_swap<KL>(w, kp[d]);
#pragma unroll
					for (int e = 0; e < KL / 2; e++) w[2 * e + 2] = cb(w[2 * e + 2], cm(lf[d][e], w[0]));
					a.Y[(row0 + j) * P + pc] = w[0];
#pragma unroll
					for (int i = 0; i < KL; i++) w[i] = w[i + 1];
					w[KL] = rn[d];
				}
			}
		}
	};
	// backward substitution with U (KV super-diagonals) from Y; emit(j, x, d)
	// per row, j = N-1 .. 0, returns beta x for the pole sums when SUM;
	// pre(d, j) issues the emitter's loads for row j (j >= 0 or not) with the
	// batch's, d = the row's place in the batch
	auto bwd = [&](auto emit, auto sum_tag, auto pre) __attribute__((always_inline)) {
		constexpr bool SUM = decltype(sum_tag)::value;
		double2 v[KV + 1];
#pragma unroll
		for (int i = 0; i <= KV; i++) v[i] = (N - 1 - i >= 0) ? a.Y[(row0 + N - 1 - i) * P + pc] : make_double2(0, 0);
		for (int jt = N - 1; jt >= 0; jt -= RD) {
			double2 uf[RD][KL + 1], yn[RD], bx[RD];
#pragma unroll
			for (int d = 0; d < RD; d++) {
				const int j = max(jt - d, 0);
				const FT *f = a.fac + (row0 + j) * (size_t)NE * P + pc;
#pragma unroll
				for (int e = 0; e <= KL; e++) uf[d][e] = ldf(f + (size_t)e * P);
				const int jy = jt - d - KV - 1;
				yn[d] = (jy >= 0) ? a.Y[(row0 + jy) * P + pc] : make_double2(0, 0);
				pre(d, jt - d);
			}
#pragma unroll
			for (int d = 0; d < RD; d++) {
				const int j = jt - d;
				bx[d] = make_double2(0, 0);
				if (j >= 0) {
					const double2 x = cm(v[0], uf[d][0]);
#pragma unroll
					for (int e = 1; e <= KL; e++) v[2 * e] = cb(v[2 * e], cm(uf[d][e], x));
					bx[d] = emit(j, x, d);
#pragma unroll
					for (int i = 0; i < KV; i++) v[i] = v[i + 1];
					v[KV] = yn[d];
				}
			}
			if constexpr (SUM) {
				// pole sums of the batch's rows
				wave_sum_n<RD>(bx);
				if (threadIdx.x == 0)
#pragma unroll
					for (int d = 0; d < RD; d++)
						if (jt - d >= 0) a.part[((size_t)blockIdx.x * 3 + 0) * a.rows + row0 + jt - d] = bx[d];
			}
		}
	};
	// outputs of phi(j): i m phi at j, D' phi at j+1 (needs phi(j) and phi(j+2)); beta phi
	auto out = [&](int j, double2 x) __attribute__((always_inline)) -> double2 {
		if (pv) a.X[(row0 + j) * 2 * P + p] = make_double2(-m * x.y, m * x.x);
		phi[2] = phi[1];
		phi[1] = phi[0];
		phi[0] = x;
		if (j + 1 < N) {
			// row j+1: phi(j) = x, phi(j+2) = phi[2]
			const int n = am + j + 1;
			const double dn = n, dm = m;
			const double c1 = (-dn + 1.0) * idR(dn - 1, dm), c2 = (dn + 2.0) * idS(dn + 1, dm);
			if (pv) a.X[(row0 + j + 1) * 2 * P + P + p] = ca(cs(x, c1), cs(phi[2], c2));
		}
		return pv ? cm(be, x) : make_double2(0, 0);
	};
	auto nopre = [](int, int) {};
	if constexpr (!REF) {
		fwd(a.Y, (size_t)P);
		bwd([&](int j, double2 x, int) __attribute__((always_inline)) { return out(j, x); }, std::true_type{}, nopre);
	} else {
		const size_t rs = 2 * (size_t)P;
		const double2 al2 = cm(al, al), al4 = cm(al2, al2), ial = cinv(al);
		// r(row q) = b(q) - sum over d = -4, -2, .., 4 of A(q, q+d) x0(q+d), in place in R;
		// xw[k] = x0(q - KL + k)
		double2 xw[KV + 1];
#pragma unroll
		for (int k = 0; k <= KV; k++) xw[k] = make_double2(0, 0);
		// rq[d]: R of row q = j + KL for the row j at place d of a batch
		double2 rq[RD];
		auto resid = [&](int q, double2 r) __attribute__((always_inline)) {
			if (q < 0 || q >= N) return;
			const double2 *Kq = a.K + (row0 + q) * 20;
#pragma unroll
			for (int e = 0; e <= KL; e++) {
				const int c = q + 2 * e - KL;
				if (c < 0 || c >= N) continue;
				double2 A = ca(ca(Kq[e], cm(al2, Kq[5 + e])), ca(cm(al, Kq[10 + e]), cm(ial, Kq[15 + e])));
				if (2 * e == KL) A = ca(A, al4);
				r = cb(r, cm(A, xw[2 * e]));
			}
			a.R[(row0 + q) * rs + pc] = r;
		};
		fwd(a.R, rs);
		bwd([&](int j, double2 x, int d) __attribute__((always_inline)) -> double2 {
			if (pv) a.X[(row0 + j) * 2 * P + p] = x;
#pragma unroll
			for (int k = KV; k > 0; k--) xw[k] = xw[k - 1];
			xw[0] = x;
			resid(j + KL, rq[d]);
			return make_double2(0, 0);
		}, std::false_type{}, [&](int d, int j) __attribute__((always_inline)) {
			const int q = j + KL;
			rq[d] = (j >= 0 && q < N) ? a.R[(row0 + q) * rs + pc] : make_double2(0, 0);
		});
		for (int j = -1; j >= -KL; j--) {
#pragma unroll
			for (int k = KV; k > 0; k--) xw[k] = xw[k - 1];
			xw[0] = make_double2(0, 0);
			const int q = j + KL;
			resid(q, q < N ? a.R[(row0 + q) * rs + pc] : make_double2(0, 0));
		}
		fwd(a.R, rs);
		// x0 of the batch's rows
		double2 xq[RD];
		bwd([&](int j, double2 dx, int d) __attribute__((always_inline)) -> double2 {
			return out(j, ca(xq[d], dx));
		}, std::true_type{}, [&](int d, int j) __attribute__((always_inline)) {
			xq[d] = j >= 0 ? a.X[(row0 + j) * 2 * P + pc] : make_double2(0, 0);
		});
	}
	// row 0: phi(-1) = 0, phi(1) = phi[1]
	{
		const int n = am;
		const double dn = n, dm = m;
		const double c2 = (dn + 2.0) * idS(dn + 1, dm);
		if (pv) a.X[row0 * 2 * P + P + p] = cs(phi[1], c2);
	}
}

template <typename FT>
struct VelArgs {
	const double2 *alpha, *beta;
	const int *moff;
	const FT *fac;
	const unsigned char *piv;
	const double2 *basis;
	const double2 *G;         // [rows][2P]: G_lon | G_lat
	double2 *Y;               // scratch [2][rows][P]
	double2 *part;
	int L, P, ncplx;
	double two_omega, ir;
	int coriolis;
	size_t rows;
	// iterative refinement: right-hand sides, then residuals, of both systems
	// in R ([rows][2P]: a | b); x0 in X ([rows][2P], over G once the forward
	// sweep has read it); A = alpha^2 I + Kv, Kv[row][3] at offsets -2, 0, 2
	double2 *R, *X;
	const double2 *Kv;
};

template <int KL, typename FT, bool REF, int RD>
__global__ __launch_bounds__(64) void k_rexi_vel(VelArgs<FT> a)
{
	constexpr int KV = 2 * KL, NE = rexi_nec(KL);
	const int m = m_of(blockIdx.y);
	const int am = abs(m);
	const int L = a.L, P = a.P;
	const int N = L + 1 - am;
	const int p = blockIdx.x * 64 + threadIdx.x;
	const bool pv = p < P;
	const int pc = pv ? p : P - 1;
	const size_t row0 = a.moff[m + L];
	const double2 al = a.alpha[pc], be = a.beta[pc];
	const double W = a.two_omega;
	const double2 *B = a.basis;

	// a(n) = u0 + G_lon / r, b(n) = v0 + G_lat / r
	auto ab_at = [&](int j, double2 &av, double2 &bv) __attribute__((always_inline)) {
		if (j < 0 || j >= N) { av = bv = make_double2(0, 0); return; }
		const int n = am + j;
		const long k = (long)n * n + n + m;
		const double2 gl = a.G[(row0 + j) * 2 * P + pc], gt = a.G[(row0 + j) * 2 * P + P + pc];
		av = ca(B[B_U0 * (long)a.ncplx + k], cs(gl, a.ir));
		bv = ca(B[B_V0 * (long)a.ncplx + k], cs(gt, a.ir));
	};
	const double2 ial = cinv(al);
	if (!a.coriolis) {
		// u = (1/alpha) (u0 + ir grad_lon phi), v likewise (SWERexiSPH.hpp:232-233)
		for (int j = 0; j < N; j++) {
			double2 av, bv;
			ab_at(j, av, bv);
			const double2 u = cm(ial, av), v = cm(ial, bv);
			double2 su = wave_sum(pv ? cm(be, u) : make_double2(0, 0));
			double2 sv = wave_sum(pv ? cm(be, v) : make_double2(0, 0));
			if (threadIdx.x == 0) {
				a.part[((size_t)blockIdx.x * 3 + 1) * a.rows + row0 + j] = su;
				a.part[((size_t)blockIdx.x * 3 + 2) * a.rows + row0 + j] = sv;
			}
		}
		return;
	}
	// a, b at row j from G and basis values already loaded
	auto ab_pre = [&](int j, double2 gl, double2 gt, double2 u0, double2 v0, double2 &av, double2 &bv)
		__attribute__((always_inline)) {
		if (j < 0 || j >= N) { av = bv = make_double2(0, 0); return; }
		av = ca(u0, cs(gl, a.ir));
		bv = ca(v0, cs(gt, a.ir));
	};
	// rhs_a = alpha a - W mu(b), rhs_b = W mu(a) + alpha b, mu = R(n-1) x_{n-1} + S(n+1) x_{n+1}
	double2 aw[3], bw[3];   // (j-1, j, j+1)
	aw[0] = bw[0] = make_double2(0, 0);
	ab_at(0, aw[1], bw[1]);
	ab_at(1, aw[2], bw[2]);
	// uses aw/bw = values at (j-1, j, j+1), then shifts in (an, bn) of row j+2
	auto rhs_next = [&](int j, double2 &ra, double2 &rb, double2 an, double2 bn) __attribute__((always_inline)) {
		const int n = am + j;
		const double dn = n, dm = m;
		const double r = idR(dn - 1, dm), s = idS(dn + 1, dm);
		const double2 mub = ca(cs(bw[0], r), cs(bw[2], s));
		const double2 mua = ca(cs(aw[0], r), cs(aw[2], s));
		ra = cb(cm(al, aw[1]), cs(mub, W));
		rb = ca(cs(mua, W), cm(al, bw[1]));
		aw[0] = aw[1]; aw[1] = aw[2]; aw[2] = an;
		bw[0] = bw[1]; bw[1] = bw[2]; bw[2] = bn;
	};
	double2 wa[KL + 1], wb[KL + 1];
#pragma unroll
	for (int i = 0; i <= KL; i++) {
		if (i < N) {
			double2 an, bn;
			ab_at(i + 2, an, bn);
			rhs_next(i, wa[i], wb[i], an, bn);
		} else {
			wa[i] = wb[i] = make_double2(0, 0);
		}
	}
	// forward substitution; row j makes the right-hand side of row j + KL + 1,
	// which shifts in a, b of row j + KL + 3 (REF: the right-hand sides are
	// also stored in R)
	const size_t rs = 2 * (size_t)P;
	double2 *Ya = a.Y, *Yb = a.Y + a.rows * P;
	auto keep = [&](int q, double2 ra, double2 rb) __attribute__((always_inline)) {
		if constexpr (REF) {
			a.R[(row0 + q) * rs + pc] = ra;
			a.R[(row0 + q) * rs + P + pc] = rb;
		}
	};
#pragma unroll
	for (int i = 0; i <= KL; i++)
		if (i < N) keep(i, wa[i], wb[i]);
	for (int jb = 0; jb < N; jb += RD) {
		double2 lf[RD][KL / 2], gl[RD], gt[RD], u0[RD], v0[RD];
		int kp[RD];
#pragma unroll
		for (int d = 0; d < RD; d++) {
			const int j = min(jb + d, N - 1);
			const FT *f = a.fac + (row0 + j) * (size_t)NE * P + pc;
#pragma unroll
			for (int e = 0; e < KL / 2; e++) lf[d][e] = ldf(f + (size_t)(KL + 1 + e) * P);
			kp[d] = a.piv[(row0 + j) * P + pc];
			const int jr = min(jb + d + KL + 3, N - 1);
			gl[d] = a.G[(row0 + jr) * 2 * P + pc];
			gt[d] = a.G[(row0 + jr) * 2 * P + P + pc];
			const long nr = am + jr, kr = nr * nr + nr + m;
			u0[d] = B[B_U0 * (long)a.ncplx + kr];
			v0[d] = B[B_V0 * (long)a.ncplx + kr];
		}
#pragma unroll
		for (int d = 0; d < RD; d++) {
			const int j = jb + d;
			if (j < N) {
				pivot_swap<KL>(wa, kp[d]);
				pivot_swap<KL>(wb, kp[d]);
#pragma unroll
				for (int e = 0; e < KL / 2; e++) {
					wa[2 * e + 2] = cb(wa[2 * e + 2], cm(lf[d][e], wa[0]));
					wb[2 * e + 2] = cb(wb[2 * e + 2], cm(lf[d][e], wb[0]));
				}
				Ya[(row0 + j) * P + pc] = wa[0];
				Yb[(row0 + j) * P + pc] = wb[0];
#pragma unroll
				for (int i = 0; i < KL; i++) { wa[i] = wa[i + 1]; wb[i] = wb[i + 1]; }
				const int jn = j + KL + 1;
				if (jn < N) {
					double2 an, bn;
					ab_pre(jn + 2, gl[d], gt[d], u0[d], v0[d], an, bn);
					rhs_next(jn, wa[KL], wb[KL], an, bn);
					keep(jn, wa[KL], wb[KL]);
				} else {
					wa[KL] = wb[KL] = make_double2(0, 0);
				}
			}
		}
	}
	// backward substitution with U from Ya, Yb; emit(j, xa, xb, sa, sb, d) per
	// row, j = N-1 .. 0, sets beta x for the pole sums when SUM; pre(d, j) as
	// for phi
	auto bwd = [&](auto emit, auto sum_tag, auto pre) __attribute__((always_inline)) {
		constexpr bool SUM = decltype(sum_tag)::value;
		double2 va[KV + 1], vb[KV + 1];
#pragma unroll
		for (int i = 0; i <= KV; i++) {
			const bool in = N - 1 - i >= 0;
			va[i] = in ? Ya[(row0 + N - 1 - i) * P + pc] : make_double2(0, 0);
			vb[i] = in ? Yb[(row0 + N - 1 - i) * P + pc] : make_double2(0, 0);
		}
		for (int jt = N - 1; jt >= 0; jt -= RD) {
			double2 uf[RD][KL + 1], ya[RD], yb[RD], sx[2 * RD];
#pragma unroll
			for (int d = 0; d < RD; d++) {
				const int j = max(jt - d, 0);
				const FT *f = a.fac + (row0 + j) * (size_t)NE * P + pc;
#pragma unroll
				for (int e = 0; e <= KL; e++) uf[d][e] = ldf(f + (size_t)e * P);
				const int jy = jt - d - KV - 1;
				ya[d] = (jy >= 0) ? Ya[(row0 + jy) * P + pc] : make_double2(0, 0);
				yb[d] = (jy >= 0) ? Yb[(row0 + jy) * P + pc] : make_double2(0, 0);
				pre(d, jt - d);
			}
#pragma unroll
			for (int d = 0; d < RD; d++) {
				const int j = jt - d;
				sx[2 * d] = sx[2 * d + 1] = make_double2(0, 0);
				if (j >= 0) {
					const double2 xa = cm(va[0], uf[d][0]), xb = cm(vb[0], uf[d][0]);
#pragma unroll
					for (int e = 1; e <= KL; e++) {
						va[2 * e] = cb(va[2 * e], cm(uf[d][e], xa));
						vb[2 * e] = cb(vb[2 * e], cm(uf[d][e], xb));
					}
					emit(j, xa, xb, sx[2 * d], sx[2 * d + 1], d);
#pragma unroll
					for (int i = 0; i < KV; i++) { va[i] = va[i + 1]; vb[i] = vb[i + 1]; }
					va[KV] = ya[d];
					vb[KV] = yb[d];
				}
			}
			if constexpr (SUM) {
				// pole sums of the batch's rows
				wave_sum_n<2 * RD>(sx);
				if (threadIdx.x == 0)
#pragma unroll
					for (int d = 0; d < RD; d++)
						if (jt - d >= 0) {
							a.part[((size_t)blockIdx.x * 3 + 1) * a.rows + row0 + jt - d] = sx[2 * d];
							a.part[((size_t)blockIdx.x * 3 + 2) * a.rows + row0 + jt - d] = sx[2 * d + 1];
						}
			}
		}
	};
	auto sums = [&](int, double2 xa, double2 xb, double2 &sa, double2 &sb, int) __attribute__((always_inline)) {
		if (pv) {
			sa = cm(be, xa);
			sb = cm(be, xb);
		}
	};
	if constexpr (!REF) {
		bwd(sums, std::true_type{}, [](int, int) {});
	} else {
		const double2 al2 = cm(al, al);
		// x0 into X; r = b - A x0 in place in R, row q once x0(q-2 .. q+2) is known;
		// xw[k] = x0(q - KL + k)
		double2 xwa[KV + 1], xwb[KV + 1];
#pragma unroll
		for (int k = 0; k <= KV; k++) xwa[k] = xwb[k] = make_double2(0, 0);
		// R and the K pieces of row q = j + KL for the row j at place d of a batch
		double2 rqa[RD], rqb[RD], kq[RD][3];
		auto resid = [&](int q, double2 ra, double2 rb, const double2 (&Kq)[3]) __attribute__((always_inline)) {
			if (q < 0 || q >= N) return;
#pragma unroll
			for (int e = 0; e <= 2; e++) {
				const int c = q + 2 * e - KL;
				if (c < 0 || c >= N) continue;
				const double2 A = (2 * e == KL) ? ca(Kq[e], al2) : Kq[e];
				ra = cb(ra, cm(A, xwa[2 * e]));
				rb = cb(rb, cm(A, xwb[2 * e]));
			}
			a.R[(row0 + q) * rs + pc] = ra;
			a.R[(row0 + q) * rs + P + pc] = rb;
		};
		// loads of row q's R and K pieces (zero outside the band)
		auto rk_load = [&](int q, double2 &ra, double2 &rb, double2 (&Kq)[3]) __attribute__((always_inline)) {
			const bool in = q >= 0 && q < N;
			ra = in ? a.R[(row0 + q) * rs + pc] : make_double2(0, 0);
			rb = in ? a.R[(row0 + q) * rs + P + pc] : make_double2(0, 0);
#pragma unroll
			for (int e = 0; e <= 2; e++) Kq[e] = in ? a.Kv[(row0 + q) * 3 + e] : make_double2(0, 0);
		};
		bwd([&](int j, double2 xa, double2 xb, double2 &, double2 &, int d) __attribute__((always_inline)) {
			if (pv) {
				a.X[(row0 + j) * rs + p] = xa;
				a.X[(row0 + j) * rs + P + p] = xb;
			}
#pragma unroll
			for (int k = KV; k > 0; k--) { xwa[k] = xwa[k - 1]; xwb[k] = xwb[k - 1]; }
			xwa[0] = xa;
			xwb[0] = xb;
			resid(j + KL, rqa[d], rqb[d], kq[d]);
		}, std::false_type{}, [&](int d, int j) __attribute__((always_inline)) {
			rk_load(j >= 0 ? j + KL : -1, rqa[d], rqb[d], kq[d]);
		});
		for (int j = -1; j >= -KL; j--) {
#pragma unroll
			for (int k = KV; k > 0; k--) { xwa[k] = xwa[k - 1]; xwb[k] = xwb[k - 1]; }
			xwa[0] = xwb[0] = make_double2(0, 0);
			double2 ra, rb, Kq[3];
			rk_load(j + KL, ra, rb, Kq);
			resid(j + KL, ra, rb, Kq);
		}
		// correction: forward substitution of the residuals, then x = x0 + dx
		{
			double2 ca_[KL + 1], cb_[KL + 1];
#pragma unroll
			for (int i = 0; i <= KL; i++) {
				ca_[i] = (i < N) ? a.R[(row0 + i) * rs + pc] : make_double2(0, 0);
				cb_[i] = (i < N) ? a.R[(row0 + i) * rs + P + pc] : make_double2(0, 0);
			}
			for (int jb = 0; jb < N; jb += RD) {
				double2 lf[RD][KL / 2], rna[RD], rnb[RD];
				int kp[RD];
#pragma unroll
				for (int d = 0; d < RD; d++) {
					const int j = min(jb + d, N - 1);
					const FT *f = a.fac + (row0 + j) * (size_t)NE * P + pc;
#pragma unroll
					for (int e = 0; e < KL / 2; e++) lf[d][e] = ldf(f + (size_t)(KL + 1 + e) * P);
					kp[d] = a.piv[(row0 + j) * P + pc];
					const int jn = jb + d + KL + 1;
					rna[d] = (jn < N) ? a.R[(row0 + jn) * rs + pc] : make_double2(0, 0);
					rnb[d] = (jn < N) ? a.R[(row0 + jn) * rs + P + pc] : make_double2(0, 0);
				}
#pragma unroll
				for (int d = 0; d < RD; d++) {
					const int j = jb + d;
					if (j < N) {
						pivot_swap<KL>(ca_, kp[d]);
						pivot_swap<KL>(cb_, kp[d]);
#pragma unroll
						for (int e = 0; e < KL / 2; e++) {
							ca_[2 * e + 2] = cb(ca_[2 * e + 2], cm(lf[d][e], ca_[0]));
							cb_[2 * e + 2] = cb(cb_[2 * e + 2], cm(lf[d][e], cb_[0]));
						}
						Ya[(row0 + j) * P + pc] = ca_[0];
						Yb[(row0 + j) * P + pc] = cb_[0];
#pragma unroll
						for (int i = 0; i < KL; i++) { ca_[i] = ca_[i + 1]; cb_[i] = cb_[i + 1]; }
						ca_[KL] = rna[d];
						cb_[KL] = rnb[d];
					}
				}
			}
		}
		// x0 of the batch's rows
		double2 xqa[RD], xqb[RD];
		bwd([&](int j, double2 dxa, double2 dxb, double2 &sa, double2 &sb, int d) __attribute__((always_inline)) {
			sums(j, ca(xqa[d], dxa), ca(xqb[d], dxb), sa, sb, d);
		}, std::true_type{}, [&](int d, int j) __attribute__((always_inline)) {
			xqa[d] = j >= 0 ? a.X[(row0 + j) * rs + pc] : make_double2(0, 0);
			xqb[d] = j >= 0 ? a.X[(row0 + j) * rs + P + pc] : make_double2(0, 0);
		});
	}
}


// the pole sums c (complex layout, as k_rexi_sum) reduced to the real spectrum
// of Re(S(c)): the reference synthesises c, keeps the real part and analyses
// it (AppSWE.hpp:519-521); on band-limited fields that is the spectral map
// a(l, 0) = Re c(l, 0), a(l, m) = (c(l, m) + (-1)^m conj(c(l, -m))) / 2
__global__ void k_rexi_sum_real(const double2 *__restrict__ part, double2 *__restrict__ out,
				const int *__restrict__ mofr, const int *__restrict__ moff, int L, int ngroups,
				size_t rows, long nlm)
{
	const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (q >= nlm) return;
	const int m = mofr[q], l = (int)(q - lm_index(L, 0, m));
	const size_t rp = moff[m + L] + (l - m), rn = moff[-m + L] + (l - m);
	const double sg = (m & 1) ? -1.0 : 1.0;
#pragma unroll
	for (int f = 0; f < 3; f++) {
		double2 sp = make_double2(0, 0), sn = make_double2(0, 0);
		for (int g = 0; g < ngroups; g++) {
			const double2 v = part[((size_t)g * 3 + f) * rows + rp];
			sp.x += v.x;
			sp.y += v.y;
			if (m > 0) {
				const double2 w = part[((size_t)g * 3 + f) * rows + rn];
				sn.x += w.x;
				sn.y += w.y;
			}
		}
		out[f * nlm + q] = m == 0 ? make_double2(sp.x, 0.0)
					  : make_double2(0.5 * (sp.x + sg * sn.x), 0.5 * (sp.y - sg * sn.y));
	}
}
__global__ void k_fill_mu(double *__restrict__ o, const double *__restrict__ ct, int nlat, long n)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i < n) o[i] = ct[i % nlat];
}
// The fields the REXI basis transforms are real: their complex spectra are
// c(l, m) = a(l, m) and c(l, -m) = (-1)^m conj(a(l, m)), with a the real
// spectrum (m >= 0), for the transform pair of SHTns' complex and real
// interfaces (SPHDataComplex.hpp:179-216, SPHData.hpp:190-227), so they run as
// real transforms between these two layout changes.
// real layout (nf fields, stride nlm) -> complex layout (stride nc)
__global__ void k_spec_r2cx(const double2 *__restrict__ a, double2 *__restrict__ c, const int *__restrict__ lofc,
			    int L, long nlm, long nc, int nf)
{
	const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= nc * nf) return;
	const long f = i / nc, k = i - f * nc;
	const int l = lofc[k], m = (int)(k - (long)l * l - l), am = abs(m);
	const double2 v = a[f * nlm + lm_index(L, l, am)];
	const double sg = (am & 1) ? -1.0 : 1.0;
	c[i] = m >= 0 ? v : make_double2(sg * v.x, -sg * v.y);
}
// complex layout of a real field's spectrum -> real layout (the m >= 0 half)
__global__ void k_spec_c2rx(const double2 *__restrict__ c, double2 *__restrict__ a, const int *__restrict__ mofr,
			    int L, long nlm, long nc, int nf)
{
	const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i >= nlm * nf) return;
	const long f = i / nlm, q = i - f * nlm;
	const int m = mofr[q], l = (int)(q - lm_index(L, 0, m));
	a[i] = c[f * nc + (long)l * l + l + m];
}
// o[f] = g * x[f] (real pointwise, g one field, x nf fields)
__global__ void k_mul_bcast(const double *__restrict__ g, const double *__restrict__ x, double *__restrict__ o,
			    long nspat, int nf)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i < nspat * nf) o[i] = g[i % nspat] * x[i];
}
// sum / difference of spectral pairs with a scale: o = s*(x + sgn*y)
__global__ void k_comb(const double2 *__restrict__ x, const double2 *__restrict__ y, double2 *__restrict__ o, double sgn,
		       double s, long n)
{
	long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (i < n) o[i] = make_double2(s * (x[i].x + sgn * y[i].x), s * (x[i].y + sgn * y[i].y));
}

inline int g256(long n) { return (int)((n + 255) / 256); }

// ---- gradient operator ------------------------------------------------------
// grad_lon / grad_lat of a single-m field without the FFT pair: synthesis to
// (m, ring), scaling by 2 pi / sqrt(1 - mu^2), analysis with the Gauss weights.
// For each m that is the real matrix
//   M_m[l][l'] = sum_j w_j (2 pi / sqrt(1 - mu_j^2)) Pbar_l^m(mu_j) Pbar_l'^m(mu_j),
// symmetric, zero unless l - l' is even (the rings pair up as mu, -mu), and the
// same for -m (the (-1)^m of the negative-m slots cancels).  It is built once
// per REXI object by running the pole-layout synthesis / analysis on unit
// vectors, and applied per step as a batched GEMM over the 4P real columns of
// the poles (columns i m phi | D' phi, re / im).

// 16-row tiles and 16-column strips per wavefront of the apply kernel: 4 x 4 in
// FP64 (measured faster than 8 x 2: T1023 REXI step 44.4 -> 42.6 ms), 8 x 2 in
// FP32 (4 x 4 measured slower there)
template <typename T>
constexpr int gop_rt() { return sizeof(T) == 8 ? 4 : 8; }

typedef double gd4v __attribute__((ext_vector_type(4)));
typedef float gf4v __attribute__((ext_vector_type(4)));

// unit vectors: column cc of a chunk is e_{c0 + cc} in every m >= 0 block
__global__ void k_gop_unit(double2 *__restrict__ x, const int *__restrict__ moff, int L, int CB, int c0)
{
	const int m = (int)blockIdx.y - L, am = abs(m);
	const int N = L + 1 - am;
	const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (t >= (long)N * CB) return;
	const int i = (int)(t / CB), cc = (int)(t % CB);
	x[((long)moff[m + L] + i) * CB + cc] = make_double2((m >= 0 && i == c0 + cc) ? 1.0 : 0.0, 0.0);
}

// the chunk's columns of M(|m|, parity) from the analysed unit vectors
__global__ void k_gop_pack(const double2 *__restrict__ g, double *__restrict__ M, const long *__restrict__ goff,
			   const int *__restrict__ moff, int L, int CB, int c0)
{
	const int am = blockIdx.y;
	const int N = L + 1 - am;
	const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (t >= (long)N * CB) return;
	const int i = (int)(t / CB), cc = (int)(t % CB), c = c0 + cc;
	if (c >= N || ((i ^ c) & 1)) return;
	const int par = i & 1, np = (N + 1 - par) / 2;
	M[goff[2 * am + par] + (long)(i >> 1) * np + (c >> 1)] = g[((long)moff[am + L] + i) * CB + cc].x;
}

// symmetrise (the pair (a, b), a <= b, handled by one thread) and round to FP32
__global__ void k_gop_sym(double *__restrict__ M, float *__restrict__ M32, const long *__restrict__ goff, int L)
{
	const int am = blockIdx.y >> 1, par = blockIdx.y & 1;
	const int N = L + 1 - am, np = (N + 1 - par) / 2;
	const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
	if (t >= (long)np * np) return;
	const int a = (int)(t / np), b = (int)(t % np);
	if (a > b) return;
	double *Mm = M + goff[2 * am + par];
	float *Mm32 = M32 + goff[2 * am + par];
	const double v = 0.5 * (Mm[(long)a * np + b] + Mm[(long)b * np + a]);
	Mm[(long)a * np + b] = v;
	Mm[(long)b * np + a] = v;
	Mm32[(long)a * np + b] = (float)v;
	Mm32[(long)b * np + a] = (float)v;
}

// Y[rows of (m, parity)][cols] = M(|m|, parity) X[...]: one wavefront per
// (m, parity, up to RT 16-row tiles, CS strips of 16 real columns).  Each
// k step loads RT A values and CS B values per lane and issues
// RT * CS = 16 MFMAs on them (FP64: 128 accumulator registers, 2 waves
// per SIMD; FP32: 64, 3 waves), so M is read once per CS strips and every X
// row once per wavefront.  A = M (symmetric: the 16 rows of one column k are
// contiguous), B = X rows.  FP64: v_mfma_f64_16x16x4f64, D[i][n] in lane
// n + 16 (i % 4), element i / 4.  FP32: v_mfma_f32_16x16x4f32 with M and X
// rounded, D[i][n] in lane n + 16 (i / 4), element i % 4.  A[i][k] in lane
// i + 16 k and B[k][n] in lane n + 16 k for both.
template <typename T>
constexpr int gop_cs() { return sizeof(T) == 8 ? 4 : 2; }

template <typename T>
__global__ __launch_bounds__(64) void k_gop_apply(const T *__restrict__ M, const long *__restrict__ goff,
						  const int4 *__restrict__ tiles, const int *__restrict__ moff,
						  const double *__restrict__ X, double *__restrict__ Y, int L, int ncol)
{
	constexpr bool F64 = sizeof(T) == 8;
	constexpr int CS = gop_cs<T>(), GOP_RT = gop_rt<T>();
	typedef std::conditional_t<F64, gd4v, gf4v> acc_t;
	const int4 tl = tiles[blockIdx.x];
	const int m = tl.x, par = tl.y, r0 = tl.z;
	const int am = abs(m), N = L + 1 - am, np = (N + 1 - par) / 2;
	const int nt = min(GOP_RT, (np - r0 + 15) / 16);   // row tiles of this wavefront (uniform)
	const T *Mm = M + goff[2 * am + par];
	const long xrow0 = (long)moff[m + L] + par;   // row of parity index 0; parity index q is row xrow0 + 2 q
	const int lane = threadIdx.x, i = lane & 15, k = lane >> 4;
	int col[CS];
	bool cv[CS];
#pragma unroll
	for (int c = 0; c < CS; c++) {
		col[c] = (blockIdx.y * CS + c) * 16 + i;
		cv[c] = col[c] < ncol;
	}
	acc_t acc[GOP_RT][CS];
#pragma unroll
	for (int t = 0; t < GOP_RT; t++)
#pragma unroll
		for (int c = 0; c < CS; c++) acc[t][c] = acc_t{0, 0, 0, 0};
	// UN k steps per pass: their loads are issued together
	constexpr int UN = 2;
	for (int kk = 0; kk < np; kk += 4 * UN) {
		T b[UN][CS], av[UN][GOP_RT];
#pragma unroll
		for (int u = 0; u < UN; u++) {
			const int q = kk + 4 * u + k;
			const bool qv = q < np;
			const double *xq = X + (xrow0 + 2 * (long)(qv ? q : 0)) * ncol;
#pragma unroll
			for (int c = 0; c < CS; c++) b[u][c] = (qv && cv[c]) ? (T)xq[col[c]] : (T)0;
			const T *mq = Mm + (long)(qv ? q : 0) * np + r0 + i;
#pragma unroll
			for (int t = 0; t < GOP_RT; t++)
				av[u][t] = (t < nt && qv && r0 + 16 * t + i < np) ? mq[16 * t] : (T)0;
		}
#pragma unroll
		for (int u = 0; u < UN; u++)
#pragma unroll
			for (int t = 0; t < GOP_RT; t++) {
				if (t < nt) {
#pragma unroll
					for (int c = 0; c < CS; c++) {
						if constexpr (F64)
							acc[t][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][t], b[u][c], acc[t][c], 0, 0, 0);
						else
							acc[t][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u][t], b[u][c], acc[t][c], 0, 0, 0);
					}
				}
			}
	}
#pragma unroll
	for (int t = 0; t < GOP_RT; t++) {
		if (t >= nt) break;
#pragma unroll
		for (int e = 0; e < 4; e++) {
			const int row = r0 + 16 * t + (F64 ? k + 4 * e : 4 * k + e);
			if (row >= np) continue;
			double *yr = Y + (xrow0 + 2 * (long)row) * ncol;
#pragma unroll
			for (int c = 0; c < CS; c++)
				if (cv[c]) yr[col[c]] = (double)acc[t][c][e];
		}
	}
}


template <typename T>
int dev_alloc(sphswe_ctx *ctx, T **p, size_t n)
{
	if (hipMalloc((void **)p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess)
		return sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "rexi: allocation of %zu bytes failed", sizeof(T) * n);
	return SPHSWE_OK;
}

}  // namespace

// the context is being destroyed: the object keeps its own memory (freed by
// sphswe_rexi_destroy, which must still be called) and accepts no other call
void sphswe_rexi_detach(sphswe_rexi *r) { r->ctx = nullptr; }

extern "C" int sphswe_rexi_destroy(sphswe_rexi *r)
{
	if (!r) return SPHSWE_OK;
	if (r->ctx) {
		hipStreamSynchronize(r->ctx->stream);
		auto &v = r->ctx->rexis;
		for (size_t i = 0; i < v.size(); i++)
			if (v[i] == r) { v.erase(v.begin() + i); break; }
	}
	void *ptrs[] = {r->d_alpha, r->d_beta, r->d_moff, r->d_fphi, r->d_fvel, r->d_pphi, r->d_pvel, r->d_oscale,
			r->d_X, r->d_G, r->d_Y, r->d_part, r->d_basis, r->d_cphys, r->d_cspec,
			r->d_gop, r->d_gop32, r->d_goff, r->d_gtiles, r->d_kphi, r->d_kvel,
			r->d_in};
	for (void *p : ptrs)
		if (p) hipFree(p);
	delete r;
	return SPHSWE_OK;
}

// build the gradient operator of a REXI object (d_moff and d_oscale on the
// device): unit vectors through the FP64 pole-layout synthesis and analysis in
// chunks of up to 64 columns; the tile table of the per-step GEMM
static int rexi_gradient_operator(sphswe_ctx *ctx, sphswe_rexi *r)
{
	const int L = ctx->L;
	hipStream_t st = ctx->stream;
	std::vector<long> goff(2 * (L + 1));
	long tot = 0;
	for (int am = 0; am <= L; am++)
		for (int par = 0; par < 2; par++) {
			const long np = (L + 1 - am + 1 - par) / 2;
			goff[2 * am + par] = tot;
			tot += np * np;
		}
	// row tiles per wavefront of the apply kernel this object will launch
	const int rt = r->prm.precision == SPHSWE_PREC_FP64 ? gop_rt<double>() : gop_rt<float>();
	std::vector<int4> tiles;
	for (int am = 0; am <= L; am++)
		for (int sg = (am ? -1 : 1); sg <= 1; sg += 2)
			for (int par = 0; par < 2; par++) {
				const int np = (L + 1 - am + 1 - par) / 2;
				for (int r0 = 0; r0 < np; r0 += 16 * rt) tiles.push_back(make_int4(sg * am, par, r0, 0));
			}
	r->ngtiles = (int)tiles.size();
	int rc;
	if ((rc = dev_alloc(ctx, &r->d_gop, tot)) || (rc = dev_alloc(ctx, &r->d_gop32, tot)) ||
	    (rc = dev_alloc(ctx, &r->d_goff, goff.size())) || (rc = dev_alloc(ctx, &r->d_gtiles, tiles.size())))
		return rc;
	SPHSWE_HIP(ctx, hipMemcpy(r->d_goff, goff.data(), sizeof(long) * goff.size(), hipMemcpyHostToDevice));
	SPHSWE_HIP(ctx, hipMemcpy(r->d_gtiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
	const int CB = std::min(L + 1, 64);
	double2 *xid = nullptr, *four = nullptr, *gout = nullptr;
	auto release = [&]() {
		hipStreamSynchronize(st);
		if (xid) hipFree(xid);
		if (four) hipFree(four);
		if (gout) hipFree(gout);
	};
	if ((rc = dev_alloc(ctx, &xid, r->rows * CB)) || (rc = dev_alloc(ctx, &gout, r->rows * CB)) ||
	    (rc = dev_alloc(ctx, &four, (size_t)CB * (2 * L + 1) * ctx->nlat))) {
		release();
		return rc;
	}
	for (int c0 = 0; c0 <= L; c0 += CB) {
		k_gop_unit<<<dim3((unsigned)(((long)(L + 1) * CB + 255) / 256), 2 * L + 1), 256, 0, st>>>(xid, r->d_moff, L,
													   CB, c0);
		if ((rc = sphswe_leg_synth_pole(ctx, xid, four, CB, CB, r->d_moff, r->d_oscale, SPHSWE_PREC_FP64)) ||
		    (rc = sphswe_leg_anal_pole(ctx, four, gout, CB, CB, r->d_moff, SPHSWE_PREC_FP64))) {
			release();
			return rc;
		}
		k_gop_pack<<<dim3((unsigned)(((long)(L + 1) * CB + 255) / 256), L + 1), 256, 0, st>>>(gout, r->d_gop,
												       r->d_goff, r->d_moff,
												       L, CB, c0);
	}
	const long npmax = (L + 2) / 2;
	k_gop_sym<<<dim3((unsigned)((npmax * npmax + 255) / 256), 2 * (L + 1)), 256, 0, st>>>(r->d_gop, r->d_gop32,
											      r->d_goff, L);
	hipError_t e = hipGetLastError();
	release();
	if (e != hipSuccess) return sphswe_set_err(ctx, SPHSWE_ERR_HIP, "rexi: gradient operator: %s", hipGetErrorString(e));
	return SPHSWE_OK;
}

// SWERexiSPH::setup for every pole of (alpha0, beta0): band assembly, pivoted LU
// factors, device buffers
static int rexi_build(sphswe_ctx *ctx, const sphswe_rexi_params *prm, const std::vector<cplx> &alpha0,
		      const std::vector<cplx> &beta0, sphswe_rexi **out)
{
	using namespace sphswe_host;
	*out = nullptr;
	if (!(prm->dt > 0)) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI needs a positive time step (AppSWE.hpp:185-190)");
	if (alpha0.empty() || alpha0.size() != beta0.size()) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: empty pole set");
	if (prm->band_precision != SPHSWE_PREC_FP64 && prm->band_precision != SPHSWE_PREC_FP32)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: band_precision must be FP64 or FP32");
	if (prm->precision != SPHSWE_PREC_FP64 && prm->precision != SPHSWE_PREC_FP32)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: precision must be FP64 or FP32");
	if (prm->refine != 0 && prm->refine != 1) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: refine must be 0 or 1");
	if (prm->robert_functions != 0 && prm->robert_functions != 1)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: robert_functions must be 0 or 1");
	if (prm->robert_functions && prm->refine)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: the refinement sweep is not available with Robert functions");
	sphswe_rexi *r = new sphswe_rexi();
	r->ctx = ctx;
	ctx->rexis.push_back(r);
	r->prm = *prm;
	r->coriolis = prm->coriolis_omega != 0.0;
	r->robert = prm->robert_functions;
	r->alpha0 = alpha0;
	r->beta0 = beta0;
	const int P = r->P = (int)r->alpha0.size();
	const int L = ctx->L;
	r->rows = (size_t)(L + 1) * (L + 1);
	r->ngroups = (P + 63) / 64;
	std::vector<double2> al(P), be(P);
	std::vector<cplx> alc(P);
	for (int i = 0; i < P; i++) {
		const cplx a = r->alpha0[i] / prm->dt, b = r->beta0[i] / prm->dt;
		alc[i] = a;
		al[i] = make_double2(a.real(), a.imag());
		be[i] = make_double2(b.real(), b.imag());
	}
	std::vector<int> moff(2 * L + 1);
	for (int m = -L; m <= L; m++) moff[m + L] = ncompact(L, std::abs(m), m);

	// the band matrices of every pole (SWERexiSPH::setup, SWERexiSPH.hpp:104-129)
	// grouped by their power of alpha,
	//   phi:   A(alpha) = alpha^4 I + K0 + alpha^2 K2 + alpha K1 + K-1 / alpha,
	//   winds: alpha^2 I + W^2 mu^2,
	// with the pole-independent pieces assembled once on the host; the pivoted LU
	// of every (pole, m) runs on the device (k_rexi_factor).  Every component
	// couples degrees of equal parity (offsets 0, +-2, +-4), so each m block splits
	// into two independent parity classes and only the even offsets are kept
	// (checked here on the pieces).
	const double rr = prm->earth_radius, W = 2.0 * prm->coriolis_omega, G = prm->h0 * prm->gravitation;
	const int NCp = rexi_nec(4), NCv = rexi_nec(2);       // kept factor entries per row
	r->fac32 = prm->band_precision == SPHSWE_PREC_FP32;
	r->refine = prm->refine && r->coriolis;   // without rotation the solves are diagonal
	std::vector<double2> hK, hKv;
	if (r->coriolis) {
		BandRows K0(L, 4), K2(L, 4), K1(L, 4), Km1(L, 4), Kv(L, 2);
		band_add(K0, 13, (W * W) * (W * W), rr);
		band_add(K0, r->robert ? 21 : 16, G * 2.0 * W * W, rr);
		band_add(K0, 18, -G * W * W, rr);
		band_add(K2, 12, 2.0 * W * W, rr);
		band_add(K2, 17, -G, rr);
		band_add(K1, r->robert ? 19 : 14, -G * W, rr);
		band_add(Km1, r->robert ? 20 : 15, G * W * W * W, rr);
		band_add(Kv, 12, W * W, rr);
		hK.resize(r->rows * 20);
		hKv.resize(r->rows * 3);
		const BandRows *kinds[4] = {&K0, &K2, &K1, &Km1};
		int parity_bad = 0;
		for (size_t q = 0; q < r->rows; q++) {
			for (int k = 0; k < 4; k++) {
				for (int e = 0; e < 5; e++) {
					const cplx v = kinds[k]->data[q * 9 + 2 * e];
					hK[q * 20 + k * 5 + e] = make_double2(v.real(), v.imag());
				}
				for (int e = 0; e < 4; e++)
					if (kinds[k]->data[q * 9 + 2 * e + 1] != cplx(0, 0)) parity_bad = 1;
			}
			for (int e = 0; e < 3; e++) {
				const cplx v = Kv.data[q * 5 + 2 * e];
				hKv[q * 3 + e] = make_double2(v.real(), v.imag());
			}
			if (Kv.data[q * 5 + 1] != cplx(0, 0) || Kv.data[q * 5 + 3] != cplx(0, 0)) parity_bad = 1;
		}
		if (parity_bad) {
			delete r;
			return sphswe_set_err(ctx, SPHSWE_ERR_INTERNAL, "REXI bands: nonzero entry at an odd offset");
		}
	}
	std::vector<double> osc(ctx->nlat);
	// the metric division of grad_lon / grad_lat, times the 2 pi that the forward
	// FFT (normalised by 2 pi / nphi) would give the Fourier coefficients the
	// Legendre analysis expects; the pole path goes from synthesis to analysis
	// without the FFT pair
	for (int j = 0; j < ctx->nlat; j++) osc[j] = 2.0 * M_PI / sqrt(1.0 - ctx->ct[j] * ctx->ct[j]);

	int rc = SPHSWE_OK;
	auto up = [&](void *d, const void *h, size_t b) -> int {
		if (hipMemcpy(d, h, b, hipMemcpyHostToDevice) != hipSuccess)
			return sphswe_set_err(ctx, SPHSWE_ERR_HIP, "rexi: upload failed");
		return SPHSWE_OK;
	};
	// factor storage on the device, FP64 or FP32, zero-filled
	auto alloc_fac = [&](void **d, size_t n) -> int {
		int rc2;
		if (!r->fac32) {
			double2 *p2 = nullptr;
			if ((rc2 = dev_alloc(ctx, &p2, n))) return rc2;
			*d = p2;
		} else {
			float2 *p2 = nullptr;
			if ((rc2 = dev_alloc(ctx, &p2, n))) return rc2;
			*d = p2;
		}
		if (hipMemsetAsync(*d, 0, n * (r->fac32 ? sizeof(float2) : sizeof(double2)), ctx->stream) != hipSuccess)
			return sphswe_set_err(ctx, SPHSWE_ERR_HIP, "rexi: clearing the factors failed");
		return SPHSWE_OK;
	};
	const size_t nsp = (size_t)ctx->nspat;
	if ((rc = dev_alloc(ctx, &r->d_alpha, P)) || (rc = dev_alloc(ctx, &r->d_beta, P)) ||
	    (rc = dev_alloc(ctx, &r->d_moff, 2 * L + 1)) || (rc = alloc_fac(&r->d_fphi, r->rows * NCp * P)) ||
	    (rc = dev_alloc(ctx, &r->d_oscale, ctx->nlat)) ||
	    (rc = dev_alloc(ctx, &r->d_X, r->rows * 2 * P)) || (rc = dev_alloc(ctx, &r->d_G, r->rows * 2 * P)) ||
	    (rc = dev_alloc(ctx, &r->d_Y, r->rows * 2 * P)) ||
	    (rc = dev_alloc(ctx, &r->d_part, (size_t)r->ngroups * 3 * r->rows)) ||
	    (rc = dev_alloc(ctx, &r->d_basis, (size_t)NBASIS * ctx->ncplx)) ||
	    (rc = dev_alloc(ctx, &r->d_cphys, 8 * nsp)) || (rc = dev_alloc(ctx, &r->d_cspec, (size_t)16 * ctx->ncplx + 8 * ctx->nlm)) ||
	    (rc = dev_alloc(ctx, &r->d_in, (size_t)3 * ctx->nlm))) {
		sphswe_rexi_destroy(r);
		return rc;
	}
	if ((rc = up(r->d_alpha, al.data(), sizeof(double2) * P)) || (rc = up(r->d_beta, be.data(), sizeof(double2) * P)) ||
	    (rc = up(r->d_moff, moff.data(), sizeof(int) * moff.size())) ||
	    (rc = up(r->d_oscale, osc.data(), sizeof(double) * osc.size()))) {
		sphswe_rexi_destroy(r);
		return rc;
	}
	// the LU factors of every (pole, m) on the device
	int *d_info = nullptr;
	if ((rc = dev_alloc(ctx, &d_info, 1))) {
		sphswe_rexi_destroy(r);
		return rc;
	}
	const int no_info = 0x7fffffff;
	if (hipMemcpy(d_info, &no_info, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
		hipFree(d_info);
		sphswe_rexi_destroy(r);
		return sphswe_set_err(ctx, SPHSWE_ERR_HIP, "rexi: upload failed");
	}
	if (r->coriolis) {
		if ((rc = alloc_fac(&r->d_fvel, r->rows * NCv * P)) || (rc = dev_alloc(ctx, &r->d_pphi, r->rows * P)) ||
		    (rc = dev_alloc(ctx, &r->d_pvel, r->rows * P)) || (rc = dev_alloc(ctx, &r->d_kphi, hK.size())) ||
		    (rc = dev_alloc(ctx, &r->d_kvel, hKv.size())) ||
		    (rc = up(r->d_kphi, hK.data(), sizeof(double2) * hK.size())) ||
		    (rc = up(r->d_kvel, hKv.data(), sizeof(double2) * hKv.size()))) {
			hipFree(d_info);
			sphswe_rexi_destroy(r);
			return rc;
		}
		dim3 fg(r->ngroups, 2 * (2 * L + 1));
		if (r->fac32) {
			k_rexi_factor<2, float2><<<fg, 64, 0, ctx->stream>>>(r->d_alpha, r->d_moff, r->d_kphi, (float2 *)r->d_fphi,
									   r->d_pphi, d_info, L, P);
			k_rexi_factor<1, float2><<<fg, 64, 0, ctx->stream>>>(r->d_alpha, r->d_moff, r->d_kvel, (float2 *)r->d_fvel,
									   r->d_pvel, d_info, L, P);
		} else {
			k_rexi_factor<2, double2><<<fg, 64, 0, ctx->stream>>>(r->d_alpha, r->d_moff, r->d_kphi, (double2 *)r->d_fphi,
									    r->d_pphi, d_info, L, P);
			k_rexi_factor<1, double2><<<fg, 64, 0, ctx->stream>>>(r->d_alpha, r->d_moff, r->d_kvel, (double2 *)r->d_fvel,
									    r->d_pvel, d_info, L, P);
		}
	} else {
		dim3 dg((unsigned)(((long)(L + 1) * P + 255) / 256), 2 * L + 1);
		if (r->fac32)
			k_rexi_diag<float2><<<dg, 256, 0, ctx->stream>>>(r->d_alpha, r->d_moff, (float2 *)r->d_fphi, L, P, NCp,
									 G / (rr * rr));
		else
			k_rexi_diag<double2><<<dg, 256, 0, ctx->stream>>>(r->d_alpha, r->d_moff, (double2 *)r->d_fphi, L, P, NCp,
									  G / (rr * rr));
	}
	{
		hipError_t e = hipGetLastError();
		int info = no_info;
		if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
		if (e == hipSuccess) e = hipMemcpy(&info, d_info, sizeof(int), hipMemcpyDeviceToHost);
		hipFree(d_info);
		if (e != hipSuccess) {
			sphswe_rexi_destroy(r);
			return sphswe_set_err(ctx, SPHSWE_ERR_HIP, "rexi: band factorisation: %s", hipGetErrorString(e));
		}
		if (info != no_info) {
			sphswe_rexi_destroy(r);
			return sphswe_set_err(ctx, SPHSWE_ERR_SINGULAR, "zgbsv-equivalent factorisation: INFO = %d", info);
		}
	}
	// the pieces stay for the refinement sweep only
	if (r->coriolis && !r->refine) {
		hipFree(r->d_kphi);
		hipFree(r->d_kvel);
		r->d_kphi = r->d_kvel = nullptr;
	}
	// Robert functions: the gradients of phi are the spectral i m phi and D' phi
	// that the phi sweep writes, so there is no gradient operator
	if (!r->robert && (rc = rexi_gradient_operator(ctx, r))) {
		sphswe_rexi_destroy(r);
		return rc;
	}
	*out = r;
	return SPHSWE_OK;
}

extern "C" int sphswe_rexi_create(sphswe_ctx *ctx, const sphswe_rexi_params *prm, sphswe_rexi **out)
{
	if (!ctx || !prm || !out) return SPHSWE_ERR_ARG;
	*out = nullptr;
	if (prm->rexi_M < 1 || !(prm->rexi_h > 0)) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: invalid h or M");
	std::vector<cplx> al, be;
	sphswe_host::rexi_coefficients(prm->rexi_h, prm->rexi_M, prm->rexi_half != 0, al, be);
	return rexi_build(ctx, prm, al, be, out);
}

extern "C" int sphswe_rexi_create_poles(sphswe_ctx *ctx, const sphswe_rexi_params *prm, int npoles, const double *alpha,
					const double *beta, sphswe_rexi **out)
{
	if (!ctx || !prm || !out || npoles < 1 || !alpha || !beta) return SPHSWE_ERR_ARG;
	*out = nullptr;
	std::vector<cplx> al(npoles), be(npoles);
	for (int i = 0; i < npoles; i++) {
		al[i] = cplx(alpha[2 * i], alpha[2 * i + 1]);
		be[i] = cplx(beta[2 * i], beta[2 * i + 1]);
		if (al[i] == cplx(0, 0)) return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "REXI: pole %d is zero", i);
	}
	return rexi_build(ctx, prm, al, be, out);
}

extern "C" int sphswe_rexi_coefficients(double h, int M, int half, int *npoles, double *alpha, double *beta_re,
					double *beta_im)
{
	if (!npoles || M < 1 || !(h > 0)) return SPHSWE_ERR_ARG;
	std::vector<cplx> al, br, bi;
	sphswe_host::rexi_coefficients(h, M, half != 0, al, br, beta_im ? &bi : nullptr);
	*npoles = (int)al.size();
	for (size_t i = 0; i < al.size(); i++) {
		if (alpha) { alpha[2 * i] = al[i].real(); alpha[2 * i + 1] = al[i].imag(); }
		if (beta_re) { beta_re[2 * i] = br[i].real(); beta_re[2 * i + 1] = br[i].imag(); }
		if (beta_im) { beta_im[2 * i] = bi[i].real(); beta_im[2 * i + 1] = bi[i].imag(); }
	}
	return SPHSWE_OK;
}

extern "C" int sphswe_rexi_poles(const sphswe_rexi *r, int *npoles, double *alpha, double *beta)
{
	if (!r || !npoles) return SPHSWE_ERR_ARG;
	*npoles = r->P;
	for (int i = 0; i < r->P; i++) {
		if (alpha) { alpha[2 * i] = r->alpha0[i].real(); alpha[2 * i + 1] = r->alpha0[i].imag(); }
		if (beta) { beta[2 * i] = r->beta0[i].real(); beta[2 * i + 1] = r->beta0[i].imag(); }
	}
	return SPHSWE_OK;
}

// pole-independent fields of one step from the complex physical phi0, u0, v0
// the first 3 real fields of r->d_cphys: the synthesis of (phi0, u0, v0)
static int rexi_basis(sphswe_ctx *ctx, sphswe_rexi *r)
{
	const long nsp = ctx->nspat, nc = ctx->ncplx, nlm = ctx->nlm;
	const int L = ctx->L;
	double2 *cs_ = r->d_cspec;     // complex spectral scratch, 16 fields, then real spectral scratch, 8 fields
	double2 *rs_ = cs_ + 16 * nc;
	double *rp_ = (double *)r->d_cphys;   // real physical scratch, 16 fields
	double2 *B = r->d_basis;
	const double ir = 1.0 / r->prm.earth_radius, W = 2.0 * r->prm.coriolis_omega;
	hipStream_t st = ctx->stream;
	int rc;
	auto SPEC = [&](int k) { return (double *)(cs_ + (size_t)k * nc); };
	auto RS = [&](int k) { return (double *)(rs_ + (size_t)k * nlm); };
	auto RP = [&](int k) { return rp_ + (size_t)k * nsp; };
	auto BF = [&](int k) { return (double *)(B + (size_t)k * nc); };
	// real spectra (nf fields from RS(k0)) into the complex layout at c, and back
	auto to_cplx = [&](int k0, double *c, int nf) -> int {
		k_spec_r2cx<<<g256(nc * nf), 256, 0, st>>>(rs_ + (size_t)k0 * nlm, (double2 *)c, ctx->d_lofcidx, L, nlm,
							  nc, nf);
		SPHSWE_LAUNCH_CHECK(ctx);
		return SPHSWE_OK;
	};
	auto to_real = [&](const double *c, int k0, int nf) -> int {
		k_spec_c2rx<<<g256(nlm * nf), 256, 0, st>>>((const double2 *)c, rs_ + (size_t)k0 * nlm, ctx->d_mofidx, L,
							   nlm, nc, nf);
		SPHSWE_LAUNCH_CHECK(ctx);
		return SPHSWE_OK;
	};
	// level A: spec[0..4] = A(phi0), A(u0), A(v0), A(v0 st), A(mu) by the real
	// analysis
	if ((rc = sphswe_spat_op(ctx, SPHSWE_SPAT_MUL_ST, RP(2), nullptr, RP(3), 1.0, 1))) return rc;
	k_fill_mu<<<g256(nsp), 256, 0, st>>>(RP(4), ctx->d_ct, ctx->nlat, nsp);
	if ((rc = sphswe_sht_analysis(ctx, RP(0), RS(0), 5, SPHSWE_PREC_FP64))) return rc;
	if ((rc = to_cplx(0, SPEC(0), 5))) return rc;
	// phi0c, u0c, v0c
	SPHSWE_HIP(ctx, hipMemcpyAsync(BF(B_PHI), SPEC(0), sizeof(double2) * nc, hipMemcpyDeviceToDevice, st));
	SPHSWE_HIP(ctx, hipMemcpyAsync(BF(B_U0), SPEC(1), sizeof(double2) * nc, hipMemcpyDeviceToDevice, st));
	SPHSWE_HIP(ctx, hipMemcpyAsync(BF(B_V0), SPEC(2), sizeof(double2) * nc, hipMemcpyDeviceToDevice, st));
	// level B (spectral): spec[5..9] = im u0c, D' A(v0 st), im v0c, D' u0c, D' A(mu)
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_DIFF_LON, SPEC(1), SPEC(5), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_ONE_MINUS_MU2_DMU, SPEC(3), SPEC(6), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_DIFF_LON, SPEC(2), SPEC(7), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_ONE_MINUS_MU2_DMU, SPEC(1), SPEC(8), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_ONE_MINUS_MU2_DMU, SPEC(4), SPEC(9), 1, 1, 0, 0, 0, 0))) return rc;
	if (r->robert) {
		// SWERexiSPHRobert.hpp solve: div0 = ir robert_div(u0, v0), one division by
		// cos^2 phi of i m u0c + D' v0c (SPHOperatorsComplex.hpp:530-550); eta0 =
		// ir robert_vort = ir (i m v0c - D' u0c), spectral; the Coriolis multiplier
		// robert_grad_lat(mu) = D' A(mu), spectral
		if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_ONE_MINUS_MU2_DMU, SPEC(2), SPEC(6), 1, 1, 0, 0, 0, 0))) return rc;
		k_comb<<<g256(nc), 256, 0, st>>>(cs_ + 5 * nc, cs_ + 6 * nc, cs_ + 10 * nc, 1.0, 1.0, nc);
		SPHSWE_LAUNCH_CHECK(ctx);
		if ((rc = to_real(SPEC(10), 0, 1))) return rc;
		if ((rc = sphswe_sht_synthesis(ctx, RS(0), RP(0), 1, SPHSWE_PREC_FP64))) return rc;
		if ((rc = sphswe_spat_op(ctx, SPHSWE_SPAT_DIV_ST2, RP(0), nullptr, RP(0), 1.0, 1))) return rc;
		if ((rc = sphswe_sht_analysis(ctx, RP(0), RS(1), 1, SPHSWE_PREC_FP64))) return rc;
		if ((rc = to_cplx(1, SPEC(11), 1))) return rc;
		if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_IDENTITY, SPEC(11), BF(B_DIV), 1, ir, 0, 0, 0, 0))) return rc;
		k_comb<<<g256(nc), 256, 0, st>>>(cs_ + 7 * nc, cs_ + 8 * nc, cs_ + 15 * nc, -1.0, ir, nc);   // eta0 in spec[15]
		SPHSWE_LAUNCH_CHECK(ctx);
		if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_IDENTITY, SPEC(9), SPEC(5), 1, W * ir, 0, 0, 0, 0))) return rc;   // gms
	} else {
	// synthesis, metric divisions: grad_lon / grad_lat divide by sqrt(1-mu^2); the
	// complex div_lat further divides by cos(lat) (SPHOperatorsComplex.hpp:326-333)
	if ((rc = to_real(SPEC(5), 0, 5))) return rc;
	if ((rc = sphswe_sht_synthesis(ctx, RS(0), RP(0), 5, SPHSWE_PREC_FP64))) return rc;
	if ((rc = sphswe_spat_op(ctx, SPHSWE_SPAT_DIV_COSPHI, RP(0), nullptr, RP(0), 1.0, 5))) return rc;
	if ((rc = sphswe_spat_op(ctx, SPHSWE_SPAT_DIV_COSLAT, RP(1), nullptr, RP(1), 1.0, 1))) return rc;
	// level C: spec[10..14] = analyses
	if ((rc = sphswe_sht_analysis(ctx, RP(0), RS(0), 5, SPHSWE_PREC_FP64))) return rc;
	if ((rc = to_cplx(0, SPEC(10), 5))) return rc;
	// div0 = ir (A1 + A2), eta0 = ir (A3 - A4), gms = (W ir) A5
	k_comb<<<g256(nc), 256, 0, st>>>(cs_ + 10 * nc, cs_ + 11 * nc, B + B_DIV * nc, 1.0, ir, nc);
	k_comb<<<g256(nc), 256, 0, st>>>(cs_ + 12 * nc, cs_ + 13 * nc, cs_ + 15 * nc, -1.0, ir, nc);   // eta0 in spec[15]
	SPHSWE_LAUNCH_CHECK(ctx);
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_IDENTITY, SPEC(14), SPEC(5), 1, W * ir, 0, 0, 0, 0))) return rc;   // gms -> spec[5]
	}
	// mu / mu2 images
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_MU, SPEC(15), BF(B_MUETA), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_MU2, BF(B_PHI), BF(B_MU2PHI), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_MU2, BF(B_DIV), BF(B_MU2DIV), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_MU2, BF(B_MUETA), BF(B_MU2MUETA), 1, 1, 0, 0, 0, 0))) return rc;
	if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_MU2, BF(B_MU2PHI), BF(B_MU2MU2PHI), 1, 1, 0, 0, 0, 0))) return rc;
	if (r->coriolis) {
		// Coriolis forcing pieces: S of gms, u0c, mu2 u0c, mu v0c; products with g; analyses
		SPHSWE_HIP(ctx, hipMemcpyAsync(SPEC(6), SPEC(1), sizeof(double2) * nc, hipMemcpyDeviceToDevice, st));
		if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_MU2, SPEC(1), SPEC(7), 1, 1, 0, 0, 0, 0))) return rc;
		if ((rc = sphswe_spec_op_cplx(ctx, SPHSWE_OP_MU, SPEC(2), SPEC(8), 1, 1, 0, 0, 0, 0))) return rc;
		if ((rc = to_real(SPEC(5), 0, 4))) return rc;
		if ((rc = sphswe_sht_synthesis(ctx, RS(0), RP(0), 4, SPHSWE_PREC_FP64))) return rc;
		k_mul_bcast<<<g256(3 * nsp), 256, 0, st>>>(RP(0), RP(1), RP(4), nsp, 3);
		SPHSWE_LAUNCH_CHECK(ctx);
		if ((rc = sphswe_sht_analysis(ctx, RP(4), RS(4), 3, SPHSWE_PREC_FP64))) return rc;
		if ((rc = to_cplx(4, BF(B_W1), 3))) return rc;
	} else {
		SPHSWE_HIP(ctx, hipMemsetAsync(BF(B_W1), 0, sizeof(double2) * 3 * nc, st));
	}
	return SPHSWE_OK;
}

// sum over the poles of Re(beta_n x_n) for real spectra (phi0, u0, v0) in `in`;
// result as real spectra in `out` (may alias `in`)
static int rexi_pole_sum(sphswe_ctx *ctx, sphswe_rexi *r, const double2 *in, double2 *out)
{
	const long nsp = ctx->nspat, nc = ctx->ncplx;
	const int L = ctx->L, P = r->P;
	const double a = r->prm.earth_radius, g = r->prm.gravitation;
	hipStream_t st = ctx->stream;
	int rc;
	// real synthesis of (phi, u, v); complex physical copies
	if ((rc = sphswe_sht_synthesis(ctx, (const double *)in, (double *)r->d_cphys, 3, SPHSWE_PREC_FP64))) return rc;
	if ((rc = rexi_basis(ctx, r))) return rc;
	dim3 grid(r->ngroups, 2 * L + 1);
	// phi solves over (m, pole)
	auto phi = [&](auto *fac) {
		typedef std::remove_const_t<std::remove_pointer_t<decltype(fac)>> FT;
		PhiArgs<FT> pa;
		pa.alpha = r->d_alpha; pa.beta = r->d_beta; pa.moff = r->d_moff;
		pa.fac = fac; pa.piv = r->d_pphi; pa.basis = r->d_basis;
		pa.Y = r->d_Y; pa.X = r->d_X; pa.part = r->d_part;
		pa.L = L; pa.P = P; pa.ncplx = (int)nc;
		pa.two_omega = 2.0 * r->prm.coriolis_omega; pa.avg_geo = r->prm.h0 * g;
		pa.coriolis = r->coriolis; pa.rows = r->rows;
		pa.R = r->d_G; pa.K = r->d_kphi;
		if (!r->refine) k_rexi_phi<4, FT, false, 8><<<grid, 64, 0, st>>>(pa);
		else if (sweep_rows(true, L) == 8) k_rexi_phi<4, FT, true, 8><<<grid, 64, 0, st>>>(pa);
		else k_rexi_phi<4, FT, true, 4><<<grid, 64, 0, st>>>(pa);
	};
	if (r->coriolis) {
		// right-hand sides into Y (solved in place), or into the first half of G
		// (kept for the residual) when refining
		const long nc2 = ctx->ncplx;
		k_rexi_phi_rhs<<<dim3((unsigned)(((long)(L + 1) * P + 255) / 256), 2 * L + 1), 256, 0, st>>>(
			r->d_alpha, r->d_moff, r->d_basis, r->refine ? r->d_G : r->d_Y, L, P, r->refine ? 2L * P : (long)P, nc2,
			r->prm.h0 * g, 2.0 * r->prm.coriolis_omega);
		SPHSWE_LAUNCH_CHECK(ctx);
	}
	if (r->fac32) phi((const float2 *)r->d_fphi);
	else phi((const double2 *)r->d_fphi);
	SPHSWE_LAUNCH_CHECK(ctx);
	// gradients of phi, A(S(x) / sqrt(1-mu^2)) per m: the precomputed operator
	// applied to the 2P columns (i m phi | D' phi) of every pole at once (Robert
	// functions: the columns themselves are the gradients)
	if (!r->robert) {
		const int ncol = 4 * P;   // real columns
		const int cs = r->prm.precision == SPHSWE_PREC_FP64 ? gop_cs<double>() : gop_cs<float>();
		dim3 gg(r->ngtiles, (ncol + 16 * cs - 1) / (16 * cs));
		if (r->prm.precision == SPHSWE_PREC_FP64)
			k_gop_apply<double><<<gg, 64, 0, st>>>(r->d_gop, r->d_goff, r->d_gtiles, r->d_moff,
							       (const double *)r->d_X, (double *)r->d_G, L, ncol);
		else
			k_gop_apply<float><<<gg, 64, 0, st>>>(r->d_gop32, r->d_goff, r->d_gtiles, r->d_moff,
							      (const double *)r->d_X, (double *)r->d_G, L, ncol);
		SPHSWE_LAUNCH_CHECK(ctx);
	}
	// velocity solves
	auto vel = [&](auto *fac) {
		typedef std::remove_const_t<std::remove_pointer_t<decltype(fac)>> FT;
		VelArgs<FT> va;
		va.alpha = r->d_alpha; va.beta = r->d_beta; va.moff = r->d_moff;
		va.fac = fac; va.piv = r->d_pvel; va.basis = r->d_basis; va.G = r->robert ? r->d_X : r->d_G;
		va.Y = r->d_Y; va.part = r->d_part;
		va.L = L; va.P = P; va.ncplx = (int)nc;
		va.two_omega = 2.0 * r->prm.coriolis_omega; va.ir = 1.0 / a;
		va.coriolis = r->coriolis; va.rows = r->rows;
		va.R = r->d_X; va.X = r->d_G; va.Kv = r->d_kvel;
		if (!r->refine) k_rexi_vel<2, FT, false, 8><<<grid, 64, 0, st>>>(va);
		else if (sweep_rows(true, L) == 8) k_rexi_vel<2, FT, true, 8><<<grid, 64, 0, st>>>(va);
		else k_rexi_vel<2, FT, true, 4><<<grid, 64, 0, st>>>(va);
	};
	if (r->fac32) vel((const float2 *)r->d_fvel);
	else vel((const double2 *)r->d_fvel);
	SPHSWE_LAUNCH_CHECK(ctx);
	// pole sum, real part (spat_RealToSPHData + accumulation, AppSWE.hpp:519-521)
	k_rexi_sum_real<<<g256(ctx->nlm), 256, 0, st>>>(r->d_part, out, ctx->d_mofidx, r->d_moff, L, r->ngroups,
							 r->rows, ctx->nlm);
	SPHSWE_LAUNCH_CHECK(ctx);
	return SPHSWE_OK;
}

extern "C" int sphswe_rexi_apply(sphswe_ctx *ctx, sphswe_rexi *r, const double *d_in, double *d_out)
{
	if (!ctx || !r || !d_in || !d_out || r->ctx != ctx) return SPHSWE_ERR_ARG;
	return rexi_pole_sum(ctx, r, (const double2 *)d_in, (double2 *)d_out);
}

extern "C" int sphswe_rexi_step(sphswe_ctx *ctx, sphswe_rexi *r, double *d_state, int nsteps)
{
	if (!ctx || !r || !d_state || nsteps < 0 || r->ctx != ctx) return SPHSWE_ERR_ARG;
	const long nlm = ctx->nlm;
	const double g = r->prm.gravitation, a = r->prm.earth_radius;
	hipStream_t st = ctx->stream;
	double2 *X = (double2 *)d_state;
	int rc;
	for (int step = 0; step < nsteps; step++) {
		// phi = g h; (u, v) as they are (AppSWE.hpp:436-440)
		double2 *sp = r->d_in;
		if ((rc = sphswe_spec_axpby(ctx, nlm, g, 0, (double *)X, 0, 0, nullptr, (double *)sp))) return rc;
		SPHSWE_HIP(ctx, hipMemcpyAsync(sp + nlm, X + nlm, sizeof(double2) * 2 * nlm, hipMemcpyDeviceToDevice, st));
		if ((rc = rexi_pole_sum(ctx, r, sp, X))) return rc;
		// h = phi / g
		if ((rc = sphswe_spec_axpby(ctx, nlm, 1.0 / g, 0, (double *)X, 0, 0, nullptr, (double *)X))) return rc;
		// implicit viscosity (1 - dt nu D2) x_new = x (AppSWE.hpp:544-555)
		if (r->prm.viscosity != 0.0) {
			const double sc = r->prm.viscosity * r->prm.dt;
			if ((rc = sphswe_spec_op(ctx, SPHSWE_OP_HELMHOLTZ, (double *)X, (double *)X, 3, 1.0, 0.0, -sc, a))) return rc;
		}
	}
	return SPHSWE_OK;
}

// Lie splitting: the linear equations by REXI (with its implicit viscosity),
// then the nonlinear remainder (flux of h - h0, advection) by one RK4 step
extern "C" int sphswe_swe_split(sphswe_ctx *ctx, sphswe_rexi *r, const sphswe_swe_params *p, double *d_state,
				int nsteps)
{
	if (!ctx || !r || !p || !d_state || nsteps < 0 || r->ctx != ctx) return SPHSWE_ERR_ARG;
	const sphswe_rexi_params &q = r->prm;
	if (p->dt != q.dt)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "swe_split: RK4 dt %g differs from the REXI dt %g", p->dt, q.dt);
	if (p->earth_radius != q.earth_radius || p->gravitation != q.gravitation ||
	    p->coriolis_omega != q.coriolis_omega || p->h0 != q.h0)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "swe_split: a, g, Omega and h0 must equal the REXI object's");
	sphswe_swe_params nl = *p;
	nl.nonlinear = 2;
	nl.viscosity = 0.0;
	nl.robert_functions = 0;
	for (int step = 0; step < nsteps; step++) {
		if (int rc = sphswe_rexi_step(ctx, r, d_state, 1)) return rc;
		if (int rc = sphswe_swe_rk4(ctx, &nl, d_state, 1, 1)) return rc;
	}
	return SPHSWE_OK;
}
