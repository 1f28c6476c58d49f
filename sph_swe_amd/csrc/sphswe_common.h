// Internal declarations shared by the libsphswe translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <unordered_map>

#include "../../include/sphswe.h"

#define SPHSWE_MAX_TIMERS 16
#define SPHSWE_MAX_RADIX 16

struct FftPlan {
	int n;
	int npass;
	int radix[SPHSWE_MAX_RADIX];
};

struct sphswe_ctx {
	int L, nlat, nphi, nlm, ncplx, nspat, NJ;
	int device;
	hipStream_t stream;
	std::string err;

	std::vector<double> ct, st, wg;

	// device tables
	double *d_ct = nullptr;       // mu = cos(theta) [nlat]
	double *d_st = nullptr;       // sin(theta) [nlat]
	double *d_sq = nullptr;       // sqrt(1 - mu^2) [nlat]
	double *d_coslat = nullptr;   // cos(lat[j]) [nlat]
	double *d_wg = nullptr;       // Gauss weights [nlat]
	double2 *d_coef = nullptr;    // recurrence (a_lm, a_lm*b_lm) [nlm], index of the real spectrum
	double2 *d_seed = nullptr;    // (Pbar_s, Pbar_s+1) at each degree-segment start s [nseg_total * NJ]
	int *d_segoff = nullptr;      // first segment of each m [L+2]
	double2 *d_seed16 = nullptr;  // as d_seed at every 16-degree segment start (4x4x4 analysis)
	int *d_segoff16 = nullptr;
	int *d_unit_m = nullptr;      // m of each segment [nseg_total]
	int2 *d_ablk = nullptr;       // (m, first degree) of each 64-degree block of an order
	int nablk = 0;
	int seg = 16;                 // degrees per segment
	int nseg_total = 0;
	int max_lds = 65536;          // LDS bytes a workgroup may allocate
	// launch choices of the autotuner (0: the built-in default)
	int fft_nseq_r2c = 0;         // complex sequences per workgroup of the real FFTs
	int fft_nseq_c2r = 0;
	int synth_mfma_min = 0;       // FP64 real synthesis on the matrix cores from this many fields (0: never)
	int synth_shape = 0;          // FP64 vector-unit synthesis: 0 = 8 wavefronts x 3 rings per lane, 1 = 4 x 2
	int anal_variant = 0;         // FP64 real analysis: 0 = 4x4x4 kernel for <= 4 fields, 1 = 16x16x4 kernel always
	int tab_min_synth = 0;        // FP32 real syntheses from stored tables from this many fields (0: never)
	int tab_min_anal = 0;         // FP32 real analyses likewise
	// FP32 associated-Legendre tables (built on first use, sphswe_legendre_tab.hip)
	float *d_tab_s = nullptr;     // per (m, parity): [degree][ring]
	float *d_tab_a = nullptr;     // per (m, parity): [ring][degree]
	long *d_tab_off = nullptr;    // offset of (m, parity), index 2m + parity
	int2 *d_tab_tiles_s = nullptr, *d_tab_tiles_a = nullptr;
	int tab_ntiles_s = 0, tab_ntiles_a = 0;
	// bf16 hi/lo associated-Legendre tables in MFMA operand order (built on
	// first use, sphswe_legendre_bt.hip)
	uint4 *d_bt_s = nullptr;      // synthesis fragments (rows rings, contraction degrees)
	uint4 *d_bt_a = nullptr;      // analysis fragments (rows degrees, contraction rings)
	int *d_bt_soff = nullptr;     // first fragment of (m, parity), index 2m + parity
	int *d_bt_aoff = nullptr;
	int *d_bt_kboff = nullptr;    // first 32-degree block of (m, parity) among all blocks
	int2 *d_bt_kbl = nullptr;     // (2m + parity, block) of every block
	int bt_nkb = 0;               // number of blocks
	int2 *d_bt_mkb = nullptr;     // (m, 64-degree block) units of the operand split
	int bt_nmkb = 0;
	double2 *d_bt_dmu = nullptr;  // (1 - mu^2) d/dmu band per real spectral index: (on l - 1, on l + 1)
	int bt_tmax = 0;              // most degree tiles of an order pair (m, L - m)
	size_t bt_bytes = 0;
	int tuned = 0;
	int *d_mofidx = nullptr;      // m of real spectral index [nlm]
	int *d_lofcidx = nullptr;     // l of complex spectral index [(L+1)^2]
	double2 *d_tw = nullptr;      // e^{-2 pi i k / nphi} [nphi]
	float2 *d_tw32 = nullptr;     // the same rounded to FP32 (the FP32 transforms of the mixed form)
	FftPlan plan;

	// scratch (grown on demand, never shrunk)
	void *d_scratch = nullptr;
	size_t scratch_bytes = 0;
	void *d_escratch = nullptr;   // analysis operands (weighted Fourier rows, operand order)
	size_t escratch_bytes = 0;
	double *d_flag = nullptr;     // small reduction scratch (64 doubles)
	// while set (by the no-allocation entry points), the Legendre analyses take
	// their operand buffer from here, never grow d_escratch, and skip the
	// stored-table path (whose tables are allocated on first use)
	void *ext_escratch = nullptr;
	size_t ext_escratch_bytes = 0;
	// device address of the library's fixed scratch array on this context's
	// device (sphswe_noalloc.hip), looked up at create
	unsigned char *noalloc_scratch = nullptr;

	std::unordered_map<void *, size_t> allocs;

	// bumped whenever a buffer or kernel choice that a captured step may refer
	// to changes (scratch regrowth, FP32 tables freed, tuning changed); a
	// captured step records the value and is rebuilt when it differs
	unsigned long buf_gen = 0;
	// REXI and band objects created on this context: detached (ctx = null) by
	// sphswe_destroy, so that a later destroy of the object touches no freed context
	std::vector<sphswe_rexi *> rexis;
	std::vector<sphswe_band *> bands;

	hipEvent_t ev[SPHSWE_MAX_TIMERS];
};

// error helpers ---------------------------------------------------------------

int sphswe_set_err(sphswe_ctx *ctx, int code, const char *fmt, ...);

#define SPHSWE_HIP(ctx, call)                                                              \
	do {                                                                               \
		hipError_t _e = (call);                                                    \
		if (_e != hipSuccess)                                                      \
			return sphswe_set_err((ctx), SPHSWE_ERR_HIP, "%s: %s (%s:%d)", #call,  \
					      hipGetErrorString(_e), __FILE__, __LINE__);  \
	} while (0)

#define SPHSWE_LAUNCH_CHECK(ctx)                                                           \
	do {                                                                               \
		hipError_t _e = hipGetLastError();                                         \
		if (_e != hipSuccess)                                                      \
			return sphswe_set_err((ctx), SPHSWE_ERR_HIP, "kernel launch: %s (%s:%d)", \
					      hipGetErrorString(_e), __FILE__, __LINE__);  \
	} while (0)

// scratch of at least `bytes`, valid until the next call that asks for scratch
int sphswe_scratch(sphswe_ctx *ctx, size_t bytes, void **p);
int sphswe_escratch(sphswe_ctx *ctx, size_t bytes, void **p);

// internal launchers ----------------------------------------------------------

int sphswe_fft_r2c(sphswe_ctx *ctx, const double *spat, double2 *four, int nf);
int sphswe_fft_nseq_candidates(const sphswe_ctx *ctx, int *out, int max);
int sphswe_leg_synth_f64_mfma(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf);
int sphswe_fft_c2r(sphswe_ctx *ctx, const double2 *four, double *spat, int nf);
// FP32 real transforms between ring-major FP32 fields and FP32 Fourier rows
int sphswe_fft_r2c32(sphswe_ctx *ctx, const float *spat, float2 *four, int nf);
int sphswe_fft_c2r32(sphswe_ctx *ctx, const float2 *four, float *spat, int nf);
bool sphswe_fft32_available(const sphswe_ctx *ctx);
int sphswe_fft_c2c_fwd(sphswe_ctx *ctx, const double2 *spat, double2 *four, int nf);
int sphswe_fft_c2c_bwd(sphswe_ctx *ctx, const double2 *four, double2 *spat, int nf);
int sphswe_leg_anal(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec);
int sphswe_leg_synth(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec);
int sphswe_leg_anal_cplx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec);
int sphswe_leg_synth_cplx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec);
int sphswe_leg_anal_reserve(sphswe_ctx *ctx, int nf, int prec);
// bytes of analysis operands per group of eight real fields in a precision
size_t sphswe_leg_anal_group_bytes(const sphswe_ctx *ctx, int prec);
int sphswe_noalloc_init(sphswe_ctx *ctx);
void sphswe_swe_release(sphswe_ctx *ctx);
void sphswe_rexi_detach(sphswe_rexi *r);
void sphswe_band_detach(sphswe_band *b);
int sphswe_leg_anal_mx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec);
int sphswe_leg_synth_mx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec);
int sphswe_leg_anal_cplx_mx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int prec);
int sphswe_leg_synth_cplx_mx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int prec);
int sphswe_leg_tables(sphswe_ctx *ctx);
void sphswe_leg_tables_free(sphswe_ctx *ctx);
int sphswe_leg_synth_tab(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf);
int sphswe_leg_bt_tables(sphswe_ctx *ctx);
void sphswe_leg_bt_tables_free(sphswe_ctx *ctx);
int sphswe_leg_synth_bt(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf);
int sphswe_leg_anal_bt(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf);
// the pieces of the table form: pre-split column operands (synthesis SB from
// FP64 spectra, analysis G from Fourier rows) and the contractions on them
size_t sphswe_bt_sb_bytes(const sphswe_ctx *ctx, int nf);
size_t sphswe_bt_g_bytes(const sphswe_ctx *ctx, int nf);
int sphswe_bt_sprep(sphswe_ctx *ctx, const double2 *spec, void *SB, int nf);
// the SWE levels' operands straight from the spectra: 9 fields per member from
// (h, u, v) (the fields, i m and D' of each), 2 per member from the level-1
// analyses (i m of the first, D' of the second of each member's four)
int sphswe_bt_sprep_swe9(sphswe_ctx *ctx, const double2 *X, void *SB, int nmem);
int sphswe_bt_sprep_swe6(sphswe_ctx *ctx, const double2 *X, void *SB, int nmem);
int sphswe_bt_sprep_swe2(sphswe_ctx *ctx, const double2 *S4, void *SB, int nmem);
int sphswe_bt_gprep(sphswe_ctx *ctx, const double2 *four, void *G, int nf);
int sphswe_bt_gprep32(sphswe_ctx *ctx, const float2 *four, void *G, int nf);
int sphswe_bt_synth(sphswe_ctx *ctx, const void *SB, double2 *four, int nf);
int sphswe_bt_synth32(sphswe_ctx *ctx, const void *SB, float2 *four, int nf);
// the divergence level of the mixed SWE tendency: synthesis of (i m P0, D' P1)
// per member straight to the analysis operand of -(P0 / sq + P1 / (sq st)) scale
int sphswe_bt_synth_div(sphswe_ctx *ctx, const void *SB, void *G1, int nmem, float scale);
int sphswe_bt_anal(sphswe_ctx *ctx, const void *G, double2 *spec, int nf);
int sphswe_leg_anal_tab(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf);
int sphswe_leg_synth_pole_mx(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int P, const int *d_moff,
			     const double *oscale, int prec);
int sphswe_leg_anal_pole_mx(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int P, const int *d_moff,
			    int prec);
int sphswe_leg_synth_pole(sphswe_ctx *ctx, const double2 *spec, double2 *four, int nf, int P, const int *d_moff,
			  const double *oscale, int prec);
int sphswe_leg_anal_pole(sphswe_ctx *ctx, const double2 *four, double2 *spec, int nf, int P, const int *d_moff, int prec);

// device helpers --------------------------------------------------------------

__host__ __device__ inline double2 cmul(double2 a, double2 b)
{
	return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b)
{
	return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// real spectral index (SPHConfig.hpp:106-116)
__host__ __device__ inline int lm_index(int L, int l, int m) { return (m * (2 * L - m + 1)) / 2 + l; }

// SPHIdentities (SPHIdentities.hpp:17-104) -------------------------------------
__host__ __device__ inline double idR(double k, double m)
{
	double n = k + 1;
	if (n < 0) return 0;
	if (n < fabs(m)) return 0;
	if (n * n - m * m < 0) return 0;
	return sqrt((n * n - m * m) / (4.0 * n * n - 1.0));
}
__host__ __device__ inline double idS(double k, double m)
{
	double n = k - 1;
	if (n < 0) return 0;
	if (n < fabs(m)) return 0;
	if (n * n - m * m < 0) return 0;
	return sqrt(((n + 1.0) * (n + 1.0) - m * m) / ((2.0 * n + 1.0) * (2.0 * n + 3.0)));
}
__host__ __device__ inline double idA(double k, double m) { double n = k + 2.0; return idR(n - 1, m) * idR(n - 2, m); }
__host__ __device__ inline double idB(double n, double m) { return idR(n - 1, m) * idS(n, m) + idS(n + 1, m) * idR(n, m); }
__host__ __device__ inline double idC(double k, double m) { double n = k - 2.0; return idS(n + 1, m) * idS(n + 2, m); }
__host__ __device__ inline double idD(double k, double m)
{
	double n = k + 1;
	return (2.0 * n + 1.0) * sqrt((n * n - m * m) / (4.0 * n * n - 1.0));
}
__host__ __device__ inline double idE(double n, double) { return -n; }
__host__ __device__ inline double idG(double k, double m) { double n = k + 1; return idD(n - 1, m) + idE(n, m) * idR(n - 1, m); }
__host__ __device__ inline double idH(double k, double m) { double n = k - 1; return idE(n, m) * idS(n + 1, m); }

// XCD-aware block remap: hardware places block b on XCD (b mod 8); consecutive
// logical blocks are gathered on one XCD so that neighbours share its L2.
__device__ inline int xcd_remap(int b, int nblocks_padded)
{
	int per = nblocks_padded >> 3;
	return (b & 7) * per + (b >> 3);
}
