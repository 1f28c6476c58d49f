/*
 * sphswe.h -- C ABI of libsphswe.so, the gfx950 (MI355X) spectral shallow-water
 * hot path.  Plain pointers and sizes only.  Every call returns an int status
 * (SPHSWE_OK = 0); the message of the last failure on a context is available
 * from sphswe_last_error().  Nothing here calls exit() or throws.
 *
 * All "d_" pointers are device pointers (from sphswe_malloc or any HIP
 * allocation on the context's device).  Calls are asynchronous on the
 * context's stream unless stated otherwise and are not thread-safe per
 * context, as the reference's SHTNS plan and solvers are not.
 *
 * Data layouts (identical to the reference):
 *   physical field   double[nphi][nlat], index i*nlat + j, j = 0 northernmost
 *                    ring (SHT_THETA_CONTIGUOUS), longitude 2*pi*i/nphi
 *   real spectrum    complex double[nlm], index m*(2L-m+1)/2 + l (m-major)
 *   complex spectrum complex double[(L+1)^2], index l*l + l + m
 * Complex numbers are interleaved (re, im) doubles.
 * Normalisation: orthonormal with Condon-Shortley phase (sht_orthonormal).
 */
#ifndef SPHSWE_H
#define SPHSWE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sphswe_ctx sphswe_ctx;
typedef struct sphswe_rexi sphswe_rexi;

enum {
	SPHSWE_OK = 0,
	SPHSWE_ERR_ARG = 1,       /* invalid argument or size */
	SPHSWE_ERR_HIP = 2,       /* HIP runtime error */
	SPHSWE_ERR_NOMEM = 3,     /* device allocation failed */
	SPHSWE_ERR_SINGULAR = 4,  /* banded factorisation met a zero pivot (zgbsv INFO > 0) */
	SPHSWE_ERR_NAN = 5,       /* NaN/Inf detected */
	SPHSWE_ERR_INTERNAL = 6   /* a structural check of the library failed */
};

/* precision of the Legendre contraction; the associated-Legendre recurrence is
 * always FP64.  The reduced modes run the contraction on the matrix cores with
 * FP32 accumulation:
 *   FP32    operands rounded to FP32;
 *   BF16    operands split into bf16 hi + lo, products hi*hi + hi*lo + lo*hi;
 *   BF16X3  operands split into bf16 hi + mid + lo (exact for FP32 values),
 *           the six products of total order <= 2;
 *   MIXED   as BF16, with the associated-Legendre values generated once per
 *           context by the FP64 recurrence and stored as bf16 hi/lo tables
 *           (two tables, 441 MB at T511, allocated on first use), contracted on
 *           the double-K bf16 matrix cores; real transforms only.  In the SWE
 *           stepper this mode also keeps the Fourier rows, FFTs and physical
 *           fields in FP32 (the state and the RK4 combinations stay FP64).
 * The Fourier stage of the transform entry points is FP64 in every mode.  The
 * real transforms take every mode, the complex ones FP64 .. BF16X3; the
 * pole-layout transforms inside the REXI step are FP64. */
enum {
	SPHSWE_PREC_FP64 = 0,
	SPHSWE_PREC_FP32 = 1,
	SPHSWE_PREC_BF16 = 2,
	SPHSWE_PREC_BF16X3 = 3,
	SPHSWE_PREC_MIXED = 4
};

/* ---- configuration (replaces shtns_create / shtns_set_grid{,_auto} and
 *      SPHConfig::setup / setupAutoPhysicalSpace, src/include/sph/SPHConfig.hpp:324-399) */

/* grid of shtns_set_grid_auto(..., nl_order, &nlat, &nphi) (SPHConfig.hpp:387-396) */
int sphswe_grid_auto(int lmax, int nl_order, int *nlat, int *nphi);

/* lmax must equal mmax (SPHConfig.hpp:171-176); nlat even, nphi >= 2*lmax+1 with
 * prime factors in {2,3,5,7}; device = HIP device ordinal */
int sphswe_create(int lmax, int mmax, int nlat, int nphi, int device, sphswe_ctx **ctx);
int sphswe_destroy(sphswe_ctx *ctx);

/* shtns->nlat, nphi, nlm, (L+1)^2 (SPHConfig.hpp:160-168) */
int sphswe_get_sizes(const sphswe_ctx *ctx, int *lmax, int *nlat, int *nphi, int *nlm, int *ncplx);
/* host copies of shtns->ct, shtns->st and the Gauss weights (SPHConfig.hpp:303-319) */
int sphswe_get_grid(const sphswe_ctx *ctx, double *ct, double *st, double *wg);
const char *sphswe_last_error(const sphswe_ctx *ctx);
/* wait for all work on the context stream */
int sphswe_sync(sphswe_ctx *ctx);
/* the context stream (hipStream_t) for callers that order their own work */
void *sphswe_stream(sphswe_ctx *ctx);

/* ---- device memory (replaces MemBlockAlloc::alloc/free, MemBlockAlloc.hpp:398-513;
 *      fftw_malloc of SHTNS buffers) */
int sphswe_malloc(sphswe_ctx *ctx, size_t bytes, void **d_ptr);
int sphswe_free(sphswe_ctx *ctx, void *d_ptr);
/* synchronous copies */
int sphswe_memcpy_h2d(sphswe_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int sphswe_memcpy_d2h(sphswe_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
/* asynchronous on the context stream */
int sphswe_memcpy_d2d(sphswe_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
int sphswe_fill_zero(sphswe_ctx *ctx, void *d_ptr, size_t bytes);

/* ---- transforms.  nfields fields stored back to back (stride nspat doubles,
 *      2*nlm doubles or 2*(L+1)^2 doubles).  spat_to_SH / SH_to_spat /
 *      spat_cplx_to_SH / SH_to_spat_cplx (call sites SPHData.hpp:201,221;
 *      SPHDataComplex.hpp:190,210).  Unlike SHTNS the input is not destroyed. */
int sphswe_sht_analysis(sphswe_ctx *ctx, const double *d_spat, double *d_spec, int nfields, int precision);
int sphswe_sht_synthesis(sphswe_ctx *ctx, const double *d_spec, double *d_spat, int nfields, int precision);
int sphswe_sht_analysis_cplx(sphswe_ctx *ctx, const double *d_spat, double *d_spec, int nfields, int precision);
int sphswe_sht_synthesis_cplx(sphswe_ctx *ctx, const double *d_spec, double *d_spat, int nfields, int precision);

/* sphswe_sht_analysis that allocates nothing, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The Fourier rows
 * and the Legendre analysis operands live in one fixed device array of the
 * library, sphswe_noalloc_scratch_bytes() long (96 MiB), held on every device
 * the library runs on until the process ends; a call whose nfields at this
 * truncation need more returns SPHSWE_ERR_ARG (FP64: up to 7 fields at T511,
 * none at T1023).  The array is shared by every context on a device and keeps
 * the last call's contents: calls that overlap on one device race.  FP64
 * results equal sphswe_sht_analysis bit for bit; FP32 never uses the stored
 * tables (sphswe_set_fp32_tables), so it may differ from sphswe_sht_analysis
 * in the order of its FP32 sums. */
int sphswe_sht_analysis_noalloc(sphswe_ctx *ctx, const double *d_spat, double *d_spec, int nfields, int precision);
size_t sphswe_noalloc_scratch_bytes(void);

/* the two stages of the real pair on their own (for measurement and fusion).
 * Fourier layout: complex double[nfields][L+1][nlat], entry (m, j) is
 * (2 pi / nphi) sum_i f(i,j) e^{-2 pi i m i / nphi}. */
int sphswe_fft_analysis(sphswe_ctx *ctx, const double *d_spat, double *d_four, int nfields);
int sphswe_fft_synthesis(sphswe_ctx *ctx, const double *d_four, double *d_spat, int nfields);
int sphswe_legendre_analysis(sphswe_ctx *ctx, const double *d_four, double *d_spec, int nfields, int precision);
int sphswe_legendre_synthesis(sphswe_ctx *ctx, const double *d_spec, double *d_four, int nfields, int precision);

/* ---- spectral operators on nfields fields: out = s * op(in), s complex
 *      (SPHOperators.hpp:30-498 and SPHOperatorsComplex.hpp:27-470; identities
 *      SPHIdentities.hpp:17-104).  in and out must not alias. */
enum {
	SPHSWE_OP_IDENTITY = 0,      /* s * x */
	SPHSWE_OP_DIFF_LON = 1,      /* diff_lon: i m x */
	SPHSWE_OP_ONE_MINUS_MU2_DMU = 2, /* spec_one_minus_mu_squared_diff_lat_mu */
	SPHSWE_OP_MU = 3,            /* mu */
	SPHSWE_OP_MU2 = 4,           /* mu2 (A/B/C band, not mu o mu) */
	SPHSWE_OP_LAPLACE = 5,       /* -n(n+1) x */
	SPHSWE_OP_HELMHOLTZ = 6      /* x / (a - (b/r^2) n(n+1)), spec_solve_helmholtz; s = a, (p0,p1)=(b, r) */
};
int sphswe_spec_op(sphswe_ctx *ctx, int op, const double *d_in, double *d_out, int nfields,
		   double s_re, double s_im, double p0, double p1);
int sphswe_spec_op_cplx(sphswe_ctx *ctx, int op, const double *d_in, double *d_out, int nfields,
			double s_re, double s_im, double p0, double p1, double p2);
/* out = a*x + b*y over n complex entries (a, b complex); y may be NULL */
int sphswe_spec_axpby(sphswe_ctx *ctx, size_t n, double a_re, double a_im, const double *d_x,
		      double b_re, double b_im, const double *d_y, double *d_out);
/* nfields real spectra of src_ctx's truncation copied into dst_ctx's: orders
 * and degrees beyond the smaller truncation are dropped or zero-filled
 * (SPHData::spec_copyToDifferentModes, SPHData.hpp:108-187, used with
 * SPHConfig::setupAdditionalModes, SPHConfig.hpp:403-417, for the extended
 * REXI modes of AppSWE.hpp:338-348,434-445,524-537).  Both contexts on one
 * device; the copy runs on dst_ctx's stream after the work queued on
 * src_ctx's stream. */
int sphswe_spec_copy_modes(sphswe_ctx *src_ctx, const double *d_src, sphswe_ctx *dst_ctx, double *d_dst,
			   int nfields);
/* spectral (0,0) += v*sqrt(4 pi) (SPHData.hpp:427-440) */
int sphswe_spec_add_scalar(sphswe_ctx *ctx, double *d_spec, double v_re, double v_im);

/* real <-> complex physical fields (SPHDataComplex::spat_fromSPHData and
 * spat_RealToSPHData / spat_ImagToSPHData, SPHDataComplex.hpp:129-175):
 * nfields fields of nspat points; part 0 = real part, 1 = imaginary part */
int sphswe_spat_real_to_cplx(sphswe_ctx *ctx, const double *d_real, double *d_cplx, int nfields);
int sphswe_spat_cplx_part(sphswe_ctx *ctx, const double *d_cplx, double *d_real, int part, int nfields);


/* ---- physical pointwise operators on real fields (spat_update_lambda*,
 *      operator* , SPHData.hpp:334-353,645-746) */
enum {
	SPHSWE_SPAT_MUL = 0,          /* out = a * b */
	SPHSWE_SPAT_DIV_COSPHI = 1,   /* out = a / sqrt(1 - mu^2) */
	SPHSWE_SPAT_MUL_ST = 2,       /* out = a * st */
	SPHSWE_SPAT_DIV_ST = 3,       /* out = a / st */
	SPHSWE_SPAT_MUL_MU = 4,       /* out = a * mu */
	SPHSWE_SPAT_SCALE = 5,        /* out = s * a */
	SPHSWE_SPAT_DIV_COSLAT = 6,   /* out = a / cos(lat[j]) */
	SPHSWE_SPAT_DIV_ST2 = 7       /* out = a / (st * st): the cos^2 phi of the Robert divergences
				       * (SPHOperators.hpp:129-172) */
};
int sphswe_spat_op(sphswe_ctx *ctx, int op, const double *d_a, const double *d_b, double *d_out,
		   double s, int nfields);
int sphswe_spat_op_cplx(sphswe_ctx *ctx, int op, const double *d_a, const double *d_b, double *d_out,
			double s, int nfields);
/* host flag: 1 if any entry is NaN/Inf (SPHData.hpp:560-569); synchronous */
int sphswe_any_nan(sphswe_ctx *ctx, const double *d_x, size_t n, int *h_flag);
/* min / max / max|x| of a physical field (SPHData.hpp:800-875); synchronous */
int sphswe_reduce_minmax(sphswe_ctx *ctx, const double *d_x, size_t n, double *h_min, double *h_max,
			 double *h_absmax);

/* ---- nonlinear SWE, RK4 (AppSWE.hpp:574-659 + TimesteppingRK.hpp:310-392) */
typedef struct {
	double earth_radius;     /* a */
	double gravitation;      /* g */
	double coriolis_omega;   /* Omega */
	double h0;               /* average height */
	double viscosity;        /* nu (explicit nu*laplace/a^2) */
	double dt;
	int nonlinear;           /* 1: nonlinear equations, 0: linear; 2: the nonlinear remainder
				  * (full minus linear tendency), for a split scheme whose linear
				  * part is integrated by REXI (no counterpart in the reference) */
	int precision;           /* SPHSWE_PREC_* for the Legendre stage */
	int robert_functions;    /* SimVars::use_robert_functions: with nonlinear = 0, u and v are
				  * Robert functions (u cos phi, v cos phi) and the tendency is
				  * AppSWE.hpp:606-620; ignored by the nonlinear equations, as there */
} sphswe_swe_params;

/* Advance nmembers independent states by nsteps RK4 steps.  d_state holds, per
 * member, the spectral fields h, u, v back to back (3*nlm complex).  The state
 * must be spectral (the caller analyses a physical initial condition first). */
int sphswe_swe_rk4(sphswe_ctx *ctx, const sphswe_swe_params *p, double *d_state, int nmembers, int nsteps);
/* The same steps with one RK4 step captured into a HIP graph and replayed
 * nsteps times (the small launches of low truncations and of ensembles stop
 * being launch-bound).  The graph is kept per context and rebuilt when p,
 * d_state or nmembers change; the first call of a new shape synchronises the
 * context stream.  Results are identical to sphswe_swe_rk4. */
int sphswe_swe_rk4_graph(sphswe_ctx *ctx, const sphswe_swe_params *p, double *d_state, int nmembers, int nsteps);

/* ---- linear SWE by REXI (AppSWE.hpp:336-565; SWERexiSPH.hpp:68-244;
 *      SPHSolverComplex.hpp:88-511; BandedMatrixSolver.hpp:235-434) */
typedef struct {
	double earth_radius;
	double gravitation;
	double coriolis_omega;   /* 0: no Coriolis, diagonal Helmholtz solves */
	double h0;               /* average height; Phi = g h0 */
	double viscosity;        /* implicit viscosity after each step (0: none) */
	double dt;               /* must be > 0 */
	double rexi_h;           /* REXI sampling width (default 0.2) */
	int rexi_M;              /* REXI approximation area (default 128) */
	int rexi_half;           /* 1: half the poles (default) */
	/* the fields below have no counterpart in the reference (0 = FP64 = its arithmetic) */
	int precision;           /* SPHSWE_PREC_FP64 or SPHSWE_PREC_FP32: the per-pole gradient
				  * operator (a GEMM on the FP64 or FP32 matrix cores) */
	int band_precision;      /* SPHSWE_PREC_FP64 or SPHSWE_PREC_FP32: storage of the LU factors
				  * (factorised in FP64 on the host; the solves run in FP64) */
	int refine;              /* FP32 factors: sweeps of FP64 iterative refinement (0 or 1) */
	int robert_functions;    /* 1: SWERexiSPHRobert (SWERexiSPHRobert.hpp): u, v are Robert
				  * functions; z4..z6robert, robert_div / robert_vort right-hand
				  * sides, spectral gradients (no gradient operator); refine = 0 */
} sphswe_rexi_params;

/* REXI::setup for (h, M, half) and SWERexiSPH::setup for every pole: pole
 * coefficients, band assembly z1..z8 and the pivoted LU factorisation of every
 * per-m system (factors cached on the device; SPHSWE_ERR_SINGULAR if a pivot
 * is zero, the reference's zgbsv INFO > 0). */
int sphswe_rexi_create(sphswe_ctx *ctx, const sphswe_rexi_params *p, sphswe_rexi **rexi);
int sphswe_rexi_destroy(sphswe_rexi *rexi);
/* number of poles and, if non-NULL, alpha / beta_re of REXI::setup (re, im pairs) */
int sphswe_rexi_poles(const sphswe_rexi *rexi, int *npoles, double *alpha, double *beta);
/* nsteps REXI steps of the state (h, u, v spectral, 3*nlm complex), in place */
int sphswe_rexi_step(sphswe_ctx *ctx, sphswe_rexi *rexi, double *d_state, int nsteps);

/* REXI::setup(h, M, 0, half) on the host (REXI.hpp:52-131), no context needed:
 * *npoles = 2(M+11)+1 (full) or M+12 (half); alpha, beta_re and, if non-NULL,
 * beta_im receive npoles (re, im) pairs each.  With all three arrays NULL only
 * the count is returned. */
int sphswe_rexi_coefficients(double h, int M, int half, int *npoles, double *alpha, double *beta_re,
			     double *beta_im);

/* As sphswe_rexi_create, for a caller-given pole set: alpha / beta as REXI::setup
 * returns them (before the division by dt that SWERexiSPH::setup applies,
 * SWERexiSPH.hpp:82-83).  One pole is one SWERexiSPH object. */
int sphswe_rexi_create_poles(sphswe_ctx *ctx, const sphswe_rexi_params *p, int npoles, const double *alpha,
			     const double *beta, sphswe_rexi **rexi);

/* The pole sum of one step: out = sum_n Re(beta_n x_n), x_n the solution of
 * SWERexiSPH::solve for pole n (SWERexiSPH.hpp:137-244) and the accumulation of
 * AppSWE.hpp:519-521, without the phi = g h conversion and the viscosity.
 * d_in: real spectra (phi0, u0, v0), 3*nlm complex; d_out: 3*nlm complex; they
 * may be the same buffer. */
int sphswe_rexi_apply(sphswe_ctx *ctx, sphswe_rexi *rexi, const double *d_in, double *d_out);

/* nsteps steps of a split scheme for the nonlinear SWE: per step, the REXI
 * step of the linear equations (sphswe_rexi_step: pole sum and implicit
 * viscosity), then one RK4 step of the nonlinear remainder (sphswe_swe_rk4
 * with nonlinear = 2 and no explicit viscosity; p->nonlinear and p->viscosity
 * are ignored).  p->dt must be the REXI object's dt and a, g, Omega and h0 must
 * agree (SPHSWE_ERR_ARG otherwise); p->precision selects the Legendre stage of
 * the RK4 part, the REXI object its own precision forms.  No counterpart in
 * the reference, whose REXI integrates the linear equations only
 * (AppSWE.hpp:185-192, 336-565); its parts replace SWERexiSPH::solve and
 * p_run_euler_timestep_update (AppSWE.hpp:574-659).  d_state: (h, u, v)
 * spectral, 3*nlm complex, in place. */
int sphswe_swe_split(sphswe_ctx *ctx, sphswe_rexi *rexi, const sphswe_swe_params *p, double *d_state, int nsteps);

/* ---- banded complex solver (SPHSolverComplex<T> with SPHMatrixComplex storage,
 *      SPHSolverComplex.hpp:50-511; BandedMatrixSolver::solve_diagBandedInverse_Carray
 *      + zgbsv, BandedMatrixSolver.hpp:235-434).  Rows in NCompact order, halo 1..4. */
typedef struct sphswe_band sphswe_band;
enum {
	SPHSWE_BAND_SCALAR = 0,           /* solver_component_scalar_phi */
	SPHSWE_BAND_MU = 1,               /* solver_component_mu_phi */
	SPHSWE_BAND_ONE_MINUS_MU2_DMU = 2, /* solver_component_one_minus_mu_mu_diff_mu_phi */
	SPHSWE_BAND_REXI_Z1 = 11,         /* solver_component_rexi_z1 .. z8 = 11 .. 18 */
	SPHSWE_BAND_REXI_Z4ROBERT = 19    /* solver_component_rexi_z4robert .. z6robert = 19 .. 21 */
};
int sphswe_band_create(sphswe_ctx *ctx, int halo, sphswe_band **band);
int sphswe_band_destroy(sphswe_band *band);
/* add scalar * component (radius r for z4..z8) to the matrix (host side) */
int sphswe_band_add(sphswe_band *band, int component, double s_re, double s_im, double r);
/* SPHSolverComplex::solve: x = A^-1 rhs for every m, complex spectra ((L+1)^2
 * complex, index l^2+l+m).  The matrix is factorised (partial pivoting, as
 * zgbsv) on the first solve after a change; SPHSWE_ERR_SINGULAR on a zero pivot. */
int sphswe_band_solve(sphswe_ctx *ctx, sphswe_band *band, const double *d_rhs, double *d_x);

/* ---- autotuner (no counterpart in the reference: launch shapes are this
 * library's).  sphswe_autotune times the variants available for the context's
 * sizes on synthetic data (about a second at T511; allocates and frees its own
 * buffers, synchronises the stream) and keeps the fastest:
 *   fft_nseq_r2c / _c2r  complex sequences per workgroup of the real FFTs
 *                        (0: built-in default);
 *   synth_mfma_min       FP64 real syntheses of at least this many fields run
 *                        on the FP64 matrix cores (0: never).
 * All variants are FP64 and differ only in the order of the sums.
 * sphswe_set_tuning pins the choices (e.g. from an earlier run). */
int sphswe_autotune(sphswe_ctx *ctx);
int sphswe_set_tuning(sphswe_ctx *ctx, int fft_nseq_r2c, int fft_nseq_c2r, int synth_mfma_min);
int sphswe_get_tuning(const sphswe_ctx *ctx, int *fft_nseq_r2c, int *fft_nseq_c2r, int *synth_mfma_min, int *tuned);
/* kernel shapes of the FP64 real Legendre transforms (no counterpart in the
 * reference; every choice computes the same sums in another order):
 *   synth_shape   vector-unit synthesis workgroups: 0 = 8 wavefronts x 3 rings
 *                 per lane, fields in groups of three (default); 1 = 4 x 2, fields
 *                 in groups of four and three;
 *   anal_variant  0 = up to four fields on v_mfma_f64_4x4x4f64, more fields in
 *                 groups of eight on v_mfma_f64_16x16x4f64 (default); 1 = always
 *                 the latter.  The two kernels agree to FP64 round-off. 
 * sphswe_autotune times both and keeps the faster. */
int sphswe_set_legendre_tuning(sphswe_ctx *ctx, int synth_shape, int anal_variant);
int sphswe_get_legendre_tuning(const sphswe_ctx *ctx, int *synth_shape, int *anal_variant);

/* FP32 real Legendre syntheses / analyses (SPHSWE_PREC_FP32) of at least
 * synth_min_fields / anal_min_fields fields run as batched GEMMs on the FP32
 * matrix cores from stored FP32 associated-Legendre tables (two tables of
 * nlm * nlat / 2 floats each, 201 MB apiece at T511, allocated on the first
 * such transform); fewer fields use the on-the-fly recurrence kernels.  0 (the
 * default) disables a direction.  Both forms round the same FP64 recurrence
 * values to FP32 and differ in the order of the FP32 sums only.
 * sphswe_autotune sets them (16 where the tables are faster at 16 fields).
 * No counterpart in the reference. */
int sphswe_set_fp32_tables(sphswe_ctx *ctx, int synth_min_fields, int anal_min_fields);
int sphswe_get_fp32_tables(const sphswe_ctx *ctx, int *synth_min_fields, int *anal_min_fields);

/* ---- timing on the context stream: ms between two sphswe_timer_mark calls */
int sphswe_timer_mark(sphswe_ctx *ctx, int slot);
int sphswe_timer_elapsed(sphswe_ctx *ctx, int slot0, int slot1, float *ms);

#ifdef __cplusplus
}
#endif
#endif
