/*
 * sphswe_oracle.c -- CPU FP64 restatement of the sph_swe spectral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is loaded by tests/,
 * by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py, and only
 * there, as the checker.  The product (sph_swe_amd, libsphswe.so) never links,
 * loads or calls it.
 *
 * It restates, in plain C and double precision, what the reference computes:
 *   - Gauss grid + orthonormal Condon-Shortley P_l^m (replaces SHTNS' tables,
 *     reference call sites src/include/sph/SPHConfig.hpp:324-399)
 *   - real / complex SHT pair (SHTNS spat_to_SH, SH_to_spat, spat_cplx_to_SH,
 *     SH_to_spat_cplx; call sites SPHData.hpp:190-227, SPHDataComplex.hpp:179-216)
 *   - the lazy dual-representation field semantics (SPHData.hpp:190-443,
 *     SPHDataComplex.hpp:95-545) and the operators (SPHOperators.hpp:30-529,
 *     SPHOperatorsComplex.hpp:27-550, SPHIdentities.hpp:17-104)
 *   - the nonlinear/linear SWE tendency (AppSWE.hpp:574-659) and RK4
 *     (TimesteppingRK.hpp:310-392)
 *   - REXI coefficients (REXI.hpp:52-131, GaussianApproximation.hpp:51-81,
 *     ExponentialApproximation.hpp:59-123)
 *   - band assembly z1..z8 (SPHSolverComplex.hpp:88-420, SPHMatrixComplex.hpp:79-206)
 *     and a banded LU with partial pivoting (the algorithm of LAPACK zgbsv,
 *     BandedMatrixSolver.hpp:235-434)
 *   - SWERexiSPH::solve (SWERexiSPH.hpp:137-244) and the REXI time step
 *     (AppSWE.hpp:415-565)
 *   - the Robert-function variants: the linear tendency (AppSWE.hpp:606-620),
 *     the robert_* operators (SPHOperators.hpp:124-208, SPHOperatorsComplex.hpp:
 *     341-550), the z4..z6robert components (SPHSolverComplex.hpp:241-370) and
 *     SWERexiSPHRobert::solve (SWERexiSPHRobert.hpp)
 *
 * Conventions (SHTNS sht_orthonormal, Condon-Shortley phase on):
 *   real field:    f = sum_l q_l0 Y_l0 + 2 Re sum_{m>0} q_lm Y_lm ; Im q_l0 ignored
 *   complex field: f = sum_l sum_{m=-l..l} q_lm Y_lm ; Y_l,-m = (-1)^m conj(Y_lm)
 *   Y_lm = Pbar_lm(mu) e^{i m phi},  2 pi int Pbar_lm^2 dmu = 1
 *   physical layout [i*nlat + j], j = 0 northernmost, phi_i = 2 pi i / nphi
 *   real spectrum index  m(2L-m+1)/2 + l  (SPHConfig.hpp:106-116)
 *   complex spectrum index  l*l + l + m   (SPHConfig.hpp:120-130)
 *
 * Parity pins: see tests/test_oracle_*.py (Gauss nodes vs numpy leggauss,
 * Pbar vs scipy sph_legendre_p and closed forms, FFT vs numpy.fft, band LU vs
 * scipy zgbsv, REXI coefficients vs the reference's own headers compiled into
 * oracle/_ref, operators vs analytic derivatives, SWE linear mode vs closed
 * form frequency; the Robert tendency vs the cos-weighted standard one, the
 * Robert REXI step to first order in dt vs that tendency).
 * Parity unpinned: component 21 (z6robert).  No closed form of it is known
 * here, a sign change of it is invisible to the first-order check, and the
 * reference's solver headers cannot be built without SHTNS.
 */
#define _GNU_SOURCE
#ifdef _OPENMP
#include <omp.h>
#endif
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

typedef double complex cplx;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* configuration                                                            */
/* ------------------------------------------------------------------------ */

typedef struct {
	int L, nlat, nphi, nlm, ncplx, nspat;
	double *ct, *st, *w, *lat;
	cplx *tw;          /* e^{-2 pi i k / nphi} */
	double *ptab;      /* optional cache of Pbar [m][j][l-m] (small configurations) */
	size_t *pofs;      /* offset of (m, j=0) in ptab */
	long n_anal, n_synth, n_canal, n_csynth;   /* transform counters */
} ocfg;

static int small_prime_factor(int n)
{
	for (int p = 2; p * p <= n; p++)
		if (n % p == 0) return p;
	return n;
}

static int max_prime_factor(int n)
{
	int mp = 1;
	for (int p = 2; n > 1; p++)
		while (n % p == 0) { mp = p; n /= p; }
	return mp;
}

/* Grid rule of shtns_set_grid_auto with nl_order = 2 (SPHConfig.hpp:387-396):
 * nlat = (3L)/2 + 1 rounded up to even, nphi = smallest n >= 3L+1 whose prime
 * factors are all <= 7. */
void orc_auto_grid(int L, int *nlat, int *nphi)
{
	int nl = (3 * L) / 2 + 1;
	nl += (nl & 1);
	int np = 3 * L + 1;
	while (max_prime_factor(np) > 7) np++;
	*nlat = nl;
	*nphi = np;
}

/* Gauss-Legendre nodes by Newton iteration on P_n (textbook), ordered with
 * ct[0] closest to +1 (north pole first). */
static void gauss_nodes(int n, double *x, double *w)
{
	for (int i = 0; i < (n + 1) / 2; i++) {
		double z = cos(M_PI * (i + 0.75) / (n + 0.5));
		double pp = 0;
		for (int it = 0; it < 100; it++) {
			double p1 = 1.0, p2 = 0.0;
			for (int k = 1; k <= n; k++) {
				double p3 = p2; p2 = p1;
				p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k;
			}
			pp = n * (z * p1 - p2) / (z * z - 1.0);
			double z1 = z;
			z = z1 - p1 / pp;
			if (fabs(z - z1) < 1e-16) {
				/* one more for full accuracy */
				p1 = 1.0; p2 = 0.0;
				for (int k = 1; k <= n; k++) {
					double p3 = p2; p2 = p1;
					p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k;
				}
				pp = n * (z * p1 - p2) / (z * z - 1.0);
				break;
			}
		}
		x[i] = z;
		x[n - 1 - i] = -z;
		w[i] = 2.0 / (((1.0 - z) * (1.0 + z)) * pp * pp);
		w[n - 1 - i] = w[i];
	}
	if (n & 1) x[n / 2] = 0.0;
}

void orc_plm_column(int L, int m, double ct, double st, double *P);

ocfg *orc_create(int L, int nlat, int nphi)
{
	ocfg *c = (ocfg *)calloc(1, sizeof(ocfg));
	c->L = L; c->nlat = nlat; c->nphi = nphi;
	c->nlm = (L + 1) * (L + 2) / 2;
	c->ncplx = (L + 1) * (L + 1);
	c->nspat = nlat * nphi;
	c->ct = (double *)malloc(sizeof(double) * nlat);
	c->st = (double *)malloc(sizeof(double) * nlat);
	c->w = (double *)malloc(sizeof(double) * nlat);
	c->lat = (double *)malloc(sizeof(double) * nlat);
	gauss_nodes(nlat, c->ct, c->w);
	for (int j = 0; j < nlat; j++) {
		c->st[j] = sqrt((1.0 - c->ct[j]) * (1.0 + c->ct[j]));
		c->lat[j] = M_PI * 0.5 - acos(c->ct[j]);      /* SPHConfig.hpp:309-310 */
	}
	c->tw = (cplx *)malloc(sizeof(cplx) * nphi);
	for (int k = 0; k < nphi; k++)
		c->tw[k] = cexp(-2.0 * M_PI * I * (double)k / (double)nphi);
	c->ptab = NULL;
	c->pofs = NULL;
	if ((size_t)c->nlm * nlat <= ((size_t)1 << 25)) {
		double *P = (double *)malloc(sizeof(double) * (L + 1));
		c->ptab = (double *)malloc(sizeof(double) * (size_t)c->nlm * nlat);
		c->pofs = (size_t *)malloc(sizeof(size_t) * (L + 1));
		size_t o = 0;
		for (int m = 0; m <= L; m++) {
			c->pofs[m] = o;
			for (int j = 0; j < nlat; j++) {
				orc_plm_column(L, m, c->ct[j], c->st[j], P);
				memcpy(c->ptab + o, P + m, sizeof(double) * (L + 1 - m));
				o += L + 1 - m;
			}
		}
		free(P);
	}
	return c;
}

static void fdata_release(ocfg *c);

void orc_destroy(ocfg *c)
{
	if (!c) return;
	fdata_release(c);
	free(c->ct); free(c->st); free(c->w); free(c->lat); free(c->tw);
	free(c->ptab); free(c->pofs);
	free(c);
}

void orc_get_grid(ocfg *c, double *ct, double *st, double *w, double *lat)
{
	memcpy(ct, c->ct, sizeof(double) * c->nlat);
	memcpy(st, c->st, sizeof(double) * c->nlat);
	memcpy(w, c->w, sizeof(double) * c->nlat);
	memcpy(lat, c->lat, sizeof(double) * c->nlat);
}

void orc_get_counters(ocfg *c, long *out4)
{
	out4[0] = c->n_anal; out4[1] = c->n_synth; out4[2] = c->n_canal; out4[3] = c->n_csynth;
}

/* ------------------------------------------------------------------------ */
/* associated Legendre functions                                            */
/* ------------------------------------------------------------------------ */

/* Pbar_l^m(ct) for l = m..L into P[l] (P[0..m-1] untouched).  Scaled
 * recurrence: value = x * 2^E, so tiny starting values near the poles do not
 * underflow before the recurrence grows them. */
void orc_plm_column(int L, int m, double ct, double st, double *P)
{
	double lg = 0.5 * log((2.0 * m + 1.0) / (4.0 * M_PI));
	for (int k = 1; k <= m; k++)
		lg += 0.5 * log((2.0 * k - 1.0) / (2.0 * k));
	if (m > 0) lg += m * log(st);
	int E = (int)floor(lg / M_LN2);
	double x0 = exp(lg - E * M_LN2);
	if (m & 1) x0 = -x0;      /* Condon-Shortley */
	P[m] = ldexp(x0, E);
	if (m == L) return;
	double x1 = sqrt(2.0 * m + 3.0) * ct * x0;
	P[m + 1] = ldexp(x1, E);
	for (int l = m + 2; l <= L; l++) {
		double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
		double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
		double x2 = a * (ct * x1 - b * x0);
		x0 = x1; x1 = x2;
		if (fabs(x1) > 0x1p200) {
			x0 = ldexp(x0, -200); x1 = ldexp(x1, -200); E += 200;
		}
		P[l] = ldexp(x1, E);
	}
}

/* column Pbar_l^m(mu_j), l = m..L, returned as a pointer p with p[l] valid
 * (cached table when present, else computed into buf) */
static const double *plm_col(ocfg *c, int m, int j, double *buf)
{
	if (c->ptab) return c->ptab + c->pofs[m] + (size_t)j * (c->L + 1 - m) - m;
	orc_plm_column(c->L, m, c->ct[j], c->st[j], buf);
	return buf;
}

/* full table [m][j][l - m] for tests (size sum_m (L-m+1) * nlat) */
void orc_plm_table(ocfg *c, double *out)
{
	double *P = (double *)malloc(sizeof(double) * (c->L + 1));
	size_t o = 0;
	for (int m = 0; m <= c->L; m++)
		for (int j = 0; j < c->nlat; j++) {
			orc_plm_column(c->L, m, c->ct[j], c->st[j], P);
			for (int l = m; l <= c->L; l++) out[o++] = P[l];
		}
	free(P);
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed radix (smallest prime factor first), table twiddles  */
/* ------------------------------------------------------------------------ */

/* y[k] = sum_t x[t*xs] e^{sign 2 pi i k t / n}, n | N, twiddle table tw of size N
 * holds e^{-2 pi i k / N}. */
static void fft_rec(int N, const cplx *tw, int n, const cplx *x, int xs, cplx *y, int sign, cplx *tmp)
{
	if (n == 1) { y[0] = x[0]; return; }
	int p = small_prime_factor(n);
	int q = n / p;
	for (int r = 0; r < p; r++)
		fft_rec(N, tw, q, x + (size_t)r * xs, xs * p, y + (size_t)r * q, sign, tmp + n);
	int step = N / n;
	if (p == 2) {
		for (int k = 0; k < q; k++) {
			cplx t = tw[(size_t)k * step];
			if (sign > 0) t = conj(t);
			cplx a = y[k], b = y[q + k] * t;
			tmp[k] = a + b;
			tmp[k + q] = a - b;
		}
	} else {
		for (int s = 0; s < p; s++)
			for (int k = 0; k < q; k++) {
				int kk = k + q * s;
				long t1 = ((long)kk * step) % N;   /* twiddle index of r = 1 */
				long e = 0;
				cplx acc = 0;
				for (int r = 0; r < p; r++) {
					cplx t = tw[e];
					if (sign > 0) t = conj(t);
					acc += y[r * q + k] * t;
					e += t1;
					if (e >= N) e -= N;
				}
				tmp[kk] = acc;
			}
	}
	memcpy(y, tmp, sizeof(cplx) * n);
}

/* exported for tests: forward (sign=-1) or backward (sign=+1), unnormalised */
void orc_fft(ocfg *c, const double *in_ri, double *out_ri, int sign)
{
	int n = c->nphi;
	cplx *x = (cplx *)malloc(sizeof(cplx) * n);
	cplx *y = (cplx *)malloc(sizeof(cplx) * n);
	cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4 * n);
	for (int i = 0; i < n; i++) x[i] = in_ri[2 * i] + I * in_ri[2 * i + 1];
	fft_rec(n, c->tw, n, x, 1, y, sign, tmp);
	for (int i = 0; i < n; i++) { out_ri[2 * i] = creal(y[i]); out_ri[2 * i + 1] = cimag(y[i]); }
	free(x); free(y); free(tmp);
}

/* ------------------------------------------------------------------------ */
/* SHT                                                                      */
/* ------------------------------------------------------------------------ */

static inline int lm_real(const ocfg *c, int l, int m) { return (m * (2 * c->L - m + 1)) / 2 + l; }
static inline int lm_cplx(int l, int m) { return l * l + l + m; }

/* threads of the OpenMP loops below (the rings of the FFT stage, the orders of
 * the Legendre stage; every output is written by one thread in the serial
 * order, so results do not depend on the thread count).  The reference runs
 * SHTNS with OpenMP over all cores (SPHConfig.hpp:332). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
	omp_set_num_threads(n > 0 ? n : 1);
#else
	(void)n;
#endif
}

/* real analysis: spat[nphi*nlat] -> spec[nlm] */
void orc_sht_analysis(ocfg *c, const double *spat, cplx *spec)
{
	int L = c->L, nlat = c->nlat, nphi = c->nphi;
	cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)(L + 1) * nlat);
	double norm = 2.0 * M_PI / nphi;
#pragma omp parallel
	{
		cplx *x = (cplx *)malloc(sizeof(cplx) * nphi);
		cplx *y = (cplx *)malloc(sizeof(cplx) * nphi);
		cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4 * nphi);
		double *P = (double *)malloc(sizeof(double) * (L + 1));
#pragma omp for schedule(static)
		for (int j = 0; j < nlat; j++) {
			for (int i = 0; i < nphi; i++) x[i] = spat[(size_t)i * nlat + j];
			fft_rec(nphi, c->tw, nphi, x, 1, y, -1, tmp);
			for (int m = 0; m <= L; m++) F[(size_t)m * nlat + j] = y[m] * norm;
			F[j] = creal(F[j]);
		}
#pragma omp for schedule(dynamic, 4)
		for (int m = 0; m <= L; m++) {
			cplx *q = spec + lm_real(c, m, m) - m;   /* q[l] */
			for (int l = m; l <= L; l++) q[l] = 0;
			for (int j = 0; j < nlat; j++) {
				const double *Pc = plm_col(c, m, j, P);
				cplx f = F[(size_t)m * nlat + j] * c->w[j];
				for (int l = m; l <= L; l++) q[l] += Pc[l] * f;
			}
		}
		free(x); free(y); free(tmp); free(P);
	}
	c->n_anal++;
	free(F);
}

/* real synthesis: spec[nlm] -> spat[nphi*nlat] */
void orc_sht_synthesis(ocfg *c, const cplx *spec, double *spat)
{
	int L = c->L, nlat = c->nlat, nphi = c->nphi;
	cplx *G = (cplx *)malloc(sizeof(cplx) * (size_t)(L + 1) * nlat);
#pragma omp parallel
	{
		cplx *x = (cplx *)malloc(sizeof(cplx) * nphi);
		cplx *y = (cplx *)malloc(sizeof(cplx) * nphi);
		cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4 * nphi);
		double *P = (double *)malloc(sizeof(double) * (L + 1));
#pragma omp for schedule(dynamic, 4)
		for (int m = 0; m <= L; m++) {
			const cplx *q = spec + lm_real(c, m, m) - m;
			for (int j = 0; j < nlat; j++) {
				const double *Pc = plm_col(c, m, j, P);
				cplx g = 0;
				if (m == 0) for (int l = 0; l <= L; l++) g += creal(q[l]) * Pc[l];
				else for (int l = m; l <= L; l++) g += q[l] * Pc[l];
				G[(size_t)m * nlat + j] = g;
			}
		}
#pragma omp for schedule(static)
		for (int j = 0; j < nlat; j++) {
			for (int i = 0; i < nphi; i++) x[i] = 0;
			x[0] = creal(G[j]);
			for (int m = 1; m <= L; m++) {
				cplx g = G[(size_t)m * nlat + j];
				x[m] += g;
				x[nphi - m] += conj(g);
			}
			fft_rec(nphi, c->tw, nphi, x, 1, y, +1, tmp);
			for (int i = 0; i < nphi; i++) spat[(size_t)i * nlat + j] = creal(y[i]);
		}
		free(x); free(y); free(tmp); free(P);
	}
	c->n_synth++;
	free(G);
}

/* complex analysis: spat cplx[nphi*nlat] -> spec[(L+1)^2] */
void orc_sht_analysis_cplx(ocfg *c, const cplx *spat, cplx *spec)
{
	int L = c->L, nlat = c->nlat, nphi = c->nphi;
	int nm = 2 * L + 1;
	cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)nm * nlat);
	cplx *x = (cplx *)malloc(sizeof(cplx) * nphi);
	cplx *y = (cplx *)malloc(sizeof(cplx) * nphi);
	cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4 * nphi);
	double *P = (double *)malloc(sizeof(double) * (L + 1));
	double norm = 2.0 * M_PI / nphi;
	for (int j = 0; j < nlat; j++) {
		for (int i = 0; i < nphi; i++) x[i] = spat[(size_t)i * nlat + j];
		fft_rec(nphi, c->tw, nphi, x, 1, y, -1, tmp);
		for (int m = -L; m <= L; m++)
			F[(size_t)(m + L) * nlat + j] = y[(m + nphi) % nphi] * norm;
	}
	for (int i = 0; i < c->ncplx; i++) spec[i] = 0;
	for (int am = 0; am <= L; am++) {
		for (int j = 0; j < nlat; j++) {
			const double *Pc = plm_col(c, am, j, P);
			cplx fp = F[(size_t)(am + L) * nlat + j] * c->w[j];
			cplx fn = F[(size_t)(-am + L) * nlat + j] * c->w[j] * ((am & 1) ? -1.0 : 1.0);
			for (int l = am; l <= L; l++) {
				spec[lm_cplx(l, am)] += Pc[l] * fp;
				if (am > 0) spec[lm_cplx(l, -am)] += Pc[l] * fn;
			}
		}
	}
	c->n_canal++;
	free(F); free(x); free(y); free(tmp); free(P);
}

void orc_sht_synthesis_cplx(ocfg *c, const cplx *spec, cplx *spat)
{
	int L = c->L, nlat = c->nlat, nphi = c->nphi;
	int nm = 2 * L + 1;
	cplx *G = (cplx *)malloc(sizeof(cplx) * (size_t)nm * nlat);
	cplx *x = (cplx *)malloc(sizeof(cplx) * nphi);
	cplx *y = (cplx *)malloc(sizeof(cplx) * nphi);
	cplx *tmp = (cplx *)malloc(sizeof(cplx) * 4 * nphi);
	double *P = (double *)malloc(sizeof(double) * (L + 1));
	for (int am = 0; am <= L; am++) {
		for (int j = 0; j < nlat; j++) {
			const double *Pc = plm_col(c, am, j, P);
			cplx gp = 0, gn = 0;
			for (int l = am; l <= L; l++) {
				gp += spec[lm_cplx(l, am)] * Pc[l];
				if (am > 0) gn += spec[lm_cplx(l, -am)] * Pc[l];
			}
			G[(size_t)(am + L) * nlat + j] = gp;
			if (am > 0) G[(size_t)(-am + L) * nlat + j] = gn * ((am & 1) ? -1.0 : 1.0);
		}
	}
	for (int j = 0; j < nlat; j++) {
		for (int i = 0; i < nphi; i++) x[i] = 0;
		for (int m = -L; m <= L; m++) x[(m + nphi) % nphi] += G[(size_t)(m + L) * nlat + j];
		fft_rec(nphi, c->tw, nphi, x, 1, y, +1, tmp);
		for (int i = 0; i < nphi; i++) spat[(size_t)i * nlat + j] = y[i];
	}
	c->n_csynth++;
	free(G); free(x); free(y); free(tmp); free(P);
}

/* ------------------------------------------------------------------------ */
/* SPHIdentities (SPHIdentities.hpp:17-104)                                 */
/* ------------------------------------------------------------------------ */

static double idR(double k, double m)
{
	double n = k + 1;
	if (n < 0) return 0;
	if (n < fabs(m)) return 0;
	if (n * n - m * m < 0) return 0;
	return sqrt((n * n - m * m) / (4.0 * n * n - 1.0));
}
static double idS(double k, double m)
{
	double n = k - 1;
	if (n < 0) return 0;
	if (n < fabs(m)) return 0;
	if (n * n - m * m < 0) return 0;
	return sqrt(((n + 1.0) * (n + 1.0) - m * m) / ((2.0 * n + 1.0) * (2.0 * n + 3.0)));
}
static double idA(double k, double m) { double n = k + 2.0; return idR(n - 1, m) * idR(n - 2, m); }
static double idB(double n, double m) { return idR(n - 1, m) * idS(n, m) + idS(n + 1, m) * idR(n, m); }
static double idC(double k, double m) { double n = k - 2.0; return idS(n + 1, m) * idS(n + 2, m); }
static double idD(double k, double m) { double n = k + 1; return (2.0 * n + 1.0) * sqrt((n * n - m * m) / (4.0 * n * n - 1.0)); }
static double idE(double n, double m) { (void)m; return -n; }
static double idG(double k, double m) { double n = k + 1; return idD(n - 1, m) + idE(n, m) * idR(n - 1, m); }
static double idH(double k, double m) { double n = k - 1; return idE(n, m) * idS(n + 1, m); }

void orc_identities(double n, double m, double *out9)
{
	out9[0] = idR(n, m); out9[1] = idS(n, m); out9[2] = idA(n, m);
	out9[3] = idB(n, m); out9[4] = idC(n, m); out9[5] = idD(n, m); out9[6] = idE(n, m);
	out9[7] = idG(n, m); out9[8] = idH(n, m);
}

/* ------------------------------------------------------------------------ */
/* lazy fields (SPHData / SPHDataComplex semantics)                         */
/* ------------------------------------------------------------------------ */

typedef struct {
	ocfg *c;
	cplx *spec;       /* nlm (real) or ncplx (complex) */
	double *spat;     /* real: nspat */
	cplx *cspat;      /* complex: nspat */
	int spec_valid, spat_valid, is_cplx;
} ofield;

/* simple arena of temporaries; every op result comes from here */
#define ARENA_MAX 4096
typedef struct { ofield *f[ARENA_MAX]; int n; } oarena;
static oarena g_arena;

static ofield *field_alloc(ocfg *c, int is_cplx, int temp)
{
	ofield *f = (ofield *)calloc(1, sizeof(ofield));
	f->c = c; f->is_cplx = is_cplx;
	f->spec = (cplx *)calloc(is_cplx ? c->ncplx : c->nlm, sizeof(cplx));
	if (is_cplx) f->cspat = (cplx *)calloc(c->nspat, sizeof(cplx));
	else f->spat = (double *)calloc(c->nspat, sizeof(double));
	if (temp) {
		if (g_arena.n >= ARENA_MAX) { fprintf(stderr, "oracle arena overflow\n"); exit(1); }
		g_arena.f[g_arena.n++] = f;
	}
	return f;
}

static void field_free(ofield *f)
{
	if (!f) return;
	free(f->spec); free(f->spat); free(f->cspat); free(f);
}

static void arena_reset(void)
{
	for (int i = 0; i < g_arena.n; i++) field_free(g_arena.f[i]);
	g_arena.n = 0;
}

static ofield *T(ocfg *c) { return field_alloc(c, 0, 1); }
static ofield *TC(ocfg *c) { return field_alloc(c, 1, 1); }

static int nspec(const ofield *f) { return f->is_cplx ? f->c->ncplx : f->c->nlm; }

static void req_spec(ofield *f)
{
	if (f->spec_valid) return;
	if (!f->spat_valid) { fprintf(stderr, "oracle: field with no valid data\n"); exit(1); }
	if (f->is_cplx) orc_sht_analysis_cplx(f->c, f->cspat, f->spec);
	else orc_sht_analysis(f->c, f->spat, f->spec);
	f->spat_valid = 0; f->spec_valid = 1;
}

static void req_spat(ofield *f)
{
	if (f->spat_valid) return;
	if (!f->spec_valid) { fprintf(stderr, "oracle: field with no valid data\n"); exit(1); }
	if (f->is_cplx) orc_sht_synthesis_cplx(f->c, f->spec, f->cspat);
	else orc_sht_synthesis(f->c, f->spec, f->spat);
	f->spec_valid = 0; f->spat_valid = 1;
}

/* operator= (SPHData.hpp:87-103): copy valid representations and flags */
static void assign(ofield *dst, const ofield *src)
{
	if (src->spat_valid) {
		if (dst->is_cplx) memcpy(dst->cspat, src->cspat, sizeof(cplx) * dst->c->nspat);
		else memcpy(dst->spat, src->spat, sizeof(double) * dst->c->nspat);
	}
	if (src->spec_valid) memcpy(dst->spec, src->spec, sizeof(cplx) * nspec(dst));
	dst->spec_valid = src->spec_valid;
	dst->spat_valid = src->spat_valid;
}

static ofield *copy_of(ofield *a)
{
	ofield *o = field_alloc(a->c, a->is_cplx, 1);
	assign(o, a);
	return o;
}

static ofield *add(ofield *a, ofield *b)
{
	req_spec(a); req_spec(b);
	ofield *o = field_alloc(a->c, a->is_cplx, 1);
	for (int i = 0; i < nspec(a); i++) o->spec[i] = a->spec[i] + b->spec[i];
	o->spec_valid = 1;
	return o;
}

static ofield *sub(ofield *a, ofield *b)
{
	req_spec(a); req_spec(b);
	ofield *o = field_alloc(a->c, a->is_cplx, 1);
	for (int i = 0; i < nspec(a); i++) o->spec[i] = a->spec[i] - b->spec[i];
	o->spec_valid = 1;
	return o;
}

static void iadd(ofield *a, ofield *b)
{
	req_spec(a); req_spec(b);
	for (int i = 0; i < nspec(a); i++) a->spec[i] += b->spec[i];
	a->spec_valid = 1; a->spat_valid = 0;
}

static void isub(ofield *a, ofield *b)
{
	req_spec(a); req_spec(b);
	for (int i = 0; i < nspec(a); i++) a->spec[i] -= b->spec[i];
	a->spec_valid = 1; a->spat_valid = 0;
}

static ofield *neg(ofield *a)
{
	req_spec(a);
	ofield *o = field_alloc(a->c, a->is_cplx, 1);
	for (int i = 0; i < nspec(a); i++) o->spec[i] = -a->spec[i];
	o->spec_valid = 1;
	return o;
}

static ofield *scale(ofield *a, cplx s)
{
	req_spec(a);
	ofield *o = field_alloc(a->c, a->is_cplx, 1);
	for (int i = 0; i < nspec(a); i++) o->spec[i] = a->spec[i] * s;
	o->spec_valid = 1;
	return o;
}

static void iscale(ofield *a, cplx s)
{
	req_spec(a);
	for (int i = 0; i < nspec(a); i++) a->spec[i] *= s;
}

/* physical product (SPHData.hpp:334-353; complex: full product, the
 * reference's SPHDataComplex.hpp:333 loop bound is not reproduced) */
static ofield *mul(ofield *a, ofield *b)
{
	req_spat(a); req_spat(b);
	ofield *o = field_alloc(a->c, a->is_cplx, 1);
	if (a->is_cplx)
		for (int i = 0; i < a->c->nspat; i++) o->cspat[i] = b->cspat[i] * a->cspat[i];
	else
		for (int i = 0; i < a->c->nspat; i++) o->spat[i] = b->spat[i] * a->spat[i];
	o->spat_valid = 1;
	return o;
}

/* SPHDataComplex(const SPHData&) via spat_fromSPHData (SPHDataComplex.hpp:162-175) */
static ofield *cplx_from_real(ofield *r)
{
	req_spat(r);
	ofield *o = TC(r->c);
	for (int i = 0; i < r->c->nspat; i++) o->cspat[i] = r->spat[i];
	o->spat_valid = 1;
	return o;
}

/* spat_RealToSPHData (SPHDataComplex.hpp:129-142) */
static void real_part_to(ofield *cf, ofield *r)
{
	req_spat(cf);
	for (int i = 0; i < cf->c->nspat; i++) r->spat[i] = creal(cf->cspat[i]);
	r->spec_valid = 0; r->spat_valid = 1;
}

/* spat_update_lambda* helpers: kind 0 = divide by sqrt(1-mu^2) (gaussian
 * grid, mu = ct), 1 = multiply by st, 2 = divide by st, 3 = divide by
 * cos(lat[j]) (SPHOperatorsComplex.hpp:326-333), 4 = divide by st*st (the
 * cos-phi grid of the Robert divergences, SPHOperators.hpp:129-172;
 * spat_update_lambda_cosphi_grid is the cogaussian grid, SPHData.hpp:756-761) */
static void spat_metric(ofield *f, int kind)
{
	ocfg *c = f->c;
	if (f->spec_valid) req_spat(f);
	for (int i = 0; i < c->nphi; i++)
		for (int j = 0; j < c->nlat; j++) {
			double mu = c->ct[j], s;
			size_t k = (size_t)i * c->nlat + j;
			if (kind == 0) s = 1.0 / sqrt(1.0 - mu * mu);
			else if (kind == 1) s = c->st[j];
			else if (kind == 2) s = 1.0 / c->st[j];
			else s = 1.0 / cos(c->lat[j]);
			if (kind == 0 || kind == 2 || kind == 3 || kind == 4) {
				double d = (kind == 0) ? sqrt(1.0 - mu * mu)
					   : (kind == 2 ? c->st[j] : (kind == 4 ? c->st[j] * c->st[j] : cos(c->lat[j])));
				if (f->is_cplx) f->cspat[k] /= d; else f->spat[k] /= d;
			} else {
				if (f->is_cplx) f->cspat[k] *= s; else f->spat[k] *= s;
			}
		}
	f->spat_valid = 1; f->spec_valid = 0;
}

/* spec_get semantics: zero outside the triangle */
static cplx sget(const ofield *f, int n, int m)
{
	const ocfg *c = f->c;
	if (f->is_cplx) {
		if (n < 0 || n > c->L || abs(m) > c->L || abs(m) > n) return 0;
		return f->spec[lm_cplx(n, m)];
	}
	if (n < 0 || m < 0 || n > c->L || m > c->L || m > n) return 0;
	return f->spec[lm_real(c, n, m)];
}

static int m_lo(const ofield *f) { return f->is_cplx ? -f->c->L : 0; }
static int sidx(const ofield *f, int n, int m) { return f->is_cplx ? lm_cplx(n, m) : lm_real(f->c, n, m); }

/* spectral band operators; op: 0 diff_lon, 1 (1-mu^2) d/dmu, 2 mu, 3 mu2, 4 laplace */
static ofield *spec_op(ofield *a, int op)
{
	req_spec(a);
	ocfg *c = a->c;
	ofield *o = field_alloc(c, a->is_cplx, 1);
	for (int m = m_lo(a); m <= c->L; m++)
		for (int n = abs(m); n <= c->L; n++) {
			cplx v;
			double dn = n, dm = m;
			switch (op) {
			case 0: v = sget(a, n, m) * (I * dm); break;
			case 1: v = ((-dn + 1.0) * idR(dn - 1, dm)) * sget(a, n - 1, m)
				  + ((dn + 2.0) * idS(dn + 1, dm)) * sget(a, n + 1, m); break;
			case 2: v = idR(dn - 1, dm) * sget(a, n - 1, m) + idS(dn + 1, dm) * sget(a, n + 1, m); break;
			case 3: v = idA(dn - 2, dm) * sget(a, n - 2, m) + idB(dn, dm) * sget(a, n, m)
				  + idC(dn + 2, dm) * sget(a, n + 2, m); break;
			default: v = sget(a, n, m) * (-dn * (dn + 1.0)); break;
			}
			o->spec[sidx(a, n, m)] = v;
		}
	o->spec_valid = 1;
	return o;
}

/* laplace: copy then spec_update_lambda (SPHOperators.hpp:482-498) */
static ofield *op_laplace(ofield *a)
{
	req_spec(a);
	return spec_op(a, 4);
}

/* grad_lon (SPHOperators.hpp:100-119) */
static ofield *grad_lon(ofield *a)
{
	req_spec(a);
	ofield *o = spec_op(a, 0);
	req_spat(o);
	spat_metric(o, 0);
	return o;
}

/* grad_lat (SPHOperators.hpp:339-388) */
static ofield *grad_lat(ofield *a)
{
	ofield *o = spec_op(a, 1);
	req_spat(o);
	spat_metric(o, 0);
	return o;
}

/* div_lat (SPHOperators.hpp:411-439; complex: SPHOperatorsComplex.hpp:300-336) */
static ofield *div_lat(ofield *a)
{
	ofield *o = copy_of(a);
	spat_metric(o, 1);
	ofield *g = grad_lat(o);
	spat_metric(g, a->is_cplx ? 3 : 2);
	return g;
}

static ofield *div_lon(ofield *a) { return grad_lon(a); }

/* Robert-function operators (SPHOperators.hpp:124-208): the divergences
 * divide by cos^2 phi in physical space, the gradients stay spectral */
static ofield *robert_div_lon(ofield *a)
{
	ofield *o = spec_op(a, 0);
	req_spat(o);
	spat_metric(o, 4);
	return o;
}

static ofield *robert_div_lat(ofield *a)
{
	ofield *o = spec_op(a, 1);
	req_spat(o);
	spat_metric(o, 4);
	return o;
}

static ofield *robert_grad_lon(ofield *a) { return spec_op(a, 0); }
static ofield *robert_grad_lat(ofield *a) { return spec_op(a, 1); }

/* vort = grad_lon(v) - grad_lat(u) (SPHOperators.hpp:507-513) */
static ofield *op_vort(ofield *u, ofield *v)
{
	ofield *a = grad_lon(v);
	ofield *b = grad_lat(u);
	return sub(a, b);
}

/* div = div_lon(u) + div_lat(v) (SPHOperators.hpp:523-529) */
static ofield *op_div(ofield *u, ofield *v)
{
	ofield *a = div_lon(u);
	ofield *b = div_lat(v);
	return add(a, b);
}

/* exported single-op entry points for tests: real field in spectral form */
void orc_spec_op(ocfg *c, int op, const cplx *in, cplx *out)
{
	ofield *a = T(c);
	memcpy(a->spec, in, sizeof(cplx) * c->nlm);
	a->spec_valid = 1;
	ofield *o = spec_op(a, op);
	memcpy(out, o->spec, sizeof(cplx) * c->nlm);
	arena_reset();
}

void orc_spec_op_cplx(ocfg *c, int op, const cplx *in, cplx *out)
{
	ofield *a = TC(c);
	memcpy(a->spec, in, sizeof(cplx) * c->ncplx);
	a->spec_valid = 1;
	ofield *o = spec_op(a, op);
	memcpy(out, o->spec, sizeof(cplx) * c->ncplx);
	arena_reset();
}

/* physical-result operators on a physical input (for analytic tests):
 * which: 0 grad_lon, 1 grad_lat, 2 div_lat, 3 laplace(->physical) */
void orc_phys_op(ocfg *c, int which, const double *in, double *out)
{
	ofield *a = T(c);
	memcpy(a->spat, in, sizeof(double) * c->nspat);
	a->spat_valid = 1;
	ofield *o;
	if (which == 0) o = grad_lon(a);
	else if (which == 1) o = grad_lat(a);
	else if (which == 2) o = div_lat(a);
	else o = op_laplace(a);
	req_spat(o);
	memcpy(out, o->spat, sizeof(double) * c->nspat);
	arena_reset();
}

/* ------------------------------------------------------------------------ */
/* SWE tendency + RK4 (AppSWE.hpp:574-659, TimesteppingRK.hpp:310-392)      */
/* ------------------------------------------------------------------------ */

typedef struct {
	double a, g, omega, h0, visc;
	int nonlinear;
	int robert;    /* SimVars::use_robert_functions; the linear equations only */
} oswe;

static ofield *g_fdata;

/* f(x) = fdata*x*2.0*omega, x passed by value (AppSWE.hpp:132-136) */
static ofield *coriolis(ofield *x, double omega)
{
	ofield *xc = copy_of(x);
	ofield *p = mul(g_fdata, xc);
	ofield *q = scale(p, 2.0);
	return scale(q, omega);
}

static void tendency(const oswe *P, ofield *ih, ofield *iu, ofield *iv, ofield *oh, ofield *ou, ofield *ov)
{
	ocfg *c = ih->c;
	if (P->nonlinear) {
		/* o_h_t = -(div_lon(h*u) + div_lat(h*v)) * (1/a) */
		ofield *t1 = mul(ih, iu);
		ofield *t2 = div_lon(t1);
		ofield *t3 = mul(ih, iv);
		ofield *t4 = div_lat(t3);
		assign(oh, scale(neg(add(t2, t4)), 1.0 / P->a));
		/* velocity, linear terms */
		assign(ou, scale(neg(grad_lon(ih)), P->g / P->a));
		assign(ov, scale(neg(grad_lat(ih)), P->g / P->a));
		if (P->omega != 0) {
			iadd(ou, coriolis(iv, P->omega));
			isub(ov, coriolis(iu, P->omega));
		}
		/* non-linear terms */
		{
			ofield *g1 = grad_lon(iu);
			ofield *p1 = mul(iu, g1);
			ofield *g2 = grad_lat(iu);
			ofield *p2 = mul(iv, g2);
			isub(ou, scale(add(p1, p2), 1.0 / P->a));
		}
		{
			ofield *g3 = grad_lon(iv);
			ofield *p3 = mul(iu, g3);
			ofield *g4 = grad_lat(iv);
			ofield *p4 = mul(iv, g4);
			isub(ov, scale(add(p3, p4), 1.0 / P->a));
		}
	} else if (P->robert) {
		/* linear, velocities as Robert functions (AppSWE.hpp:606-620) */
		assign(oh, scale(neg(add(robert_div_lon(iu), robert_div_lat(iv))), P->h0 / P->a));
		assign(ou, scale(neg(robert_grad_lon(ih)), P->g / P->a));
		assign(ov, scale(neg(robert_grad_lat(ih)), P->g / P->a));
		if (P->omega != 0) {
			iadd(ou, coriolis(iv, P->omega));
			isub(ov, coriolis(iu, P->omega));
		}
	} else {
		/* linear (AppSWE.hpp:591-603) */
		assign(oh, scale(neg(add(div_lon(iu), div_lat(iv))), P->h0 / P->a));
		assign(ou, scale(neg(grad_lon(ih)), P->g / P->a));
		assign(ov, scale(neg(grad_lat(ih)), P->g / P->a));
		if (P->omega != 0) {
			iadd(ou, coriolis(iv, P->omega));
			isub(ov, coriolis(iu, P->omega));
		}
	}
	if (P->visc != 0) {
		double s = P->visc / (P->a * P->a);
		iadd(oh, scale(op_laplace(ih), s));
		iadd(ou, scale(op_laplace(iu), s));
		iadd(ov, scale(op_laplace(iv), s));
	}
	(void)c;
}

/* io + dt*(a*k) built as in TimesteppingRK.hpp:343-385 */
static ofield *stage_input(ofield *io, ofield *k, double a, double dt)
{
	ofield *ak = scale(k, a);
	ofield *dak = scale(ak, dt);
	return add(io, dak);
}

static void rk4_step(const oswe *P, ofield *h, ofield *u, ofield *v, ofield *k[4][3], double dt)
{
	tendency(P, h, u, v, k[0][0], k[0][1], k[0][2]);
	arena_reset();
	{
		ofield *sh = stage_input(h, k[0][0], 0.5, dt);
		ofield *su = stage_input(u, k[0][1], 0.5, dt);
		ofield *sv = stage_input(v, k[0][2], 0.5, dt);
		tendency(P, sh, su, sv, k[1][0], k[1][1], k[1][2]);
		arena_reset();
	}
	{
		ofield *sh = stage_input(h, k[1][0], 0.5, dt);
		ofield *su = stage_input(u, k[1][1], 0.5, dt);
		ofield *sv = stage_input(v, k[1][2], 0.5, dt);
		tendency(P, sh, su, sv, k[2][0], k[2][1], k[2][2]);
		arena_reset();
	}
	{
		ofield *sh = stage_input(h, k[2][0], 1.0, dt);
		ofield *su = stage_input(u, k[2][1], 1.0, dt);
		ofield *sv = stage_input(v, k[2][2], 1.0, dt);
		tendency(P, sh, su, sv, k[3][0], k[3][1], k[3][2]);
		arena_reset();
	}
	const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
	ofield *io[3] = {h, u, v};
	for (int f = 0; f < 3; f++) {
		ofield *s = add(add(add(scale(k[0][f], b[0]), scale(k[1][f], b[1])), scale(k[2][f], b[2])), scale(k[3][f], b[3]));
		iadd(io[f], scale(s, dt));
		arena_reset();
	}
}

static void make_fdata(ocfg *c)
{
	if (g_fdata && g_fdata->c == c) return;
	if (g_fdata) field_free(g_fdata);
	g_fdata = field_alloc(c, 0, 0);
	for (int i = 0; i < c->nphi; i++)
		for (int j = 0; j < c->nlat; j++) g_fdata->spat[(size_t)i * c->nlat + j] = c->ct[j];
	g_fdata->spat_valid = 1;
}

/* the cached mu field belongs to one configuration: drop it with that
 * configuration, so that a later one allocated at the same address never
 * finds a field of another grid */
static void fdata_release(ocfg *c)
{
	if (g_fdata && g_fdata->c == c) {
		field_free(g_fdata);
		g_fdata = NULL;
	}
}

/*
 * Run nsteps of RK4 from a physical initial state (h, u, v physical arrays).
 * params: [a, g, omega, h0, visc, nonlinear, dt]
 * Outputs (spectral, nlm complex each): h, u, v, and the diagnostics
 * phi = g*h, vort = vort(u,v)/a, div = div(u,v)/a (AppSWE.hpp:63-95).
 */
void orc_swe_rk4_run(ocfg *c, const double *params, const double *h0s, const double *u0s, const double *v0s,
		     int nsteps, cplx *out_h, cplx *out_u, cplx *out_v, cplx *out_phi, cplx *out_vort, cplx *out_div)
{
	oswe P = {params[0], params[1], params[2], params[3], params[4], (int)params[5], (int)params[7]};
	double dt = params[6];
	make_fdata(c);
	ofield *h = field_alloc(c, 0, 0), *u = field_alloc(c, 0, 0), *v = field_alloc(c, 0, 0);
	memcpy(h->spat, h0s, sizeof(double) * c->nspat); h->spat_valid = 1;
	memcpy(u->spat, u0s, sizeof(double) * c->nspat); u->spat_valid = 1;
	memcpy(v->spat, v0s, sizeof(double) * c->nspat); v->spat_valid = 1;
	ofield *k[4][3];
	for (int s = 0; s < 4; s++) for (int f = 0; f < 3; f++) k[s][f] = field_alloc(c, 0, 0);
	for (int n = 0; n < nsteps; n++)
		rk4_step(&P, h, u, v, k, dt);
	req_spec(h); req_spec(u); req_spec(v);
	memcpy(out_h, h->spec, sizeof(cplx) * c->nlm);
	memcpy(out_u, u->spec, sizeof(cplx) * c->nlm);
	memcpy(out_v, v->spec, sizeof(cplx) * c->nlm);
	if (out_phi) {
		ofield *phi = scale(h, P.g);
		memcpy(out_phi, phi->spec, sizeof(cplx) * c->nlm);
		ofield *vo = scale(op_vort(u, v), 1.0 / P.a);
		memcpy(out_vort, vo->spec, sizeof(cplx) * c->nlm);
		ofield *dv = scale(op_div(u, v), 1.0 / P.a);
		memcpy(out_div, dv->spec, sizeof(cplx) * c->nlm);
		arena_reset();
	}
	for (int s = 0; s < 4; s++) for (int f = 0; f < 3; f++) field_free(k[s][f]);
	field_free(h); field_free(u); field_free(v);
}

/* one tendency evaluation from spectral inputs (for unit tests) */
void orc_swe_tendency(ocfg *c, const double *params, const cplx *h, const cplx *u, const cplx *v,
		      cplx *oh, cplx *ou, cplx *ov)
{
	oswe P = {params[0], params[1], params[2], params[3], params[4], (int)params[5], (int)params[7]};
	make_fdata(c);
	ofield *fh = T(c), *fu = T(c), *fv = T(c);
	memcpy(fh->spec, h, sizeof(cplx) * c->nlm); fh->spec_valid = 1;
	memcpy(fu->spec, u, sizeof(cplx) * c->nlm); fu->spec_valid = 1;
	memcpy(fv->spec, v, sizeof(cplx) * c->nlm); fv->spec_valid = 1;
	ofield *th = field_alloc(c, 0, 0), *tu = field_alloc(c, 0, 0), *tv = field_alloc(c, 0, 0);
	tendency(&P, fh, fu, fv, th, tu, tv);
	req_spec(th); req_spec(tu); req_spec(tv);
	memcpy(oh, th->spec, sizeof(cplx) * c->nlm);
	memcpy(ou, tu->spec, sizeof(cplx) * c->nlm);
	memcpy(ov, tv->spec, sizeof(cplx) * c->nlm);
	field_free(th); field_free(tu); field_free(tv);
	arena_reset();
}

/* ------------------------------------------------------------------------ */
/* REXI coefficients (REXI.hpp:52-131)                                      */
/* ------------------------------------------------------------------------ */

/* Gaussian approximation weights for L = 11 (Haut et al.; the table at
 * GaussianApproximation.hpp:53-81) */
static const double ga_mu = -4.315321510875024;
static const double ga_a[23][2] = {
	{-1.0845749544592896e-7, 2.77075431662228e-8},
	{1.858753344202957e-8, -9.105375434750162e-7},
	{3.6743713227243024e-6, 7.073284346322969e-7},
	{-2.7990058083347696e-6, 0.0000112564827639346},
	{0.000014918577548849352, -0.0000316278486761932},
	{-0.0010751767283285608, -0.00047282220513073084},
	{0.003816465653840016, 0.017839810396560574},
	{0.12124105653274578, -0.12327042473830248},
	{-0.9774980792734348, -0.1877130220537587},
	{1.3432866123333178, 3.2034715228495942},
	{4.072408546157305, -6.123755543580666},
	{-9.442699917778205, 0.},
	{4.072408620272648, 6.123755841848161},
	{1.3432860877712938, -3.2034712658530275},
	{-0.9774985292598916, 0.18771238018072134},
	{0.1212417070363373, 0.12326987628935386},
	{0.0038169724770333343, -0.017839242222443888},
	{-0.0010756025812659208, 0.0004731874917343858},
	{0.000014713754789095218, 0.000031358475831136815},
	{-2.659323898804944e-6, -0.000011341571201752273},
	{3.6970377676364553e-6, -6.517457477594937e-7},
	{3.883933649142257e-9, 9.128496023863376e-7},
	{-1.0816457995911385e-7, -2.954309729192276e-8},
};

/* returns number of poles written; alpha/beta as interleaved re,im */
int orc_rexi_coeffs(double h, int M, int half, double *alpha, double *beta)
{
	const int L = 11;
	int N = M + L;
	int n2 = 2 * N + 1;
	cplx *al = (cplx *)calloc(n2, sizeof(cplx));
	cplx *be = (cplx *)calloc(n2, sizeof(cplx));
	for (int l = -L; l <= L; l++)
		for (int m = -M; m <= M; m++) {
			int n = l + m;
			al[n + N] = h * (ga_mu + I * (double)n);
			/* b_m = e^{h^2} e^{-i m h} (ExponentialApproximation.hpp:122) */
			cplx bm = exp(h * h) * cexp(-I * ((double)m * h));
			cplx a = ga_a[l + L][0] + I * ga_a[l + L][1];
			be[n + N] += creal(bm) * h * a;
		}
	int np = n2;
	if (half) {
		np = N + 1;
		for (int i = 0; i < N; i++) be[i] *= 2.0;
	}
	for (int i = 0; i < np; i++) {
		alpha[2 * i] = creal(al[i]); alpha[2 * i + 1] = cimag(al[i]);
		beta[2 * i] = creal(be[i]); beta[2 * i + 1] = cimag(be[i]);
	}
	free(al); free(be);
	return np;
}

/* ------------------------------------------------------------------------ */
/* banded matrices (SPHMatrixComplex / SPHSolverComplex)                    */
/* ------------------------------------------------------------------------ */

typedef struct {
	ocfg *c;
	int halo, nd;
	cplx *data;     /* NCompact rows x nd */
} omat;

static int ncompact(const ocfg *c, int n, int m)
{
	int L = c->L;
	if (m < 0) {
		int minc = L + m;
		int row_idx = (((minc - 1) * minc) >> 1) + minc;
		return row_idx + (n + m);
	}
	int rel = (m * (2 * L - m + 1) >> 1) + n;
	return ((L * (L + 1)) >> 1) + rel;
}

omat *orc_mat_create(ocfg *c, int halo)
{
	omat *M = (omat *)calloc(1, sizeof(omat));
	M->c = c; M->halo = halo; M->nd = 2 * halo + 1;
	M->data = (cplx *)calloc((size_t)c->ncplx * M->nd, sizeof(cplx));
	return M;
}

void orc_mat_destroy(omat *M) { if (M) { free(M->data); free(M); } }

/* rowElement_add (SPHMatrixComplex.hpp:181-206): note the n < row_m test
 * (not |row_m|) is kept as written */
static void row_add(omat *M, int n, int m, int rel, cplx v)
{
	if (n < m) return;
	int nn = n + rel;
	if (nn < 0 || nn < m || nn > M->c->L) return;
	M->data[(size_t)ncompact(M->c, n, m) * M->nd + rel + M->halo] += v;
}

/* component ids: 0 scalar, 1 mu, 2 (1-mu^2)d/dmu, 11..18 rexi z1..z8 */
void orc_mat_add(omat *M, int comp, double sre, double sim, double r)
{
	ocfg *c = M->c;
	cplx s = sre + I * sim;
	double ir2 = 1.0 / (r * r);
	for (int m = -c->L; m <= c->L; m++)
		for (int n = abs(m); n <= c->L; n++) {
			double dn = n, dm = m;
			switch (comp) {
			case 0: case 11: row_add(M, n, m, 0, s); break;
			case 1:
				row_add(M, n, m, -1, idR(dn - 1, dm) * s);
				row_add(M, n, m, +1, idS(dn + 1, dm) * s); break;
			case 2:
				row_add(M, n, m, -1, (-dn + 1.0) * idR(dn - 1, dm));
				row_add(M, n, m, +1, (dn + 2.0) * idS(dn + 1, dm)); break;
			case 12:
				row_add(M, n, m, -2, s * idA(dn - 2, dm));
				row_add(M, n, m, 0, s * idB(dn, dm));
				row_add(M, n, m, +2, s * idC(dn + 2, dm)); break;
			case 13:
				row_add(M, n, m, -4, s * (idA(dn - 2, dm) * idA(dn - 4.0, dm)));
				row_add(M, n, m, -2, s * (idA(dn - 2, dm) * idB(dn - 2, dm) + idB(dn, dm) * idA(dn - 2, dm)));
				row_add(M, n, m, 0, s * (idA(dn - 2, dm) * idC(dn, dm) + idB(dn, dm) * idB(dn, dm) + idC(dn + 2, dm) * idA(dn, dm)));
				row_add(M, n, m, +2, s * (idB(dn, dm) * idC(dn + 2, dm) + idC(dn + 2, dm) * idB(dn + 2, dm)));
				row_add(M, n, m, +4, s * (idC(dn + 2, dm) * idC(dn + 4, dm))); break;
			case 14:
				row_add(M, n, m, 0, ir2 * s * (I * dm)); break;
			case 15: {
				cplx fac = ir2 * s * (I * dm);
				row_add(M, n, m, -2, fac * idA(dn - 2, dm));
				row_add(M, n, m, 0, fac * idB(dn + 0, dm));
				row_add(M, n, m, +2, fac * idC(dn + 2, dm)); break;
			}
			case 16:
				row_add(M, n, m, -2, -ir2 * s * (idD(dn - 1, dm) * idR(dn - 2, dm) + (idE(dn, dm) - 3.0) * idA(dn - 2, dm)));
				row_add(M, n, m, 0, -ir2 * s * (1.0 + idD(dn - 1, dm) * idS(dn, dm) + (idE(dn, dm) - 3.0) * idB(dn, dm)));
				row_add(M, n, m, +2, -ir2 * s * (idE(dn, dm) - 3.0) * idC(dn + 2, dm)); break;
			case 17:
				row_add(M, n, m, 0, -ir2 * s * dn * (dn + 1.0)); break;
			case 18: {
				row_add(M, n, m, 0, ir2 * s * 2.0);
				double fac = -ir2 * (dm * dm - dn * (dn + 1.0));
				row_add(M, n, m, -2, s * fac * idA(dn - 2, dm));
				row_add(M, n, m, 0, s * fac * idB(dn, dm));
				row_add(M, n, m, +2, s * fac * idC(dn + 2, dm)); break;
			}
			/* Robert-function components (SPHSolverComplex.hpp:241-370) */
			case 19: {   /* z4robert */
				cplx fac = ir2 * s * (I * dm);
				row_add(M, n, m, -2, -fac * idA(dn - 2, dm));
				row_add(M, n, m, 0, -fac * idB(dn, dm) + fac);
				row_add(M, n, m, +2, -fac * idC(dn + 2, dm)); break;
			}
			case 20: {   /* z5robert: every term on the -2 diagonal, as the reference adds them (:305-309) */
				cplx fac = ir2 * s * (I * dm);
				row_add(M, n, m, -2, -fac * (idA(dn - 2, dm) * idA(dn - 4, dm)));
				row_add(M, n, m, -2, -fac * (idA(dn - 2, dm) * idB(dn - 2, dm) + idB(dn, dm) * idA(dn - 2, dm))
						      + fac * idA(dn - 2, dm));
				row_add(M, n, m, -2, -fac * (idA(dn - 2, dm) * idC(dn, dm) + idB(dn, dm) * idB(dn, dm)
							     + idC(dn + 2, dm) * idA(dn, dm)) + fac * idB(dn, dm));
				row_add(M, n, m, -2, -fac * (idB(dn, dm) * idC(dn + 2, dm) + idC(dn + 2, dm) * idB(dn + 2, dm))
						      + fac * idC(dn + 2, dm));
				row_add(M, n, m, -2, -fac * (idC(dn + 2, dm) * idC(dn + 4, dm))); break;
			}
			case 21: {   /* z6robert */
				cplx s6 = ir2 * s * dm;
				double gr = idG(dn - 1, dm) * idR(dn - 2, dm);
				double gs = idG(dn - 1, dm) * idS(dn, dm) + idH(dn + 1, dm) * idR(dn, dm);
				double hs = idH(dn + 1, dm) * idS(dn + 2, dm);
				row_add(M, n, m, -2, s6 * gr);
				row_add(M, n, m, 0, s6 * gs);
				row_add(M, n, m, +2, s6 * hs);
				row_add(M, n, m, -4, s6 * (-idA(dn - 4, dm) * gr));
				row_add(M, n, m, -2, s6 * (-idB(dn - 2, dm) * gr));
				row_add(M, n, m, 0, s6 * (-idC(dn, dm) * gr));
				row_add(M, n, m, -2, s6 * (-idA(dn - 2, dm) * gs));
				row_add(M, n, m, 0, s6 * (-idB(dn, dm) * gs));
				row_add(M, n, m, +2, s6 * (-idC(dn + 2, dm) * gs));
				row_add(M, n, m, 0, s6 * (-idA(dn, dm) * hs));
				row_add(M, n, m, +2, s6 * (-idB(dn + 2, dm) * hs));
				row_add(M, n, m, +4, s6 * (-idC(dn + 4, dm) * hs)); break;
			}
			default: break;
			}
		}
}

void orc_mat_get(omat *M, double *out)
{
	memcpy(out, M->data, sizeof(cplx) * (size_t)M->c->ncplx * M->nd);
}

/*
 * Banded LU with partial pivoting and solve -- the algorithm of LAPACK
 * zgbtf2 + zgbtrs as used by zgbsv (BandedMatrixSolver.hpp:398-434).
 * AB: LAPACK band storage, ldab = 2*kl+ku+1, column-major (column j holds
 * A(i,j) at AB[kl+ku+i-j + j*ldab]).  Returns INFO (0 ok, >0 singular pivot).
 */
int orc_gbsv(int n, int kl, int ku, cplx *AB, int ldab, int *ipiv, cplx *b)
{
	int kv = ku + kl;
#define ABX(i, j) AB[(size_t)(kv + (i) - (j)) + (size_t)(j) * ldab]
	/* zero the fill-in area */
	for (int j = 0; j < n; j++)
		for (int i = 0; i < kl; i++) AB[i + (size_t)j * ldab] = 0;
	int info = 0;
	int ju = 0;
	for (int j = 0; j < n; j++) {
		int km = kl < (n - 1 - j) ? kl : (n - 1 - j);
		/* pivot search in column j, rows j..j+km */
		int jp = 0;
		double amax = -1;
		for (int i = 0; i <= km; i++) {
			double v = fabs(creal(ABX(j + i, j))) + fabs(cimag(ABX(j + i, j)));
			if (v > amax) { amax = v; jp = i; }
		}
		ipiv[j] = j + jp;
		if (ABX(j + jp, j) != 0) {
			int jmax = j + ku + jp;
			if (jmax > n - 1) jmax = n - 1;
			if (jmax > ju) ju = jmax;
			if (jp != 0) {
				for (int jj = j; jj <= ju; jj++) {
					cplx t = ABX(j + jp, jj); ABX(j + jp, jj) = ABX(j, jj); ABX(j, jj) = t;
				}
			}
			if (km > 0) {
				cplx piv = 1.0 / ABX(j, j);
				for (int i = 1; i <= km; i++) ABX(j + i, j) *= piv;
				for (int jj = j + 1; jj <= ju; jj++) {
					cplx t = ABX(j, jj);
					if (t != 0)
						for (int i = 1; i <= km; i++) ABX(j + i, jj) -= ABX(j + i, j) * t;
				}
			}
		} else if (info == 0) {
			info = j + 1;
		}
	}
	if (info) return info;
	/* forward: L with row interchanges */
	for (int j = 0; j < n - 1; j++) {
		int km = kl < (n - 1 - j) ? kl : (n - 1 - j);
		int l = ipiv[j];
		if (l != j) { cplx t = b[l]; b[l] = b[j]; b[j] = t; }
		for (int i = 1; i <= km; i++) b[j + i] -= ABX(j + i, j) * b[j];
	}
	/* backward: U has kv super-diagonals */
	for (int j = n - 1; j >= 0; j--) {
		b[j] /= ABX(j, j);
		int i0 = j - kv; if (i0 < 0) i0 = 0;
		for (int i = i0; i < j; i++) b[i] -= ABX(i, j) * b[j];
	}
#undef ABX
	return 0;
}

/* SPHSolverComplex::solve (SPHSolverComplex.hpp:462-511) on spectral data */
int orc_mat_solve(omat *M, const cplx *rhs, cplx *x)
{
	ocfg *c = M->c;
	int h = M->halo, nd = M->nd, ldab = 3 * h + 1;
	cplx *AB = (cplx *)malloc(sizeof(cplx) * (size_t)ldab * (c->L + 1));
	cplx *b = (cplx *)malloc(sizeof(cplx) * (c->L + 1));
	int *ipiv = (int *)malloc(sizeof(int) * (c->L + 1));
	int info_all = 0;
	for (int m = -c->L; m <= c->L; m++) {
		int n0 = abs(m), size = c->L + 1 - abs(m);
		const cplx *A = M->data + (size_t)ncompact(c, n0, m) * nd;
		for (int k = 0; k < size; k++) b[k] = rhs[lm_cplx(n0 + k, m)];
		for (size_t k = 0; k < (size_t)ldab * size; k++) AB[k] = 0;
		/* C row-band -> LAPACK band (BandedMatrixSolver.hpp:344-368) */
		for (int j = 0; j < size; j++)
			for (int i = 0; i < nd; i++) {
				int si = j + (h - i), sj = j;
				if (si < 0 || si >= size) continue;
				AB[(nd + si - sj - 1) + (size_t)sj * ldab] = A[(size_t)(j - i + h) * nd + i];
			}
		int info = orc_gbsv(size, h, h, AB, ldab, ipiv, b);
		if (info && !info_all) info_all = info;
		for (int k = 0; k < size; k++) x[lm_cplx(n0 + k, m)] = b[k];
	}
	free(AB); free(b); free(ipiv);
	return info_all;
}

/* ------------------------------------------------------------------------ */
/* SWERexiSPH (SWERexiSPH.hpp:68-244) and the REXI time step                */
/* ------------------------------------------------------------------------ */

typedef struct {
	cplx alpha, beta;
	double r, ir, two_omega, avg_geo;
	int coriolis;
	int robert;    /* SWERexiSPHRobert (SWERexiSPHRobert.hpp) */
	omat *Mphi, *Mvel;
} orexi_pole;

static void pole_setup(orexi_pole *p, ocfg *c, cplx alpha, cplx beta, double r, double omega,
		       double avg_geo, double dt, int coriolis, int robert)
{
	p->coriolis = coriolis;
	p->robert = robert;
	p->alpha = alpha / dt;
	p->beta = beta / dt;
	p->r = r; p->ir = 1.0 / r;
	p->two_omega = 2.0 * omega;
	p->avg_geo = avg_geo;
	cplx a = p->alpha;
	double to = p->two_omega, G = avg_geo;
	p->Mphi = orc_mat_create(c, 4);
	cplx s;
#define ADD(M, comp, val) do { s = (val); orc_mat_add(M, comp, creal(s), cimag(s), r); } while (0)
	ADD(p->Mphi, 11, (a * a) * (a * a));
	if (coriolis) {
		ADD(p->Mphi, 12, 2.0 * to * to * a * a);
		ADD(p->Mphi, 13, (to * to) * (to * to));
		/* z4..z6, or their Robert forms (SWERexiSPHRobert.hpp setup) */
		ADD(p->Mphi, robert ? 19 : 14, -G * a * to);
		ADD(p->Mphi, robert ? 20 : 15, G / a * to * to * to);
		ADD(p->Mphi, robert ? 21 : 16, G * 2.0 * to * to);
	}
	ADD(p->Mphi, 17, -G * a * a);
	if (coriolis) ADD(p->Mphi, 18, -G * to * to);
	p->Mvel = orc_mat_create(c, 2);
	ADD(p->Mvel, 11, a * a);
	if (coriolis) ADD(p->Mvel, 12, to * to);
#undef ADD
}

static ofield *solve_field(omat *M, ofield *rhs)
{
	req_spec(rhs);
	ofield *o = TC(rhs->c);
	if (orc_mat_solve(M, rhs->spec, o->spec)) { fprintf(stderr, "oracle zgbsv INFO != 0\n"); exit(1); }
	o->spec_valid = 1;
	return o;
}

/* spec_solve_helmholtz (SPHDataComplex.hpp:523-545) */
static ofield *helmholtz_c(ofield *x, cplx a, cplx b, double r)
{
	ofield *o = copy_of(x);
	req_spec(o);
	cplx bb = b / (r * r);
	ocfg *c = x->c;
	for (int n = 0; n <= c->L; n++)
		for (int m = -n; m <= n; m++)
			o->spec[lm_cplx(n, m)] /= (a + (-bb * (double)n * ((double)n + 1.0)));
	o->spat_valid = 0; o->spec_valid = 1;
	return o;
}

/* one pole of SWERexiSPH::solve (SWERexiSPH.hpp:137-244) up to, not including,
 * the beta scaling and the real part: complex phi, u, v */
static void pole_solve_c(orexi_pole *p, ofield *i_phi0, ofield *i_u0, ofield *i_v0,
			 ofield **o_phi, ofield **o_u, ofield **o_v)
{
	ocfg *c = i_phi0->c;
	ofield *mu = TC(c);
	for (int i = 0; i < c->nphi; i++)
		for (int j = 0; j < c->nlat; j++) mu->cspat[(size_t)i * c->nlat + j] = c->ct[j];
	mu->spat_valid = 1;
	ofield *phi0 = cplx_from_real(i_phi0);
	ofield *u0 = cplx_from_real(i_u0);
	ofield *v0 = cplx_from_real(i_v0);
	ofield *div0, *eta0;
	if (p->robert) {
		/* robert_div: one division by cos^2 phi of the spectral sum
		 * (SPHOperatorsComplex.hpp:530-550); robert_vort stays spectral (:496-502) */
		ofield *d = add(spec_op(u0, 0), spec_op(v0, 1));
		req_spat(d);
		spat_metric(d, 4);
		div0 = scale(d, p->ir);
		eta0 = scale(sub(spec_op(v0, 0), spec_op(u0, 1)), p->ir);
	} else {
		div0 = scale(op_div(u0, v0), p->ir);
		eta0 = scale(op_vort(u0, v0), p->ir);
	}
	cplx a = p->alpha;
	double to = p->two_omega, G = p->avg_geo;
	ofield *phi, *u, *v;
	if (p->coriolis) {
		/* Fc_k = ((two_omega*ir)*grad_lat(mu)) * ( -(a^2 u0 - to^2 mu2(u0)) + 2 a to mu(v0) ):
		 * the scalar product projects grad_lat(mu) to spectral space before the
		 * physical product (left-associative evaluation of SWERexiSPH.hpp:183-186) */
		ofield *gm = scale(p->robert ? robert_grad_lat(mu) : grad_lat(mu), to * p->ir);
		ofield *br = add(neg(sub(scale(u0, a * a), scale(spec_op(u0, 3), to * to))),
				 scale(spec_op(v0, 2), 2.0 * a * to));
		ofield *Fc = mul(gm, br);
		/* foo = G*(div0 - to*(1/a)*mu(eta0)) + (a*phi0 + to^2*(1/a)*mu2(phi0)) */
		ofield *t1 = scale(sub(div0, scale(spec_op(eta0, 2), to * (1.0 / a))), G);
		ofield *aphi = scale(cplx_from_real(i_phi0), a);
		ofield *t2 = add(aphi, scale(spec_op(cplx_from_real(i_phi0), 3), to * to * (1.0 / a)));
		ofield *foo = add(t1, t2);
		ofield *rhs = add(add(scale(foo, a * a), scale(spec_op(foo, 3), to * to)), scale(Fc, G / a));
		phi = solve_field(p->Mphi, rhs);
		ofield *A_ = add(u0, scale(p->robert ? robert_grad_lon(phi) : grad_lon(phi), p->ir));
		ofield *B_ = add(v0, scale(p->robert ? robert_grad_lat(phi) : grad_lat(phi), p->ir));
		ofield *rhsa = sub(scale(A_, a), scale(spec_op(B_, 2), to));
		ofield *rhsb = add(scale(spec_op(A_, 2), to), scale(B_, a));
		u = solve_field(p->Mvel, rhsa);
		v = solve_field(p->Mvel, rhsb);
	} else {
		ofield *rhs = add(scale(div0, G), scale(cplx_from_real(i_phi0), a));
		phi = helmholtz_c(rhs, a * a, -G, p->r);
		u = scale(add(u0, scale(p->robert ? robert_grad_lon(phi) : grad_lon(phi), p->ir)), 1.0 / a);
		v = scale(add(v0, scale(p->robert ? robert_grad_lat(phi) : grad_lat(phi), p->ir)), 1.0 / a);
	}
	(void)phi0;
	*o_phi = phi; *o_u = u; *o_v = v;
}

static void pole_solve(orexi_pole *p, ofield *i_phi0, ofield *i_u0, ofield *i_v0,
		       ofield *o_phi, ofield *o_u, ofield *o_v)
{
	ofield *phi, *u, *v;
	pole_solve_c(p, i_phi0, i_u0, i_v0, &phi, &u, &v);
	iscale(phi, p->beta);
	iscale(u, p->beta);
	iscale(v, p->beta);
	real_part_to(phi, o_phi);
	real_part_to(u, o_u);
	real_part_to(v, o_v);
}

/*
 * One pole solve for a given alpha (dt = 1, beta = 1), for residual checks.
 * params: [a, g, omega, h0]; inputs spectral (nlm) real fields phi0 (geopotential),
 * u0, v0; outputs the complex physical fields phi, u, v (nspat each).
 */
void orc_rexi_pole_solve(ocfg *c, const double *params, double alpha_re, double alpha_im,
			 const cplx *phi_in, const cplx *u_in, const cplx *v_in,
			 cplx *out_phi, cplx *out_u, cplx *out_v)
{
	double a = params[0], g = params[1], omega = params[2], h0 = params[3];
	orexi_pole pole;
	memset(&pole, 0, sizeof(pole));
	pole_setup(&pole, c, alpha_re + I * alpha_im, 1.0, a, omega, h0 * g, 1.0, omega != 0, 0);
	ofield *f[3];
	const cplx *in[3] = {phi_in, u_in, v_in};
	for (int k = 0; k < 3; k++) {
		f[k] = field_alloc(c, 0, 0);
		memcpy(f[k]->spec, in[k], sizeof(cplx) * c->nlm);
		f[k]->spec_valid = 1;
	}
	ofield *phi, *u, *v;
	pole_solve_c(&pole, f[0], f[1], f[2], &phi, &u, &v);
	req_spat(phi); req_spat(u); req_spat(v);
	memcpy(out_phi, phi->cspat, sizeof(cplx) * c->nspat);
	memcpy(out_u, u->cspat, sizeof(cplx) * c->nspat);
	memcpy(out_v, v->cspat, sizeof(cplx) * c->nspat);
	arena_reset();
	for (int k = 0; k < 3; k++) field_free(f[k]);
	orc_mat_destroy(pole.Mphi);
	orc_mat_destroy(pole.Mvel);
}

/*
 * REXI time stepping (AppSWE.hpp:336-565), linear equations.
 * params: [a, g, omega, h0, visc, rexi_h, rexi_M, half, dt, robert]
 * Input state: spectral h, u, v (nlm each).  Output spectral after nsteps.
 */
int orc_swe_rexi_run(ocfg *c, const double *params, const cplx *h_in, const cplx *u_in, const cplx *v_in,
		     int nsteps, cplx *out_h, cplx *out_u, cplx *out_v)
{
	double a = params[0], g = params[1], omega = params[2], h0 = params[3], visc = params[4];
	double rh = params[5]; int rM = (int)params[6], half = (int)params[7];
	double dt = params[8];
	int robert = (int)params[9];
	int maxp = 2 * (rM + 11) + 1;
	double *al = (double *)malloc(sizeof(double) * 2 * maxp), *be = (double *)malloc(sizeof(double) * 2 * maxp);
	int np = orc_rexi_coeffs(rh, rM, half, al, be);
	int coriolis = omega != 0;
	orexi_pole *poles = (orexi_pole *)calloc(np, sizeof(orexi_pole));
	for (int i = 0; i < np; i++)
		pole_setup(&poles[i], c, al[2 * i] + I * al[2 * i + 1], be[2 * i] + I * be[2 * i + 1],
			   a, omega, h0 * g, dt, coriolis, robert);
	ofield *h = field_alloc(c, 0, 0), *u = field_alloc(c, 0, 0), *v = field_alloc(c, 0, 0);
	memcpy(h->spec, h_in, sizeof(cplx) * c->nlm); h->spec_valid = 1;
	memcpy(u->spec, u_in, sizeof(cplx) * c->nlm); u->spec_valid = 1;
	memcpy(v->spec, v_in, sizeof(cplx) * c->nlm); v->spec_valid = 1;
	ofield *tp = field_alloc(c, 0, 0), *tu = field_alloc(c, 0, 0), *tv = field_alloc(c, 0, 0);
	ofield *ap = field_alloc(c, 0, 0), *au = field_alloc(c, 0, 0), *av = field_alloc(c, 0, 0);
	ofield *pphi = field_alloc(c, 0, 0), *pu = field_alloc(c, 0, 0), *pv = field_alloc(c, 0, 0);
	for (int s = 0; s < nsteps; s++) {
		assign(pphi, scale(h, g));
		assign(pu, u);
		assign(pv, v);
		arena_reset();
		memset(ap->spat, 0, sizeof(double) * c->nspat); ap->spat_valid = 1; ap->spec_valid = 0;
		memset(au->spat, 0, sizeof(double) * c->nspat); au->spat_valid = 1; au->spec_valid = 0;
		memset(av->spat, 0, sizeof(double) * c->nspat); av->spat_valid = 1; av->spec_valid = 0;
		for (int i = 0; i < np; i++) {
			pole_solve(&poles[i], pphi, pu, pv, tp, tu, tv);
			iadd(ap, tp);
			iadd(au, tu);
			iadd(av, tv);
			arena_reset();
		}
		assign(h, scale(ap, 1.0 / g));
		assign(u, au);
		assign(v, av);
		arena_reset();
		if (visc != 0) {
			double sc = visc * dt;
			ofield *fl[3] = {h, u, v};
			for (int f = 0; f < 3; f++) {
				ofield *o = copy_of(fl[f]);
				req_spec(o);
				double bb = -sc / (a * a);
				for (int m = 0; m <= c->L; m++)
					for (int n = m; n <= c->L; n++)
						o->spec[lm_real(c, n, m)] /= (1.0 + (-bb * (double)n * ((double)n + 1.0)));
				assign(fl[f], o);
				arena_reset();
			}
		}
	}
	req_spec(h); req_spec(u); req_spec(v);
	memcpy(out_h, h->spec, sizeof(cplx) * c->nlm);
	memcpy(out_u, u->spec, sizeof(cplx) * c->nlm);
	memcpy(out_v, v->spec, sizeof(cplx) * c->nlm);
	for (int i = 0; i < np; i++) { orc_mat_destroy(poles[i].Mphi); orc_mat_destroy(poles[i].Mvel); }
	free(poles); free(al); free(be);
	field_free(h); field_free(u); field_free(v);
	field_free(tp); field_free(tu); field_free(tv);
	field_free(ap); field_free(au); field_free(av);
	field_free(pphi); field_free(pu); field_free(pv);
	return np;
}
