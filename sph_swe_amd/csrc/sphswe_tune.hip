// Autotuner: times the launch variants available for this context's sizes on
// synthetic data and keeps the fastest in the context.  Every variant computes
// the same transform in the same precision (they differ in the order of the
// sums only), so tuning changes results by round-off of that precision at most.
//   - real FFTs: complex sequences per workgroup (rings per workgroup, LDS per
//     workgroup), per direction;
//   - FP64 real synthesis: the vector-unit kernel against the FP64 matrix-core
//     kernel, as the number of fields from which the latter is used, and the
//     workgroup shape of the vector-unit kernel;
//   - FP64 real analysis of up to four fields: the 4x4x4 matrix-instruction
//     kernel or the 16x16x4 kernel;
//   - FP32 real synthesis and analysis: the stored-table GEMMs against the
//     recurrence kernels, per direction, at 16 fields.
#include "sphswe_common.h"

namespace {

template <typename F>
int time_ms(sphswe_ctx *ctx, int reps, F fn, float *ms)
{
	hipEvent_t a, b;
	SPHSWE_HIP(ctx, hipEventCreate(&a));
	SPHSWE_HIP(ctx, hipEventCreate(&b));
	int rc = fn();   // warm-up
	if (!rc) {
		SPHSWE_HIP(ctx, hipEventRecord(a, ctx->stream));
		for (int i = 0; i < reps && !rc; i++) rc = fn();
		SPHSWE_HIP(ctx, hipEventRecord(b, ctx->stream));
		SPHSWE_HIP(ctx, hipEventSynchronize(b));
		SPHSWE_HIP(ctx, hipEventElapsedTime(ms, a, b));
	}
	(void)hipEventDestroy(a);
	(void)hipEventDestroy(b);
	return rc;
}

}  // namespace

extern "C" int sphswe_set_tuning(sphswe_ctx *ctx, int fft_nseq_r2c, int fft_nseq_c2r, int synth_mfma_min)
{
	if (!ctx || fft_nseq_r2c < 0 || fft_nseq_c2r < 0 || synth_mfma_min < 0) return SPHSWE_ERR_ARG;
	ctx->fft_nseq_r2c = fft_nseq_r2c;
	ctx->fft_nseq_c2r = fft_nseq_c2r;
	ctx->synth_mfma_min = synth_mfma_min;
	ctx->buf_gen++;   // a captured step may use the previous choices
	return SPHSWE_OK;
}

extern "C" int sphswe_set_legendre_tuning(sphswe_ctx *ctx, int synth_shape, int anal_variant)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (synth_shape < 0 || synth_shape > 1 || anal_variant < 0 || anal_variant > 1)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "legendre tuning: synthesis shape %d, analysis variant %d",
				      synth_shape, anal_variant);
	ctx->synth_shape = synth_shape;
	ctx->anal_variant = anal_variant;
	ctx->buf_gen++;
	return SPHSWE_OK;
}

extern "C" int sphswe_get_legendre_tuning(const sphswe_ctx *ctx, int *synth_shape, int *anal_variant)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (synth_shape) *synth_shape = ctx->synth_shape;
	if (anal_variant) *anal_variant = ctx->anal_variant;
	return SPHSWE_OK;
}

extern "C" int sphswe_get_tuning(const sphswe_ctx *ctx, int *fft_nseq_r2c, int *fft_nseq_c2r, int *synth_mfma_min,
				 int *tuned)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (fft_nseq_r2c) *fft_nseq_r2c = ctx->fft_nseq_r2c;
	if (fft_nseq_c2r) *fft_nseq_c2r = ctx->fft_nseq_c2r;
	if (synth_mfma_min) *synth_mfma_min = ctx->synth_mfma_min;
	if (tuned) *tuned = ctx->tuned;
	return SPHSWE_OK;
}

extern "C" int sphswe_set_fp32_tables(sphswe_ctx *ctx, int synth_min_fields, int anal_min_fields)
{
	if (!ctx || synth_min_fields < 0 || anal_min_fields < 0) return SPHSWE_ERR_ARG;
	ctx->tab_min_synth = synth_min_fields;
	ctx->tab_min_anal = anal_min_fields;
	ctx->buf_gen++;
	return SPHSWE_OK;
}

extern "C" int sphswe_get_fp32_tables(const sphswe_ctx *ctx, int *synth_min_fields, int *anal_min_fields)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	if (synth_min_fields) *synth_min_fields = ctx->tab_min_synth;
	if (anal_min_fields) *anal_min_fields = ctx->tab_min_anal;
	return SPHSWE_OK;
}

extern "C" int sphswe_autotune(sphswe_ctx *ctx)
{
	if (!ctx) return SPHSWE_ERR_ARG;
	constexpr int NF = 16, REPS = 5;
	const size_t spat_b = sizeof(double) * (size_t)ctx->nspat * NF;
	const size_t four_b = sizeof(double2) * (size_t)(ctx->L + 1) * ctx->nlat * NF;
	const size_t spec_b = sizeof(double2) * (size_t)ctx->nlm * NF;
	void *spat = nullptr, *four = nullptr, *spec = nullptr;
	int rc = SPHSWE_OK;
	if (hipMalloc(&spat, spat_b) != hipSuccess || hipMalloc(&four, four_b) != hipSuccess ||
	    hipMalloc(&spec, spec_b) != hipSuccess) {
		rc = sphswe_set_err(ctx, SPHSWE_ERR_NOMEM, "autotune buffers");
	}
	if (!rc && (hipMemsetAsync(spat, 0, spat_b, ctx->stream) != hipSuccess ||
		    hipMemsetAsync(four, 0, four_b, ctx->stream) != hipSuccess ||
		    hipMemsetAsync(spec, 0, spec_b, ctx->stream) != hipSuccess))
		rc = sphswe_set_err(ctx, SPHSWE_ERR_HIP, "autotune buffer fill");
	const int nf_fft = 8;
	// real FFTs
	int cand[8];
	const int nc = sphswe_fft_nseq_candidates(ctx, cand, 8);
	for (int dir = 0; dir < 2 && !rc && nc > 1; dir++) {
		float best = 1e30f;
		int pick = 0;
		for (int i = 0; i < nc && !rc; i++) {
			(dir ? ctx->fft_nseq_c2r : ctx->fft_nseq_r2c) = cand[i];
			float ms = 0;
			rc = time_ms(ctx, REPS, [&]() {
				return dir ? sphswe_fft_c2r(ctx, (const double2 *)four, (double *)spat, nf_fft)
					   : sphswe_fft_r2c(ctx, (const double *)spat, (double2 *)four, nf_fft);
			}, &ms);
			if (!rc && ms < best) {
				best = ms;
				pick = cand[i];
			}
		}
		(dir ? ctx->fft_nseq_c2r : ctx->fft_nseq_r2c) = pick;
	}
	// FP64 real synthesis: vector units or matrix cores, by field count
	const int counts[] = {1, 3, 8, 16};
	int threshold = 0;
	for (int ci = 3; ci >= 0 && !rc; ci--) {
		const int nf = counts[ci];
		float t_valu = 0, t_mfma = 0;
		ctx->synth_mfma_min = 0;
		rc = time_ms(ctx, REPS, [&]() {
			return sphswe_leg_synth(ctx, (const double2 *)spec, (double2 *)four, nf, SPHSWE_PREC_FP64);
		}, &t_valu);
		if (rc) break;
		ctx->synth_mfma_min = 1;
		rc = time_ms(ctx, REPS, [&]() {
			return sphswe_leg_synth(ctx, (const double2 *)spec, (double2 *)four, nf, SPHSWE_PREC_FP64);
		}, &t_mfma);
		if (rc) break;
		if (t_mfma < t_valu) threshold = nf;   // the matrix cores win from here up
		else break;
	}
	ctx->synth_mfma_min = threshold;
	// FP64 real Legendre kernel shapes: synthesis of 3 + 9 fields, analysis of
	// 3 + 4 fields (a round trip; a SWE step)
	for (int which = 0; which < 2 && !rc; which++) {
		float t[2] = {0, 0};
		for (int v = 0; v < 2 && !rc; v++) {
			(which ? ctx->anal_variant : ctx->synth_shape) = v;
			float a = 0, b = 0;
			if (which == 0) {
				rc = time_ms(ctx, REPS, [&]() { return sphswe_leg_synth(ctx, (const double2 *)spec, (double2 *)four, 3, SPHSWE_PREC_FP64); }, &a);
				if (!rc) rc = time_ms(ctx, REPS, [&]() { return sphswe_leg_synth(ctx, (const double2 *)spec, (double2 *)four, 9, SPHSWE_PREC_FP64); }, &b);
			} else {
				rc = time_ms(ctx, REPS, [&]() { return sphswe_leg_anal(ctx, (const double2 *)four, (double2 *)spec, 3, SPHSWE_PREC_FP64); }, &a);
				if (!rc) rc = time_ms(ctx, REPS, [&]() { return sphswe_leg_anal(ctx, (const double2 *)four, (double2 *)spec, 4, SPHSWE_PREC_FP64); }, &b);
			}
			t[v] = a + b;
		}
		(which ? ctx->anal_variant : ctx->synth_shape) = t[1] < t[0] ? 1 : 0;
	}
	// FP32 real transforms of NF fields: stored tables or the recurrence kernels,
	// per direction; the tables are released when neither direction uses them
	for (int dir = 0; dir < 2 && !rc; dir++) {
		float t_rec = 0, t_tab = 0;
		int &thr = dir ? ctx->tab_min_anal : ctx->tab_min_synth;
		auto run = [&]() {
			return dir ? sphswe_leg_anal(ctx, (const double2 *)four, (double2 *)spec, NF, SPHSWE_PREC_FP32)
				   : sphswe_leg_synth(ctx, (const double2 *)spec, (double2 *)four, NF, SPHSWE_PREC_FP32);
		};
		thr = 0;
		rc = time_ms(ctx, REPS, run, &t_rec);
		if (rc) break;
		thr = NF;
		rc = time_ms(ctx, REPS, run, &t_tab);
		if (rc) break;
		thr = t_tab < t_rec ? NF : 0;
	}
	if (!rc && !ctx->tab_min_synth && !ctx->tab_min_anal) {
		(void)hipStreamSynchronize(ctx->stream);
		sphswe_leg_tables_free(ctx);
	}
	if (spat) (void)hipFree(spat);
	if (four) (void)hipFree(four);
	if (spec) (void)hipFree(spec);
	if (!rc) ctx->tuned = 1;
	ctx->buf_gen++;   // the choices changed and the timing buffers were freed
	return rc;
}
