// Real analysis that allocates nothing: the Fourier rows and the Legendre
// analysis operands live in one fixed device array of this library, so the
// call makes no hipMalloc / hipFree (for callers that cannot allocate: code
// under stream capture, tight serving loops).  No counterpart in the
// reference, whose spat_to_SH (SPHData.hpp:190-207) works on SHTNS-owned
// buffers.
//
// The array exists once per device and is shared by every context and call on
// that device, and it keeps the contents of the last call; calls that overlap
// on one device (two contexts on two streams) race on it.  Every byte a call
// reads is written earlier in the same call (the FFT writes every Fourier row
// entry, the operand pass every operand it later reads), so nothing needs
// clearing.
#include "sphswe_common.h"

#include <mutex>

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)96 << 20;

// the host-side name is a stand-in, not a device address: host code uses the
// address looked up by hipGetSymbolAddress, per device
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
std::unordered_map<int, unsigned char *> g_ptr;   // device -> address of scratch on it

}  // namespace

// looks up the array's address on the context's device (once per device); the
// caller (sphswe_create) has made that device current
int sphswe_noalloc_init(sphswe_ctx *ctx)
{
	std::lock_guard<std::mutex> lock(g_mu);
	auto it = g_ptr.find(ctx->device);
	if (it == g_ptr.end()) {
		unsigned char *p = nullptr;
		const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			return sphswe_set_err(ctx, SPHSWE_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
		it = g_ptr.emplace(ctx->device, p).first;
	}
	ctx->noalloc_scratch = it->second;
	return SPHSWE_OK;
}

extern "C" size_t sphswe_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

extern "C" int sphswe_sht_analysis_noalloc(sphswe_ctx *ctx, const double *d_spat, double *d_spec, int nfields,
					   int precision)
{
	if (!ctx || !d_spat || !d_spec || nfields < 0) return SPHSWE_ERR_ARG;
	if (precision < SPHSWE_PREC_FP64 || precision > SPHSWE_PREC_BF16X3)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG, "analysis_noalloc: precision %d not available", precision);
	if (nfields == 0) return SPHSWE_OK;
	if (!ctx->noalloc_scratch) return sphswe_set_err(ctx, SPHSWE_ERR_INTERNAL, "analysis_noalloc: no scratch address");
	const size_t four_b = sizeof(double2) * (size_t)(ctx->L + 1) * ctx->nlat * nfields;
	const size_t e_off = (four_b + 255) & ~(size_t)255;
	const size_t group_b = sphswe_leg_anal_group_bytes(ctx, precision);
	if (e_off + group_b > NOALLOC_SCRATCH_BYTES)
		return sphswe_set_err(ctx, SPHSWE_ERR_ARG,
				      "analysis_noalloc: %d fields at T%d need %zu bytes of scratch, the limit is %zu",
				      nfields, ctx->L, e_off + group_b, NOALLOC_SCRATCH_BYTES);
	unsigned char *base = ctx->noalloc_scratch;
	if (int rc = sphswe_fft_r2c(ctx, d_spat, (double2 *)base, nfields)) return rc;
	ctx->ext_escratch = base + e_off;
	ctx->ext_escratch_bytes = NOALLOC_SCRATCH_BYTES - e_off;
	const int rc = sphswe_leg_anal(ctx, (const double2 *)base, (double2 *)d_spec, nfields, precision);
	ctx->ext_escratch = nullptr;
	ctx->ext_escratch_bytes = 0;
	return rc;
}
