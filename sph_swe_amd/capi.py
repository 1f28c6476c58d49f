"""ctypes binding of libsphswe.so (the C ABI declared in include/sphswe.h).

This is a thin loader used by the Python test-suite and bench.py; it has no
fallback: if the HIP library is missing or cannot be loaded, every entry point
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsphswe.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sphswe.h")

OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_SINGULAR, ERR_NAN = range(6)
PREC_FP64, PREC_FP32, PREC_BF16, PREC_BF16X3, PREC_MIXED = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16": PREC_BF16, "bf16x3": PREC_BF16X3, "mixed": PREC_MIXED}
# spectral operators (sphswe_spec_op / _cplx)
OP_IDENTITY, OP_DIFF_LON, OP_ONE_MINUS_MU2_DMU, OP_MU, OP_MU2, OP_LAPLACE, OP_HELMHOLTZ = range(7)
# physical operators (sphswe_spat_op / _cplx)
SPAT_MUL, SPAT_DIV_COSPHI, SPAT_MUL_ST, SPAT_DIV_ST, SPAT_MUL_MU, SPAT_SCALE, SPAT_DIV_COSLAT, SPAT_DIV_ST2 = range(8)

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_sz = ctypes.c_size_t
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class SweParams(ctypes.Structure):
    _fields_ = [("earth_radius", _d), ("gravitation", _d), ("coriolis_omega", _d), ("h0", _d),
                ("viscosity", _d), ("dt", _d), ("nonlinear", _i), ("precision", _i), ("robert_functions", _i)]


class RexiParams(ctypes.Structure):
    _fields_ = [("earth_radius", _d), ("gravitation", _d), ("coriolis_omega", _d), ("h0", _d),
                ("viscosity", _d), ("dt", _d), ("rexi_h", _d), ("rexi_M", _i), ("rexi_half", _i),
                ("precision", _i), ("band_precision", _i), ("refine", _i), ("robert_functions", _i)]


_SIGS = {
    "sphswe_grid_auto": [_i, _i, _ip, _ip],
    "sphswe_create": [_i, _i, _i, _i, _i, ctypes.POINTER(_vp)],
    "sphswe_destroy": [_vp],
    "sphswe_get_sizes": [_vp, _ip, _ip, _ip, _ip, _ip],
    "sphswe_get_grid": [_vp, _dp, _dp, _dp],
    "sphswe_sync": [_vp],
    "sphswe_malloc": [_vp, _sz, ctypes.POINTER(_vp)],
    "sphswe_free": [_vp, _vp],
    "sphswe_memcpy_h2d": [_vp, _vp, _vp, _sz],
    "sphswe_memcpy_d2h": [_vp, _vp, _vp, _sz],
    "sphswe_memcpy_d2d": [_vp, _vp, _vp, _sz],
    "sphswe_fill_zero": [_vp, _vp, _sz],
    "sphswe_sht_analysis": [_vp, _vp, _vp, _i, _i],
    "sphswe_sht_synthesis": [_vp, _vp, _vp, _i, _i],
    "sphswe_sht_analysis_cplx": [_vp, _vp, _vp, _i, _i],
    "sphswe_sht_synthesis_cplx": [_vp, _vp, _vp, _i, _i],
    "sphswe_sht_analysis_noalloc": [_vp, _vp, _vp, _i, _i],
    "sphswe_fft_analysis": [_vp, _vp, _vp, _i],
    "sphswe_fft_synthesis": [_vp, _vp, _vp, _i],
    "sphswe_legendre_analysis": [_vp, _vp, _vp, _i, _i],
    "sphswe_legendre_synthesis": [_vp, _vp, _vp, _i, _i],
    "sphswe_spec_op": [_vp, _i, _vp, _vp, _i, _d, _d, _d, _d],
    "sphswe_spec_op_cplx": [_vp, _i, _vp, _vp, _i, _d, _d, _d, _d, _d],
    "sphswe_spec_axpby": [_vp, _sz, _d, _d, _vp, _d, _d, _vp, _vp],
    "sphswe_spec_copy_modes": [_vp, _vp, _vp, _vp, _i],
    "sphswe_spec_add_scalar": [_vp, _vp, _d, _d],
    "sphswe_spat_op": [_vp, _i, _vp, _vp, _vp, _d, _i],
    "sphswe_spat_op_cplx": [_vp, _i, _vp, _vp, _vp, _d, _i],
    "sphswe_any_nan": [_vp, _vp, _sz, _ip],
    "sphswe_reduce_minmax": [_vp, _vp, _sz, _dp, _dp, _dp],
    "sphswe_swe_rk4": [_vp, ctypes.POINTER(SweParams), _vp, _i, _i],
    "sphswe_swe_rk4_graph": [_vp, ctypes.POINTER(SweParams), _vp, _i, _i],
    "sphswe_autotune": [_vp],
    "sphswe_set_tuning": [_vp, _i, _i, _i],
    "sphswe_get_tuning": [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)],
    "sphswe_set_legendre_tuning": [_vp, _i, _i],
    "sphswe_get_legendre_tuning": [_vp, _ip, _ip],
    "sphswe_set_fp32_tables": [_vp, _i, _i],
    "sphswe_get_fp32_tables": [_vp, _ip, _ip],
    "sphswe_rexi_create": [_vp, ctypes.POINTER(RexiParams), ctypes.POINTER(_vp)],
    "sphswe_rexi_destroy": [_vp],
    "sphswe_rexi_poles": [_vp, _ip, _dp, _dp],
    "sphswe_rexi_step": [_vp, _vp, _vp, _i],
    "sphswe_rexi_coefficients": [_d, _i, _i, _ip, _dp, _dp, _dp],
    "sphswe_rexi_create_poles": [_vp, ctypes.POINTER(RexiParams), _i, _dp, _dp, ctypes.POINTER(_vp)],
    "sphswe_rexi_apply": [_vp, _vp, _vp, _vp],
    "sphswe_swe_split": [_vp, _vp, ctypes.POINTER(SweParams), _vp, _i],
    "sphswe_spat_real_to_cplx": [_vp, _vp, _vp, _i],
    "sphswe_spat_cplx_part": [_vp, _vp, _vp, _i, _i],
    "sphswe_band_create": [_vp, _i, ctypes.POINTER(_vp)],
    "sphswe_band_destroy": [_vp],
    "sphswe_band_add": [_vp, _i, _d, _d, _d],
    "sphswe_band_solve": [_vp, _vp, _vp, _vp],
    "sphswe_timer_mark": [_vp, _i],
    "sphswe_timer_elapsed": [_vp, _i, _i, _fp],
}

_lib = None


def load():
    """Load libsphswe.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsphswe.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _i
    lib.sphswe_last_error.argtypes = [_vp]
    lib.sphswe_last_error.restype = ctypes.c_char_p
    lib.sphswe_noalloc_scratch_bytes.argtypes = []
    lib.sphswe_noalloc_scratch_bytes.restype = _sz
    lib.sphswe_stream.argtypes = [_vp]
    lib.sphswe_stream.restype = _vp
    _lib = lib
    return lib


def header_functions():
    """Names of every function declared in include/sphswe.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sphswe_[a-z0-9_]+)\s*\(", txt)))


class SphSweError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sphswe error %d: %s" % (code, msg))
        self.code = code


def grid_auto(lmax, nl_order=2):
    nl, nph = _i(), _i()
    rc = load().sphswe_grid_auto(lmax, nl_order, ctypes.byref(nl), ctypes.byref(nph))
    if rc:
        raise SphSweError(rc, "grid_auto")
    return nl.value, nph.value


class DeviceArray:
    """A device buffer owned by a Context (pooled allocation)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        ctx._check(ctx.lib.sphswe_malloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            self.ctx.lib.sphswe_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One SHT configuration on one GPU (the device-side SPHConfig)."""

    def __init__(self, lmax, nlat=None, nphi=None, device=0):
        self.lib = load()
        if nlat is None or nphi is None:
            nlat, nphi = grid_auto(lmax)
        h = _vp()
        rc = self.lib.sphswe_create(lmax, lmax, nlat, nphi, device, ctypes.byref(h))
        if rc:
            msg = self.lib.sphswe_last_error(h).decode() if h.value else "create failed"
            if h.value:
                self.lib.sphswe_destroy(h)
            raise SphSweError(rc, msg)
        self.h = h.value
        self.L, self.nlat, self.nphi = lmax, nlat, nphi
        self.nlm = (lmax + 1) * (lmax + 2) // 2
        self.ncplx = (lmax + 1) ** 2
        self.nspat = nlat * nphi
        ct, st, wg = np.empty(nlat), np.empty(nlat), np.empty(nlat)
        self._check(self.lib.sphswe_get_grid(self.h, ct.ctypes.data_as(_dp), st.ctypes.data_as(_dp),
                                             wg.ctypes.data_as(_dp)))
        self.ct, self.st, self.wg = ct, st, wg

    def close(self):
        if getattr(self, "h", None):
            self.lib.sphswe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise SphSweError(rc, self.lib.sphswe_last_error(self.h).decode())

    # memory --------------------------------------------------------------
    def empty(self, nbytes):
        return DeviceArray(self, nbytes)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.nbytes)
        self._check(self.lib.sphswe_memcpy_h2d(self.h, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def to_host(self, d, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        self._check(self.lib.sphswe_memcpy_d2h(self.h, out.ctypes.data, d.ptr, out.nbytes))
        return out

    def sync(self):
        self._check(self.lib.sphswe_sync(self.h))

    # transforms (host in / host out convenience) -------------------------
    def analysis(self, spat, precision=PREC_FP64):
        spat = np.ascontiguousarray(spat, dtype=np.float64).reshape(-1, self.nspat)
        nf = spat.shape[0]
        d_in = self.to_device(spat)
        d_out = self.empty(nf * self.nlm * 16)
        self._check(self.lib.sphswe_sht_analysis(self.h, d_in.ptr, d_out.ptr, nf, precision))
        return self.to_host(d_out, np.complex128, (nf, self.nlm))

    def synthesis(self, spec, precision=PREC_FP64):
        spec = np.ascontiguousarray(spec, dtype=np.complex128).reshape(-1, self.nlm)
        nf = spec.shape[0]
        d_in = self.to_device(spec)
        d_out = self.empty(nf * self.nspat * 8)
        self._check(self.lib.sphswe_sht_synthesis(self.h, d_in.ptr, d_out.ptr, nf, precision))
        return self.to_host(d_out, np.float64, (nf, self.nspat))

    def analysis_cplx(self, spat, precision=PREC_FP64):
        spat = np.ascontiguousarray(spat, dtype=np.complex128).reshape(-1, self.nspat)
        nf = spat.shape[0]
        d_in = self.to_device(spat)
        d_out = self.empty(nf * self.ncplx * 16)
        self._check(self.lib.sphswe_sht_analysis_cplx(self.h, d_in.ptr, d_out.ptr, nf, precision))
        return self.to_host(d_out, np.complex128, (nf, self.ncplx))

    def synthesis_cplx(self, spec, precision=PREC_FP64):
        spec = np.ascontiguousarray(spec, dtype=np.complex128).reshape(-1, self.ncplx)
        nf = spec.shape[0]
        d_in = self.to_device(spec)
        d_out = self.empty(nf * self.nspat * 16)
        self._check(self.lib.sphswe_sht_synthesis_cplx(self.h, d_in.ptr, d_out.ptr, nf, precision))
        return self.to_host(d_out, np.complex128, (nf, self.nspat))

    def spec_op(self, op, spec, s=1.0, p0=0.0, p1=1.0, p2=0.0, cplx=False):
        spec = np.ascontiguousarray(spec, dtype=np.complex128)
        n = self.ncplx if cplx else self.nlm
        spec2 = spec.reshape(-1, n)
        nf = spec2.shape[0]
        d_in = self.to_device(spec2)
        d_out = self.empty(spec2.nbytes)
        s = complex(s)
        if cplx:
            self._check(self.lib.sphswe_spec_op_cplx(self.h, op, d_in.ptr, d_out.ptr, nf, s.real, s.imag, p0, p1, p2))
        else:
            self._check(self.lib.sphswe_spec_op(self.h, op, d_in.ptr, d_out.ptr, nf, s.real, s.imag, p0, p1))
        return self.to_host(d_out, np.complex128, spec.shape)

    def copy_modes_to(self, dst_ctx, spec):
        """Real spectra of this context's truncation (nf, nlm) as spectra of
        dst_ctx's truncation (SPHData::spec_copyToDifferentModes)."""
        spec2 = np.ascontiguousarray(spec, dtype=np.complex128).reshape(-1, self.nlm)
        nf = spec2.shape[0]
        d_in = self.to_device(spec2)
        d_out = dst_ctx.empty(nf * dst_ctx.nlm * 16)
        dst_ctx._check(self.lib.sphswe_spec_copy_modes(self.h, d_in.ptr, dst_ctx.h, d_out.ptr, nf))
        return dst_ctx.to_host(d_out, np.complex128, (nf, dst_ctx.nlm))

    def spat_op(self, op, a, b=None, s=1.0, cplx=False):
        dt = np.complex128 if cplx else np.float64
        a = np.ascontiguousarray(a, dtype=dt)
        nf = a.size // self.nspat
        d_a = self.to_device(a)
        d_b = self.to_device(np.ascontiguousarray(b, dtype=dt)) if b is not None else None
        d_o = self.empty(a.nbytes)
        fn = self.lib.sphswe_spat_op_cplx if cplx else self.lib.sphswe_spat_op
        self._check(fn(self.h, op, d_a.ptr, d_b.ptr if d_b else None, d_o.ptr, s, nf))
        return self.to_host(d_o, dt, a.shape)

    def swe_rk4(self, params, state, nmembers, nsteps):
        """state: device buffer holding nmembers*(h,u,v) spectral fields."""
        self._check(self.lib.sphswe_swe_rk4(self.h, ctypes.byref(params), state.ptr, nmembers, nsteps))

    def swe_rk4_graph(self, params, state, nmembers, nsteps):
        """As swe_rk4, one captured step replayed nsteps times."""
        self._check(self.lib.sphswe_swe_rk4_graph(self.h, ctypes.byref(params), state.ptr, nmembers, nsteps))

    # checks ------------------------------------------------------------------
    def any_nan(self, d, n):
        """1 if any of the first n doubles of device buffer d is NaN or Inf."""
        f = _i()
        self._check(self.lib.sphswe_any_nan(self.h, d.ptr, n, ctypes.byref(f)))
        return f.value

    def minmax(self, d, n):
        """(min, max, max |x|) of the first n doubles of device buffer d."""
        a, b, c = _d(), _d(), _d()
        self._check(self.lib.sphswe_reduce_minmax(self.h, d.ptr, n, ctypes.byref(a), ctypes.byref(b),
                                                  ctypes.byref(c)))
        return a.value, b.value, c.value

    # autotuner ----------------------------------------------------------------
    def autotune(self):
        """Times the launch variants of this context's sizes and keeps the
        fastest; returns the choices (see tuning())."""
        self._check(self.lib.sphswe_autotune(self.h))
        return self.tuning()

    def tuning(self):
        v = [_i(0) for _ in range(4)]
        self._check(self.lib.sphswe_get_tuning(self.h, *[ctypes.byref(x) for x in v]))
        ts, ta = self.fp32_tables()
        return {"fft_nseq_r2c": v[0].value, "fft_nseq_c2r": v[1].value, "synth_mfma_min": v[2].value,
                "fp32_table_synth_min": ts, "fp32_table_anal_min": ta, "tuned": bool(v[3].value)}

    def set_legendre_tuning(self, synth_shape=0, anal_variant=0):
        self._check(self.lib.sphswe_set_legendre_tuning(self.h, synth_shape, anal_variant))

    def legendre_tuning(self):
        v = [_i(0) for _ in range(2)]
        self._check(self.lib.sphswe_get_legendre_tuning(self.h, *[ctypes.byref(x) for x in v]))
        return {"synth_shape": v[0].value, "anal_variant": v[1].value}

    def set_tuning(self, fft_nseq_r2c=0, fft_nseq_c2r=0, synth_mfma_min=0):
        self._check(self.lib.sphswe_set_tuning(self.h, fft_nseq_r2c, fft_nseq_c2r, synth_mfma_min))

    # REXI -------------------------------------------------------------------
    def rexi(self, params, alpha=None, beta=None):
        return Rexi(self, params, alpha, beta)

    def band(self, halo):
        return Band(self, halo)

    def set_fp32_tables(self, synth_min_fields, anal_min_fields=None):
        """FP32 real syntheses / analyses of at least that many fields use the
        stored FP32 Legendre tables (0: never); one value sets both."""
        if anal_min_fields is None:
            anal_min_fields = synth_min_fields
        self._check(self.lib.sphswe_set_fp32_tables(self.h, int(synth_min_fields), int(anal_min_fields)))

    def fp32_tables(self):
        a, b = _i(), _i()
        self._check(self.lib.sphswe_get_fp32_tables(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    # timing ----------------------------------------------------------------
    def mark(self, slot):
        self._check(self.lib.sphswe_timer_mark(self.h, slot))

    def elapsed_ms(self, s0, s1):
        ms = ctypes.c_float()
        self._check(self.lib.sphswe_timer_elapsed(self.h, s0, s1, ctypes.byref(ms)))
        return ms.value


def rexi_coefficients(h, M, half=True, with_beta_im=False):
    """REXI::setup on the host (no GPU): alpha, beta_re[, beta_im] as complex arrays."""
    lib = load()
    n = _i()
    _check_rc(lib.sphswe_rexi_coefficients(h, M, int(half), ctypes.byref(n), None, None, None))
    a, br, bi = (np.empty(2 * n.value) for _ in range(3))
    _check_rc(lib.sphswe_rexi_coefficients(h, M, int(half), ctypes.byref(n), a.ctypes.data_as(_dp),
                                           br.ctypes.data_as(_dp),
                                           bi.ctypes.data_as(_dp) if with_beta_im else None))
    out = (a.view(np.complex128), br.view(np.complex128))
    return out + (bi.view(np.complex128),) if with_beta_im else out


def _check_rc(rc):
    if rc:
        raise SphSweError(rc, "sphswe call failed")


class Rexi:
    """Device REXI integrator (poles, cached band factors) for one context.
    With `alpha`/`beta` given, the integrator uses that pole set."""

    def __init__(self, ctx, params, alpha=None, beta=None):
        self.ctx = ctx
        h = _vp()
        if alpha is None:
            ctx._check(ctx.lib.sphswe_rexi_create(ctx.h, ctypes.byref(params), ctypes.byref(h)))
        else:
            a = np.ascontiguousarray(alpha, dtype=np.complex128).view(np.float64)
            b = np.ascontiguousarray(beta, dtype=np.complex128).view(np.float64)
            ctx._check(ctx.lib.sphswe_rexi_create_poles(ctx.h, ctypes.byref(params), a.size // 2,
                                                        a.ctypes.data_as(_dp), b.ctypes.data_as(_dp),
                                                        ctypes.byref(h)))
        self.h = h.value
        n = _i()
        ctx._check(ctx.lib.sphswe_rexi_poles(self.h, ctypes.byref(n), None, None))
        self.npoles = n.value

    def poles(self):
        a = np.empty(2 * self.npoles)
        b = np.empty(2 * self.npoles)
        self.ctx._check(self.ctx.lib.sphswe_rexi_poles(self.h, ctypes.byref(_i()), a.ctypes.data_as(_dp),
                                                        b.ctypes.data_as(_dp)))
        return a.view(np.complex128), b.view(np.complex128)

    def step(self, state, nsteps=1):
        self.ctx._check(self.ctx.lib.sphswe_rexi_step(self.ctx.h, self.h, state.ptr, nsteps))

    def split_step(self, params, state, nsteps=1):
        """nsteps split steps: this REXI step of the linear equations, then one
        RK4 step of the nonlinear remainder with params (SweParams)."""
        self.ctx._check(self.ctx.lib.sphswe_swe_split(self.ctx.h, self.h, ctypes.byref(params), state.ptr, nsteps))

    def apply(self, d_in, d_out):
        """sum_n Re(beta_n x_n) for real spectra (phi0, u0, v0) -> (phi, u, v)."""
        self.ctx._check(self.ctx.lib.sphswe_rexi_apply(self.ctx.h, self.h, d_in.ptr, d_out.ptr))

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.sphswe_rexi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rexi_params(a=6.37122e6, g=9.80616, omega=7.292e-5, h0=10000.0, visc=1e5, dt=3600.0, rexi_h=0.2, rexi_M=128,
                half=1, precision=PREC_FP64, band_precision=PREC_FP64, refine=0, robert=0):
    return RexiParams(a, g, omega, h0, visc, dt, rexi_h, rexi_M, half, precision, band_precision, refine, robert)


def swe_params(a=6.37122e6, g=9.80616, omega=7.292e-5, h0=10000.0, visc=1e5, dt=120.0, nonlinear=1,
               precision=PREC_FP64, robert=0):
    return SweParams(a, g, omega, h0, visc, dt, nonlinear, precision, robert)


class Band:
    """SPHSolverComplex on the device: components added on the host, per-m
    pivoted LU, solves on complex spectra."""

    def __init__(self, ctx, halo):
        self.ctx = ctx
        h = _vp()
        ctx._check(ctx.lib.sphswe_band_create(ctx.h, halo, ctypes.byref(h)))
        self.h = h.value

    def add(self, component, scalar=1.0, r=1.0):
        s = complex(scalar)
        self.ctx._check(self.ctx.lib.sphswe_band_add(self.h, component, s.real, s.imag, r))

    def solve(self, rhs):
        """rhs: complex spectrum ((L+1)^2) on the host; returns the solution on the host."""
        d = self.ctx.to_device(np.ascontiguousarray(rhs, dtype=np.complex128))
        x = self.ctx.empty(d.nbytes)
        self.ctx._check(self.ctx.lib.sphswe_band_solve(self.ctx.h, self.h, d.ptr, x.ptr))
        return self.ctx.to_host(x, np.complex128, rhs.shape)

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.sphswe_band_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
