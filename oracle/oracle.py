"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  The product path (sph_swe_amd) never does.

The C restatement lives in oracle/sphswe_oracle.c; every function there cites
the reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle_sphswe.so")
_REF_DUMP = os.path.join(_HERE, "_ref", "ref_rexi_dump")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build():
    """Compile the oracle (and, when the reference is present, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    vp = ctypes.c_void_p
    sig = {
        "orc_create": (vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "orc_destroy": (None, [vp]),
        "orc_set_threads": (None, [ctypes.c_int]),
        "orc_auto_grid": (None, [ctypes.c_int, _ip, _ip]),
        "orc_get_grid": (None, [vp, _dp, _dp, _dp, _dp]),
        "orc_get_counters": (None, [vp, _lp]),
        "orc_plm_column": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]),
        "orc_plm_table": (None, [vp, _dp]),
        "orc_fft": (None, [vp, _dp, _dp, ctypes.c_int]),
        "orc_sht_analysis": (None, [vp, _dp, _dp]),
        "orc_sht_synthesis": (None, [vp, _dp, _dp]),
        "orc_sht_analysis_cplx": (None, [vp, _dp, _dp]),
        "orc_sht_synthesis_cplx": (None, [vp, _dp, _dp]),
        "orc_identities": (None, [ctypes.c_double, ctypes.c_double, _dp]),
        "orc_spec_op": (None, [vp, ctypes.c_int, _dp, _dp]),
        "orc_spec_op_cplx": (None, [vp, ctypes.c_int, _dp, _dp]),
        "orc_phys_op": (None, [vp, ctypes.c_int, _dp, _dp]),
        "orc_swe_rk4_run": (None, [vp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_swe_tendency": (None, [vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "orc_rexi_coeffs": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp]),
        "orc_mat_create": (vp, [vp, ctypes.c_int]),
        "orc_mat_destroy": (None, [vp]),
        "orc_mat_add": (None, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
        "orc_mat_get": (None, [vp, _dp]),
        "orc_mat_solve": (ctypes.c_int, [vp, _dp, _dp]),
        "orc_gbsv": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, _dp]),
        "orc_swe_rexi_run": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]),
        "orc_rexi_pole_solve": (None, [vp, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                                       _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def auto_grid(L):
    nl, npz = ctypes.c_int(), ctypes.c_int()
    lib().orc_auto_grid(L, ctypes.byref(nl), ctypes.byref(npz))
    return nl.value, npz.value


# operator ids shared with the C side
OP_DIFF_LON, OP_ONE_MINUS_MU2_DMU, OP_MU, OP_MU2, OP_LAPLACE = 0, 1, 2, 3, 4


class Oracle:
    """One truncation / grid (the role of SPHConfig in the reference)."""

    def __init__(self, L, nlat=None, nphi=None):
        if nlat is None or nphi is None:
            nlat, nphi = auto_grid(L)
        self.L, self.nlat, self.nphi = L, nlat, nphi
        self.nlm = (L + 1) * (L + 2) // 2
        self.ncplx = (L + 1) ** 2
        self.nspat = nlat * nphi
        self.h = lib().orc_create(L, nlat, nphi)
        ct, st, w, lat = (np.empty(nlat) for _ in range(4))
        lib().orc_get_grid(self.h, _d(ct), _d(st), _d(w), _d(lat))
        self.ct, self.st, self.w, self.lat = ct, st, w, lat

    def __del__(self):
        try:
            lib().orc_destroy(self.h)
        except Exception:
            pass

    # index helpers (SPHConfig.hpp:106-130)
    def idx(self, l, m):
        return (m * (2 * self.L - m + 1)) // 2 + l

    @staticmethod
    def cidx(l, m):
        return l * l + l + m

    def lon(self):
        return np.arange(self.nphi) / self.nphi * 2.0 * np.pi

    def counters(self):
        out = np.zeros(4, dtype=np.int64)
        lib().orc_get_counters(self.h, out.ctypes.data_as(_lp))
        return out

    def plm_table(self):
        n = sum(self.L - m + 1 for m in range(self.L + 1)) * self.nlat
        out = np.empty(n)
        lib().orc_plm_table(self.h, _d(out))
        return out

    def fft(self, x, sign=-1):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.empty_like(x)
        lib().orc_fft(self.h, _d(x.view(np.float64)), _d(out.view(np.float64)), sign)
        return out

    def analysis(self, spat):
        spat = np.ascontiguousarray(spat, dtype=np.float64).reshape(self.nspat)
        out = np.empty(self.nlm, dtype=np.complex128)
        lib().orc_sht_analysis(self.h, _d(spat), _d(out.view(np.float64)))
        return out

    def synthesis(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128)
        out = np.empty(self.nspat)
        lib().orc_sht_synthesis(self.h, _d(spec.view(np.float64)), _d(out))
        return out

    def analysis_cplx(self, spat):
        spat = np.ascontiguousarray(spat, dtype=np.complex128).reshape(self.nspat)
        out = np.empty(self.ncplx, dtype=np.complex128)
        lib().orc_sht_analysis_cplx(self.h, _d(spat.view(np.float64)), _d(out.view(np.float64)))
        return out

    def synthesis_cplx(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128)
        out = np.empty(self.nspat, dtype=np.complex128)
        lib().orc_sht_synthesis_cplx(self.h, _d(spec.view(np.float64)), _d(out.view(np.float64)))
        return out

    def spec_op(self, op, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128)
        out = np.empty_like(spec)
        fn = lib().orc_spec_op_cplx if spec.size == self.ncplx and self.ncplx != self.nlm else lib().orc_spec_op
        fn(self.h, op, _d(spec.view(np.float64)), _d(out.view(np.float64)))
        return out

    def phys_op(self, which, spat):
        spat = np.ascontiguousarray(spat, dtype=np.float64).reshape(self.nspat)
        out = np.empty(self.nspat)
        lib().orc_phys_op(self.h, which, _d(spat), _d(out))
        return out

    def swe_rk4_run(self, params, h, u, v, nsteps):
        """params dict: a, g, omega, h0, visc, nonlinear, dt, and optionally robert
        (1: the linear equations with Robert-function velocities); h/u/v physical."""
        p = np.array([params["a"], params["g"], params["omega"], params["h0"], params["visc"],
                      float(params["nonlinear"]), params["dt"], float(params.get("robert", 0))])
        outs = [np.empty(self.nlm, dtype=np.complex128) for _ in range(6)]
        h, u, v = (np.ascontiguousarray(x, dtype=np.float64).reshape(self.nspat) for x in (h, u, v))
        lib().orc_swe_rk4_run(self.h, _d(p), _d(h), _d(u), _d(v), nsteps,
                              *[_d(o.view(np.float64)) for o in outs])
        return dict(zip(["h", "u", "v", "phi", "vort", "div"], outs))

    def swe_tendency(self, params, h, u, v):
        p = np.array([params["a"], params["g"], params["omega"], params["h0"], params["visc"],
                      float(params["nonlinear"]), params.get("dt", 0.0), float(params.get("robert", 0))])
        ins = [np.ascontiguousarray(x, dtype=np.complex128) for x in (h, u, v)]
        outs = [np.empty(self.nlm, dtype=np.complex128) for _ in range(3)]
        lib().orc_swe_tendency(self.h, _d(p), *[_d(x.view(np.float64)) for x in ins],
                               *[_d(o.view(np.float64)) for o in outs])
        return outs

    def swe_rexi_run(self, params, h, u, v, nsteps):
        """params dict: a, g, omega, h0, visc, rexi_h, rexi_M, half, dt; h/u/v spectral."""
        p = np.array([params["a"], params["g"], params["omega"], params["h0"], params["visc"],
                      params["rexi_h"], float(params["rexi_M"]), float(params["half"]), params["dt"],
                      float(params.get("robert", 0))])
        ins = [np.ascontiguousarray(x, dtype=np.complex128) for x in (h, u, v)]
        outs = [np.empty(self.nlm, dtype=np.complex128) for _ in range(3)]
        lib().orc_swe_rexi_run(self.h, _d(p), *[_d(x.view(np.float64)) for x in ins], nsteps,
                               *[_d(o.view(np.float64)) for o in outs])
        return outs

    def rexi_pole_solve(self, params, alpha, phi0, u0, v0):
        """One SWERexiSPH pole (beta = 1, dt = 1, complex result before the real part).
        params dict: a, g, omega, h0; phi0/u0/v0 real spectral.  Returns complex
        physical phi, u, v of shape (nphi, nlat)."""
        p = np.array([params["a"], params["g"], params["omega"], params["h0"]])
        ins = [np.ascontiguousarray(x, dtype=np.complex128) for x in (phi0, u0, v0)]
        outs = [np.empty(self.nspat, dtype=np.complex128) for _ in range(3)]
        lib().orc_rexi_pole_solve(self.h, _d(p), float(np.real(alpha)), float(np.imag(alpha)),
                                  *[_d(x.view(np.float64)) for x in ins],
                                  *[_d(o.view(np.float64)) for o in outs])
        return [o.reshape(self.nphi, self.nlat) for o in outs]


class Matrix:
    """SPHMatrixComplex + SPHSolverComplex restatement (row-band storage)."""

    def __init__(self, orc, halo):
        self.orc, self.halo = orc, halo
        self.h = lib().orc_mat_create(orc.h, halo)

    def __del__(self):
        try:
            lib().orc_mat_destroy(self.h)
        except Exception:
            pass

    # component ids: 0 scalar, 1 mu, 2 (1-mu^2) d/dmu, 11..18 rexi z1..z8
    def add(self, comp, scalar=1.0, r=1.0):
        s = complex(scalar)
        lib().orc_mat_add(self.h, comp, s.real, s.imag, r)

    def data(self):
        nd = 2 * self.halo + 1
        out = np.empty(self.orc.ncplx * nd, dtype=np.complex128)
        lib().orc_mat_get(self.h, _d(out.view(np.float64)))
        return out.reshape(self.orc.ncplx, nd)

    def solve(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.complex128)
        out = np.empty_like(rhs)
        info = lib().orc_mat_solve(self.h, _d(rhs.view(np.float64)), _d(out.view(np.float64)))
        return out, info


def rexi_coeffs(h, M, half=True):
    n = 2 * (M + 11) + 1
    a = np.empty(2 * n)
    b = np.empty(2 * n)
    np_ = lib().orc_rexi_coeffs(h, M, int(half), _d(a), _d(b))
    a = a[: 2 * np_].view(np.complex128)
    b = b[: 2 * np_].view(np.complex128)
    return a.copy(), b.copy()


def gbsv(AB, kl, ku, b):
    """Banded LU with partial pivoting; AB in LAPACK band layout (ldab, n)."""
    AB = np.asfortranarray(AB, dtype=np.complex128).copy(order="F")
    ldab, n = AB.shape
    b = np.ascontiguousarray(b, dtype=np.complex128).copy()
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().orc_gbsv(n, kl, ku, AB.ctypes.data_as(_dp), ldab, ipiv.ctypes.data_as(_ip),
                          _d(b.view(np.float64)))
    return b, ipiv, info


def identities(n, m):
    out = np.empty(9)
    lib().orc_identities(float(n), float(m), _d(out))
    return dict(zip("RSABCDEGH", out))


def ref_available():
    return os.path.exists(_REF_DUMP)


def ref_rexi(h, M, half=True):
    """REXI coefficients from the reference's own headers compiled into oracle/_ref."""
    out = subprocess.run([_REF_DUMP, "rexi", repr(h), str(M), str(int(half))],
                         check=True, capture_output=True, text=True).stdout
    arr = np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()])
    return arr[:, 0] + 1j * arr[:, 1], arr[:, 2] + 1j * arr[:, 3]


def ref_identities(L):
    out = subprocess.run([_REF_DUMP, "ident", str(L)], check=True, capture_output=True, text=True).stdout
    return np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()])


def set_threads(n):
    """Threads of the oracle's OpenMP loops (the CPU baseline times 1 and all)."""
    lib().orc_set_threads(int(n))
