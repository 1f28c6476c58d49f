"""Seeded inputs of the benchmark workloads, built with the library's own
operators (no host reference code).

- `dam_state`: resting state with a Gaussian dam in height,
  h = h0 + amp * exp(-(d / width)^2), d the great-circle angle to (lat_c, lon_c);
- `seeded_spectra`: complex coefficients with amplitude l^-p, uniform phase,
  l = 0 zeroed and m = 0 real;
- `rexi_state`: the linear REXI start state: seeded (phi, vort, div) spectra mapped
  to (h, u, v) with h = phi / g, psi = Lap^-1 vort, chi = Lap^-1 div and
  u = (grad_lon chi - grad_lat psi) / a, v = (grad_lon psi + grad_lat chi) / a,
  grad_lon / grad_lat as in SPHOperators (src/include/sph/SPHOperators.hpp:100-119,
  339-388): spectral derivative, synthesis, division by sqrt(1 - mu^2), analysis.
"""
import numpy as np

from . import capi

A_EARTH, GRAV, OMEGA, H0 = 6.37122e6, 9.80616, 7.292e-5, 10000.0


def real_degrees(L):
    """Degree l of each real spectral index, m-major (m(2L-m+1)/2 + l)."""
    return np.concatenate([np.arange(m, L + 1) for m in range(L + 1)])


def seeded_spectra(L, nfields, seed, power=1.5):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nfields):
        qs = []
        for m in range(L + 1):
            l = np.arange(m, L + 1)
            amp = np.where(l > 0, np.maximum(l, 1).astype(float) ** -power, 0.0)
            q = amp * np.exp(1j * rng.uniform(0, 2 * np.pi, l.size))
            if m == 0:
                q = q.real.astype(complex)
            qs.append(q)
        out.append(np.concatenate(qs))
    return np.stack(out)


def dam_height(lon, ct, lat_c=np.pi / 4, lon_c=np.pi, amp=500.0, width=0.1, h0=H0):
    phi1 = np.arcsin(ct)[None, :]
    c = np.sin(phi1) * np.sin(lat_c) + np.cos(phi1) * np.cos(lat_c) * np.cos(np.asarray(lon)[:, None] - lon_c)
    d = np.arccos(np.clip(c, -1, 1))
    return (h0 + amp * np.exp(-(d / width) ** 2)).reshape(-1)


def dam_state(ctx, **kw):
    """Spectral (h, u, v) of the Gaussian dam at rest, shape (3, nlm)."""
    lon = np.arange(ctx.nphi) / ctx.nphi * 2 * np.pi
    h = dam_height(lon, ctx.ct, **kw)
    z = np.zeros_like(h)
    return ctx.analysis(np.stack([h, z, z]))


def _grad(ctx, op, spec):
    x = ctx.synthesis(ctx.spec_op(op, spec))
    return ctx.analysis(ctx.spat_op(capi.SPAT_DIV_COSPHI, x))


def rexi_state(ctx, seed=2024, a=A_EARTH, g=GRAV):
    """Spectral (h, u, v) of the seeded (phi, vort, div) start state, shape (3, nlm)."""
    L = ctx.L
    phi, zeta, delta = seeded_spectra(L, 3, seed)
    phi = phi * (g * 100.0)          # ~100 m of height at l = 1
    zeta = zeta * 1e-5
    delta = delta * 1e-6
    l = real_degrees(L).astype(float)
    inv_lap = np.where(l > 0, -(a * a) / np.maximum(l * (l + 1.0), 1.0), 0.0)
    psi, chi = zeta * inv_lap, delta * inv_lap
    gl = lambda x: _grad(ctx, capi.OP_DIFF_LON, x)
    gt = lambda x: _grad(ctx, capi.OP_ONE_MINUS_MU2_DMU, x)
    u = (gl(chi) - gt(psi)) / a
    v = (gl(psi) + gt(chi)) / a
    return np.stack([phi[None, :] / g, u, v]).reshape(3, -1)


def galewsky_state(ctx, seed=None, pert_amp=1.0, a=A_EARTH, g=GRAV, omega=OMEGA, h_avg=10000.0):
    """(h, u, v) spectra of the Galewsky et al. (2004) balanced zonal jet without
    its bump, set up as the reference does (src/include/benchmarks/
    BenchmarkGalewsky.hpp:101-246):
      u(phi) = u_max / e_n * exp(1 / ((phi - phi0)(phi - phi1))) on phi0 < phi < phi1
      (u_max = 80 m/s, phi0 = pi/7, phi1 = pi/2 - phi0, e_n = exp(-4 / (phi1 - phi0)^2)), v = 0;
      g h(lat_j) = - integral from -pi/2 to lat_j of a u (2 Omega sin phi + tan(phi) u / a),
      accumulated from the south pole through the Gauss latitudes, then shifted so
      that its area mean, with weights 2 pi (sin(end) - sin(start)) per interval
      and the value at each interval's end, is h_avg.
    With a seed, a height perturbation with spectrum ~ l^-2 (l = 0 zeroed), scaled
    to a largest value of pert_amp metres on the grid, is added."""
    from scipy.integrate import quad

    phi0 = np.pi / 7.0
    phi1 = np.pi / 2.0 - phi0
    u_max = 80.0
    en = np.exp(-4.0 / (phi1 - phi0) ** 2)

    def u_of(phi):
        if phi <= phi0 or phi >= phi1:
            return 0.0
        return u_max / en * np.exp(1.0 / ((phi - phi0) * (phi - phi1)))

    def integrand(phi):
        up = u_of(phi)
        return a * up * (2.0 * omega * np.sin(phi) + (np.tan(phi) / a) * up)

    def integral(s, e):
        # the integrand vanishes outside (phi0, phi1) and is smooth inside
        s2, e2 = max(s, phi0), min(e, phi1)
        if e2 <= s2:
            return 0.0
        return quad(integrand, s2, e2, epsabs=1e-13, epsrel=1e-13, limit=200)[0]

    lat = np.arcsin(np.asarray(ctx.ct))   # j = 0 northernmost
    nlat = lat.size
    hg = np.empty(nlat)
    hg_sum = area = 0.0
    start = -np.pi / 2.0
    prev = 0.0
    for j in range(nlat - 1, -1, -1):
        end = lat[j]
        prev = prev - integral(start, end)
        hg[j] = prev
        mterm = (np.sin(end) - np.sin(start)) * 2.0 * np.pi
        hg_sum += prev * mterm
        area += mterm
        start = end
    last = prev - integral(start, np.pi / 2.0)
    mterm = (np.sin(np.pi / 2.0) - np.sin(start)) * 2.0 * np.pi
    hg_sum += last * mterm
    area += mterm
    h_lat = hg / g + (h_avg - (hg_sum / g) / area)
    u_lat = np.array([u_of(p) for p in lat])
    nphi = ctx.nphi
    h = np.tile(h_lat, nphi)      # [i * nlat + j]
    u = np.tile(u_lat, nphi)
    v = np.zeros_like(h)
    x = ctx.analysis(np.stack([h, u, v]))
    if seed is not None:
        q = seeded_spectra(ctx.L, 1, seed, power=1.0)   # amplitude l^-1: spectrum ~ l^-2
        pert = ctx.synthesis(q)[0]
        x[0] += q[0] * (pert_amp / np.max(np.abs(pert)))
    return x
