"""Shared helpers of the test-suite: seeded inputs and error norms."""
import numpy as np

A_EARTH, G, OMEGA, H0 = 6.37122e6, 9.80616, 7.292e-5, 10000.0


def rel_l2(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    n = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (n if n > 0 else 1.0)


def rel_linf(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    n = np.max(np.abs(ref))
    return np.max(np.abs(x - ref)) / (n if n > 0 else 1.0)


def real_spec_index(L):
    """(l, m) of every real spectral index (m-major)."""
    ls, ms = [], []
    for m in range(L + 1):
        for l in range(m, L + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


def cplx_spec_index(L):
    ls, ms = [], []
    for l in range(L + 1):
        for m in range(-l, l + 1):
            ls.append(l)
            ms.append(m)
    return np.array(ls), np.array(ms)


def random_real_spec(L, seed, power=-1.5, zero_l0=True):
    """Seeded coefficients with amplitude ~ l^power, uniform phase, m = 0 real."""
    rng = np.random.default_rng(seed)
    l, m = real_spec_index(L)
    amp = np.where(l > 0, np.maximum(l, 1).astype(float) ** power, 0.0 if zero_l0 else 1.0)
    ph = rng.uniform(0, 2 * np.pi, l.size)
    q = amp * np.exp(1j * ph) * rng.uniform(0.5, 1.5, l.size)
    q[m == 0] = q[m == 0].real
    return q.astype(np.complex128)


def random_cplx_spec(L, seed, power=-1.5):
    rng = np.random.default_rng(seed)
    l, m = cplx_spec_index(L)
    amp = np.maximum(l, 1).astype(float) ** power
    return (amp * (rng.standard_normal(l.size) + 1j * rng.standard_normal(l.size))).astype(np.complex128)


def gaussian_dam(lon, ct, lat_c=np.pi / 4, lon_c=np.pi, amp=500.0, width=0.1, h0=H0):
    """h = h0 + amp*exp(-(d/width)^2), d the great-circle angle (d*a / (a*width))."""
    phi1 = np.arcsin(ct)[None, :]
    lam = np.asarray(lon)[:, None]
    c = np.sin(phi1) * np.sin(lat_c) + np.cos(phi1) * np.cos(lat_c) * np.cos(lam - lon_c)
    d = np.arccos(np.clip(c, -1.0, 1.0))
    return (h0 + amp * np.exp(-(d / width) ** 2)).reshape(-1)


def swe_prm(dt, visc=1e5, omega=OMEGA, nonlinear=1):
    return dict(a=A_EARTH, g=G, omega=omega, h0=H0, visc=visc, nonlinear=nonlinear, dt=dt)
