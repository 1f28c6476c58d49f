"""Pins of the CPU oracle's grid, Legendre functions, FFT and SHT pair.

Each check compares the oracle with an independent source: numpy's Gauss-
Legendre rule, scipy's spherical harmonics, numpy.fft, closed forms of the
reference's test solutions (src/include/benchmarks/SphereTestSolutions_SPH.hpp:
25-91) and exact identities of the transform pair.
"""
import numpy as np
import pytest
from scipy.special import sph_harm_y

from oracle import oracle as O
from tests._util import random_cplx_spec, random_real_spec, rel_l2


def test_auto_grid_matches_published_grids():
    # grids quoted for T127, T511, T1023 plus the rule's T64 / T256 values
    assert O.auto_grid(127) == (192, 384)
    assert O.auto_grid(511) == (768, 1536)
    assert O.auto_grid(1023) == (1536, 3072)
    assert O.auto_grid(64) == (98, 196)
    assert O.auto_grid(256) == (386, 784)


@pytest.mark.parametrize("n", [8, 98, 192, 386])
def test_gauss_nodes_vs_numpy(n):
    o = O.Oracle(5, n, 16)
    x, w = np.polynomial.legendre.leggauss(n)
    # north first: ct decreasing
    assert np.all(np.diff(o.ct) < 0)
    assert np.max(np.abs(o.ct - x[::-1])) < 2e-15
    assert np.max(np.abs(o.w - w[::-1])) < 3e-14
    assert abs(o.w.sum() - 2.0) < 1e-13
    # the rule integrates mu^k exactly for k <= 2n-1
    for k in range(0, min(2 * n - 1, 60), 2):
        assert abs(np.sum(o.w * o.ct ** k) - 2.0 / (k + 1)) < 1e-13


def test_plm_vs_scipy():
    L = 40
    o = O.Oracle(L, 64, 128)
    tab = o.plm_table()
    th = np.arccos(o.ct)
    k = 0
    worst = 0.0
    for m in range(L + 1):
        for j in range(o.nlat):
            col = tab[k:k + L - m + 1]
            k += L - m + 1
            ref = sph_harm_y(np.arange(m, L + 1), m, th[j], 0.0).real
            worst = max(worst, np.max(np.abs(col - ref)))
    assert worst < 1e-13


def test_plm_high_degree_orthonormality():
    # Gauss quadrature is exact for Pbar_l Pbar_l' (degree <= 2L < 2 nlat): 2 pi sum w P P = delta
    L = 511
    nlat = 768
    P = np.empty(L + 1)
    o = O.Oracle(L, nlat, 1536)
    for m in (0, 1, 255, 400, 511):
        M = np.zeros((nlat, L + 1 - m))
        for j in range(nlat):
            O.lib().orc_plm_column(L, m, o.ct[j], o.st[j], P.ctypes.data_as(O._dp))
            M[j] = P[m:]
        G = 2 * np.pi * (M.T * o.w) @ M
        assert np.max(np.abs(G - np.eye(L + 1 - m))) < 1e-12


@pytest.mark.parametrize("nphi", [196, 384, 784, 1536, 30, 3072])
def test_fft_vs_numpy(nphi):
    o = O.Oracle(5, 8, nphi)
    rng = np.random.default_rng(nphi)
    x = rng.standard_normal(nphi) + 1j * rng.standard_normal(nphi)
    assert np.max(np.abs(o.fft(x, -1) - np.fft.fft(x))) < 1e-12 * np.sqrt(nphi)
    assert np.max(np.abs(o.fft(x, +1) - np.fft.ifft(x) * nphi)) < 1e-12 * np.sqrt(nphi)


def _closed_forms(n, m, lon, mu):
    # Re(Y_n^m) closed forms (SphereTestSolutions_SPH.hpp:25-91), orthonormal with
    # Condon-Shortley phase
    s = np.sqrt(1 - mu * mu)
    f = {
        (1, 0): np.sqrt(6.0) * mu / 2.0,
        (2, 0): np.sqrt(10.0) / 4.0 * (3 * mu * mu - 1.0),
        (2, 1): -mu / 2.0 * np.sqrt(-15.0 * mu * mu + 15.0),
        (2, 2): np.sqrt(15.0) / 4.0 * (-mu * mu + 1.0),
        (3, 0): np.sqrt(14.0) / 4.0 * mu * (5 * mu * mu - 3.0),
        (3, 1): np.sqrt(-42.0 * mu * mu + 42.0) * (-5.0 * mu * mu / 8.0 + 1.0 / 8.0),
        (3, 2): np.sqrt(105.0) * mu / 4.0 * (-mu * mu + 1.0),
        (3, 3): -np.sqrt(70.0) / 8.0 * s ** 3,
    }[(n, m)]
    return f * np.cos(lon * m) / np.sqrt(2 * np.pi)


@pytest.mark.parametrize("nm", [(1, 0), (2, 0), (2, 1), (2, 2), (3, 0), (3, 1), (3, 2), (3, 3)])
def test_synthesis_closed_forms(nm):
    n, m = nm
    o = O.Oracle(16)
    q = np.zeros(o.nlm, complex)
    # real field: f = q_n0 Y_n0 + 2 Re(q_nm Y_nm)
    q[o.idx(n, m)] = 1.0 if m == 0 else 0.5
    f = o.synthesis(q).reshape(o.nphi, o.nlat)
    ref = _closed_forms(n, m, o.lon()[:, None], o.ct[None, :])
    assert np.max(np.abs(f - ref)) < 1e-14


def test_constant_field_convention():
    # operator+(double) adds v*sqrt(4 pi) to (0,0) (SPHData.hpp:427-440): Y_00 = 1/sqrt(4 pi)
    o = O.Oracle(8)
    q = np.zeros(o.nlm, complex)
    q[0] = 3.5 * np.sqrt(4 * np.pi)
    assert np.max(np.abs(o.synthesis(q) - 3.5)) < 1e-14
    assert abs(o.analysis(np.full(o.nspat, 3.5))[0] - 3.5 * np.sqrt(4 * np.pi)) < 1e-13


@pytest.mark.parametrize("L", [16, 31, 64])
def test_real_pair_round_trip_and_scipy(L):
    o = O.Oracle(L)
    q = random_real_spec(L, seed=L)
    f = o.synthesis(q)
    assert rel_l2(o.analysis(f), q) < 1e-13
    # synthesis vs a direct scipy sum
    th = np.arccos(o.ct)[None, :]
    ph = o.lon()[:, None]
    ref = np.zeros((o.nphi, o.nlat))
    for m in range(L + 1):
        for l in range(m, L + 1):
            Y = sph_harm_y(l, m, th, ph)
            c = q[o.idx(l, m)]
            ref += (c * Y).real if m == 0 else 2 * (c * Y).real
    assert np.max(np.abs(f.reshape(o.nphi, o.nlat) - ref)) < 1e-12 * np.max(np.abs(ref))


def test_m0_imaginary_part_ignored():
    o = O.Oracle(12)
    q = random_real_spec(12, seed=3)
    q2 = q.copy()
    for l in range(13):
        q2[o.idx(l, 0)] += 1j * (l + 1)
    assert np.max(np.abs(o.synthesis(q2) - o.synthesis(q))) == 0.0


@pytest.mark.parametrize("L", [16, 33])
def test_complex_pair(L):
    o = O.Oracle(L)
    q = random_cplx_spec(L, seed=L)
    f = o.synthesis_cplx(q)
    assert rel_l2(o.analysis_cplx(f), q) < 1e-13
    th = np.arccos(o.ct)[None, :]
    ph = o.lon()[:, None]
    ref = np.zeros((o.nphi, o.nlat), complex)
    for l in range(L + 1):
        for m in range(-l, l + 1):
            ref += q[o.cidx(l, m)] * sph_harm_y(l, m, th, ph)
    assert np.max(np.abs(f.reshape(o.nphi, o.nlat) - ref)) < 1e-12 * np.max(np.abs(ref))


def test_complex_analysis_of_real_field():
    # a real field's complex spectrum obeys q_{l,-m} = (-1)^m conj(q_lm) and its
    # m >= 0 part equals the real spectrum
    L = 20
    o = O.Oracle(L)
    q = random_real_spec(L, seed=5)
    f = o.synthesis(q)
    qc = o.analysis_cplx(f.astype(complex))
    for l in range(L + 1):
        for m in range(0, l + 1):
            assert abs(qc[o.cidx(l, m)] - q[o.idx(l, m)]) < 1e-13
            assert abs(qc[o.cidx(l, -m)] - (-1) ** m * np.conj(q[o.idx(l, m)])) < 1e-13
