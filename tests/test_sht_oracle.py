"""Pins of the SHT oracle (oracle/sht_oracle.c), independent of it.

Ring geometry: every pixel centre against Gorski et al. (2005) eqs. 4-9 as written in test_oracle.py;
Legendre functions: scipy.special.sph_harm_y (l <= 80) and Unsold's theorem
sum_m |Y_lm|^2 = (2l+1)/(4 pi) to l = 3000, polar rings included; map2alm: the defining sum
(4 pi/Npix) sum_p f_p Y*_lm(p) with scipy's Y_lm at every pixel; alm2cl: the defining sum over m;
alm2map: scipy's Y_lm at every pixel and the adjoint identity with the pinned map2alm; the beam window:
the reference's FWHM2SIGMA literal; the JoinT noise: its distribution (variance sigma2 / 2, the
reference's modulus-times-cosine draw) and its independence across pixels and channels.
"""
import math

import numpy as np
import pytest
from scipy.special import sph_harm_y

import oracle
from test_oracle import pix2ang_ring


def pixel_angles(nside):
    th, ph = [], []
    for p in range(12 * nside * nside):
        z, phi = pix2ang_ring(nside, p)
        th.append(math.acos(z))
        ph.append(phi)
    return np.array(th), np.array(ph)


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 8])
def test_ring_info_matches_pixel_centres(nside):
    seen = 0
    for ir in range(1, 4 * nside):
        nphi, first, z, st, s = oracle.ring_info(nside, ir)
        assert first == seen
        assert abs(st - math.sqrt(max(0.0, 1 - z * z))) < 1e-15
        for j in range(nphi):
            zc, phic = pix2ang_ring(nside, first + j)
            assert abs(zc - z) < 1e-15
            assert abs(phic - math.pi * (2 * j + s) / nphi) < 1e-13
        seen += nphi
    assert seen == 12 * nside * nside


def test_legendre_vs_scipy():
    rng = np.random.default_rng(2)
    lmax = 80
    for theta in np.r_[rng.uniform(0, np.pi, 6), 1e-3, np.pi / 2]:
        x, st = math.cos(theta), math.sin(theta)
        for m in (0, 1, 2, 7, 30, 80):
            lam = oracle.legendre(lmax, m, x, st)
            ref = np.array([sph_harm_y(l, m, theta, 0.0).real for l in range(m, lmax + 1)])
            assert np.abs(lam - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("lmax", [383, 1535, 3071])
def test_legendre_unsold(lmax):
    # polar ring of Nside 1024, a mid-latitude ring and the equator
    for x in (1 - 1 / (3 * 1024.0 ** 2), 0.3, 0.0):
        st = math.sqrt((1 - x) * (1 + x))
        tot = np.zeros(lmax + 1)
        for m in range(lmax + 1):
            lam = oracle.legendre(lmax, m, x, st)
            tot[m:] += (1 if m == 0 else 2) * lam ** 2
        ls = np.arange(lmax + 1)
        assert np.abs(tot * 4 * np.pi / (2 * ls + 1) - 1).max() < 1e-11


@pytest.mark.parametrize("nside,lmax", [(2, 7), (4, 11), (4, 16)])
def test_map2alm_defining_sum(nside, lmax):
    rng = np.random.default_rng(nside + lmax)
    npix = 12 * nside * nside
    maps = rng.normal(size=(2, npix))
    got = oracle.map2alm(maps, nside, lmax)
    th, ph = pixel_angles(nside)
    for l in range(lmax + 1):
        for m in range(l + 1):
            y = sph_harm_y(l, m, th, ph)
            ref = (4 * np.pi / npix) * (maps * np.conj(y)).sum(axis=1)
            i = oracle.alm_index(l, m, lmax)
            assert np.abs(got[:, i] - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_alm_index_layout():
    lmax = 9
    idx = [oracle.alm_index(l, m, lmax) for m in range(lmax + 1) for l in range(m, lmax + 1)]
    assert idx == list(range((lmax + 1) * (lmax + 2) // 2))


def test_alm2cl_defining_sum():
    rng = np.random.default_rng(8)
    lmax, nm = 12, 3
    nalm = (lmax + 1) * (lmax + 2) // 2
    alms = rng.normal(size=(nm, nalm)) + 1j * rng.normal(size=(nm, nalm))
    for l in range(lmax + 1):  # m = 0 coefficients of real maps are real
        alms[:, oracle.alm_index(l, 0, lmax)] = alms[:, oracle.alm_index(l, 0, lmax)].real
    cl = oracle.alm2cl(alms, lmax)
    for a in range(nm):
        for b in range(nm):
            for l in range(lmax + 1):
                s = sum((alms[a, oracle.alm_index(l, m, lmax)] * np.conj(alms[b, oracle.alm_index(l, m, lmax)])
                         ).real * (1 if m == 0 else 2) for m in range(l + 1))
                assert abs(cl[a, b, l] - s / (2 * l + 1)) < 1e-12 * abs(s)


def _real_alms(rng, nm, lmax):
    nalm = (lmax + 1) * (lmax + 2) // 2
    a = rng.normal(size=(nm, nalm)) + 1j * rng.normal(size=(nm, nalm))
    for l in range(lmax + 1):   # a real map has real a_l0
        a[:, oracle.alm_index(l, 0, lmax)] = a[:, oracle.alm_index(l, 0, lmax)].real
    return a


@pytest.mark.parametrize("nside,lmax", [(1, 3), (2, 7), (4, 11), (4, 16), (3, 9)])
def test_alm2map_defining_sum(nside, lmax):
    """f(p) = sum_l a_l0 Y_l0 + 2 Re sum_{m>0} a_lm Y_lm(p) with scipy's Y_lm at every pixel centre."""
    rng = np.random.default_rng(100 + nside + lmax)
    alms = _real_alms(rng, 2, lmax)
    got = oracle.alm2map(alms, nside, lmax)
    th, ph = pixel_angles(nside)
    ref = np.zeros_like(got)
    for l in range(lmax + 1):
        for m in range(l + 1):
            y = sph_harm_y(l, m, th, ph)
            term = (alms[:, oracle.alm_index(l, m, lmax)][:, None] * y[None, :]).real
            ref += term if m == 0 else 2 * term
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("nside,lmax", [(8, 23), (16, 47), (16, 20)])
def test_alm2map_adjoint_of_map2alm(nside, lmax):
    """(4 pi/Npix) <f, alm2map(a)> = sum_lm c_m Re(a_lm conj map2alm(f)_lm), c_0 = 1, c_m = 2: the two
    transforms are adjoint, so a slip in either (a phase sign, a dropped m, a parity) breaks it."""
    rng = np.random.default_rng(nside * lmax)
    npix = 12 * nside * nside
    alms = _real_alms(rng, 1, lmax)
    f = rng.normal(size=(1, npix))
    lhs = (4 * np.pi / npix) * (f * oracle.alm2map(alms, nside, lmax)).sum()
    af = oracle.map2alm(f, nside, lmax)
    cm = np.concatenate([np.full(lmax - m + 1, 1.0 if m == 0 else 2.0) for m in range(lmax + 1)])
    rhs = (cm * (alms[0] * np.conj(af[0])).real).sum()
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_beam_window_constant():
    """he_generate_beam_window: the reference's FWHM2SIGMA literal (src/healpix_extra.c:389) equals
    (pi/10800)/sqrt(8 ln 2); b_l = exp(-l(l+1) sigma^2/2) has b_0 = 1 and halves near the expected l."""
    fwhm = 30.0
    b = oracle.beam_window(3000, fwhm)
    sigma = 0.00012352884853326381 * fwhm
    ls = np.arange(3001)
    assert np.abs(b - np.exp(-0.5 * ls * (ls + 1) * sigma * sigma)).max() < 1e-15
    assert b[0] == 1.0
    lh = np.searchsorted(-b, -0.5)
    assert abs(lh * (lh + 1) * sigma * sigma - 2 * np.log(2)) < 2 * (2 * lh + 2) * sigma * sigma


def test_noise_distribution():
    """JoinT's draw sqrt(-sigma2 ln(1-u2)) cos(2 pi u1) is Gaussian with variance sigma2/2 (the modulus
    has E[mod^2] = sigma2, E[cos^2] = 1/2); channels and pixels are independent draws."""
    from scipy import stats
    npix, s2 = 200000, np.array([4.0, 0.25])
    n = oracle.add_noise(np.zeros((2, npix)), s2, 77)
    for c in range(2):
        z = n[c] / np.sqrt(s2[c] / 2)
        assert abs(z.mean()) < 5 / np.sqrt(npix)
        assert abs(z.var() - 1) < 5 * np.sqrt(2 / npix)
        assert stats.kstest(z, "norm").pvalue > 1e-3
    r = np.corrcoef(n[0] / 2, n[1] / 0.5)[0, 1]
    assert abs(r) < 5 / np.sqrt(npix)
    assert abs(np.corrcoef(n[0][:-1], n[0][1:])[0, 1]) < 5 / np.sqrt(npix)
    # same seed, same noise; another seed, other noise; the input map is added to, not replaced
    assert np.array_equal(oracle.add_noise(np.zeros((2, npix)), s2, 77), n)
    assert not np.allclose(oracle.add_noise(np.zeros((2, npix)), s2, 78), n)
    assert np.array_equal(oracle.add_noise(np.ones((2, npix)), s2, 77), n + 1)
