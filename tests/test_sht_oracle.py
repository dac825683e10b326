"""Pins of the SHT oracle (oracle/sht_oracle.c), independent of it.

Ring geometry: every pixel centre against Gorski et al. (2005) eqs. 4-9 as written in test_oracle.py;
Legendre functions: scipy.special.sph_harm_y (l <= 80) and Unsold's theorem
sum_m |Y_lm|^2 = (2l+1)/(4 pi) to l = 3000, polar rings included; map2alm: the defining sum
(4 pi/Npix) sum_p f_p Y*_lm(p) with scipy's Y_lm at every pixel; alm2cl: the defining sum over m.
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
