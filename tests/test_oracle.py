"""Pins of the FP64 oracle (oracle/gethi_oracle.c) that do not depend on the oracle itself.

FFT: brute-force DFT and numpy.fft (pocketfft), including the c2r treatment of non-Hermitian kx=0,N/2
planes; Philox: Random123's known-answer vectors; vec2pix_ring: pixel centres from Gorski et al.
(2005) eqs. 4-9 written here independently, poles, equal-area counts; channels: brute-force scan;
radial velocity: the exact central difference of single Fourier modes; lognormal: E[exp(g d - g^2 s^2/2)]
= 1; maps: mass conservation and the closed-form mK factor.
"""
import math

import numpy as np
import pytest

import crime_amd
import oracle
from conftest import small_params


# ---------------------------------------------------------------------------------------------
# FFT
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("n", [4, 6, 8, 12])
def test_ifft3d_bruteforce(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))
    got = oracle.ifft3d(a)
    j = np.arange(n)
    W = np.exp(2j * np.pi * np.outer(j, j) / n)
    ref = np.einsum("zk,yj,xi,kji->zyx", W, W, W, a)  # sum_k a[kz,ky,kx] e^{+i k.x}
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [16, 24, 48])
def test_ifft3d_vs_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))
    got = oracle.ifft3d(a)
    ref = np.fft.ifftn(a) * n ** 3
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [8, 12, 32])
def test_c2r_non_hermitian_planes(n):
    # the reference fills kx=0 and kx=N/2 without Hermitian symmetry (fourier.c:275-299); a c2r
    # transform uses their Hermitian part, as numpy's irfftn does
    rng = np.random.default_rng(7 + n)
    h = rng.normal(size=(n, n, n // 2 + 1)) + 1j * rng.normal(size=(n, n, n // 2 + 1))
    got = oracle.c2r_3d(h)
    ref = np.fft.irfftn(h, s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


# ---------------------------------------------------------------------------------------------
# Philox-4x32-10 known answers (Random123 kat_vectors)
# ---------------------------------------------------------------------------------------------
def test_philox_known_answers():
    assert oracle.philox([0, 0, 0, 0], 0, 0) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, 0xFFFFFFFF, 0xFFFFFFFF) == [0x408F276D, 0x41C83B0E,
                                                                      0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], 0xA4093822, 0x299F31D0) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ---------------------------------------------------------------------------------------------
# HEALPix RING pixelisation
# ---------------------------------------------------------------------------------------------
def pix2ang_ring(nside, p):
    """Centre (z, phi) of RING pixel p (Gorski et al. 2005, sec. 4), written from the paper."""
    npix = 12 * nside * nside
    ncap = 2 * nside * (nside - 1)
    if p < ncap:  # north polar cap
        i = int((1 + math.isqrt(1 + 2 * p)) // 2)
        while 2 * i * (i - 1) > p:
            i -= 1
        while 2 * (i + 1) * i <= p:
            i += 1
        j = p - 2 * i * (i - 1) + 1
        z = 1 - i * i / (3.0 * nside * nside)
        phi = math.pi / (2 * i) * (j - 0.5)
    elif p < npix - ncap:  # equatorial belt
        q = p - ncap
        i = q // (4 * nside) + nside
        j = q % (4 * nside) + 1
        shift = 1.0 if (i + nside) % 2 else 0.5  # rings alternate between phi = 0 and a half step
        z = 4.0 / 3.0 - 2.0 * i / (3.0 * nside)
        phi = math.pi / (2 * nside) * (j - shift)
    else:  # south polar cap
        q = npix - p - 1
        i = int((1 + math.isqrt(1 + 2 * q)) // 2)
        while 2 * i * (i - 1) > q:
            i -= 1
        while 2 * (i + 1) * i <= q:
            i += 1
        j = 4 * i - (q - 2 * i * (i - 1))
        z = -1 + i * i / (3.0 * nside * nside)
        phi = math.pi / (2 * i) * (j - 0.5)
    return z, phi


@pytest.mark.parametrize("nside", [1, 2, 3, 4, 8, 16, 32])
def test_vec2pix_pixel_centres(nside):
    for p in range(12 * nside * nside):
        z, phi = pix2ang_ring(nside, p)
        s = math.sqrt(max(0.0, 1 - z * z))
        r = 3.7  # non-unit vectors are accepted
        got = oracle.vec2pix_ring(nside, r * s * math.cos(phi), r * s * math.sin(phi), r * z)
        assert got == p, (nside, p, got)


def test_vec2pix_poles_and_range():
    for nside in (1, 4, 512):
        assert oracle.vec2pix_ring(nside, 0.0, 0.0, 1.0) == 0
        assert oracle.vec2pix_ring(nside, 0.0, 0.0, -2.0) == 12 * nside * nside - 4
    rng = np.random.default_rng(3)
    v = rng.normal(size=(20000, 3))
    for nside in (7, 64):
        pix = np.array([oracle.vec2pix_ring(nside, *x) for x in v])
        assert pix.min() >= 0 and pix.max() < 12 * nside * nside


def test_vec2pix_equal_area():
    nside = 4
    npix = 12 * nside * nside
    rng = np.random.default_rng(11)
    v = rng.normal(size=(200000, 3))
    pix = np.array([oracle.vec2pix_ring(nside, *x) for x in v])
    counts = np.bincount(pix, minlength=npix)
    exp = len(v) / npix
    chi2 = ((counts - exp) ** 2 / exp).sum()
    assert chi2 < npix + 6 * math.sqrt(2 * npix)  # 6-sigma bound of a chi^2 with npix dof


def test_vec2pix_nearest_centre():
    # a direction belongs to a pixel whose centre is within about one pixel size of it
    nside = 8
    rng = np.random.default_rng(5)
    v = rng.normal(size=(3000, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    size = math.sqrt(4 * math.pi / (12 * nside * nside))
    for x in v:
        p = oracle.vec2pix_ring(nside, *x)
        z, phi = pix2ang_ring(nside, p)
        s = math.sqrt(1 - z * z)
        c = np.array([s * math.cos(phi), s * math.sin(phi), z])
        assert math.acos(min(1.0, float(np.dot(c, x)))) < 1.5 * size


# ---------------------------------------------------------------------------------------------
# channels, velocity, lognormal, maps
# ---------------------------------------------------------------------------------------------
def test_get_inu_walk_equals_scan(cfgdir):
    par = small_params(cfgdir, "inu", n_nu=37)
    d = par.desc()
    nu0 = np.ctypeslib.as_array(d.nu0, shape=(d.n_nu,))
    nuf = np.ctypeslib.as_array(d.nuf, shape=(d.n_nu,))
    rng = np.random.default_rng(1)
    vals = np.r_[rng.uniform(nu0[0] - 20, nuf[-1] + 20, 3000), nu0, nuf]
    for nu in vals:
        scan = [i for i in range(d.n_nu) if nu0[i] <= nu < nuf[i]]
        exp = scan[0] if scan else (-1 if nu < nu0[0] else d.n_nu)
        for start in (0, d.n_nu - 1, int(rng.integers(0, d.n_nu))):
            assert oracle.get_inu(d, float(nu), start) == exp


def test_uniform_channels_truncation(cfgdir):
    # (int) truncation of the non-table build: frequencies just below nu_min land in channel 0
    par = small_params(cfgdir, "unif", n_nu=10, nu_lo=1150.0, nu_hi=1250.0, uniform_nu=True)
    d = par.desc()
    assert oracle.get_inu(d, 1149.5, 0) == 0
    assert oracle.get_inu(d, 1140.5, 0) == 0   # t = -0.95 truncates to 0
    assert oracle.get_inu(d, 1139.5, 0) == -1  # t = -1.05
    assert oracle.get_inu(d, 1249.99, 0) == 9
    assert oracle.get_inu(d, 1250.0, 0) == -1


def test_radial_velocity_single_mode(cfgdir):
    n = 16
    par = small_params(cfgdir, "vr", n_grid=n)
    d = par.desc()
    L = d.l_box
    dx = L / n
    x = dx * np.arange(n)
    m = (2, 3, 1)  # mode numbers along x, y, z
    k = [2 * np.pi * mi / L for mi in m]
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    phase = k[0] * X + k[1] * Y + k[2] * Z
    vpot = np.cos(phase)
    got = oracle.radial_velocity(d, vpot)
    # central difference of cos: -sin(phase) sin(k dx)/dx per axis (exact)
    g = [-np.sin(phase) * math.sin(ki * dx) / dx for ki in k]
    cx_, cy_, cz_ = [dx * (np.arange(n) + 0.5) - d.pos_obs[i] for i in range(3)]
    CZ, CY, CX = np.meshgrid(cz_, cy_, cx_, indexing="ij")
    rr = np.sqrt(CX ** 2 + CY ** 2 + CZ ** 2)
    ref = (g[0] * CX + g[1] * CY + g[2] * CZ) / rr
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_sigma2_vs_numpy():
    rng = np.random.default_rng(2)
    a = rng.normal(0.3, 1.7, size=(24, 24, 24))
    s2, m = oracle.sigma2(a)
    assert abs(m - a.mean()) < 1e-14
    assert abs(s2 - a.var()) < 1e-12


def test_lognormal_mean_and_zero_field(cfgdir):
    n = 32
    par = small_params(cfgdir, "ln", n_grid=n, bias_model="const", bias_b0=1.3)
    d = par.desc()
    dx = d.l_box / n
    rng = np.random.default_rng(4)
    s2 = 0.6
    dens = rng.normal(0, math.sqrt(s2), size=(n, n, n))
    zero = np.zeros((n, n, n))
    mass0, dz0 = oracle.get_hi(d, zero, zero, 0.0)
    mass, _ = oracle.get_hi(d, dens, zero, s2)
    assert np.all(dz0 == 0)
    # delta = 0, sigma^2 = 0: mass = dx^3 x_HI(z(r)) exactly
    c = dx * (np.arange(n) + 0.5) - d.pos_obs[0]
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    z = np.array([oracle.z_of_r(d, float(v)) for v in r.ravel()]).reshape(r.shape)
    assert np.allclose(mass0, dx ** 3 * 0.008 * (1 + z) ** 0.6, rtol=1e-14, atol=0)
    # E[exp(g d - g^2 s^2 / 2)] = 1 for Gaussian d of variance s^2 (g = D(r) b)
    ratio = mass / mass0
    assert abs(ratio.mean() - 1) < 6 * ratio.std() / math.sqrt(ratio.size)


def test_maps_mass_conservation_and_factor(cfgdir):
    par = small_params(cfgdir, "cons", n_grid=32, n_side=16)
    d = par.desc()
    dens, vpot, s2, _ = oracle.fields(d)
    rv = oracle.radial_velocity(d, vpot)
    mass, dz = oracle.get_hi(d, dens, rv, s2)
    maps = oracle.mk_T_maps(d, mass, dz)
    pref = oracle.channel_factors(d)
    inu, ipix = oracle.sample_indices(d, dz, 0, 32 ** 3)
    msub = (mass.ravel() / d.n_sub)[:, None] * np.ones((1, d.n_sub))
    ok = inu >= 0
    per_ch = np.bincount(inu[ok], weights=msub[ok], minlength=d.n_nu)
    assert np.allclose((maps / pref[:, None]).sum(axis=1), per_ch, rtol=1e-12, atol=0)
    # mK factor 90.057156 Omega_b h Npix / (4 pi) / (r(nu_c)^2 dnu) (pixelize.c:155, 246-256)
    nu0 = np.ctypeslib.as_array(d.nu0, shape=(d.n_nu,))
    nuf = np.ctypeslib.as_array(d.nuf, shape=(d.n_nu,))
    npix = 12 * d.n_side ** 2
    h = crime_amd._ffi.host_lib()
    for i in (0, d.n_nu // 2, d.n_nu - 1):
        r = h.gh_r_of_z(par._p, 1420.40575177 / (0.5 * (nu0[i] + nuf[i])) - 1)
        exp = 90.057156 * d.omega_b * d.hhub * npix / (4 * math.pi) / (r * r * (nuf[i] - nu0[i]))
        assert abs(pref[i] - exp) < 1e-13 * exp


def test_power_spectrum_of_oracle_fields(cfgdir):
    """The oracle's Gaussian field carries the input P(k) with the estimator's normalisation
    |DFT|^2 L^3/N^6: shell means within cosmic variance of <P>, and the two mode counts (folded FFT
    indices, signed lattice enumeration) agree exactly."""
    from pk_ref import pk_expect, pk_model_grid, pk_numpy, zscores

    par = small_params(cfgdir, "pk_orc", n_grid=64, l_box=1200.0)
    d = par.desc()
    dens, _, _, _ = oracle.fields(d)
    dk = 2 * np.pi / d.l_box
    nb = 30
    k, pk, cnt = pk_numpy(dens, d.l_box, nb, dk)
    ecnt, ep = pk_expect(d, nb, dk)
    assert np.array_equal(cnt, ecnt)
    z = zscores(pk, cnt, ep)
    assert np.abs(z).max() < 5.5
    assert abs(z.mean()) < 4 / np.sqrt(len(z))
    assert np.all((k[1:] >= np.arange(1, nb) * dk) & (k[1:] < np.arange(2, nb + 1) * dk))
    # per-mode power over the input P: 1 off the kx = 0, N/2 planes and 1/2 on them (the real inverse
    # transform keeps the Hermitian part there); ~4000 and ~250000 modes give ~2% and ~0.3% errors
    f = np.fft.fftn(dens)
    ratio = (np.abs(f) ** 2 * d.l_box ** 3 / 64.0 ** 6)[1:, 1:, :] / pk_model_grid(d)[1:, 1:, :]
    plane = np.zeros(ratio.shape, bool)
    plane[..., 0] = plane[..., 32] = True
    assert abs(ratio[plane].mean() - 0.5) < 0.04
    assert abs(ratio[~plane].mean() - 1.0) < 0.01


def test_pk_numpy_bruteforce():
    """pk_numpy against an explicit DFT sum on a 6^3 grid."""
    from pk_ref import pk_numpy

    rng = np.random.default_rng(3)
    n, L = 6, 100.0
    g = rng.normal(size=(n, n, n))
    x = np.arange(n)
    e = np.exp(-2j * np.pi * np.outer(x, x) / n)
    f = np.zeros((n, n, n), complex)
    for a in range(n):
        for b in range(n):
            for c in range(n):
                f[a, b, c] = np.sum(g * e[a][:, None, None] * e[b][None, :, None] * e[c][None, None, :])
    p = np.abs(f) ** 2 * L ** 3 / n ** 6
    k, pk, cnt = pk_numpy(g, L, 6, 2 * np.pi / L)
    # shell 1 = |k| in [1, 2) dk: the 6 axis modes, 12 face diagonals (sqrt 2) and 8 corners (sqrt 3)
    assert cnt[1] == 26
    idx = [(a, b, c) for a in range(n) for b in range(n) for c in range(n)
           if 1 <= math.sqrt(sum(min(v, n - v) ** 2 for v in (a, b, c))) < 2]
    assert np.isclose(pk[1], np.mean([p[i] for i in idx]), rtol=1e-12)
