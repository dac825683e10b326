"""HEALPix synthesis (alm2map) and JoinT's beam, noise and mask stages on the GPU (include/crime_sht.h)
against the FP64 oracle (oracle/sht_oracle.c: alm2map by the defining sum, the beam window, the Philox
noise draw), and size-independent properties at the production Nside.

Bars: maps to 1e-11 of the largest |f| (FP64 throughout; the GPU folds the spectrum and transforms each
ring by Stockham FFTs or Bluestein convolutions, the oracle sums every pixel directly); smoothed maps to
1e-10 (FP64) or 2e-6 (FP32 maps) of the largest |f|; noise to 1e-12 of sigma (FP64) and to the FP32
rounding of the map; the mask exactly.  At Nside 512 (lmax 1535): the closed forms of Y_00, Y_10, Y_11
at every pixel and the adjoint identity with map2alm to 1e-11.
"""
import numpy as np
import pytest

import crime_amd
import oracle
from conftest import small_params

pytestmark = pytest.mark.gpu


def random_alms(rng, nmaps, lmax):
    nalm = (lmax + 1) * (lmax + 2) // 2
    return rng.normal(size=(nmaps, nalm)) + 1j * rng.normal(size=(nmaps, nalm))


@pytest.mark.parametrize("nside,lmax,nmaps", [(1, 3, 2), (2, 7, 3), (4, 11, 1), (4, 15, 2), (8, 23, 3),
                                              (16, 47, 2), (24, 71, 2), (32, 95, 2), (48, 143, 3), (64, 191, 2),
                                              (64, 100, 1), (1024, 24, 1)])
def test_alm2map_vs_oracle(nside, lmax, nmaps):
    """Random alms, Im a_l0 included (both sides ignore it for a real map)."""
    rng = np.random.default_rng(7 * nside + lmax)
    alms = random_alms(rng, nmaps, lmax)
    with crime_amd.Sht(nside, lmax, max_maps=nmaps) as sh:
        got = sh.alm2map(alms)
        st = sh.stats()
    print(f"\nalm2map Nside {nside} lmax {lmax} x{nmaps}: {st}", flush=True)
    ref = oracle.alm2map(alms, nside, lmax)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  max err / max|f| = {err:.2e}", flush=True)
    assert err < 1e-11


def test_production_alm2map_closed_forms_and_adjoint():
    nside, lmax = 512, 1535
    npix = 12 * nside * nside
    nalm = (lmax + 1) * (lmax + 2) // 2
    idx = crime_amd.alm_index
    a = np.zeros((2, nalm), complex)
    a[0, idx(0, 0, lmax)] = 1.5
    a[0, idx(1, 0, lmax)] = -0.75
    a[0, idx(1, 1, lmax)] = 0.25 - 2.0j
    rng = np.random.default_rng(3)
    a[1] = random_alms(rng, 1, lmax)[0]
    f = rng.normal(size=(2, npix))
    with crime_amd.Sht(nside, lmax, max_maps=2) as sh:
        maps = sh.alm2map(a)
        st = sh.stats()
        af = sh.map2alm(f)
    print(f"\nalm2map Nside 512 of 2 maps: {st}", flush=True)
    z = np.empty(npix)
    phi = np.empty(npix)
    for ir in range(1, 4 * nside):
        nphi, first, zr, _, s = oracle.ring_info(nside, ir)
        z[first:first + nphi] = zr
        phi[first:first + nphi] = np.pi * (2 * np.arange(nphi) + s) / nphi
    sth = np.sqrt((1 - z) * (1 + z))
    y00 = np.sqrt(1 / (4 * np.pi))
    y10 = np.sqrt(3 / (4 * np.pi)) * z
    y11 = -np.sqrt(3 / (8 * np.pi)) * sth * np.exp(1j * phi)
    closed = 1.5 * y00 - 0.75 * y10 + 2 * ((0.25 - 2.0j) * y11).real
    assert np.abs(maps[0] - closed).max() < 1e-12 * np.abs(closed).max()
    # adjoint identity: (4 pi / Npix) <f, alm2map(a)> = sum_lm c_m Re(a_lm conj map2alm(f)_lm)
    cm = np.concatenate([np.full(lmax - m + 1, 1.0 if m == 0 else 2.0) for m in range(lmax + 1)])
    a1 = a[1].copy()
    a1[:lmax + 1] = a1[:lmax + 1].real     # m = 0 block: a real map has real a_l0
    lhs = (4 * np.pi / npix) * (f[0] * maps[1]).sum()
    rhs = (cm * (a1 * np.conj(af[0])).real).sum()
    assert abs(lhs - rhs) < 1e-11 * np.sqrt((cm * np.abs(a1) ** 2).sum() * (cm * np.abs(af[0]) ** 2).sum())


@pytest.mark.parametrize("precision", ["double", "single"])
def test_beam_noise_mask_vs_oracle(cfgdir, precision):
    """JoinT's order on a GetHI handle's resident maps: beam (map2alm, b_l, alm2map), noise, mask."""
    par = small_params(cfgdir, f"jt_{precision}", precision=precision, n_side=16, n_nu=8)
    nside, lmax = 16, 47
    fp64 = precision == "double"
    with crime_amd.GetHI(par) as gh:
        gh.run()
        before = gh.maps(np.float64)
        nch = gh.n_nu
        nu0 = np.linspace(1150.0, 1281.25, nch)
        fwhm, sig2 = crime_amd.jt_instrument(nu0, nu0 + 18.75, nside, dish_diameter=15.0, tsys=2.5e4,
                                             time_total=1000.0, n_dish=200)
        with crime_amd.Sht(nside, lmax, max_maps=nch) as sh:
            sh.instrument_of(gh, fwhm_arcmin=fwhm)
            smoothed = gh.maps(np.float64)
            sh.instrument_of(gh, sigma2=sig2, seed=99)
            noisy = gh.maps(np.float64)
            mask = np.ones(12 * nside * nside, np.uint8)
            mask[: 12 * nside * nside // 3] = 0
            sh.instrument_of(gh, mask=mask)
            masked = gh.maps(np.float64)
    alm = oracle.map2alm(before, nside, lmax)
    for c in range(nch):
        b = oracle.beam_window(lmax, fwhm[c])
        for m in range(lmax + 1):
            for l in range(m, lmax + 1):
                alm[c, oracle.alm_index(l, m, lmax)] *= b[l]
    ref = oracle.alm2map(alm, nside, lmax)
    scale = np.abs(ref).max()
    err = np.abs(smoothed - ref).max() / scale
    print(f"\nsmoothing ({precision}) max err / max|f| = {err:.2e}; fwhm {fwhm[0]:.1f}..{fwhm[-1]:.1f} arcmin")
    assert err < (1e-10 if fp64 else 2e-6)
    noise = oracle.add_noise(np.zeros_like(smoothed), sig2, 99)
    if fp64:
        assert np.abs(noisy - smoothed - noise).max() < 1e-12 * np.sqrt(sig2.max())
    else:
        expect = (smoothed.astype(np.float32).astype(np.float64) + noise).astype(np.float32).astype(np.float64)
        assert np.abs(noisy - expect).max() <= 2 * np.spacing(np.abs(expect).astype(np.float32)).max()
    assert np.all(masked[:, mask == 0] == 0)
    assert np.array_equal(masked[:, mask == 1], noisy[:, mask == 1])


def test_production_smoothing(cfgdir):
    """The 200 channels of the 1024^3 production mock smoothed in place for a 15 m dish (FP32 maps):
    at sampled pixels of three channels the result equals sum_lm b_l a_lm Y_lm(p) with the alms of the
    unsmoothed maps (map2alm, pinned separately) and the oracle's lambda_lm; and C_l of the smoothed maps
    is b_l^2 C_l of the originals at l <= Nside / 2, where the HEALPix analysis of a band-limited map is
    accurate, and b_l^2 > 1e-4, above the FP32 rounding of the smoothed maps."""
    import crime_amd.configs as cf

    par = crime_amd.Params(cf.production_1024(cfgdir))
    chans = [0, 99, 199]
    with crime_amd.GetHI(par) as gh:
        gh.run()
        nside, lmax, nch = gh.n_side, 3 * gh.n_side - 1, gh.n_nu
        nu0 = np.linspace(400.0, 798.0, nch)
        fwhm, _ = crime_amd.jt_instrument(nu0, nu0 + 2.0, nside, dish_diameter=15.0)
        with crime_amd.Sht(nside, lmax, max_maps=nch) as sh:
            p, _ = gh.device_maps()
            alm = sh.map2alm_device(p, nch, False)[chans]
            cl0 = sh.cls_of(gh, cross=False)
            sh.instrument_of(gh, fwhm_arcmin=fwhm)
            st = sh.stats()
            cl1 = sh.cls_of(gh, cross=False)
        maps = gh.maps(np.float64)[chans]
    print(f"\nbeam on {nch} channels, Nside {nside}, lmax {lmax}: {st}", flush=True)
    rng = np.random.default_rng(5)
    npix = 12 * nside * nside
    pix = np.concatenate([[0, 5, npix // 2, npix - 1], rng.integers(0, npix, 4)])
    rings = {}
    for ir in range(1, 4 * nside):
        nphi, first, z, st_, s = oracle.ring_info(nside, ir)
        rings[ir] = (nphi, first, z, st_, s)
    starts = np.array([rings[ir][1] for ir in range(1, 4 * nside)])
    for k, c in enumerate(chans):
        b = oracle.beam_window(lmax, fwhm[c])
        for q in pix:
            ir = int(np.searchsorted(starts, q, side="right"))
            nphi, first, z, st_, s = rings[ir]
            phi = np.pi * (2 * (q - first) + s) / nphi
            v = 0.0
            for m in range(lmax + 1):
                lam = oracle.legendre(lmax, m, z, st_)
                ii = [crime_amd.alm_index(l, m, lmax) for l in range(m, lmax + 1)]
                fm = (b[m:] * lam * alm[k, ii]).sum()
                v += (fm.real if m == 0 else 2 * (fm * np.exp(1j * m * phi)).real)
            assert abs(maps[k, q] - v) < 2e-6 * np.abs(maps[k]).max(), (c, q, maps[k, q], v)
    ls = np.arange(lmax + 1)
    for c in chans:
        b = oracle.beam_window(lmax, fwhm[c])
        # where b_l^2 < 1e-4 the FP32 rounding of the smoothed maps outweighs the signal
        sel = (ls >= 2) & (ls <= nside // 2) & (b ** 2 > 1e-4)
        dev = np.abs(cl1[c, sel] / (cl0[c, sel] * b[sel] ** 2) - 1).max()
        print(f"  channel {c}: fwhm {fwhm[c]:.1f} arcmin, max |C_l ratio - 1| (l <= {nside // 2}) = {dev:.2e}")
        assert dev < 1e-2
