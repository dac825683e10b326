"""HEALPix analysis on the GPU (include/crime_sht.h) against the direct-sum FP64 oracle
(oracle/sht_oracle.c) on the same maps, and closed forms at the production Nside.

Bars: alms to 1e-11 of the largest |a_lm| (FP64 throughout; the GPU and the oracle differ in summation
order, ring transform and Legendre recursion); C_l to 1e-10 relative; at Nside 512 and 1024
(lmax 3 Nside - 1) the closed forms of a_00, a_10, a_11 and ring-by-ring direct sums at sampled (l, m)
to 1e-10 of max(|a_lm|, white-noise rms).
"""
import numpy as np
import pytest

import crime_amd
import oracle
from conftest import small_params

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("nside,lmax,nmaps,max_maps", [(1, 3, 2, 2), (2, 7, 3, 5), (4, 11, 1, 1), (4, 15, 2, 3),
                                                       (8, 23, 3, 3), (16, 47, 2, 40), (32, 95, 2, 2),
                                                       (64, 191, 2, 2), (24, 71, 2, 2), (48, 143, 3, 3),
                                                       (1024, 48, 1, 1)])
def test_map2alm_vs_oracle(nside, lmax, nmaps, max_maps):
    rng = np.random.default_rng(nside * 100 + lmax)
    maps = rng.normal(size=(nmaps, 12 * nside * nside))
    maps[0] += 3.0  # a monopole
    with crime_amd.Sht(nside, lmax, max_maps=max_maps) as sh:
        got = sh.map2alm(maps)
        st = sh.stats()
    ref = oracle.map2alm(maps, nside, lmax)
    err = rel_err(got, ref)
    print(f"Nside {nside} lmax {lmax}: alm max err / max|a| = {err:.2e}; stages {st}")
    assert err < 1e-11


def test_map2alm_nside128_and_cls():
    nside, lmax = 128, 383
    rng = np.random.default_rng(5)
    maps = rng.normal(size=(3, 12 * nside * nside))
    maps[1] = 0.5 * maps[0] + maps[1]   # correlated pair
    with crime_amd.Sht(nside, lmax, max_maps=3) as sh:
        got = sh.map2alm(maps)
        cross = sh.anafast(maps, cross=True)
        auto = sh.anafast(maps, cross=False)
    ref = oracle.map2alm(maps, nside, lmax)
    assert rel_err(got, ref) < 1e-11
    rcl = oracle.alm2cl(ref, lmax)
    assert np.abs(cross - rcl).max() < 1e-10 * np.abs(rcl).max()
    for a in range(3):   # the two kernels sum over m in different orders
        assert np.allclose(auto[a], cross[a, a], rtol=1e-13, atol=0)
    assert np.allclose(cross, cross.transpose(1, 0, 2), rtol=1e-13, atol=0)


def test_cls_of_resident_fp32_maps(cfgdir):
    par = small_params(cfgdir, "sht_gh", n_grid=64, n_side=32, n_nu=6, precision="single")
    with crime_amd.GetHI(par) as gh:
        gh.run()
        maps = gh.maps().astype(np.float64)
        with crime_amd.Sht(32, 95, max_maps=6) as sh:
            cls = sh.cls_of(gh, cross=True)
            p, _ = gh.device_maps()
            alm = sh.map2alm_device(p, 6, False)
    ref = oracle.map2alm(maps, 32, 95)
    assert rel_err(alm, ref) < 1e-11
    rcl = oracle.alm2cl(ref, 95)
    assert np.abs(cls - rcl).max() < 1e-10 * np.abs(rcl).max()


def _alm_direct(maps, nside, l, m):
    """a_lm of each map by the defining sum, ring by ring: exact phases m phi_j = pi t / nphi with
    t = m (2j + s) mod 2 nphi, and the oracle's long-double lambda_lm (pinned against scipy and
    Unsold's theorem)."""
    npix = 12 * nside * nside
    out = np.zeros(maps.shape[0], complex)
    for ir in range(1, 4 * nside):
        nphi, first, z, st, s = oracle.ring_info(nside, ir)
        j = np.arange(nphi)
        t = (m * (2 * j + s)) % (2 * nphi)
        e = np.exp(-1j * np.pi * t / nphi)
        lam = oracle.legendre(l, m, z, st)[-1]
        out += lam * (maps[:, first:first + nphi] * e).sum(axis=1)
    return out * (4 * np.pi / npix)


@pytest.mark.parametrize("nside,lms", [
    (512, [(2, 1), (40, 17), (700, 3), (1000, 999), (1535, 400), (1535, 1535)]),
    # Nside 1024 (the largest mock's maps): 8192-point Bluestein convolutions for the polar rings,
    # the 4096-point equatorial FFT and a Legendre table rebuilt in m-chunks
    (1024, [(3, 2), (1500, 700), (2900, 2048), (3071, 17), (3071, 3071)])])
def test_production_closed_forms(nside, lms):
    lmax = 3 * nside - 1
    npix = 12 * nside * nside
    rng = np.random.default_rng(11)
    maps = rng.normal(size=(2, npix)) + np.array([[2.0], [0.0]])
    with crime_amd.Sht(nside, lmax, max_maps=2) as sh:
        alm = sh.map2alm(maps)
        st = sh.stats()
    print(f"Nside {nside} map2alm of 2 maps: {st}")
    idx = crime_amd.alm_index
    w = 4 * np.pi / npix
    # textbook Y_00, Y_10, Y_11 at the pixel centres
    z = np.empty(npix)
    phi = np.empty(npix)
    for ir in range(1, 4 * nside):
        nphi, first, zr, _, s = oracle.ring_info(nside, ir)
        z[first:first + nphi] = zr
        phi[first:first + nphi] = np.pi * (2 * np.arange(nphi) + s) / nphi
    sth = np.sqrt((1 - z) * (1 + z))
    closed = {(0, 0): np.sqrt(4 * np.pi) * maps.mean(axis=1),
              (1, 0): w * np.sqrt(3 / (4 * np.pi)) * (maps * z).sum(axis=1),
              (1, 1): -w * np.sqrt(3 / (8 * np.pi)) * (maps * sth * np.exp(-1j * phi)).sum(axis=1)}
    noise = 4 * np.pi / np.sqrt(npix)   # ~ rms |a_lm| of a unit white-noise map
    for (l, m), ref in closed.items():
        assert np.abs(alm[:, idx(l, m, lmax)] - ref).max() < 1e-10 * max(np.abs(ref).max(), noise), (l, m)
    for l, m in lms:
        ref = _alm_direct(maps, nside, l, m)
        assert np.abs(alm[:, idx(l, m, lmax)] - ref).max() < 1e-10 * max(np.abs(ref).max(), noise), (l, m)


def test_production_maps_cross_spectra(cfgdir):
    """C_l^{ab} of all 200 channels of the 1024^3 production mock, straight from the resident FP32 maps:
    symmetric, autos non-negative, alms of sampled channels equal to ring-by-ring direct sums, and C_l
    of a channel pair equal to the sum over m of their alms."""
    import crime_amd.configs as cf

    par = crime_amd.Params(cf.production_1024(cfgdir))
    with crime_amd.GetHI(par) as gh:
        gh.run()
        nside, lmax, nch = gh.n_side, 3 * gh.n_side - 1, gh.n_nu
        with crime_amd.Sht(nside, lmax, max_maps=nch) as sh:
            cls = sh.cls_of(gh, cross=True)
            st = sh.stats()
            p, _ = gh.device_maps()
            alm = sh.map2alm_device(p, nch, False)
        maps = gh.maps()[[0, 99, 199]].astype(np.float64)
    print(f"C_l of {nch} channels, Nside {nside}, lmax {lmax}: {st}")
    assert np.all(np.isfinite(cls))
    assert np.allclose(cls, cls.transpose(1, 0, 2), rtol=1e-12, atol=0)
    assert np.all(np.einsum("aal->al", cls) >= 0)
    idx = crime_amd.alm_index
    for l, m in [(0, 0), (10, 3), (300, 150), (1535, 700)]:
        ref = _alm_direct(maps, nside, l, m)
        got = alm[[0, 99, 199], idx(l, m, lmax)]
        assert np.abs(got - ref).max() < 1e-10 * np.abs(alm[[0, 99, 199]]).max(), (l, m)
    a, b = alm[17], alm[42]
    ms = np.concatenate([np.full(lmax - m + 1, m) for m in range(lmax + 1)])
    ls = np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)])
    prod = np.where(ms == 0, 1.0, 2.0) * (a * np.conj(b)).real
    ref_cl = np.bincount(ls, weights=prod, minlength=lmax + 1) / (2 * np.arange(lmax + 1) + 1)
    assert np.abs(cls[17, 42] - ref_cl).max() < 1e-10 * np.abs(ref_cl).max()


def test_member_sets_equal_single_calls(cfgdir):
    """Spectra of three members in one call (crime_sht_anafast_sets_device) equal each member's own."""
    hs = [crime_amd.GetHI(small_params(cfgdir, f"set{i}", seed=500 + i, n_side=32, n_nu=8)) for i in range(3)]
    try:
        for h in hs:
            h.run()
        with crime_amd.Sht(32, 95, max_maps=24) as sh:
            sets = sh.cls_of_members(hs)
            single = [sh.cls_of(h, cross=True) for h in hs]
    finally:
        for h in hs:
            h.close()
    assert sets.shape == (3, 8, 8, 96)
    for i in range(3):
        assert np.allclose(sets[i], single[i], rtol=1e-13, atol=0)
    assert not np.allclose(sets[0], sets[1])


def test_plan_beyond_device_memory_reports_enomem():
    """A plan whose spectra buffer cannot fit (max_maps far beyond HBM) fails with ENOMEM and a message,
    and the library stays usable."""
    from crime_amd import _ffi
    import ctypes as C

    lib = _ffi.gpu_lib()
    h = C.c_void_p()
    rc = lib.crime_sht_create(512, 1535, 1 << 20, 0, C.byref(h))
    assert rc == -2 and h.value is None
    assert b"hipMalloc" in lib.crime_sht_last_error(None)
    with crime_amd.Sht(8, 23, max_maps=1) as sh:
        assert sh.map2alm(np.ones((1, 768))).shape == (1, 300)
