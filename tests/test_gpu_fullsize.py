"""Parity at the workloads' own sizes.

- 256^3 FP64 workload (L=4900, 64 channels 700-800 MHz, Nside 128, 2^3 sub-cells, b=1, seed 1001) with
  the reference's single-thread MT19937 unit modes injected into both codes: maps against the FP64
  oracle, every pixel.
- 1024^3 FP32 production workload: (channel, pixel) of every sample of slabs of cells bit-exact against
  the oracle given the same RSD shifts, sigma^2 against the oracle's FP64 reduction of the same field,
  and the map total against the mass the binned samples carry.
- 2048^3 FP32 workload (400 channels, Nside 1024, 4^3 sub-cells): runs, finite maps, sigma^2.
"""
import ctypes as C
import time

import numpy as np
import pytest

import crime_amd
import crime_amd.configs as cf
import oracle
from crime_amd import _ffi

pytestmark = pytest.mark.gpu


def test_small_fp64_256_maps_vs_oracle(cfgdir):
    par = crime_amd.Params(cf.small_fp64_256(cfgdir))
    d = par.desc()
    n = d.n_grid
    um = np.empty(2 * n * n * (n // 2 + 1))
    _ffi.host_lib().gh_unit_modes_mt(n, d.seed, um.ctypes.data_as(C.POINTER(C.c_double)))
    with crime_amd.GetHI(par, mode_source=crime_amd.MODES_INJECTED, unit_modes=um) as gh:
        st = gh.run()
        m = gh.maps(np.float64)
    d.mode_source = crime_amd.MODES_INJECTED
    d.unit_modes = um.ctypes.data_as(type(d.unit_modes))
    om, s2, _ = oracle.run(d)
    assert abs(st["sigma2_gauss"] - s2) < 1e-12 * s2
    nz = om != 0
    assert nz.sum() > 0.9 * om.size          # the shell fills the sky in every channel
    assert np.array_equal(m != 0, nz)
    err = np.abs(m - om)[nz] / np.abs(om[nz])
    print(f"256^3 FP64 maps vs oracle: max rel {err.max():.3e}, rms rel {np.sqrt(np.mean(err ** 2)):.3e}")
    assert err.max() < 1e-10


@pytest.fixture(scope="module")
def production(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("prod"))
    par = crime_amd.Params(cf.production_1024(d))
    gh = crime_amd.GetHI(par)
    st = gh.run()
    yield par, gh, st
    gh.close()


def test_production_sigma2_and_mean(production):
    par, gh, st = production
    dens = gh.grid(crime_amd.GRID_DENS)
    s2, mean = oracle.sigma2(dens)
    del dens
    assert abs(st["sigma2_gauss"] - s2) < 1e-10 * s2   # same FP32 values, 1e9-term FP64 sums in another order
    assert abs(mean) < 1e-6 * np.sqrt(s2)             # no k = 0 mode


def test_production_sample_indices_bitexact(production):
    par, gh, st = production
    d = par.desc()
    n = d.n_grid
    dz = gh.grid(crime_amd.GRID_DZRSD).ravel()
    t0 = time.time()
    total = 0
    # slabs of 4 planes at radii inside the shell, at its inner and outer edges and in a box corner
    for iz in (n // 2 - 300, n // 2 - 2, n // 2 + 236, n // 2 + 400, 40):
        c0, nc = iz * n * n, 4 * n * n
        g_inu, g_pix = gh.sample_indices(c0, nc)
        o_inu, o_pix = oracle.sample_indices(d, dz, c0, nc)
        assert np.array_equal(g_inu, o_inu)
        assert np.array_equal(g_pix, o_pix)
        total += int((g_inu >= 0).sum())
    print(f"1024^3: {total} binned samples bit-exact in {time.time() - t0:.1f} s")
    assert total > 10 ** 6


def test_production_map_total(production):
    # sum over channels of maps/prefac = sum over cells of mass x (fraction of its samples binned),
    # evaluated on the GPU's own per-cell mass and per-sample channels (every sample of the box)
    par, gh, st = production
    d = par.desc()
    n = d.n_grid
    maps = gh.maps(np.float64)
    pref = oracle.channel_factors(d)
    lhs = (maps.sum(axis=1) / pref).sum()
    mass = gh.grid(crime_amd.GRID_MASS).ravel()
    rhs = 0.0
    step = 16 * n * n
    for c0 in range(0, n ** 3, step):
        inu, _ = gh.sample_indices(c0, step)
        rhs += float((mass[c0:c0 + step] * (inu >= 0).sum(axis=1)).sum()) / d.n_sub
    print(f"1024^3 map total {lhs:.9e}, binned mass {rhs:.9e}")
    assert abs(lhs - rhs) < 2e-5 * rhs   # FP32 atomics over ~5e9 samples


def test_largest_2048_runs(tmp_path):
    """The 2048^3 mock (400 channels, Nside 1024, 4^3 sub-cells) and the cross spectra of its first 64
    channels at lmax 3071 (Legendre table rebuilt in m-chunks): symmetric, non-negative autos."""
    par = crime_amd.Params(cf.largest_2048(str(tmp_path)))
    with crime_amd.GetHI(par) as gh:
        st = gh.run()
        ptr, nbytes = gh.device_maps()
        with crime_amd.Sht(gh.n_side, 3 * gh.n_side - 1, max_maps=64) as sh:
            cls = sh.cls_of(gh, cross=True, channels=64)
            sst = sh.stats()
    print(f"2048^3: {st['ms_total']:.1f} ms, sigma2 {st['sigma2_gauss']:.6f}, maps {nbytes / 2 ** 30:.1f} GiB")
    print(f"C_l of 64 channels at Nside 1024, lmax 3071: {sst}")
    assert np.isfinite(st["sigma2_gauss"]) and 0.05 < st["sigma2_gauss"] < 5
    assert np.all(np.isfinite(cls)) and cls.shape == (64, 64, 3072)
    assert np.allclose(cls, cls.transpose(1, 0, 2), rtol=1e-12, atol=0)
    assert np.all(np.einsum("aal->al", cls) >= 0) and cls[0, 0, 2:].max() > 0
