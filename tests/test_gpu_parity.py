"""GPU parity: libcrime_gh.so (through its C ABI) against the FP64 oracle on the same inputs.

Bars: indices (channel, RING pixel) bit-exact; FP64 fields and maps to ~1e-12 relative; FP32 fields to
1e-5 of the field's rms; FP32 maps to 1e-5 relative on every pixel whose samples bin identically.
"""
import numpy as np
import pytest

import crime_amd
import oracle
from conftest import small_params

pytestmark = pytest.mark.gpu


def unit_modes(par_or_desc, seed):
    """The reference's single-thread MT19937 mode stream (src/fourier.c:253-302)."""
    import ctypes as C
    from crime_amd import _ffi

    d = par_or_desc.desc() if hasattr(par_or_desc, "desc") else par_or_desc
    n = d.n_grid
    out = np.empty(2 * n * n * (n // 2 + 1))
    _ffi.host_lib().gh_unit_modes_mt(n, seed, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("n", [32, 48, 64])
def test_fields_fp64_injected(cfgdir, n):
    par = small_params(cfgdir, f"fld64_{n}", n_grid=n, r_smooth=5.0)
    um = unit_modes(par, 1001)
    with crime_amd.GetHI(par, mode_source=crime_amd.MODES_INJECTED, unit_modes=um) as gh:
        st = gh.create_d_and_vr_fields()
        dens, vpot = gh.grid(crime_amd.GRID_DENS), gh.grid(crime_amd.GRID_VPOT)
        s2, mean = gh.sigma2()
    d = par.desc()
    d.mode_source = crime_amd.MODES_INJECTED
    d.unit_modes = um.ctypes.data_as(type(d.unit_modes))
    od, ov, os2, om = oracle.fields(d)
    assert rel(dens, od) < 1e-12
    assert rel(vpot, ov) < 1e-12
    assert abs(s2 - os2) / os2 < 1e-12
    assert abs(st["sigma2_gauss"] - os2) / os2 < 1e-12
    assert abs(mean - om) < 1e-12 * np.sqrt(os2)


@pytest.mark.parametrize("n", [32, 96])
def test_fields_fp32_philox(cfgdir, n):
    par = small_params(cfgdir, f"fld32_{n}", n_grid=n, precision="single")
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        dens, vpot = gh.grid(crime_amd.GRID_DENS), gh.grid(crime_amd.GRID_VPOT)
        s2, _ = gh.sigma2()
    od, ov, os2, _ = oracle.fields(par.desc())
    assert np.abs(dens - od).max() < 1e-5 * od.std()
    assert np.abs(vpot - ov).max() < 1e-5 * ov.std()
    assert abs(s2 - os2) / os2 < 1e-5


def test_cell_fields_fp64(cfgdir):
    par = small_params(cfgdir, "cell64", n_grid=32)
    d = par.desc()
    od, ov, os2, _ = oracle.fields(d)
    with crime_amd.GetHI(par) as gh:
        gh.set_grid(crime_amd.GRID_DENS, od)
        gh.set_grid(crime_amd.GRID_VPOT, ov)
        gh.set_sigma2(os2)
        rv = gh.grid(crime_amd.GRID_RVEL)
        mass = gh.grid(crime_amd.GRID_MASS)
        dz = gh.grid(crime_amd.GRID_DZRSD)
    orv = oracle.radial_velocity(d, ov)
    omass, odz = oracle.get_hi(d, od, orv, os2)
    assert np.array_equal(rv, orv)       # +,-,*,/,sqrt only: identical operations
    assert np.array_equal(dz, odz)
    assert rel(mass, omass) < 1e-14      # exp/pow of two libms


@pytest.mark.parametrize("kw", [dict(n_subcell=0), dict(n_subcell=2, bias_model="const"),
                                dict(n_subcell=0, uniform_nu=True), dict(n_subcell=3, rsd=0)])
def test_sample_indices_bitexact(cfgdir, kw):
    name = "idx_" + "_".join(f"{k}{v}" for k, v in kw.items())
    par = small_params(cfgdir, name, n_grid=32, **kw)
    d = par.desc()
    od, ov, os2, _ = oracle.fields(d)
    orv = oracle.radial_velocity(d, ov)
    _, odz = oracle.get_hi(d, od, orv, os2)
    with crime_amd.GetHI(par) as gh:
        gh.set_grid(crime_amd.GRID_DENS, od)
        gh.set_grid(crime_amd.GRID_VPOT, ov)
        gh.set_sigma2(os2)
        inu, ipix = gh.sample_indices(0, 32 ** 3)
    oinu, oipix = oracle.sample_indices(d, odz, 0, 32 ** 3)
    assert (oinu >= 0).sum() > 1000  # the shell crosses the box
    assert np.array_equal(inu, oinu)
    assert np.array_equal(ipix, oipix)


@pytest.mark.parametrize("seed", [1001, 7, 424242])
def test_sample_indices_irregular_channels(cfgdir, seed):
    """Channels of very uneven width: the narrowest are far below the channel-lookup buckets (1/8 of
    the mean width), so several edges share a bucket and the lookup walks the edges."""
    import os
    import crime_amd.configs as cf
    name = f"idx_irregular_{seed}"
    ini = cf.make(cfgdir, name, n_grid=32, n_side=16, n_nu=16, nu_lo=1200.0, nu_hi=1350.0,
                  l_box=600.0, seed=seed, precision="double", bias_model="reference", n_subcell=0)
    e = 1350.0 - 150.0 * np.linspace(0.0, 1.0, 17)[::-1] ** 3   # widths 26.4 MHz down to 0.037 MHz
    np.savetxt(os.path.join(cfgdir, f"{name}_nu.txt"), e, fmt="%.10f")
    par = crime_amd.Params(ini)
    d = par.desc()
    width = np.ctypeslib.as_array(d.nuf, shape=(d.n_nu,)) - np.ctypeslib.as_array(d.nu0, shape=(d.n_nu,))
    assert width.min() < 0.05 < 150.0 / 128   # narrower than one bucket
    od, ov, os2, _ = oracle.fields(d)
    orv = oracle.radial_velocity(d, ov)
    _, odz = oracle.get_hi(d, od, orv, os2)
    with crime_amd.GetHI(par) as gh:
        gh.set_grid(crime_amd.GRID_DENS, od)
        gh.set_grid(crime_amd.GRID_VPOT, ov)
        gh.set_sigma2(os2)
        inu, ipix = gh.sample_indices(0, 32 ** 3)
    oinu, oipix = oracle.sample_indices(d, odz, 0, 32 ** 3)
    assert np.bincount(oinu[oinu >= 0], minlength=16).min() > 10  # the narrow channels receive samples
    assert np.array_equal(inu, oinu)
    assert np.array_equal(ipix, oipix)


def test_maps_fp64_vs_oracle(cfgdir):
    par = small_params(cfgdir, "maps64", n_grid=48, n_side=16)
    um = unit_modes(par, 1001)
    d = par.desc()
    d.mode_source = crime_amd.MODES_INJECTED
    d.unit_modes = um.ctypes.data_as(type(d.unit_modes))
    with crime_amd.GetHI(par, mode_source=crime_amd.MODES_INJECTED, unit_modes=um) as gh:
        gh.run()
        m = gh.maps(np.float64)
    om, _, _ = oracle.run(d)
    assert (om != 0).sum() > 10000
    assert np.array_equal(m != 0, om != 0)
    assert rel(m, om) < 1e-10


def test_maps_every_aggregation_table_vs_oracle(cfgdir):
    """The three LDS aggregation tables (2048, 1024, 4096 bins) on a box whose tiles touch more bins than
    the smaller tables hold: near the observer a cell covers many Nside-256 pixels, so samples overflow
    to direct HBM atomics.  Every table gives the oracle's maps."""
    par = small_params(cfgdir, "aggtab", n_grid=32, n_side=256, n_nu=8, nu_lo=1250.0, nu_hi=1400.0)
    d = par.desc()
    od, ov, os2, _ = oracle.fields(d)
    orv = oracle.radial_velocity(d, ov)
    _, odz = oracle.get_hi(d, od, orv, os2)
    inu, ipix = oracle.sample_indices(d, odz, 0, 32 ** 3)
    cell = np.arange(32 ** 3)
    tile = (cell % 32) // 16 + 2 * ((cell // 32) % 32 // 4 + 8 * (cell // 1024 // 4))
    key = np.where(inu >= 0, inu.astype(np.int64) * 12 * 256 ** 2 + ipix, -1)
    bins = [np.unique(key[tile == t][key[tile == t] >= 0]).size for t in range(tile.max() + 1)]
    assert max(bins) > 2048          # the default and the small table overflow
    om, _, _ = oracle.run(d)
    with crime_amd.GetHI(par) as gh:
        for v in (0, 1, 2):
            gh.set_schedule([0, 0, 0, v])
            gh.run()
            m = gh.maps(np.float64)
            assert np.array_equal(m != 0, om != 0), v
            assert rel(m, om) < 1e-10, v


def test_maps_fp32_vs_fp64_oracle(cfgdir):
    par = small_params(cfgdir, "maps32", n_grid=64, n_side=32, precision="single")
    with crime_amd.GetHI(par) as gh:
        gh.run()
        m = gh.maps()
    om, _, _ = oracle.run(par.desc())
    nz = om != 0
    err = np.abs(m[nz] - om[nz]) / np.abs(om[nz])
    # a sample whose channel flips between the FP32 and FP64 RSD shifts moves its mass between two
    # bins; every other bin agrees to the FP32 field accuracy
    assert np.mean(err < 1e-5) > 0.999
    assert np.sqrt(np.mean(err ** 2)) < 1e-4
    assert abs(m.sum() - om.sum()) / om.sum() < 1e-5
