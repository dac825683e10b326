"""Spectral velocity components (velocity_mode=spectral) on the GPU against the FP64 oracle.

v_j(k) = i sin(k_j dx)/dx phi_k is the Fourier symbol of the reference's central difference
(src/fourier.c:359-366), so the spectral v_r must equal the reference's stencil v_r to rounding.
Bars: FP64 components and v_r within 1e-12 relative of the oracle (its spectral components and its
stencil v_r); v_r from the oracle's components bit-identical (same operations); (channel, pixel) of every
sample bit-exact; FP64 maps within 1e-10; FP32 components within 1e-5 of their rms.
"""
import numpy as np
import pytest

import crime_amd
import oracle
from conftest import small_params
from test_gpu_parity import rel, unit_modes

pytestmark = pytest.mark.gpu


def _injected(par, um):
    d = par.desc()
    d.mode_source = crime_amd.MODES_INJECTED
    d.unit_modes = um.ctypes.data_as(type(d.unit_modes))
    return d


@pytest.mark.parametrize("n", [32, 48])
def test_spectral_fields_fp64_vs_oracle(cfgdir, n):
    print(f"spectral fields FP64 N={n}", flush=True)
    par = small_params(cfgdir, f"vel64_{n}", n_grid=n, velocity_mode="spectral", r_smooth=5.0)
    um = unit_modes(par, 1001)
    with crime_amd.GetHI(par, mode_source=crime_amd.MODES_INJECTED, unit_modes=um) as gh:
        st = gh.create_d_and_vr_fields()
        dens = gh.grid(crime_amd.GRID_DENS)
        vel = [gh.grid(w) for w in (crime_amd.GRID_VX, crime_amd.GRID_VY, crime_amd.GRID_VZ)]
        rv = gh.grid(crime_amd.GRID_RVEL)
        with pytest.raises(crime_amd.CrimeError):
            gh.grid(crime_amd.GRID_VPOT)  # not held in spectral mode
    d = _injected(par, um)
    od, ov, os2, _ = oracle.fields(d)
    ovel = oracle.velocity_spectral(d)
    assert rel(dens, od) < 1e-12
    assert abs(st["sigma2_gauss"] - os2) < 1e-12 * os2
    for j in range(3):
        assert rel(vel[j], ovel[j]) < 1e-12, j
    # against the reference's own route: central difference of phi projected on r_hat
    orv = oracle.radial_velocity(d, ov)
    err = rel(rv, orv)
    print(f"spectral v_r vs stencil v_r (FP64, N={n}): max rel {err:.2e}")
    assert err < 1e-12


def test_spectral_cell_fields_bitexact(cfgdir):
    par = small_params(cfgdir, "vel_cell", n_grid=32, velocity_mode="spectral")
    d = par.desc()
    od, _, os2, _ = oracle.fields(d)
    ovel = oracle.velocity_spectral(d)
    with crime_amd.GetHI(par) as gh:
        gh.set_grid(crime_amd.GRID_DENS, od)
        for w, v in zip((crime_amd.GRID_VX, crime_amd.GRID_VY, crime_amd.GRID_VZ), ovel):
            gh.set_grid(w, v)
        gh.set_sigma2(os2)
        rv = gh.grid(crime_amd.GRID_RVEL)
        dz = gh.grid(crime_amd.GRID_DZRSD)
        inu, ipix = gh.sample_indices(0, 32 ** 3)
    orv = oracle.radial_velocity_spectral(d, *ovel)
    _, odz = oracle.get_hi(d, od, orv, os2)
    assert np.array_equal(rv, orv)
    assert np.array_equal(dz, odz)
    oinu, oipix = oracle.sample_indices(d, odz, 0, 32 ** 3)
    assert (oinu >= 0).sum() > 1000
    assert np.array_equal(inu, oinu)
    assert np.array_equal(ipix, oipix)


def test_spectral_maps_fp64_vs_oracle(cfgdir):
    par = small_params(cfgdir, "vel_maps64", n_grid=48, n_side=16, velocity_mode="spectral")
    um = unit_modes(par, 1001)
    with crime_amd.GetHI(par, mode_source=crime_amd.MODES_INJECTED, unit_modes=um) as gh:
        gh.run()
        m = gh.maps(np.float64)
    d = _injected(par, um)
    om, _, _ = oracle.run(d)
    assert (om != 0).sum() > 10000
    assert np.array_equal(m != 0, om != 0)
    assert rel(m, om) < 1e-10
    # and the stencil run of the same modes gives the same maps up to rounding
    d.velocity_mode = 0
    om0, _, _ = oracle.run(d)
    assert rel(m, om0) < 1e-10


def test_spectral_fp32_philox(cfgdir):
    par = small_params(cfgdir, "vel32", n_grid=64, n_side=32, precision="single", velocity_mode="spectral")
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        vel = [gh.grid(w) for w in (crime_amd.GRID_VX, crime_amd.GRID_VY, crime_amd.GRID_VZ)]
        rv = gh.grid(crime_amd.GRID_RVEL)
    d = par.desc()
    ovel = oracle.velocity_spectral(d)
    for j in range(3):
        assert np.abs(vel[j] - ovel[j]).max() < 1e-5 * ovel[j].std(), j
    _, ov, _, _ = oracle.fields(d)
    orv = oracle.radial_velocity(d, ov)
    assert np.abs(rv - orv).max() < 1e-5 * orv.std()
