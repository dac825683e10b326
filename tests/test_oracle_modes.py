"""Closed-form pins of the oracle's mode fill (src/fourier.c:240-304) through its fields.

One unit mode g is injected at a chosen half-spectrum position and every other mode is zero.  The
oracle's normalised fields must then equal, at every grid point x = dx (ix, iy, iz),

    delta(x) = (sqrt(2 pi)/L)^3 c A Re(g e^{i k.x}),   A = sqrt(P(k)/dk^3 [e^{-r_s^2 k^2}])
    phi(x)   = delta(x) f0 H0 / k^2

with c = 2 off the kx = 0, N/2 planes (the mode and its Hermitian image) and c = 1 on them (a real
inverse transform keeps the Hermitian part (X(k) + X*(-k))/2 of a plane filled without symmetry).
P(k) is numpy's linear interpolation of the table in log10 k, f0 H0 and r_s^2 come from the
descriptor.  A dropped f0 or H0, k in place of k^2, an unsquared r_s or a lost sqrt fails one of these.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import small_params

N = 16


def _single_mode_fields(par, kx, ky, kz, g, smoothing=None):
    d = par.desc()
    nh = N // 2 + 1
    um = np.zeros((N, N, nh, 2))
    um[kz, ky, kx] = (g.real, g.imag)
    um = um.ravel()
    d.mode_source = 1
    d.unit_modes = um.ctypes.data_as(type(d.unit_modes))
    if smoothing is not None:
        d.do_smoothing = 1
        d.r2_smooth = smoothing
    else:
        d.do_smoothing = 0
    dens, vpot, _, _ = oracle.fields(d)
    return d, dens, vpot


def _expected(d, kx, ky, kz, g, r2=None):
    L = d.l_box
    dk = 2 * math.pi / L
    signed = [m if 2 * m <= N else m - N for m in (kx, ky, kz)]  # +Nyquist (fourier.c:264-283)
    k2 = dk * dk * sum(m * m for m in signed)
    logk = np.ctypeslib.as_array(d.logk, shape=(d.numk,))
    pk = np.ctypeslib.as_array(d.pk, shape=(d.numk,))
    p = np.interp(0.5 * math.log10(k2), logk, pk)
    amp2 = p / dk ** 3
    if r2 is not None:
        amp2 *= math.exp(-r2 * k2)
    c = 1.0 if kx in (0, N // 2) else 2.0
    i = np.arange(N)
    Z, Y, X = np.meshgrid(i, i, i, indexing="ij")
    phase = 2 * math.pi * (kx * X + ky * Y + kz * Z) / N
    norm = (math.sqrt(2 * math.pi) / L) ** 3
    delta = norm * c * math.sqrt(amp2) * np.real(g * np.exp(1j * phase))
    return delta, k2


@pytest.mark.parametrize("mode", [(3, 5, 2), (1, 0, 7), (0, 2, 3), (N // 2, 1, 4), (5, N // 2, N - 3)])
@pytest.mark.parametrize("r2", [None, 40.0 ** 2])
def test_single_mode_closed_form(cfgdir, mode, r2):
    par = small_params(cfgdir, "onemode", n_grid=N, l_box=600.0)
    g = 0.6 - 1.3j
    d, dens, vpot = _single_mode_fields(par, *mode, g, smoothing=r2)
    delta, k2 = _expected(d, *mode, g, r2)
    scale = np.abs(delta).max()
    assert scale > 0
    assert np.abs(dens - delta).max() < 1e-12 * scale
    # velocity potential: phi_k = delta_k f0 H0 / k^2 (fourier.c:297-298), cell by cell
    assert d.f0h0 > 0
    assert np.abs(vpot - delta * d.f0h0 / k2).max() < 1e-12 * scale * d.f0h0 / k2


def test_smoothing_ratio_and_sigma(cfgdir):
    """Smoothing multiplies sigma^2 by e^{-r_s^2 k^2} (fourier.c:291-295, r_s squared on input,
    io_gh.c:255-260): the field ratio is e^{-r_s^2 k^2 / 2} at every point, for modes of different k."""
    par = small_params(cfgdir, "onemode_s", n_grid=N, l_box=600.0)
    r2 = 35.0 ** 2
    for mode in [(2, 1, 1), (6, 3, 5)]:
        g = 1.0 + 0.25j
        d, off, _ = _single_mode_fields(par, *mode, g)
        _, on, _ = _single_mode_fields(par, *mode, g, smoothing=r2)
        _, k2 = _expected(d, *mode, g)
        m = np.abs(off) > 1e-3 * np.abs(off).max()
        assert np.allclose(on[m] / off[m], math.exp(-0.5 * r2 * k2), rtol=1e-12, atol=0)


def test_two_modes_superpose(cfgdir):
    """Linearity of the fill + c2r: two injected modes give the sum of their single-mode fields."""
    par = small_params(cfgdir, "twomode", n_grid=N, l_box=600.0)
    d = par.desc()
    nh = N // 2 + 1
    um = np.zeros((N, N, nh, 2))
    um[2, 5, 3] = (0.3, -0.7)
    um[9, 1, 0] = (-1.1, 0.4)
    um = um.ravel()
    d.mode_source = 1
    d.unit_modes = um.ctypes.data_as(type(d.unit_modes))
    d.do_smoothing = 0
    dens, vpot, _, _ = oracle.fields(d)
    a, ka = _expected(d, 3, 5, 2, 0.3 - 0.7j)
    b, kb = _expected(d, 0, 1, 9, -1.1 + 0.4j)
    s = np.abs(a + b).max()
    assert np.abs(dens - (a + b)).max() < 1e-12 * s
    ref_v = (a / ka + b / kb) * d.f0h0
    assert np.abs(vpot - ref_v).max() < 1e-12 * np.abs(ref_v).max()


# ---------------------------------------------------------------------------------------------
# spectral velocity components: v_j(k) = i sin(k_j dx)/dx phi_k (the symbol of fourier.c:359-366)
# ---------------------------------------------------------------------------------------------
def _central_difference(vpot, dx):
    """(phi(x + dx e_j) - phi(x - dx e_j)) / (2 dx) per axis, periodic; arrays are [z, y, x]."""
    return [(np.roll(vpot, -1, axis=a) - np.roll(vpot, 1, axis=a)) / (2 * dx) for a in (2, 1, 0)]


@pytest.mark.parametrize("mode", [(3, 5, 2), (0, 2, 3), (N // 2, 1, 4), (5, N // 2, N - 3), (1, 7, N // 2)])
def test_spectral_velocity_single_mode(cfgdir, mode):
    par = small_params(cfgdir, "onemode_v", n_grid=N, l_box=600.0)
    g = 0.6 - 1.3j
    d, dens, vpot = _single_mode_fields(par, *mode, g)
    vel = oracle.velocity_spectral(d)
    delta, k2 = _expected(d, *mode, g)
    dx = d.l_box / N
    i = np.arange(N)
    Z, Y, X = np.meshgrid(i, i, i, indexing="ij")
    phase = 2 * math.pi * (mode[0] * X + mode[1] * Y + mode[2] * Z) / N
    # delta = C Re(g e^{i phase}) with C = norm c A: v_j = C f0H0/k^2 Re(i sin(2 pi m_j/N)/dx g e^{i phase})
    C = np.abs(delta).max() / np.abs(np.real(g * np.exp(1j * phase))).max()
    scale = C * d.f0h0 / k2 / dx
    for j, m in enumerate(mode):
        ref = C * d.f0h0 / k2 * np.real(1j * math.sin(2 * math.pi * m / N) / dx * g * np.exp(1j * phase))
        assert np.abs(vel[j] - ref).max() < 1e-12 * scale, j
        if 2 * m == N:
            assert np.abs(vel[j]).max() < 1e-12 * scale  # zero at Nyquist, like the stencil
    for j, cd in enumerate(_central_difference(vpot, dx)):
        assert np.abs(vel[j] - cd).max() < 1e-12 * scale, j


@pytest.mark.parametrize("n", [16, 24])
def test_spectral_radial_velocity_equals_stencil(cfgdir, n):
    """On a full random field the spectral v_r equals the reference's stencil v_r to rounding."""
    par = small_params(cfgdir, f"vspec_{n}", n_grid=n, r_smooth=5.0)
    d = par.desc()
    dens, vpot, _, _ = oracle.fields(d)
    vel = oracle.velocity_spectral(d)
    dx = d.l_box / n
    for j, cd in enumerate(_central_difference(vpot, dx)):
        assert np.abs(vel[j] - cd).max() < 1e-12 * np.abs(cd).max(), j
    rv_spec = oracle.radial_velocity_spectral(d, *vel)
    rv_sten = oracle.radial_velocity(d, vpot)
    assert np.abs(rv_spec - rv_sten).max() < 1e-12 * np.abs(rv_sten).max()


def test_run_spectral_equals_stencil_maps(cfgdir):
    """The whole oracle pipeline gives the same maps in both velocity modes, up to samples whose
    channel flips with the last-bit difference of v_r (none expected at this size)."""
    par = small_params(cfgdir, "vspec_run", n_grid=16, n_side=8, n_nu=8)
    d = par.desc()
    a, s2a, _ = oracle.run(d)
    d.velocity_mode = 1
    b, s2b, _ = oracle.run(d)
    assert abs(s2a - s2b) < 1e-13 * s2a  # OpenMP reduction order
    assert np.abs(a - b).max() < 1e-10 * np.abs(a).max()
