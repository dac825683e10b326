"""Pins of the point-source oracle (oracle/psrc_oracle.c), independent of it.

- n(z) table of setup_psources: scipy quad of the luminosity function over log10 L in [-5, 6];
- max_Lpdf is an envelope of the normalised pdf on a fine grid (the rejection sampler is exact);
- the Poisson sampler (inversion below lambda = 10, PTRS above): chi-square against scipy's pmf, mean
  and variance, at lambda on both sides of the switch;
- the luminosity draw: Kolmogorov-Smirnov against the CDF of l_z_function / n(z);
- temp_of_l summed over the SED's two power laws equals K L / (r^2 (1+z)) [((1+z)/1000)^-0.1 nu^-2.1 +
  10 ((1+z)/1000)^-0.75 nu^-2.75], the per-pixel form the GPU accumulates;
- a whole small box with delta = 0: total sources against sum n(z) dV (Poisson), maps against the
  two-term form.
"""
import numpy as np
import pytest
from scipy import integrate, stats

import oracle
from conftest import small_params

Z_MAX = 1420.40575177 / 400.0 - 1


def test_nz_table_vs_quadrature():
    nz, mp = oracle.ps_tables(Z_MAX)
    dz = Z_MAX / oracle.PS_NZ
    for i in (0, 17, 100, 180, 255):
        z = i * dz
        ref, _ = integrate.quad(lambda x: oracle.lib().orc_ps_lfunc(10.0 ** x, z), -5, 6, limit=200)
        assert abs(nz[i] - ref) < 2e-4 * ref, (i, nz[i], ref)
        xs = -5 + (np.arange(256) + 0.5) * 11 / 256
        lf = np.array([oracle.lib().orc_ps_lfunc(10.0 ** x, z) for x in xs])
        assert abs(mp[i] - 1.1 * lf.max() / nz[i]) < 1e-14 * mp[i]
        fine = np.linspace(-5, 6, 20001)
        pdf = np.array([oracle.lib().orc_ps_lfunc(10.0 ** x, z) for x in fine]) / nz[i]
        assert pdf.max() < mp[i]


@pytest.mark.parametrize("lam", [0.3, 2.5, 9.99, 10.0, 17.0, 150.0])
def test_poisson_sampler(lam):
    n = 60000
    k = oracle.ps_poisson(lam, 12345, range(n))
    assert abs(k.mean() - lam) < 5 * np.sqrt(lam / n)
    assert abs(k.var() / lam - 1) < 5 * np.sqrt(2 / n) + 5 / np.sqrt(n * lam)
    lo, hi = int(stats.poisson.ppf(1e-4, lam)), int(stats.poisson.ppf(1 - 1e-4, lam))
    edges = np.arange(lo, hi + 2)
    obs = np.histogram(np.clip(k, lo, hi), bins=edges)[0]
    p = stats.poisson.pmf(edges[:-1], lam)
    p[0] += stats.poisson.cdf(lo - 1, lam)
    p[-1] += stats.poisson.sf(hi, lam)
    keep = p * n > 5
    chi2 = ((obs[keep] - n * p[keep]) ** 2 / (n * p[keep])).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-3


@pytest.mark.parametrize("z", [0.3, 2.0])
def test_luminosity_distribution(z):
    L = oracle.ps_luminosity(z, Z_MAX, 7, range(20000))
    assert np.all(L > 0)
    x = np.log10(L)
    grid = np.linspace(-5, 6, 40001)
    pdf = np.array([oracle.lib().orc_ps_lfunc(10.0 ** g, z) for g in grid])
    cdf = np.concatenate([[0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    assert stats.kstest(x, lambda v: np.interp(v, grid, cdf)).pvalue > 1e-3
    assert oracle.ps_luminosity(Z_MAX + 0.01, Z_MAX, 7, range(3)).tolist() == [0.0, 0.0, 0.0]


def test_temperature_two_term_form():
    rng = np.random.default_rng(4)
    h, domega = 0.7, 4 * np.pi / 3072
    K = 3.2548291e-2 * 8.35774e7 * 4 * np.pi * 0.114461 * h * h / domega
    for _ in range(200):
        L, nu, z, r = 10 ** rng.uniform(-3, 3), rng.uniform(300, 1400), rng.uniform(0, 3), rng.uniform(50, 5000)
        t = oracle.lib().orc_ps_temp(h, L, nu, z, r, domega)
        w = L / (r * r * (1 + z))
        two = K * w * (((1 + z) * 1e-3) ** -0.1 * nu ** -2.1 + 10 * ((1 + z) * 1e-3) ** -0.75 * nu ** -2.75)
        assert abs(t - two) < 1e-13 * t


def test_box_counts_and_maps(cfgdir):
    par = small_params(cfgdir, "ps_oracle", n_grid=16, n_side=8, n_nu=6, l_box=400.0)
    d = par.desc()
    n = d.n_grid
    dens, dz = np.zeros(n ** 3), np.zeros(n ** 3)
    maps, a, b, total = oracle.psources(d, dens, dz, 0.0)
    # expected count with delta = 0: sum over cells of n(z(r_centre)) dx^3 (x offset by pos_obs[1])
    nz, _ = oracle.ps_tables(1420.40575177 / d.nu_min - 1)
    dzt = (1420.40575177 / d.nu_min - 1) / 256
    dx = d.l_box / n
    idx = np.arange(n)
    zc, yc, xc = np.meshgrid((idx + 0.5) * dx - d.pos_obs[2], (idx + 0.5) * dx - d.pos_obs[1],
                             (idx + 0.5) * dx - d.pos_obs[1], indexing="ij")
    r = np.sqrt(xc ** 2 + yc ** 2 + zc ** 2).ravel()
    red = np.array([oracle.lib().orc_z_of_r(d, float(v)) for v in r])
    iz = (red / dzt).astype(int)
    lam = np.where(iz < 255, nz[np.minimum(iz, 254)] + (nz[np.minimum(iz + 1, 255)] - nz[np.minimum(iz, 254)])
                   * (red - iz * dzt) / dzt, nz[255]) * dx ** 3
    assert abs(total - lam.sum()) < 5 * np.sqrt(lam.sum())
    # maps = c1(nu) A + c2(nu) B with the channel centres of mk_psources_maps
    npix = 12 * d.n_side ** 2
    K = 3.2548291e-2 * 8.35774e7 * 4 * np.pi * 0.114461 * d.hhub ** 2 / (4 * np.pi / npix)
    nu0 = np.array([d.nu0[i] for i in range(d.n_nu)])
    nuf = np.array([d.nuf[i] for i in range(d.n_nu)])
    nu = 0.5 * (nu0 + nuf)
    two = K * (np.outer(nu ** -2.1, a) + np.outer(nu ** -2.75, b))
    assert np.abs(maps - two).max() < 1e-12 * np.abs(maps).max()
    assert total > 1000
