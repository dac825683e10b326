"""P(k) estimator of the resident Gaussian field (crime_gh_power_spectrum) against numpy's FFT of the
oracle's field on the same inputs, and, at the production size, against the input P(k).

Bars: shell mode counts exact; mean |k| to 1e-12; P per shell to 1e-10 relative (FP64 grids) or 1e-4
(FP32 grids, whose field differs from the FP64 oracle's by ~1e-6 of its rms); at 1024^3 every shell
within 5.5 sigma of the input P(k) (cosmic variance of a shell: sqrt(2/count) of its mean).
"""
import numpy as np
import pytest

import crime_amd
import crime_amd.configs as cf
import oracle
from conftest import small_params
from pk_ref import pk_expect, pk_numpy, zscores

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,prec,tol", [(32, "double", 1e-10), (48, "double", 1e-10), (96, "single", 1e-4)])
def test_pk_vs_numpy_of_oracle(cfgdir, n, prec, tol):
    par = small_params(cfgdir, f"pk_{n}_{prec}", n_grid=n, precision=prec, l_box=800.0)
    d = par.desc()
    dk = 2 * np.pi / d.l_box
    # shells one fundamental mode wide past the grid's corner: every mode of the grid is counted
    nb = int(np.ceil(np.sqrt(3.0) * n / 2)) + 1
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        k, pk, cnt = gh.power_spectrum(nb, dk)
    od = oracle.fields(d)[0]
    ek, epk, ecnt = pk_numpy(od, d.l_box, nb, dk)
    assert cnt.sum() == n ** 3
    assert np.array_equal(cnt, ecnt)
    assert np.allclose(k, ek, rtol=1e-12, atol=0)
    # shell 0 holds k = 0 only, whose power is the squared mean of a zero-mean field's rounding
    sel = np.arange(nb) > 0
    sel &= ecnt > 0
    assert np.array_equal(pk[~sel & (np.arange(nb) > 0)], np.zeros((~sel[1:]).sum()))
    err = np.abs(pk[sel] - epk[sel]) / epk[sel]
    print(f"P(k) N={n} {prec}: max rel {err.max():.2e}")
    assert err.max() < tol


def test_pk_wide_bins_and_errors(cfgdir):
    par = small_params(cfgdir, "pk_wide", n_grid=32, l_box=800.0)
    d = par.desc()
    dk = 2 * np.pi / d.l_box
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        with pytest.raises(crime_amd.CrimeError):
            gh.power_spectrum(0, dk)
        with pytest.raises(crime_amd.CrimeError):
            gh.power_spectrum(4096, dk)
        with pytest.raises(crime_amd.CrimeError):
            gh.power_spectrum(8, 0.0)
        k, pk, cnt = gh.power_spectrum(5, 2.5 * dk)
        with pytest.raises(crime_amd.CrimeError):   # the grid now holds the transform
            gh.mk_T_maps()
        gh.create_d_and_vr_fields()
        gh.mk_T_maps()
    ek, epk, ecnt = pk_numpy(oracle.fields(d)[0], d.l_box, 5, 2.5 * dk)
    assert np.array_equal(cnt, ecnt)
    assert np.allclose(pk[1:], epk[1:], rtol=1e-10)


def test_pk_production_1024_vs_input(cfgdir):
    par = crime_amd.Params(cf.production_1024(cfgdir))
    d = par.desc()
    dk = 2 * np.pi / d.l_box
    nb = 96
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        k, pk, cnt = gh.power_spectrum(nb, dk)
    ecnt, ep = pk_expect(d, nb, dk)
    assert np.array_equal(cnt, ecnt)
    z = zscores(pk, cnt, ep)
    print(f"1024^3 P(k): max |z| {np.abs(z).max():.2f}, mean z {z.mean():.3f} over {len(z)} shells")
    assert np.abs(z).max() < 5.5
    assert abs(z.mean()) < 4 / np.sqrt(len(z))
