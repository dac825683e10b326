"""crime_gh_power_spectrum_noalloc (static device scratch, no allocation) against crime_gh_power_spectrum.

Two calls in a row through the C ABI, the second on a smaller grid with fewer shells, each compared with
the allocating entry point on the same fields: mode counts equal, mean |k| and P within 1e-12 relative
(the shell sums are FP64 atomics, whose order varies between launches).  A third call asks for more
shells than the scratch holds and must fail with CRIME_GH_EINVAL.
"""
import ctypes as C

import numpy as np
import pytest

import crime_amd
from conftest import small_params

pytestmark = pytest.mark.gpu


def _spectrum(gh, fn, nbins, bin_dk):
    k = np.empty(nbins)
    pk = np.empty(nbins)
    cnt = np.empty(nbins, np.int64)
    dp = C.POINTER(C.c_double)
    rc = fn(gh._h, nbins, bin_dk, k.ctypes.data_as(dp), pk.ctypes.data_as(dp),
            cnt.ctypes.data_as(C.POINTER(C.c_int64)))
    return rc, k, pk, cnt


@pytest.mark.parametrize("prec", ["double", "single"])
def test_noalloc_matches_allocating_entry(cfgdir, prec):
    lib = crime_amd._ffi.gpu_lib()
    for n, nbins in ((64, 60), (32, 20)):
        print(f"no-alloc P(k): N={n} nbins={nbins} {prec}", flush=True)
        par = small_params(cfgdir, f"noalloc_{n}_{prec}", n_grid=n, precision=prec, l_box=900.0)
        dk = 2 * np.pi / 900.0
        with crime_amd.GetHI(par) as gh:
            gh.create_d_and_vr_fields()
            rc, k0, p0, c0 = _spectrum(gh, lib.crime_gh_power_spectrum, nbins, dk)
            assert rc == 0
            gh.create_d_and_vr_fields()  # the transform consumed the fields: the same fields again
            rc, k1, p1, c1 = _spectrum(gh, lib.crime_gh_power_spectrum_noalloc, nbins, dk)
            assert rc == 0, lib.crime_gh_last_error(gh._h)
            assert np.array_equal(c0, c1)
            assert c1.sum() > 0
            assert np.allclose(k1, k0, rtol=1e-12, atol=0)
            assert np.allclose(p1, p0, rtol=1e-12, atol=0)
            # the call consumed the fields, like the allocating one
            rc, *_ = _spectrum(gh, lib.crime_gh_power_spectrum_noalloc, nbins, dk)
            assert rc == -5  # CRIME_GH_ESTATE
            gh.create_d_and_vr_fields()
            rc, *_ = _spectrum(gh, lib.crime_gh_power_spectrum_noalloc, 2049, dk)
            assert rc == -1
            assert b"static scratch" in lib.crime_gh_last_error(gh._h)
