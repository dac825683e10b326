"""The FP32 screen of the map stage's per-sample decisions against the FP64 path, sample by sample.

The map kernels decide a sample's channel and RING pixel from FP32 arithmetic when the argument of each
decision lies farther than a certified margin from its boundary, and from FP64 otherwise.  Every sample
of a box goes through both (crime_gh_screen_check): the screened (channel, pixel, mK factor) must equal
the FP64 ones for every sample.  The fallback rates are printed.  The bit-exact index tests against the
oracle (test_gpu_parity, test_gpu_fullsize) run the same screened path through crime_gh_sample_indices.
"""
import numpy as np
import pytest

import crime_amd
import crime_amd.configs as cf
from conftest import small_params

pytestmark = pytest.mark.gpu


def _report(tag, c):
    print(f"screen {tag}: {c['samples']} samples, {c['binned']} binned; FP64 fallbacks: channel "
          f"{c['channel_fallbacks'] / c['samples']:.2e}, pixel {c['pixel_fallbacks'] / max(c['binned'], 1):.2e}; "
          f"margins {c['eps_nu_mhz']:.2e} MHz, {c['eps_pix']:.2e}; mismatches {c['mismatches']}", flush=True)


@pytest.mark.parametrize("kw", [dict(n_side=16), dict(n_side=1), dict(n_side=3, n_subcell=3), dict(n_side=64, n_nu=40),
                                dict(n_side=256, n_nu=8, nu_lo=1250.0, nu_hi=1400.0),
                                dict(n_side=16, precision="single", velocity_mode="spectral")])
def test_screen_matches_fp64_small(cfgdir, kw):
    name = "scr_" + "_".join(f"{k}{v}" for k, v in kw.items())
    par = small_params(cfgdir, name, n_grid=48, **kw)
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        c = gh.screen_check()
    _report(name, c)
    assert c["binned"] > 10000
    assert c["mismatches"] == 0


def test_screen_matches_fp64_production(tmp_path):
    par = crime_amd.Params(cf.production_1024(str(tmp_path)))
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        c = gh.screen_check()
    _report("1024^3", c)
    assert c["samples"] == 1024 ** 3 * 10
    assert c["mismatches"] == 0
    assert c["channel_fallbacks"] < 0.01 * c["samples"]
    assert c["pixel_fallbacks"] < 0.02 * c["binned"]
