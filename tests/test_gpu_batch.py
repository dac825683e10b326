"""Batched runs: one captured hipGraph over uneven members must give every member the maps of its own
single run, on every replay, with ordinary eager work interleaved between replays."""
import numpy as np
import pytest

import crime_amd
from conftest import small_params

pytestmark = pytest.mark.gpu

MEMBERS = [dict(n_grid=32, n_side=16, seed=11, omega_m=0.28, sigma_8=0.75),
           dict(n_grid=48, n_side=8, seed=12, omega_m=0.31, sigma_8=0.85, n_nu=9),
           dict(n_grid=64, n_side=16, seed=13, omega_m=0.33, sigma_8=0.8, n_subcell=2),
           dict(n_grid=24, n_side=4, seed=14, omega_m=0.26, sigma_8=0.9, bias_model="const"),
           dict(n_grid=32, n_side=32, seed=15, omega_m=0.3, sigma_8=0.7, uniform_nu=True)]


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-5)])
def test_batch_equals_single_runs(cfgdir, precision, tol):
    pars = [small_params(cfgdir, f"bm_{precision}_{i}", precision=precision, **kw) for i, kw in enumerate(MEMBERS)]
    hs = [crime_amd.GetHI(p) for p in pars]
    ref = []
    for h in hs:
        h.run()
        ref.append(h.maps(np.float64))
    extra = crime_amd.GetHI(small_params(cfgdir, f"bm_{precision}_extra", precision=precision, n_grid=16))
    with crime_amd.Batch(hs, n_streams=3) as b:
        for replay in range(3):
            for h in hs:   # stale contents must not survive into the replay
                h.set_grid(crime_amd.GRID_DENS, np.zeros(h.n ** 3))
            b.run()
            for i, h in enumerate(hs):
                m = h.maps(np.float64)
                assert np.array_equal(m != 0, ref[i] != 0), (replay, i)
                err = np.abs(m - ref[i]).max() / np.abs(ref[i]).max()
                assert err < tol, (replay, i, err)
            extra.run()  # eager work between replays
    for h in hs + [extra]:
        h.close()


def test_batch_rejects_bad_members(cfgdir):
    p = small_params(cfgdir, "bm_bad")
    h = crime_amd.GetHI(p)
    with pytest.raises(crime_amd.CrimeError, match="repeats"):
        crime_amd.Batch([h, h])
    h.close()
