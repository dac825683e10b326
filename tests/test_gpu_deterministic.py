"""Deterministic scatter (desc.deterministic = 1): bit-identical maps from run to run, in a batch graph
as in a single run, and the same maps as the floating-point scatter up to its quantisation.

The reference's maps depend on the order of its OpenMP FP32 atomics (src/pixelize.c:223-226); the
fixed-point scatter sums integers round(s mass / n_sub), whose atomic adds commute.
"""
import numpy as np
import pytest

import crime_amd
import oracle
from conftest import small_params
from test_gpu_parity import rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["double", "single"])
def test_deterministic_bit_identical_runs(cfgdir, prec):
    print(f"deterministic scatter, {prec}: repeated runs", flush=True)
    par = small_params(cfgdir, f"det_{prec}", n_grid=64, n_side=32, precision=prec,
                       extra=dict(deterministic=1))
    assert par.get("deterministic") == 1
    with crime_amd.GetHI(par) as gh:
        gh.run()
        a = gh.maps(np.float64)
        for v in (1, 2, 0):  # other LDS table sizes change the order of every atomic
            gh.set_schedule([0, 0, 0, v])
            gh.run()
            b = gh.maps(np.float64)
            assert np.array_equal(a, b), v
    assert (a != 0).sum() > 10000


def test_deterministic_vs_float_scatter_and_oracle(cfgdir):
    par = small_params(cfgdir, "det_vs", n_grid=48, n_side=16)
    om, _, _ = oracle.run(par.desc())
    with crime_amd.GetHI(par, deterministic=1) as gh:
        gh.run()
        m = gh.maps(np.float64)
    assert np.array_equal(m != 0, om != 0)
    err = rel(m, om)
    print(f"deterministic FP64 maps vs oracle: max rel {err:.2e}")
    assert err < 1e-10


def test_deterministic_batch_equals_single(cfgdir):
    pars = [small_params(cfgdir, f"det_b{i}", n_grid=n, n_side=16, seed=50 + i, precision="single",
                         extra=dict(deterministic=1))
            for i, n in enumerate((32, 48, 64))]
    hs = [crime_amd.GetHI(p) for p in pars]
    try:
        single = []
        for h in hs:
            h.run()
            single.append(h.maps())
        with crime_amd.Batch(hs, n_streams=3) as b:
            for rep in range(3):
                for h in hs:  # eager work between replays
                    h.create_d_and_vr_fields()
                b.run()
                for h, ref in zip(hs, single):
                    assert np.array_equal(h.maps(), ref), rep
    finally:
        for h in hs:
            h.close()


def test_set_seed_equals_fresh_handle(cfgdir):
    """crime_gh_set_seed re-keys the mode fill: the next run equals a handle created with that seed."""
    par = small_params(cfgdir, "det_seed", n_grid=48, n_side=16, precision="single", extra=dict(deterministic=1))
    with crime_amd.GetHI(par, seed=11) as a, crime_amd.GetHI(par, seed=12) as b:
        a.run()
        first = a.maps()
        a.set_seed(12)
        a.run()
        b.run()
        assert np.array_equal(a.maps(), b.maps())
        assert not np.array_equal(first, b.maps())


@pytest.mark.parametrize("prec", ["single", "double"])
def test_tile_culling_is_exact(cfgdir, prec):
    """The map stage skips whole tiles before any grid load when the grid's bound on |phi| (from the kx
    pass) keeps every sample out of the band.  Writing a grid from outside drops that bound, so the same
    fields then go through the per-cell test only: with the fixed-point scatter the two maps must agree
    bit for bit.  The band (1250-1300 MHz in a 1500 Mpc/h box) leaves the inner sphere and the corners
    of the box out, so most tiles are culled."""
    par = small_params(cfgdir, f"cull_{prec}", n_grid=96, n_side=32, precision=prec, l_box=1500.0,
                       nu_lo=1250.0, nu_hi=1300.0, extra=dict(deterministic=1))
    with crime_amd.GetHI(par) as gh:
        gh.run()
        culled = gh.maps(np.float64)
        dens = gh.grid(crime_amd.GRID_DENS)
        gh.set_grid(crime_amd.GRID_DENS, dens)  # the same values: only the bound is dropped
        gh.mk_T_maps()
        full = gh.maps(np.float64)
    print(f"tile culling ({prec}): {np.count_nonzero(full)} occupied bins", flush=True)
    assert np.count_nonzero(full) > 1000
    assert np.array_equal(culled, full)


@pytest.mark.parametrize("det", [0, 1])
def test_band_out_of_reach(cfgdir, det):
    """A band that no cell of the box can reach (540-560 MHz, z > 1.5, in a 1500 Mpc/h box): the tile
    range is empty, the map kernel launches no workgroup (handle stream) or only ones that return at once
    (captured graph), and the maps are zero both ways; a band inside the box afterwards still fills."""
    far = small_params(cfgdir, f"far_{det}", n_grid=64, n_side=16, precision="single", l_box=1500.0,
                       nu_lo=540.0, nu_hi=560.0, extra=dict(deterministic=det))
    with crime_amd.GetHI(far) as gh:
        gh.run()
        assert not np.any(gh.maps(np.float64))
        with crime_amd.Batch([gh], n_streams=1) as b:
            b.run()
        assert not np.any(gh.maps(np.float64))
    near = small_params(cfgdir, f"near_{det}", n_grid=64, n_side=16, precision="single", l_box=1500.0,
                        nu_lo=1250.0, nu_hi=1300.0, extra=dict(deterministic=det))
    with crime_amd.GetHI(near) as gh:
        gh.run()
        assert np.count_nonzero(gh.maps(np.float64)) > 100
