"""Kernel schedules and the autotuner: every FFT variant gives bit-identical fields, every map-stage
variant the same maps (to the FP32 atomic order), the tuner picks the fastest timed variant of each
stage, and batches capture the tuned schedule."""
import numpy as np
import pytest

import crime_amd
from conftest import small_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [128, 192])
def test_fft_variants_bit_identical(cfgdir, n):
    par = small_params(cfgdir, f"sched_{n}", n_grid=n, precision="single", n_side=32, l_box=1500.0)
    with crime_amd.GetHI(par) as gh:
        gh.run()
        ref_d, ref_v = gh.grid(crime_amd.GRID_DENS), gh.grid(crime_amd.GRID_VPOT)
        ref_m = gh.maps(np.float64)
        for v in range(3):
            for stage in range(3):
                s = [0, 0, 0, 0]
                s[stage] = v
                gh.set_schedule(s)
                assert gh.schedule() == s
                gh.create_d_and_vr_fields()
                assert np.array_equal(gh.grid(crime_amd.GRID_DENS), ref_d), (stage, v)
                assert np.array_equal(gh.grid(crime_amd.GRID_VPOT), ref_v), (stage, v)
            gh.set_schedule([0, 0, 0, v])
            gh.run()
            m = gh.maps(np.float64)
            assert np.array_equal(m != 0, ref_m != 0)
            assert np.abs(m - ref_m).max() < 2e-5 * np.abs(ref_m).max(), v


def test_autotune_picks_fastest_and_batch_uses_it(cfgdir):
    pars = [small_params(cfgdir, f"tune_{i}", n_grid=128, precision="single", n_side=32, l_box=1500.0, seed=50 + i)
            for i in range(2)]
    hs = [crime_amd.GetHI(p) for p in pars]
    ref = []
    for h in hs:
        h.run()
        ref.append(h.maps(np.float64))
    sched, table = hs[0].autotune(reps=2)
    print("schedule", sched, table)
    for i, st in enumerate(crime_amd.GetHI.STAGES):
        times = [t for t in table[st] if t >= 0]
        assert len(times) == 3
        assert table[st][sched[i]] == min(times)
    assert hs[0].schedule() == sched
    m0 = hs[0].maps(np.float64)   # the tuner leaves the handle's maps computed
    assert np.abs(m0 - ref[0]).max() < 2e-5 * np.abs(ref[0]).max()
    hs[1].set_schedule(sched)
    with crime_amd.Batch(hs, n_streams=2) as b:
        b.run()
        for h, r in zip(hs, ref):
            m = h.maps(np.float64)
            assert np.abs(m - r).max() < 2e-5 * np.abs(r).max()
    for h in hs:
        h.close()


def test_small_grids_have_one_fft_variant(cfgdir):
    par = small_params(cfgdir, "sched_small", n_grid=32)
    with crime_amd.GetHI(par) as gh:
        gh.set_schedule([2, 1, 2, 1])
        assert gh.schedule() == [0, 0, 0, 1]
        with pytest.raises(crime_amd.CrimeError):
            gh.set_schedule([3, 0, 0, 0])
        sched, table = gh.autotune(reps=1)
        assert sched[:3] == [0, 0, 0]
        assert table["fft_y"][1] == -1
