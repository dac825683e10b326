"""Kernel schedules and the autotuner: the tiling variants of the FFT stages give bit-identical fields and
the radix-8 variant the same fields to FP32 rounding, every map-stage variant the same maps (to the FP32
atomic order; bit for bit with the deterministic scatter, also where the screen's fallback queue overflows), the tuner picks the fastest timed variant of each stage, batches capture the tuned
schedule, and the schedule cache returns it without timing."""
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
        for v in range(4):
            for stage in range(3):
                s = [0, 0, 0, 0]
                s[stage] = v
                gh.set_schedule(s)
                assert gh.schedule() == s
                gh.create_d_and_vr_fields()
                d, p = gh.grid(crime_amd.GRID_DENS), gh.grid(crime_amd.GRID_VPOT)
                if v < 3:
                    assert np.array_equal(d, ref_d), (stage, v)
                    assert np.array_equal(p, ref_v), (stage, v)
                else:  # radix-8 plans: other butterflies, FP32 rounding
                    assert np.abs(d - ref_d).max() < 1e-5 * ref_d.std(), stage
                    assert np.abs(p - ref_v).max() < 1e-5 * ref_v.std(), stage
            gh.set_schedule([0, 0, 0, v])
            gh.run()
            m = gh.maps(np.float64)
            assert np.array_equal(m != 0, ref_m != 0)
            assert np.abs(m - ref_m).max() < 2e-5 * np.abs(ref_m).max(), v


def test_autotune_picks_fastest_and_batch_uses_it(cfgdir):
    pars = [small_params(cfgdir, f"tune_{i}", n_grid=128, precision="single", n_side=32, l_box=1500.0, seed=50 + i)
            for i in range(2)]
    hs = [crime_amd.GetHI(p) for p in pars]
    sched, table = hs[0].autotune(reps=2)
    print("schedule", sched, table)
    for i, st in enumerate(crime_amd.GetHI.STAGES):
        times = [t for t in table[st] if t >= 0]
        assert len(times) == 4
        assert table[st][sched[i]] == min(times)
    assert hs[0].schedule() == sched
    m0 = hs[0].maps(np.float64)   # the tuner leaves the handle's maps computed
    # references with the tuned schedule (a radix-8 variant changes the fields' rounding, and through the
    # RSD shift the channel of a few samples)
    hs[1].set_schedule(sched)
    ref = []
    for h in hs:
        h.run()
        ref.append(h.maps(np.float64))
    assert np.abs(m0 - ref[0]).max() < 2e-5 * np.abs(ref[0]).max()
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
            gh.set_schedule([4, 0, 0, 0])
        with pytest.raises(crime_amd.CrimeError):
            gh.set_schedule([0, 0, 0, 4])
        sched, table = gh.autotune(reps=1)
        assert sched[:3] == [0, 0, 0]
        assert table["fft_y"][1] == -1


def test_schedule_cache(cfgdir, tmp_path):
    """A miss tunes and appends; a hit on another handle of the same kind sets the cached schedule; another
    kind misses."""
    path = str(tmp_path / "schedules.txt")
    pars = [small_params(cfgdir, f"cache_{i}", n_grid=128, precision="single", n_side=32, l_box=1500.0, seed=70 + i)
            for i in range(2)]
    with crime_amd.GetHI(pars[0]) as a, crime_amd.GetHI(pars[1]) as b:
        sa, hit = a.autotune_cached(path, reps=1)
        assert not hit
        b.set_schedule([0, 0, 0, 0])
        sb, hit = b.autotune_cached(path, reps=1)
        assert hit and sb == sa and b.schedule() == sa
    with crime_amd.GetHI(small_params(cfgdir, "cache_other", n_grid=128, n_side=16, l_box=1500.0)) as c:
        _, hit = c.autotune_cached(path, reps=1)
        assert not hit
    lines = open(path).read().splitlines()
    assert len(lines) == 2 and all(len(l.split()) == 5 for l in lines)


@pytest.mark.parametrize("kw", [dict(), dict(n_nu=600, nu_lo=1300.0, nu_hi=1306.0)])
def test_pixelize_variants_identical_deterministic(cfgdir, kw):
    """With the fixed-point scatter the maps do not depend on the order of the additions, so every map-stage
    variant (2048 LDS bins in groups of 4 or 8 map words, 4096 and 1024 bins) must give the same maps bit
    for bit.  The second case has 0.01 MHz channels, narrower than twice the frequency margin: nearly every
    binned sample misses it, so the per-wave fallback queue (32 entries) overflows and the in-step FP64
    resolution runs too."""
    name = "pixvar_" + "_".join(f"{k}{v}" for k, v in kw.items())
    par = small_params(cfgdir, name, n_grid=64, precision="single", extra=dict(deterministic=1), **kw)
    with crime_amd.GetHI(par) as gh:
        gh.create_d_and_vr_fields()
        c = gh.screen_check()
        print(f"{name}: {c['binned']} of {c['samples']} samples binned; fallbacks per binned sample: channel "
              f"{c['channel_fallbacks'] / max(c['binned'], 1):.3f}, pixel {c['pixel_fallbacks'] / max(c['binned'], 1):.2e}",
              flush=True)
        assert c["mismatches"] == 0
        maps = []
        for v in range(4):
            gh.set_schedule([0, 0, 0, v])
            gh.run()
            maps.append(gh.maps(np.float64))
        assert np.count_nonzero(maps[0]) > 1000
        for v in range(1, 4):
            assert np.array_equal(maps[v], maps[0]), v
        if kw:
            assert c["channel_fallbacks"] > 0.1 * c["binned"]
