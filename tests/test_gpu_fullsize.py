"""Parity at the workloads' own sizes.

- 256^3 FP64 workload (L=4900, 64 channels 700-800 MHz, Nside 128, 2^3 sub-cells, b=1, seed 1001) with
  the reference's single-thread MT19937 unit modes injected into both codes: maps against the FP64
  oracle, every pixel.
- 1024^3 FP32 production workload:
  - (channel, pixel) of every sample of slabs of cells bit-exact against the oracle given the same RSD
    shifts; sigma^2 against the oracle's FP64 reduction of the same field;
  - the timed map kernel (cell screen, LDS aggregation, FP32 atomics) against an FP64 scatter of every
    sample of the same fields, bin by bin, and per channel;
  - the FP32 maps against an FP64 run of the same workload (FP64 fields, FP64 maps): the map error of
    the FP32 path, stated and bounded.
- 2048^3 FP32 workload (400 channels, Nside 1024, 4^3 sub-cells, spectral velocities): index slabs
  bit-exact against the oracle, per-channel totals of the timed kernel, the FP32 maps against an FP64
  run (stencil velocities: the spectral v_r equals the stencil's to rounding, and the FP64 spectral
  grids would not fit next to the maps), spectra of 64 channels.
"""
import ctypes as C
import time

import numpy as np
import pytest

import crime_amd
import crime_amd.configs as cf
import oracle
from crime_amd import _ffi

pytestmark = pytest.mark.gpu


def test_small_fp64_256_maps_vs_oracle(cfgdir):
    par = crime_amd.Params(cf.small_fp64_256(cfgdir))
    d = par.desc()
    n = d.n_grid
    um = np.empty(2 * n * n * (n // 2 + 1))
    _ffi.host_lib().gh_unit_modes_mt(n, d.seed, um.ctypes.data_as(C.POINTER(C.c_double)))
    with crime_amd.GetHI(par, mode_source=crime_amd.MODES_INJECTED, unit_modes=um) as gh:
        st = gh.run()
        m = gh.maps(np.float64)
    d.mode_source = crime_amd.MODES_INJECTED
    d.unit_modes = um.ctypes.data_as(type(d.unit_modes))
    om, s2, _ = oracle.run(d)
    assert abs(st["sigma2_gauss"] - s2) < 1e-12 * s2
    nz = om != 0
    assert nz.sum() > 0.9 * om.size          # the shell fills the sky in every channel
    assert np.array_equal(m != 0, nz)
    err = np.abs(m - om)[nz] / np.abs(om[nz])
    print(f"256^3 FP64 maps vs oracle: max rel {err.max():.3e}, rms rel {np.sqrt(np.mean(err ** 2)):.3e}")
    assert err.max() < 1e-10


@pytest.fixture(scope="module")
def production(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("prod"))
    par = crime_amd.Params(cf.production_1024(d))
    gh = crime_amd.GetHI(par)
    st = gh.run()
    yield par, gh, st
    gh.close()


def test_production_sigma2_and_mean(production):
    par, gh, st = production
    dens = gh.grid(crime_amd.GRID_DENS)
    s2, mean = oracle.sigma2(dens)
    del dens
    assert abs(st["sigma2_gauss"] - s2) < 1e-10 * s2   # same FP32 values, 1e9-term FP64 sums in another order
    assert abs(mean) < 1e-6 * np.sqrt(s2)             # no k = 0 mode


def test_production_sample_indices_bitexact(production):
    par, gh, st = production
    d = par.desc()
    n = d.n_grid
    dz = gh.grid(crime_amd.GRID_DZRSD).ravel()
    t0 = time.time()
    total = 0
    # slabs of 4 planes at radii inside the shell, at its inner and outer edges and in a box corner
    for iz in (n // 2 - 300, n // 2 - 2, n // 2 + 236, n // 2 + 400, 40):
        c0, nc = iz * n * n, 4 * n * n
        g_inu, g_pix = gh.sample_indices(c0, nc)
        o_inu, o_pix = oracle.sample_indices(d, dz, c0, nc)
        assert np.array_equal(g_inu, o_inu)
        assert np.array_equal(g_pix, o_pix)
        total += int((g_inu >= 0).sum())
    print(f"1024^3: {total} binned samples bit-exact in {time.time() - t0:.1f} s")
    assert total > 10 ** 6


def _bin_errors(a, b):
    """Relative errors of a against b over the bins b occupies, and the bins occupied by one only."""
    nz = b != 0
    err = np.abs(a[nz] - b[nz]) / np.abs(b[nz])
    return err, int(np.count_nonzero((a != 0) != nz))


def test_production_timed_kernel_vs_fp64_scatter(production):
    """k_pixelize's FP32 maps against an FP64 scatter of every sample of the same FP32 fields (no cell
    screen, no LDS aggregation): same occupied bins, every bin within the FP32 accumulation error, and
    per-channel totals against an independent per-channel FP64 sum."""
    par, gh, st = production
    m32 = gh.maps(np.float64)
    ref = gh.reference_maps()
    err, occ = _bin_errors(m32, ref)
    print(f"1024^3 timed kernel vs FP64 scatter: {err.size} bins, max rel {err.max():.2e}, "
          f"rms rel {np.sqrt(np.mean(err ** 2)):.2e}, occupancy mismatches {occ}")
    assert occ == 0
    assert err.max() < 1e-5
    tot, cnt = gh.channel_totals()
    ch = m32.sum(axis=1)
    cerr = np.abs(ch - tot) / tot
    print(f"1024^3 per-channel totals: max rel {cerr.max():.2e}; {cnt.sum()} binned samples")
    assert np.all(cnt > 0)
    assert cerr.max() < 1e-6
    assert np.allclose(ref.sum(axis=1), tot, rtol=1e-10, atol=0)


def test_production_fp32_vs_fp64_run(production, tmp_path):
    """The FP32 path against an FP64 run of the same workload (same Philox draws, FP64 fields and maps):
    the map error of the FP32 path.  Samples whose channel flips with the FP32 RSD shift move their mass
    between neighbouring bins; they are counted on slabs through the per-sample indices."""
    par, gh, st = production
    par64 = crime_amd.Params(cf.production_1024(str(tmp_path), precision="double"))
    m32 = gh.maps(np.float64)
    with crime_amd.GetHI(par64) as g64:
        st64 = g64.run()
        m64 = g64.maps(np.float64)
        n = par.desc().n_grid
        flips = total = 0
        for iz in (n // 2 - 300, n // 2 + 236, n // 2 + 400):
            c0, nc = iz * n * n, n * n
            a, _ = gh.sample_indices(c0, nc)
            b, _ = g64.sample_indices(c0, nc)
            flips += int(np.count_nonzero(a != b))
            total += int(np.count_nonzero(b >= 0))
    err, occ = _bin_errors(m32, m64)
    within = float(np.mean(err < 1e-5))
    med = float(np.median(err))
    # map-level RMS: the rms of the difference over the rms of the FP64 map (per-bin relative errors are
    # dominated by the few bins a channel flip moves a sample into or out of)
    rms_map = float(np.sqrt(np.mean((m32 - m64) ** 2) / np.mean(m64 ** 2)))
    print(f"1024^3 FP32 vs FP64 maps: max rel {err.max():.2e}, median rel {med:.2e}, {100 * within:.4f}% of "
          f"{err.size} bins within 1e-5, rms(diff)/rms(map) {rms_map:.2e}, occupancy mismatches {occ}; channel "
          f"flips {flips} of {total} binned samples on 3 planes; sigma2 {st['sigma2_gauss']:.9f} vs "
          f"{st64['sigma2_gauss']:.9f}")
    assert abs(st["sigma2_gauss"] - st64["sigma2_gauss"]) < 1e-5 * st64["sigma2_gauss"]
    assert within > 0.9999
    assert med < 1e-6
    # the rms is set by the ~1e-7 of binned samples whose channel flips with the FP32 RSD shift (the
    # heaviest bins, near the observer, weigh most); every other bin agrees to FP32 accuracy
    assert rms_map < 1e-3
    assert flips < 1e-4 * total
    assert abs(m32.sum() - m64.sum()) < 1e-5 * m64.sum()


@pytest.fixture(scope="module")
def largest(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("largest"))
    par = crime_amd.Params(cf.largest_2048(d))
    assert par.get("velocity_mode") == crime_amd.VEL_SPECTRAL
    gh = crime_amd.GetHI(par)
    print("2048^3: running", flush=True)
    st = gh.run()
    yield d, par, gh, st
    gh.close()


def test_largest_2048_index_slabs_bitexact(largest):
    """(channel, pixel) of every sample of partial planes at several radii of the 2048^3 box, with 4^3
    sub-cells, against the oracle given the GPU's RSD shifts of those cells."""
    d_, par, gh, st = largest
    d = par.desc()
    n = d.n_grid
    total = 0
    for iz, iy in ((n // 2 - 600, n // 2), (n // 2 + 450, 100), (n // 2 + 900, n // 2 + 50), (n // 2 + 20, n // 2)):
        c0, nc = (iz * n + iy) * n, 32 * n
        dz = gh.grid_range(crime_amd.GRID_DZRSD, c0, nc)
        g_inu, g_pix = gh.sample_indices(c0, nc)
        o_inu, o_pix = oracle.sample_indices_slab(d, dz, c0, nc)
        assert np.array_equal(g_inu, o_inu)
        assert np.array_equal(g_pix, o_pix)
        total += int((g_inu >= 0).sum())
    print(f"2048^3: {total} binned samples bit-exact")
    assert total > 10 ** 6


def test_largest_2048_channel_totals(largest):
    d_, par, gh, st = largest
    tot, cnt = gh.channel_totals()
    ch = gh.maps().sum(axis=1, dtype=np.float64)
    cerr = np.abs(ch - tot) / tot
    print(f"2048^3 timed kernel per-channel totals: max rel {cerr.max():.2e} over {cnt.sum()} binned samples")
    assert np.all(cnt > 0)
    assert cerr.max() < 1e-6


def test_largest_2048_fp32_vs_fp64_run(largest):
    """The 2048^3 FP32 maps (spectral velocities) against an FP64 run with the reference's stencil
    velocities (the FP64 spectral grids would not fit beside the FP64 maps; the two velocity routes agree
    to rounding), channel by channel."""
    d_, par, gh, st = largest
    m32 = gh.maps()
    gh.close()  # frees the FP32 grids and maps on the device
    par64 = crime_amd.Params(cf.largest_2048(d_, name="largest_2048_fp64", precision="double",
                                             velocity_mode="stencil"))
    with crime_amd.GetHI(par64) as g64:
        print("2048^3 FP64: running", flush=True)
        st64 = g64.run()
        m64 = g64.maps(np.float64)
    nbins = nwithin = nocc = 0
    emax = dsq = msq = 0.0
    meds = []
    for c in range(m64.shape[0]):
        a = m32[c].astype(np.float64)
        err, occ = _bin_errors(a, m64[c])
        nbins += err.size
        nwithin += int(np.count_nonzero(err < 1e-5))
        nocc += occ
        emax = max(emax, float(err.max()))
        meds.append(float(np.median(err)))
        dsq += float(np.sum((a - m64[c]) ** 2))
        msq += float(np.sum(m64[c] ** 2))
    rms_map = np.sqrt(dsq / msq)
    print(f"2048^3 FP32 vs FP64 maps: max rel {emax:.2e}, median rel {np.median(meds):.2e}, "
          f"{100 * nwithin / nbins:.4f}% of {nbins} bins within 1e-5, rms(diff)/rms(map) {rms_map:.2e}, occupancy "
          f"mismatches {nocc}; sigma2 {st['sigma2_gauss']:.9f} vs {st64['sigma2_gauss']:.9f}")
    assert abs(st["sigma2_gauss"] - st64["sigma2_gauss"]) < 1e-5 * st64["sigma2_gauss"]
    assert nwithin > 0.9999 * nbins
    assert rms_map < 1e-3


def test_largest_2048_spectra(tmp_path):
    """The 2048^3 mock and the cross spectra of its first 64 channels at lmax 3071 (Legendre table
    rebuilt in m-chunks): symmetric, non-negative autos."""
    par = crime_amd.Params(cf.largest_2048(str(tmp_path)))
    with crime_amd.GetHI(par) as gh:
        st = gh.run()
        ptr, nbytes = gh.device_maps()
        with crime_amd.Sht(gh.n_side, 3 * gh.n_side - 1, max_maps=64) as sh:
            cls = sh.cls_of(gh, cross=True, channels=64)
            sst = sh.stats()
    print(f"2048^3: {st['ms_total']:.1f} ms, sigma2 {st['sigma2_gauss']:.6f}, maps {nbytes / 2 ** 30:.1f} GiB")
    print(f"C_l of 64 channels at Nside 1024, lmax 3071: {sst}")
    assert np.isfinite(st["sigma2_gauss"]) and 0.05 < st["sigma2_gauss"] < 5
    assert np.all(np.isfinite(cls)) and cls.shape == (64, 64, 3072)
    assert np.allclose(cls, cls.transpose(1, 0, 2), rtol=1e-12, atol=0)
    assert np.all(np.einsum("aal->al", cls) >= 0) and cls[0, 0, 2:].max() > 0
