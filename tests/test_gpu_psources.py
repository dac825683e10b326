"""Point sources on the GPU (crime_gh_mk_psources_maps) against the oracle (oracle/psrc_oracle.c) on
the GPU's own resident delta and Delta z_RSD grids, so the comparison isolates the point-source pass.

Bars: the number of sources exactly (same Philox streams, same samplers); the per-pixel terms A, B to
1e-12 of their largest value (FP64 atomics in another order); the maps to 1e-11 (FP64 grids) or 2e-7
(FP32 maps) of their largest value -- the oracle sums temp_of_l per source and channel as the reference
does, the GPU expands A, B per channel.
"""
import numpy as np
import pytest

import crime_amd
import oracle
from conftest import small_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["double", "single"])
def test_psources_vs_oracle(cfgdir, precision):
    par = small_params(cfgdir, f"ps_{precision}", precision=precision, n_grid=32, n_side=16, n_nu=12)
    d = par.desc()
    with crime_amd.GetHI(par) as gh:
        gh.run()
        st = gh.mk_psources_maps()
        maps = gh.ps_maps()
        a, b = gh.ps_pixel_sums()
        dens = gh.grid(crime_amd.GRID_DENS).ravel()
        dz = gh.grid(crime_amd.GRID_DZRSD).ravel()
        s2, _ = gh.sigma2()
        # a second pass redraws the same sources (summed by atomics in another order)
        st2 = gh.mk_psources_maps()
        assert st2["n_sources"] == st["n_sources"]
        assert np.abs(gh.ps_maps() - maps).max() < 1e-13 * np.abs(maps).max()
    print(f"\npoint sources ({precision}): {st}", flush=True)
    ref_maps, ra, rb, total = oracle.psources(d, dens, dz, s2)
    assert st["n_sources"] == total
    assert total > 10000
    assert np.abs(a - ra).max() < 1e-12 * np.abs(ra).max()
    assert np.abs(b - rb).max() < 1e-12 * np.abs(rb).max()
    assert np.array_equal(a > 0, ra > 0)
    tol = 1e-11 if precision == "double" else 2e-7
    assert np.abs(maps - ref_maps).max() < tol * np.abs(ref_maps).max()


def test_psources_production(cfgdir):
    """The 1024^3 production box: timing and sources count, and the maps equal c1(nu) A + c2(nu) B."""
    import crime_amd.configs as cf

    par = crime_amd.Params(cf.production_1024(cfgdir))
    with crime_amd.GetHI(par) as gh:
        gh.run()
        st = gh.mk_psources_maps()
        st = gh.mk_psources_maps()
        a, b = gh.ps_pixel_sums()
        p, nb = gh.device_ps_maps()
        maps = gh.ps_maps()[[0, 199]]
        d = par.desc()
    print(f"\nproduction point sources: {st}", flush=True)
    assert st["n_sources"] > 10 ** 8
    assert nb == 200 * 12 * 512 ** 2 * 4
    assert np.all(np.isfinite(maps)) and maps.min() >= 0
    npix = 12 * 512 ** 2
    K = 3.2548291e-2 * 8.35774e7 * 4 * np.pi * 0.114461 * d.hhub ** 2 / (4 * np.pi / npix)
    for row, i in ((0, 0), (1, 199)):
        nu = 0.5 * (d.nu0[i] + d.nuf[i])
        two = K * (nu ** -2.1 * a + nu ** -2.75 * b)
        assert np.abs(maps[row] - two).max() < 2e-7 * np.abs(two).max()


def _read_healpix_fits(fn):
    """T column of a HEALPix BINTABLE written by gh_write_healpix_fits: two HDUs of 2880-byte blocks,
    the table's data big-endian FP32 after its header."""
    raw = open(fn, "rb").read()
    pos, hdus = 0, []
    while pos < len(raw):
        cards = {}
        while True:
            block = raw[pos:pos + 2880]
            pos += 2880
            done = False
            for i in range(0, 2880, 80):
                card = block[i:i + 80].decode("ascii")
                if card.startswith("END"):
                    done = True
                    break
                if "=" in card[:10]:
                    cards[card[:8].strip()] = card[10:].split("/")[0].strip().strip("'").strip()
            if done:
                break
        nbytes = int(cards.get("NAXIS1", 0)) * int(cards.get("NAXIS2", 0)) if cards.get("NAXIS") == "2" else 0
        hdus.append((cards, raw[pos:pos + nbytes]))
        pos += (nbytes + 2879) // 2880 * 2880
    cards, data = hdus[1]
    return np.frombuffer(data, dtype=">f4").astype(np.float64), cards


def test_cli_writes_point_source_maps(cfgdir):
    """gethi with do_psources = 1: the HI maps and the <prefix>_ps_###.fits maps equal those of the
    library calls on the same parameter file (src/main_gh.c:55-66, src/io_gh.c:122-127)."""
    import os
    import subprocess
    import crime_amd.configs as cf

    ini = cf.make(cfgdir, "cli_ps", n_grid=32, n_side=16, n_nu=4, nu_lo=1150.0, nu_hi=1300.0, l_box=600.0,
                  seed=77, precision="single", extra={"do_psources": 1})
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "crime_amd", "_lib", "gethi")
    print(f"\nrunning {exe} {ini}", flush=True)
    res = subprocess.run([exe, ini], capture_output=True, text=True, timeout=120)
    print(res.stdout[-600:], res.stderr[-600:], flush=True)
    assert res.returncode == 0
    assert "particles in total" in res.stdout
    par = crime_amd.Params(ini)
    with crime_amd.GetHI(par) as gh:
        gh.run()
        hi = gh.maps(np.float64)
        gh.mk_psources_maps()
        ps = gh.ps_maps()
    prefix = os.path.join(cfgdir, "cli_ps")
    for inu in range(4):
        m, cards = _read_healpix_fits(f"{prefix}_{inu + 1:03d}.fits")
        assert cards["ORDERING"] == "RING" and int(cards["NSIDE"]) == 16
        assert np.abs(m - hi[inu]).max() <= 1e-6 * np.abs(hi[inu]).max()
        p, _ = _read_healpix_fits(f"{prefix}_ps_{inu + 1:03d}.fits")
        assert np.abs(p - ps[inu]).max() <= 1e-6 * np.abs(ps[inu]).max()
        assert p.max() > 0
