#!/usr/bin/env python3
"""Benchmark of the GetHI density-to-sky path on one MI355X.

A step is one full pass of the hot path (mode fill, 3-D inverse FFT of delta and phi, sigma^2, radial
velocity, lognormal HI mass, RSD, sub-cell ang2pix and the (nu, pixel) scatter into zeroed mK maps)
over the 1024^3 production workload, with every input already resident in HBM.  value = cells/s.

Prints one JSON line.  Extra objects: "roofline" for the dominant kernel (the map kernel, priced against
the VALU issue rate that bounds it, with its HBM figures beside), "cpu_baseline" (the FP64 C oracle,
OpenMP on the host cores, over a bounded sample of the same workload), "stages" (ms per stage),
"pixelize" (samples/s, binned samples/s, global atomics/s), "end_to_end_cli" (the gethi CLI from the
.ini to the FITS files), and the other two metrics of the headline: "mocks_per_hour_batched" (the
96-member ensemble in one graph) and "emulator_err_sigma" (the C_l emulator trained and judged at full
size).  --quick skips the last three.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# VALU issue peak of the MI355X: its published FP32 vector peak, 157.3 TFLOP/s, is 7.86e13 lane-FMAs per
# second = 256 CUs x 4 SIMDs x 32 lanes x 2.4 GHz, a wave64 instruction every 2 cycles per SIMD.  Measured on
# the MI355X, only the simplest instructions issue that fast (FP32 add/mul, 32-bit add/and, moves); FP32 fma,
# compares, selects, floors and conversions take 4 cycles, FP64 arithmetic and 32-bit multiplies 4.8,
# v_rcp/v_sqrt/v_rsq 8 -- so this is an upper bound, and the roofline also reports the VALU busy share of
# the kernel's actual instruction mix
VALU_PEAK_LANE_INSTR = 256 * 4 * 32 * 2.4e9
WORKLOADS = ("production_1024", "small_fp64_256", "largest_2048")
METRIC = {"production_1024": "cells/s end-to-end at 1024^3",
          "small_fp64_256": "cells/s end-to-end at 256^3 (FP64)",
          "largest_2048": "cells/s end-to-end at 2048^3"}


def stage_bytes(n, prec):
    """Algorithmic HBM bytes per launch: complex grid element = 2 * sizeof(real)."""
    e = 2 * (8 if prec else 4)
    n3 = float(n) ** 3
    return {
        "fill_fft_z": e * n3,            # write F once (modes are generated, not read)
        "fft_y": 2 * e * n3,             # read + write F
        "fft_x": 2 * e * n3,             # read + write F (+ sigma^2 partials, negligible)
        "pixelize": e * n3,              # read (delta, phi) of every cell once
    }


KERNEL_OF_STAGE = {"fill_fft_z": "k_fill_z", "fft_y": "k_fft_y", "fft_x": "k_fft_x", "pixelize": "k_pixelize"}


def profile_values(workload, key):
    """Per-launch values of one counter-derived quantity from the committed PMC summaries
    (lines "<key> <kernel> <value>" in profiles/*_<workload>.md): kernel -> value."""
    out = {}
    pdir = os.path.join(ROOT, "profiles")
    for name in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if not name.endswith(f"_{workload}.md"):
            continue
        for line in open(os.path.join(pdir, name)):
            f = line.split()
            if len(f) == 3 and f[0] == key:
                out[f[1]] = float(f[2])
    return out


def measured_traffic(workload):
    """HBM bytes per launch from the committed PMC summary (profiles/hbm_traffic_<workload>.md)."""
    return profile_values(workload, "traffic_bytes")


def cpu_baseline(workload, n_cpu, tmp):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import crime_amd
    import crime_amd.configs as cf
    import oracle

    ini = getattr(cf, workload)(tmp, n_grid=n_cpu, name=f"cpu_{workload}")
    par = crime_amd.Params(ini)
    d = par.desc()
    t0 = time.perf_counter()
    oracle.run(d)
    dt = time.perf_counter() - t0
    return {"value": n_cpu ** 3 / dt, "unit": "cells/s", "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{workload} parameters at n_grid={n_cpu} ({n_cpu ** 3} cells, same box, channels, "
                      f"Nside and samples per cell), FP64 C oracle with OpenMP, {dt:.1f} s"}


def ensemble_rate(steps=3, streams=4, tune=True):
    """mocks/hour of the 96-member uneven ensemble in one captured graph (maps only, resident handles),
    and the plain per-mock rate on the same handles.  Returns a dict."""
    from concurrent.futures import ThreadPoolExecutor

    import crime_amd
    import crime_amd.configs as cf

    nmem = 96
    with tempfile.TemporaryDirectory() as tmp:
        inis = [cf.ensemble_member(tmp, i) for i in range(nmem)]
        with ThreadPoolExecutor(16) as ex:
            pars = list(ex.map(crime_amd.Params, inis))
        hs = [crime_amd.GetHI(p) for p in pars]
        try:
            if tune:
                schedules = {}
                for h in hs:
                    key = (h.n, h.n_side)
                    if key not in schedules:
                        schedules[key] = h.autotune(reps=2)[0]
                    h.set_schedule(schedules[key])
            for h in hs:
                h.run()
            t0 = time.perf_counter()
            for h in hs:
                h.run()
            ms_plain = (time.perf_counter() - t0) * 1e3
            with crime_amd.Batch(hs, n_streams=streams) as b:
                b.run()
                ms = min(b.run() for _ in range(steps))
        finally:
            for h in hs:
                h.close()
    return {"value": nmem * 3600e3 / ms, "unit": "mocks/hour", "ms_per_ensemble": ms, "members": nmem,
            "plain_resident_mocks_per_hour": nmem * 3600e3 / ms_plain, "speedup_vs_plain_resident": ms_plain / ms,
            "config": "32x128^3/Nside128 + 32x192^3/Nside256 + 32x256^3/Nside256, 64 channels, one hipGraph"}


def emulator_run(steps=20000, log=None):
    """The C_l emulator at full size (512 x 8 training mocks of 256^3 / Nside 256 / 64 channels, 64 x 16
    held-out): returns the result dict of crime_amd.emulator.evaluate plus rates."""
    import crime_amd.emulator as em

    spec = em.MockSpec(n_grid=256, n_side=256, n_nu=64)
    train_c = em.latin_hypercube(512, seed=1)
    test_c = em.latin_hypercube(64, seed=2)
    t0 = time.perf_counter()
    cls_tr, t_mock = em.simulate(train_c, 8, spec, seed0=100000, log=log)
    cls_te, _ = em.simulate(test_c, 16, spec, seed0=900000, log=log)
    t_sim = time.perf_counter() - t0
    t1 = time.perf_counter()
    emu, losses = em.train_emulator(train_c, cls_tr, spec, steps=steps, batch=4096, lr0=2e-3, seed=0, log=log)
    t_train = time.perf_counter() - t1
    res = em.evaluate(emu, train_c, cls_tr, test_c, cls_te, spec)
    xq = em.features(em.latin_hypercube(4096, seed=3), spec)
    emu.predict(xq)  # warm-up on the full batch: the timed call reuses its buffers, as repeated batches do
    t2 = time.perf_counter()
    emu.predict(xq)
    t_eval = (time.perf_counter() - t2) / 4096
    emu.close()
    res.update(ms_per_mock=t_mock * 1e3, us_per_emulated_cosmology=t_eval * 1e6,
               emulator_rate_vs_mock_rate=t_mock / t_eval, simulation_s=t_sim, training_s=t_train,
               training_ms_per_step=t_train * 1e3 / steps, final_loss=float(losses[-200:].mean()),
               gain_vs_nearest_cosmology=res["nearest_cosmology"]["rms_debiased"] /
               max(res["emulator"]["rms_debiased"], 1e-9))
    return res


def cli_end_to_end(workload, tmp):
    """Wall time of the gethi CLI from the parameter file to the FITS maps and nu table, in a child process
    (GPU set-up, host tables, run, maps to the host, FITS writes).  Output into tmp (removed afterwards)."""
    import subprocess

    import crime_amd.configs as cf

    ini = getattr(cf, workload)(tmp, name=f"cli_{workload}")
    exe = os.path.join(ROOT, "crime_amd", "_lib", "gethi")
    t0 = time.perf_counter()
    r = subprocess.run([exe, ini], capture_output=True, text=True)
    wall = time.perf_counter() - t0
    if r.returncode:
        return {"error": r.stderr[-500:]}
    out = {"seconds": wall}
    for line in r.stdout.splitlines():
        if "maps to the host in" in line:
            f = line.split()
            out["ms_maps_to_host"] = float(f[f.index("in") + 1])
            out["ms_fits_after_copy"] = float(f[f.index("+") + 1])
    nbytes = sum(os.path.getsize(os.path.join(tmp, f)) for f in os.listdir(tmp) if f.endswith(".fits"))
    out["fits_bytes"] = nbytes
    return out


def bench_ensemble(args):
    """mocks/hour of the 96-member uneven ensemble (32 x 128^3/Nside 128, 32 x 192^3/Nside 256,
    32 x 256^3/Nside 256; 64 channels; seeds 4000-4095; Omega_m, sigma_8 drawn per member): maps in one
    captured graph (value), then every member's 64 x 64 cross-frequency C_l, l <= 3 Nside - 1
    (mocks_per_hour_maps_and_cls)."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor

    import crime_amd
    import crime_amd.configs as cf

    nmem = 96
    with tempfile.TemporaryDirectory() as tmp:
        inis = [cf.ensemble_member(tmp, i) for i in range(nmem)]
        t0 = time.perf_counter()
        with ThreadPoolExecutor(16) as ex:
            pars = list(ex.map(crime_amd.Params, inis))
        t_host = time.perf_counter() - t0
        print(f"bench: host set-up of {nmem} members in {t_host:.2f} s (16 threads)", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        hs = [crime_amd.GetHI(p) for p in pars]
        t_create = time.perf_counter() - t0
        cells = sum(h.n ** 3 for h in hs)
        # one autotune per member kind (grid, Nside); its siblings take the schedule
        schedules = {}
        t0 = time.perf_counter()
        if not args.no_tune:
            for h in hs:
                key = (h.n, h.n_side)
                if key not in schedules:
                    schedules[key] = h.autotune(reps=2)[0]
                h.set_schedule(schedules[key])
        t_tune = time.perf_counter() - t0
        for h in hs:
            h.run()
        # plain per-mock path on resident handles: one crime_gh_run (with its own sync) per member
        t0 = time.perf_counter()
        for h in hs:
            h.run()
        ms_plain = (time.perf_counter() - t0) * 1e3
        print(f"bench: plain per-mock runs {ms_plain:.1f} ms for {nmem}", file=sys.stderr, flush=True)
        b = crime_amd.Batch(hs, n_streams=args.streams)
        for _ in range(args.warmup):
            b.run()
        ts = [b.run() for _ in range(args.steps)]
        ms = sum(ts) / len(ts)
        print(f"bench: batched graph {ms:.1f} ms per ensemble", file=sys.stderr, flush=True)
        b.close()
        # per-member cross-frequency C_l (l <= 3 Nside - 1) of the resident maps, left in HBM like the
        # maps.  Members of one Nside are analysed cls_batch at a time in one call (their 64 x 64
        # spectra each); the one-member-per-call path is timed beside it.  Stage times are summed
        # from the plans' own events.
        groups = {}
        for h in hs:
            groups.setdefault(h.n_side, []).append(h)
        nb = max(1, args.cls_batch)

        def spectra(plans, batch):
            stage, t0 = {}, time.perf_counter()
            for ns, g in groups.items():
                for i in range(0, len(g), batch):
                    plans[ns].cls_of_members(g[i:i + batch], to_host=False)
                    for k, v in plans[ns].stats().items():
                        stage[k] = stage.get(k, 0.0) + v
            return (time.perf_counter() - t0) * 1e3, stage

        plans1 = {ns: crime_amd.Sht(ns, 3 * ns - 1, max_maps=64) for ns in groups}
        spectra(plans1, 1)
        ms_cls1, stage1 = spectra(plans1, 1)
        for p in plans1.values():
            p.close()
        plans = {ns: crime_amd.Sht(ns, 3 * ns - 1, max_maps=64 * nb) for ns in groups}
        spectra(plans, nb)
        ms_cls, stage_ms = spectra(plans, nb)
        for p in plans.values():
            p.close()
        print(f"bench: per-member cross C_l of {nmem} members in {ms_cls:.1f} ms ({nb} members per call), "
              f"{ms_cls1:.1f} ms one per call", file=sys.stderr, flush=True)
        # the drop-in sequence for one mock (what one GetHI run per mock does): host set-up, device
        # set-up, run, maps to the host, release -- timed on 8 members of each kind
        sample = [i for k in range(3) for i in range(32 * k, 32 * k + 8)]
        for h in hs:
            h.close()
        t0 = time.perf_counter()
        for i in sample:
            p = crime_amd.Params(inis[i])
            with crime_amd.GetHI(p) as g:
                g.run()
                g.maps()
        ms_e2e = (time.perf_counter() - t0) * 1e3 / len(sample)
    out = {
        "metric": "mocks/hr batched 256^3",
        "value": nmem * 3600e3 / ms,
        "unit": "mocks/hour",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (Philox modes of BBKS P(k) per member cosmology; resident in HBM)",
        "config": {"workload": "ensemble_96", "members": "32x128^3/Nside128 + 32x192^3/Nside256 + "
                   "32x256^3/Nside256", "n_nu": 64, "cells_per_ensemble": cells, "streams": args.streams},
        "ms_cls_per_ensemble": ms_cls, "cls_stage_ms": stage_ms, "cls_members_per_call": nb,
        "ms_cls_one_member_per_call": ms_cls1, "cls_stage_ms_one_member_per_call": stage1,
        "mocks_per_hour_maps_and_cls": nmem * 3600e3 / (ms + ms_cls),
        "plain_resident_mocks_per_hour": nmem * 3600e3 / ms_plain,
        "plain_end_to_end_mocks_per_hour": 3600e3 / ms_e2e,
        "speedup_vs_plain_resident": ms_plain / ms,
        "speedup_vs_plain_end_to_end": (ms_e2e * nmem) / ms,
        "cells_per_s": cells / (ms * 1e-3),
        "host_setup_s_16_threads": t_host, "device_setup_s": t_create,
        "autotune": {f"{k[0]}^3/Nside{k[1]}": v for k, v in schedules.items()}, "autotune_s": t_tune,
    }
    print(json.dumps(out), flush=True)


def bench_emulator(args):
    """Emulator C_l err/sigma: 512 Latin-hypercube cosmologies x 8 realisations of 256^3 / Nside 256 /
    64-channel mocks train the network; 64 held-out cosmologies x 16 fresh realisations judge it."""
    import crime_amd.emulator as em

    def log(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    spec = em.MockSpec(n_grid=256, n_side=256, n_nu=64)
    train_c = em.latin_hypercube(512, seed=1)
    test_c = em.latin_hypercube(64, seed=2)
    t0 = time.perf_counter()
    cls_tr, t_mock = em.simulate(train_c, 8, spec, seed0=100000, log=log)
    cls_te, _ = em.simulate(test_c, 16, spec, seed0=900000, log=log)
    t_sim = time.perf_counter() - t0
    log(f"simulated {len(train_c) * 8 + len(test_c) * 16} mocks in {t_sim:.1f} s ({t_mock * 1e3:.2f} ms/mock)")
    t1 = time.perf_counter()
    emu, losses = em.train_emulator(train_c, cls_tr, spec, steps=args.emu_steps, batch=4096, lr0=2e-3, seed=0,
                                    log=log)
    t_train = time.perf_counter() - t1
    res = em.evaluate(emu, train_c, cls_tr, test_c, cls_te, spec)
    xq = em.features(em.latin_hypercube(4096, seed=3), spec)
    emu.predict(xq)  # warm-up on the full batch: the timed call reuses its buffers, as repeated batches do
    t2 = time.perf_counter()
    emu.predict(xq)
    t_eval = (time.perf_counter() - t2) / 4096
    emu.close()
    out = {
        "metric": "emulator C_l err/sigma",
        "value": res["emulator"]["rms_debiased"],
        "unit": "sigma (rms over 64 held-out cosmologies x 16 pairs x 16 l bins, truth noise removed)",
        "n_gpus": 1, "steps": args.emu_steps, "warmup": 0, "ms_per_step": t_train * 1e3 / args.emu_steps,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (mocks of this pipeline on Latin-hypercube cosmologies)",
        "config": {"workload": "emulator", "train": "512 cosmologies x 8 realisations", "test": "64 x 16",
                   "mock": "256^3, Nside 256, 64 channels 700-800 MHz", "network": "8-256x5-1 SiLU",
                   "pairs": len(spec.pairs), "l_bins": spec.nbins},
        "emulator": res["emulator"], "nearest_cosmology": res["nearest_cosmology"],
        "sigma_calibration": res["sigma_calibration"],
        "gain_vs_nearest_cosmology": res["nearest_cosmology"]["rms_debiased"] / max(res["emulator"]["rms_debiased"], 1e-9),
        "ms_per_mock": t_mock * 1e3, "us_per_emulated_cosmology": t_eval * 1e6,
        "emulator_rate_vs_mock_rate": t_mock / t_eval,
        "simulation_s": t_sim, "training_s": t_train, "final_loss": float(losses[-200:].mean()),
    }
    print(json.dumps(out), flush=True)


def bench_sht(args):
    """Angular spectra and the JoinT beam on the 200 resident FP32 channels of the 1024^3 production mock
    (Nside 512, lmax 1535).  A step is map2alm of every channel plus C_l of all 200 x 200 pairs, left in
    HBM; the beam (map2alm, b_l, alm2map, in place) is timed as its own step after them."""
    import numpy as np
    import crime_amd
    import crime_amd.configs as cf

    with tempfile.TemporaryDirectory() as tmp:
        par = crime_amd.Params(cf.production_1024(tmp))
        with crime_amd.GetHI(par) as gh:
            gh.run()
            nside, lmax, nch = gh.n_side, 3 * gh.n_side - 1, gh.n_nu
            nu0 = np.linspace(400.0, 798.0, nch)
            fwhm, _ = crime_amd.jt_instrument(nu0, nu0 + 2.0, nside, dish_diameter=15.0)
            with crime_amd.Sht(nside, lmax, max_maps=nch) as sh:
                def spectra():
                    sh.cls_of(gh, cross=True, to_host=False)
                    return sh.stats()

                def beam():
                    sh.instrument_of(gh, fwhm_arcmin=fwhm)
                    return sh.stats()

                for _ in range(args.warmup):
                    spectra()
                    beam()
                print("bench: warmup done", file=sys.stderr, flush=True)
                cl_st = [spectra() for _ in range(args.steps)]
                bm_st = [beam() for _ in range(args.steps)]
    mean = lambda sts, k: sum(st[k] for st in sts) / len(sts)
    ms_cl = mean(cl_st, "ms_total")
    ms_gemm = mean(cl_st, "ms_alm")
    nalm = (lmax + 1) * (lmax + 2) // 2
    flops = 2.0 * nalm * (2 * nside) * (2 * nch)      # alm = Lambda G: rows x rings x (re, im) columns
    achieved = flops / (ms_gemm * 1e-3) / 1e12
    out = {
        "metric": "maps/s map2alm + all-pair C_l at Nside 512", "value": nch / (ms_cl * 1e-3), "unit": "maps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_cl, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (the 200 FP32 channels of the 1024^3 production mock, resident in HBM)",
        "config": {"workload": "sht_production", "n_side": nside, "lmax": lmax, "maps": nch,
                   "pairs": nch * nch, "beam": "15 m dish, fwhm 1030610.55/(nu D) arcmin"},
        "roofline": {"bound": "mfma", "kernel": "k_alm_gemm", "achieved": achieved, "peak": 78.6,
                     "unit": "TFLOP/s", "frac": achieved / 78.6, "traffic": None,
                     "flop_per_launch": flops, "ms_per_launch": ms_gemm},
        "stages_ms": {k: mean(cl_st, k) for k in ("ms_rings", "ms_legendre", "ms_alm", "ms_cls")},
        "beam_ms": {k: mean(bm_st, k) for k in ("ms_rings", "ms_legendre", "ms_alm", "ms_cls", "ms_total")},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="production_1024",
                    choices=WORKLOADS + ("ensemble_96", "emulator", "sht_production"))
    ap.add_argument("--emu-steps", type=int, default=20000, help="training steps of the emulator workload")
    ap.add_argument("--no-tune", action="store_true", help="keep the default kernel schedule (no autotune)")
    ap.add_argument("--streams", type=int, default=4, help="concurrent streams of the ensemble graph")
    ap.add_argument("--cls-batch", type=int, default=16, help="ensemble members per spectra call")
    ap.add_argument("--cpu-n", type=int, default=768, help="grid of the CPU-baseline sample (0: skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (scale reference for PMC byte counters)")
    ap.add_argument("--quick", action="store_true",
                    help="the map workload only: no ensemble, emulator or CLI run (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the GetHI path on one GPU; multi-GPU slabs are a design note "
              "only (no collectives are built)", file=sys.stderr)
        sys.exit(2)

    import crime_amd
    import crime_amd.configs as cf

    if args.workload == "ensemble_96":
        return bench_ensemble(args)
    if args.workload == "emulator":
        return bench_emulator(args)
    if args.workload == "sht_production":
        return bench_sht(args)

    if args.calib_copy:
        import torch

        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("bench: calibration copy of 1 GiB done", file=sys.stderr, flush=True)
        del src, dst

    with tempfile.TemporaryDirectory() as tmp:
        ini = getattr(cf, args.workload)(tmp)
        par = crime_amd.Params(ini)
        d = par.desc()
        n, prec = d.n_grid, d.precision
        print(f"bench: {args.workload} N={n} Nside={d.n_side} n_nu={d.n_nu} n_sub={d.n_sub} "
              f"prec={'fp64' if prec else 'fp32'}", file=sys.stderr, flush=True)
        gh = crime_amd.GetHI(par)
        tune = None
        if not args.no_tune:
            t0 = time.perf_counter()
            sched, table = gh.autotune(reps=2)
            tune = {"schedule": sched, "ms_per_variant": table, "seconds": time.perf_counter() - t0}
            print(f"bench: autotuned schedule {sched} in {tune['seconds']:.1f} s", file=sys.stderr, flush=True)
        for i in range(args.warmup):
            gh.run()
            print(f"bench: warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
        stats = []
        t0 = time.perf_counter()
        for i in range(args.steps):
            stats.append(gh.run())
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        print(f"bench: {args.steps} steps, {ms:.2f} ms/step", file=sys.stderr, flush=True)

        # maps device -> host (PCIe), reported beside value, never as value
        import numpy as np

        t2 = time.perf_counter()
        maps = gh.maps()
        ms_d2h = (time.perf_counter() - t2) * 1e3
        binned = float(np.asarray(maps, dtype=np.float64).sum())
        del maps
        # samples the map kernel bins (an FP64 pass over every sample, after the timed region)
        _, per_channel = gh.channel_totals()
        n_binned = int(per_channel.sum())
        gh.close()

        keys = ["ms_fill_fft_z", "ms_fft_y", "ms_fft_x", "ms_sigma2", "ms_maps_clear", "ms_pixelize", "ms_total"]
        mean = {k: sum(s[k] for s in stats) / len(stats) for k in keys}
        by = stage_bytes(n, prec)
        launch = {"fill_fft_z": mean["ms_fill_fft_z"], "fft_y": mean["ms_fft_y"], "fft_x": mean["ms_fft_x"],
                  "pixelize": mean["ms_pixelize"]}
        dom = max(launch, key=launch.get)
        achieved = by[dom] / (launch[dom] * 1e-3) / 1e9
        stages = {k: {"ms": launch[k], "GB/s": by[k] / (launch[k] * 1e-3) / 1e9,
                      "frac_hbm": by[k] / (launch[k] * 1e-3) / 1e9 / HBM_PEAK_GBS} for k in launch}

        cpu = None
        if args.cpu_n > 0:
            print(f"bench: CPU baseline (oracle) at n_grid={args.cpu_n}", file=sys.stderr, flush=True)
            cpu = cpu_baseline(args.workload, args.cpu_n, tmp)

        cells = float(n) ** 3
        ms_pix = launch["pixelize"]
        atomics = profile_values(args.workload, "global_atomics").get("k_pixelize")
        valu = profile_values(args.workload, "valu_lane_instr").get("k_pixelize")
        vcyc = (profile_values(args.workload, "valu_cycles_lo").get("k_pixelize"),
                profile_values(args.workload, "valu_cycles_hi").get("k_pixelize"))
        pixelize = {"ms_per_launch": ms_pix, "samples": cells * d.n_sub, "binned_samples": n_binned,
                    "samples_per_s": cells * d.n_sub / (ms_pix * 1e-3), "binned_samples_per_s": n_binned / (ms_pix * 1e-3),
                    "global_atomics_per_launch": atomics,
                    "global_atomics_per_s": atomics / (ms_pix * 1e-3) if atomics else None,
                    "counters": f"profiles/*_{args.workload}.md"}
        if dom == "pixelize" and valu:
            roof = {"bound": "valu", "kernel": "k_pixelize", "achieved": valu / (ms_pix * 1e-3),
                    "peak": VALU_PEAK_LANE_INSTR, "unit": "VALU lane-instr/s",
                    "frac": valu / (ms_pix * 1e-3) / VALU_PEAK_LANE_INSTR,
                    "traffic": measured_traffic(args.workload).get("k_pixelize"),
                    "valu_lane_instr_per_launch": valu, "ms_per_launch": ms_pix,
                    # VALU issue cycles per launch from the committed instruction-mix profile and the measured
                    # per-instruction costs, as a share of the 1024 SIMDs' cycles (lower, upper bound)
                    "valu_busy": ([v / (1024 * 2.4e9 * ms_pix * 1e-3) for v in vcyc] if all(vcyc) else None),
                    "hbm": {"bytes_per_launch": by["pixelize"], "achieved_GB/s": achieved if dom == "pixelize" else None,
                            "frac": (achieved / HBM_PEAK_GBS) if dom == "pixelize" else None},
                    "note": "VALU lane-instructions per launch from the committed SQ_INSTS_VALU profile x 64, "
                            "divided by the live hipEvent duration; HBM bytes are the algorithmic read of the fields"}
        else:
            roof = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.workload).get(KERNEL_OF_STAGE[dom]),
                    "bytes_per_launch": by[dom], "ms_per_launch": launch[dom]}
        extras = {}
        if not args.quick and args.workload == "production_1024":
            print("bench: gethi CLI end to end", file=sys.stderr, flush=True)
            with tempfile.TemporaryDirectory() as ctmp:
                cli = cli_end_to_end(args.workload, ctmp)
            if "seconds" in cli:
                cli["cells_per_s"] = cells / cli["seconds"]
            extras["end_to_end_cli"] = cli
            print("bench: 96-member ensemble", file=sys.stderr, flush=True)
            ens = ensemble_rate()
            extras["mocks_per_hour_batched"] = ens["value"]
            extras["ensemble"] = ens
            print("bench: emulator (full size)", file=sys.stderr, flush=True)
            emu = emulator_run(log=lambda m: print(f"bench: {m}", file=sys.stderr, flush=True))
            extras["emulator_err_sigma"] = emu["emulator"]["rms_debiased"]
            extras["emulator"] = emu
        out = {
            "metric": METRIC[args.workload],
            "value": cells / (ms * 1e-3),
            "unit": "cells/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": ms,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "fp64" if prec else "fp32",
            "data": "synthetic (Philox Gaussian modes of a BBKS P(k); resident in HBM)",
            "config": {"workload": args.workload, "n_grid": n, "n_side": int(d.n_side), "n_nu": d.n_nu,
                       "samples_per_cell": d.n_sub, "l_box": d.l_box, "seed": d.seed},
            "roofline": roof,
            "cpu_baseline": cpu,
            "stages": stages,
            "pixelize": pixelize,
            "autotune": tune,
            "ms_maps_d2h": ms_d2h,
            "cells_per_s_with_d2h": cells / ((ms + ms_d2h) * 1e-3),
            "maps_sum_mK": binned,
        }
        out.update(extras)
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
