#!/usr/bin/env python3
"""Benchmark of the GetHI density-to-sky path on one MI355X.

A step is one full pass of the hot path (mode fill, 3-D inverse FFT of delta and phi, sigma^2, radial
velocity, lognormal HI mass, RSD, sub-cell ang2pix and the (nu, pixel) scatter into zeroed mK maps)
over the 1024^3 production workload, with every input already resident in HBM.  value = cells/s.

Prints one JSON line.  Extra objects: "roofline" for the dominant kernel (algorithmic bytes per
launch / its hipEvent-timed duration against 8 TB/s), "cpu_baseline" (the FP64 C oracle, OpenMP on
the host cores, over a bounded sample of the same workload), "stages" (ms per stage).
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
WORKLOADS = ("production_1024", "small_fp64_256", "largest_2048")


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


def measured_traffic(workload):
    """HBM bytes per launch from the committed PMC summary (profiles/hbm_traffic_<workload>.md)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    out = {}
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes":
                out[f[1]] = float(f[2])
    return out


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="production_1024", choices=WORKLOADS)
    ap.add_argument("--cpu-n", type=int, default=768, help="grid of the CPU-baseline sample (0: skip)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (scale reference for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the GetHI path on one GPU; multi-GPU slabs are a design note "
              "only (no collectives are built)", file=sys.stderr)
        sys.exit(2)

    import crime_amd
    import crime_amd.configs as cf

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
        out = {
            "metric": "cells/s end-to-end at 1024^3",
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
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(args.workload).get(KERNEL_OF_STAGE[dom]),
                         "bytes_per_launch": by[dom], "ms_per_launch": launch[dom]},
            "cpu_baseline": cpu,
            "stages": stages,
            "ms_maps_d2h": ms_d2h,
            "cells_per_s_with_d2h": cells / ((ms + ms_d2h) * 1e-3),
            "maps_sum_mK": binned,
        }
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
