"""Writes the small golden fixtures of tests/golden from the FP64 oracle (test infrastructure only).

gethi_16_fp64.npy / gethi_16_fp64_sigma2.npy: maps (n_nu, Npix) and sigma^2 of a 16^3 FP64 run
(Nside 8, 8 channels 1330-1410 MHz, L = 300 Mpc/h, seed 4242, reference bias, RSD, 10 shared random
sub-cell samples) -- the parameter file is rebuilt by fixture_params() below.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def fixture_params(outdir):
    sys.path.insert(0, ROOT)
    import crime_amd
    import crime_amd.configs as cf

    ini = cf.make(outdir, "golden16", n_grid=16, n_side=8, n_nu=8, nu_lo=1330.0, nu_hi=1410.0, l_box=300.0,
                  seed=4242, precision="double", bias_model="reference", n_subcell=0)
    return crime_amd.Params(ini)


def main():
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    with tempfile.TemporaryDirectory() as tmp:
        par = fixture_params(tmp)
        maps, s2, _ = oracle.run(par.desc())
    np.save(os.path.join(HERE, "gethi_16_fp64.npy"), maps)
    np.save(os.path.join(HERE, "gethi_16_fp64_sigma2.npy"), np.array([s2]))
    print(f"wrote maps {maps.shape}, sigma2 {s2:.17g}, sum {maps.sum():.17g}")


if __name__ == "__main__":
    main()
