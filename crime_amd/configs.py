"""Parameter files of the standard workloads, written in the reference's .ini grammar.

Each function writes ``<name>.ini``, its frequency-edge table and its BBKS P(k) table into ``outdir``
and returns the .ini path.  The P(k) is A k^ns T_BBKS(k)^2 on 2000 log-uniform k in [1e-5, 1e2] h/Mpc
(past every grid's k_max, so pk_linear0 never leaves the table), rescaled by the pipeline to sigma_8.
Box sides are given with the ``l_box=`` extension key; the reference would derive
L = 2 r_max (1 + 2/N) instead (src/cosmo.c:361).
"""
import os

import numpy as np

from . import write_bbks_pk, write_ini


def _edges(path, nu_lo, nu_hi, n_nu):
    e = np.linspace(nu_lo, nu_hi, n_nu + 1)
    np.savetxt(path, e, fmt="%.10f")
    return path


def make(outdir, name, *, n_grid, n_side, n_nu, nu_lo, nu_hi, l_box, seed, omega_m=0.3, omega_b=0.05,
         h=0.7, sigma_8=0.8, n_s=0.96, precision="single", bias_model="reference", bias_b0=1.0,
         bias_b1=0.0, n_subcell=0, rsd=1, modes="philox", r_smooth=-1.0, uniform_nu=False,
         velocity_mode="stencil", extra=None):
    """uniform_nu: the reference's non-_IRREGULAR_NUTABLE binning (nu_min/nu_max/n_nu keys) instead
    of a table of edges.  velocity_mode: "stencil" (the reference's central difference of phi) or
    "spectral" (three velocity-component FFTs)."""
    os.makedirs(outdir, exist_ok=True)
    pk = os.path.join(outdir, f"{name}_pk.dat")
    write_bbks_pk(pk, omega_m, omega_b, h, n_s)
    keys = dict(prefix_out=os.path.join(outdir, name), pk_filename=pk, omega_M=omega_m,
                omega_L=1 - omega_m, omega_B=omega_b, h=h, w=-1.0, ns=n_s, sigma_8=sigma_8,
                r_smooth=r_smooth)
    if uniform_nu:
        keys.update(nu_min=nu_lo, nu_max=nu_hi, n_nu=n_nu)
    else:
        nut = os.path.join(outdir, f"{name}_nu.txt")
        _edges(nut, nu_lo, nu_hi, n_nu)
        keys.update(frequencies_filename=nut)
    keys.update(n_side=n_side, n_grid=n_grid, seed=seed, do_psources=0, l_box=l_box,
                bias_model=bias_model, bias_b0=bias_b0, bias_b1=bias_b1, n_subcell=n_subcell,
                precision=precision, rsd=rsd, modes=modes)
    if velocity_mode != "stencil":
        keys.update(velocity_mode=velocity_mode)
    if extra:
        keys.update(extra)
    return write_ini(os.path.join(outdir, f"{name}.ini"), **keys)


def small_fp64_256(outdir, **kw):
    """FP64, 256^3, L=4900 Mpc/h, bias 1, RSD, 64 channels 700-800 MHz, Nside 128, 2^3 sub-cells,
    seed 1001 (the reference's CPU-runnable size)."""
    a = dict(n_grid=256, n_side=128, n_nu=64, nu_lo=700.0, nu_hi=800.0, l_box=4900.0, seed=1001,
             precision="double", bias_model="const", bias_b0=1.0, n_subcell=2)
    name = kw.pop("name", "small_fp64_256")
    a.update(kw)
    return make(outdir, name, **a)


def production_1024(outdir, **kw):
    """FP32 grids, 1024^3, L=8300 Mpc/h, b(z)=0.7+0.3(1+z), RSD, 200 channels 400-800 MHz,
    Nside 512, the reference's 10 shared random sub-cell samples, seed 2002."""
    a = dict(n_grid=1024, n_side=512, n_nu=200, nu_lo=400.0, nu_hi=800.0, l_box=8300.0, seed=2002,
             precision="single", bias_model="linear", bias_b0=0.7, bias_b1=0.3, n_subcell=0)
    name = kw.pop("name", "production_1024")
    a.update(kw)
    return make(outdir, name, **a)


def largest_2048(outdir, **kw):
    """FP32 grids, 2048^3, L=8300 Mpc/h, 400 channels 400-800 MHz, Nside 1024, 4^3 sub-cells,
    three velocity components (spectral velocities), sigma_8=0.83, n_s=0.965, seed 3003."""
    a = dict(n_grid=2048, n_side=1024, n_nu=400, nu_lo=400.0, nu_hi=800.0, l_box=8300.0, seed=3003,
             sigma_8=0.83, n_s=0.965, precision="single", n_subcell=4, velocity_mode="spectral")
    name = kw.pop("name", "largest_2048")
    a.update(kw)
    return make(outdir, name, **a)


def ensemble_member(outdir, i, **kw):
    """Member i of the uneven 96-member ensemble: 32 x 128^3 (Nside 128), 32 x 192^3 (Nside 256),
    32 x 256^3 (Nside 256); 64 channels 700-800 MHz; seed 4000+i; Omega_m ~ U[0.25,0.35] and
    sigma_8 ~ U[0.7,0.9] from numpy's PCG64 seeded with 4000+i."""
    kind = i // 32
    n_grid, n_side = [(128, 128), (192, 256), (256, 256)][kind]
    rng = np.random.default_rng(4000 + i)
    om = float(rng.uniform(0.25, 0.35))
    s8 = float(rng.uniform(0.7, 0.9))
    # l_box <= 0: the reference's own box, L = 2 r_max (1 + 2/N) for the member's cosmology
    a = dict(n_grid=n_grid, n_side=n_side, n_nu=64, nu_lo=700.0, nu_hi=800.0, l_box=-1.0,
             seed=4000 + i, omega_m=om, sigma_8=s8, precision="single")
    a.update(kw)
    return make(outdir, f"member_{i:03d}", **a)
