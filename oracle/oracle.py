"""ORACLE -- test infrastructure only: ctypes wrapper of oracle/_build/liboracle.so, the FP64 C
restatement of the reference GetHI path.  Imported by tests/, __graft_entry__.smoke() and
bench.py's CPU baseline; never by crime_amd."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

_lib = None
_dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        from crime_amd._ffi import CrimeGhDesc
        D = C.POINTER(CrimeGhDesc)
        L.orc_fill_modes.argtypes = [D, _dp, _dp]
        L.orc_ifft3d.argtypes = [C.c_int, _dp]
        L.orc_ifft3d.restype = C.c_int
        L.orc_c2r_3d.argtypes = [C.c_int, _dp, _dp]
        L.orc_c2r_3d.restype = C.c_int
        L.orc_sigma2.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_fields.argtypes = [D, _dp, _dp, _dp, _dp]
        L.orc_fields.restype = C.c_int
        L.orc_radial_velocity.argtypes = [D, _dp, _dp]
        L.orc_velocity_spectral.argtypes = [D, _dp, _dp, _dp]
        L.orc_velocity_spectral.restype = C.c_int
        L.orc_radial_velocity_spectral.argtypes = [D, _dp, _dp, _dp, _dp]
        L.orc_get_hi.argtypes = [D, _dp, _dp, C.c_double, _dp, _dp]
        L.orc_vec2pix_ring.argtypes = [C.c_int64, C.c_double, C.c_double, C.c_double]
        L.orc_vec2pix_ring.restype = C.c_int64
        L.orc_get_inu.argtypes = [D, C.c_double, C.c_int]
        L.orc_get_inu.restype = C.c_int
        L.orc_channel_factors.argtypes = [D, _dp]
        L.orc_sample_indices.argtypes = [D, _dp, C.c_int64, C.c_int64, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int64)]
        L.orc_mk_T_maps.argtypes = [D, _dp, _dp, _dp]
        L.orc_mk_T_maps.restype = C.c_int
        L.orc_run.argtypes = [D, _dp, _dp, _dp]
        L.orc_run.restype = C.c_int
        L.orc_pk_linear0.argtypes = [D, C.c_double]
        L.orc_pk_linear0.restype = C.c_double
        for f in ("orc_z_of_r", "orc_dgrowth_of_r", "orc_vgrowth_of_r", "orc_r_of_z", "orc_bias"):
            getattr(L, f).argtypes = [D, C.c_double]
            getattr(L, f).restype = C.c_double
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]
        L.orc_num_threads.restype = C.c_int
        I64 = C.c_int64
        L.orc_ring_info.argtypes = [I64, I64, C.POINTER(I64), C.POINTER(I64), _dp, _dp, C.POINTER(C.c_int)]
        L.orc_legendre.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_alm_index.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_alm_index.restype = I64
        L.orc_map2alm.argtypes = [I64, C.c_int, C.c_int, _dp, _dp]
        L.orc_map2alm.restype = C.c_int
        L.orc_alm2cl.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_ps_lfunc.argtypes = [C.c_double, C.c_double]
        L.orc_ps_lfunc.restype = C.c_double
        L.orc_ps_sed.argtypes = [C.c_double]
        L.orc_ps_sed.restype = C.c_double
        L.orc_ps_temp.argtypes = [C.c_double] * 6
        L.orc_ps_temp.restype = C.c_double
        L.orc_ps_tables.argtypes = [C.c_double, _dp, _dp]
        L.orc_ps_poisson.argtypes = [C.c_double, C.c_uint32, C.c_uint64]
        L.orc_ps_poisson.restype = C.c_uint32
        L.orc_ps_luminosity.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_uint32, C.c_uint64, C.c_uint32]
        L.orc_ps_luminosity.restype = C.c_double
        L.orc_psources.argtypes = [D, _dp, _dp, C.c_double, _dp, _dp, _dp]
        L.orc_psources.restype = C.c_int64
        L.orc_alm2map.argtypes = [I64, C.c_int, C.c_int, _dp, _dp]
        L.orc_alm2map.restype = C.c_int
        L.orc_beam_window.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_add_noise.argtypes = [I64, C.c_int, _dp, C.c_uint32, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def fill_modes(desc):
    n = desc.n_grid
    shape = (n, n, n // 2 + 1, 2)
    dk = np.empty(shape)
    pk = np.empty(shape)
    lib().orc_fill_modes(C.byref(desc), _p(dk), _p(pk))
    return dk, pk


def c2r_3d(half):
    """half: (N, N, N/2+1) complex [kz, ky, kx] -> (N, N, N) real [z, y, x]."""
    n = half.shape[0]
    h = np.ascontiguousarray(half, dtype=np.complex128)
    out = np.empty((n, n, n))
    rc = lib().orc_c2r_3d(n, h.ctypes.data_as(_dp), _p(out))
    assert rc == 0
    return out


def ifft3d(full):
    a = np.ascontiguousarray(full, dtype=np.complex128).copy()
    rc = lib().orc_ifft3d(a.shape[0], a.ctypes.data_as(_dp))
    assert rc == 0
    return a


def fields(desc):
    n = desc.n_grid
    dens = np.empty((n, n, n))
    vpot = np.empty((n, n, n))
    s2, m = C.c_double(), C.c_double()
    rc = lib().orc_fields(C.byref(desc), _p(dens), _p(vpot), C.byref(s2), C.byref(m))
    assert rc == 0
    return dens, vpot, s2.value, m.value


def sigma2(dens):
    s2, m = C.c_double(), C.c_double()
    d = np.ascontiguousarray(dens, dtype=np.float64)
    lib().orc_sigma2(d.shape[0], _p(d), C.byref(s2), C.byref(m))
    return s2.value, m.value


def radial_velocity(desc, vpot):
    out = np.empty_like(vpot, dtype=np.float64)
    lib().orc_radial_velocity(C.byref(desc), _p(np.ascontiguousarray(vpot, np.float64)), _p(out))
    return out


def velocity_spectral(desc):
    """(vx, vy, vz) from v_j(k) = i sin(k_j dx)/dx phi_k of the descriptor's modes."""
    n = desc.n_grid
    v = [np.empty((n, n, n)) for _ in range(3)]
    rc = lib().orc_velocity_spectral(C.byref(desc), *[_p(a) for a in v])
    assert rc == 0
    return v


def radial_velocity_spectral(desc, vx, vy, vz):
    out = np.empty_like(vx, dtype=np.float64)
    lib().orc_radial_velocity_spectral(C.byref(desc), *[_p(np.ascontiguousarray(a, np.float64)) for a in (vx, vy, vz)],
                                       _p(out))
    return out


def get_hi(desc, dens, rvel, sigma2):
    mass = np.empty(dens.shape)
    dz = np.empty(dens.shape)
    lib().orc_get_hi(C.byref(desc), _p(np.ascontiguousarray(dens, np.float64)),
                     _p(np.ascontiguousarray(rvel, np.float64)), sigma2, _p(mass), _p(dz))
    return mass, dz


def vec2pix_ring(nside, x, y, z):
    return lib().orc_vec2pix_ring(nside, x, y, z)


def get_inu(desc, nu, start=0):
    return lib().orc_get_inu(C.byref(desc), nu, start)


def channel_factors(desc):
    out = np.empty(desc.n_nu)
    lib().orc_channel_factors(C.byref(desc), _p(out))
    return out


def sample_indices(desc, dzrsd, cell0, ncell):
    inu = np.empty((ncell, desc.n_sub), np.int32)
    ipix = np.empty((ncell, desc.n_sub), np.int64)
    lib().orc_sample_indices(C.byref(desc), _p(np.ascontiguousarray(dzrsd, np.float64)), cell0, ncell,
                             inu.ctypes.data_as(C.POINTER(C.c_int32)), ipix.ctypes.data_as(C.POINTER(C.c_int64)))
    return inu, ipix


def sample_indices_slab(desc, dz_slab, cell0, ncell):
    """sample_indices with the RSD shifts of cells [cell0, cell0 + ncell) only (dz_slab[c] is cell
    cell0 + c): the C code reads dzrsd[cell0 .. cell0 + ncell) through a pointer moved back by cell0."""
    dz = np.ascontiguousarray(dz_slab, np.float64)
    assert dz.size == ncell
    inu = np.empty((ncell, desc.n_sub), np.int32)
    ipix = np.empty((ncell, desc.n_sub), np.int64)
    base = C.cast(dz.ctypes.data - 8 * int(cell0), _dp)
    lib().orc_sample_indices(C.byref(desc), base, cell0, ncell, inu.ctypes.data_as(C.POINTER(C.c_int32)),
                             ipix.ctypes.data_as(C.POINTER(C.c_int64)))
    return inu, ipix


def mk_T_maps(desc, mass, dz):
    maps = np.empty((desc.n_nu, 12 * desc.n_side ** 2))
    rc = lib().orc_mk_T_maps(C.byref(desc), _p(np.ascontiguousarray(mass, np.float64)),
                             _p(np.ascontiguousarray(dz, np.float64)), _p(maps))
    assert rc == 0
    return maps


def run(desc):
    maps = np.empty((desc.n_nu, 12 * desc.n_side ** 2))
    s2, m = C.c_double(), C.c_double()
    rc = lib().orc_run(C.byref(desc), _p(maps), C.byref(s2), C.byref(m))
    assert rc == 0
    return maps, s2.value, m.value


def philox(ctr, k0, k1):
    c = (C.c_uint32 * 4)(*ctr)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k0, k1, o)
    return list(o)


def pk_linear0(desc, lgk):
    return lib().orc_pk_linear0(C.byref(desc), lgk)


def z_of_r(desc, r):
    return lib().orc_z_of_r(C.byref(desc), r)


def num_threads():
    return lib().orc_num_threads()


# ---------------------------------------------------------------------------------------------
# HEALPix map2alm / C_l (sht_oracle.c)
# ---------------------------------------------------------------------------------------------
def ring_info(nside, ir):
    """(nphi, first pixel, z, sin theta, s) of RING ring ir (1..4 nside-1); phi_j = pi (2j+s)/nphi."""
    nphi, first, s = C.c_int64(), C.c_int64(), C.c_int()
    z, st = C.c_double(), C.c_double()
    lib().orc_ring_info(nside, ir, C.byref(nphi), C.byref(first), C.byref(z), C.byref(st), C.byref(s))
    return nphi.value, first.value, z.value, st.value, s.value


def legendre(lmax, m, x, st):
    """lambda_lm(x), l = m..lmax (Y_lm = lambda_lm(cos theta) e^{i m phi}, Condon-Shortley phase)."""
    out = np.empty(lmax - m + 1)
    lib().orc_legendre(lmax, m, x, st, _p(out))
    return out


def alm_index(l, m, lmax):
    return lib().orc_alm_index(l, m, lmax)


def map2alm(maps, nside, lmax):
    """maps (nmaps, 12 nside^2) RING -> alms (nmaps, (lmax+1)(lmax+2)/2) complex, m-major."""
    m = np.ascontiguousarray(np.atleast_2d(maps), dtype=np.float64)
    nalm = (lmax + 1) * (lmax + 2) // 2
    out = np.empty((m.shape[0], nalm), np.complex128)
    rc = lib().orc_map2alm(nside, lmax, m.shape[0], _p(m), out.ctypes.data_as(_dp))
    assert rc == 0
    return out


def alm2cl(alms, lmax):
    """cls[a, b, l] for every pair of the alms' maps."""
    a = np.ascontiguousarray(np.atleast_2d(alms), dtype=np.complex128)
    out = np.empty((a.shape[0], a.shape[0], lmax + 1))
    lib().orc_alm2cl(lmax, a.shape[0], a.ctypes.data_as(_dp), _p(out))
    return out


def alm2map(alms, nside, lmax):
    """alms (nmaps, nalm) complex, m-major -> maps (nmaps, 12 nside^2) RING, by the defining sum."""
    a = np.ascontiguousarray(np.atleast_2d(alms), dtype=np.complex128)
    out = np.empty((a.shape[0], 12 * nside * nside))
    rc = lib().orc_alm2map(nside, lmax, a.shape[0], a.ctypes.data_as(_dp), _p(out))
    assert rc == 0
    return out


def beam_window(lmax, fwhm_arcmin):
    """Gaussian beam b_l of he_generate_beam_window, l = 0..lmax."""
    out = np.empty(lmax + 1)
    lib().orc_beam_window(lmax, fwhm_arcmin, _p(out))
    return out


def add_noise(maps, sigma2, seed):
    """JoinT's per-pixel noise added to a copy of maps (nmaps, npix): sqrt(-sigma2 ln(1-u2)) cos(2 pi u1)."""
    m = np.array(np.atleast_2d(maps), dtype=np.float64, order="C")
    s2 = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma2, np.float64), (m.shape[0],)))
    lib().orc_add_noise(m.shape[1], m.shape[0], _p(s2), seed, _p(m))
    return m


PS_NZ = 256


def ps_tables(z_max):
    """setup_psources: n(z_i) and max_Lpdf at z_i = i z_max / 256."""
    nz, mp = np.empty(PS_NZ), np.empty(PS_NZ)
    lib().orc_ps_tables(z_max, _p(nz), _p(mp))
    return nz, mp


def ps_poisson(lam, seed, cells):
    return np.array([lib().orc_ps_poisson(float(lam), seed, int(c)) for c in cells], dtype=np.int64)


def ps_luminosity(z, z_max, seed, cells, src=0):
    nz, mp = ps_tables(z_max)
    return np.array([lib().orc_ps_luminosity(_p(nz), _p(mp), z_max, float(z), seed, int(c), src) for c in cells])


def psources(desc, dens, dzrsd, sigma2, with_maps=True):
    """Point sources of the grids: (maps (n_nu, Npix) or None, A, B, number of sources)."""
    npix = 12 * desc.n_side ** 2
    maps = np.zeros((desc.n_nu, npix)) if with_maps else None
    a, b = np.zeros(npix), np.zeros(npix)
    n = lib().orc_psources(C.byref(desc), _p(np.ascontiguousarray(dens, np.float64)),
                           _p(np.ascontiguousarray(dzrsd, np.float64)), sigma2,
                           _p(maps) if with_maps else None, _p(a), _p(b))
    return maps, a, b, n
