"""MI355X-native GetHI density-to-sky path (CRIME's GetHI hot path on gfx950).

Host-side mirror of the reference's operator sequence (src/main_gh.c:44-68)::

    par = crime_amd.Params("param_GetHI.ini")      # read_run_params (src/io_gh.c:188-296)
    gh = crime_amd.GetHI(par)                      # init_fftw + allocate_maps, on the device
    gh.create_d_and_vr_fields()                    # src/fourier.c:375-438
    gh.mk_T_maps()                                 # get_HI + mk_T_maps (src/grid_tools.c, src/pixelize.c)
    par.write_maps(gh.maps())                      # write_maps (src/io_gh.c:109-131)

Everything below the Python layer is C (host) and HIP (device) behind the C ABIs in ``include/``;
this module only marshals arguments.  Errors surface as ``CrimeError`` with the library's message.
"""
import ctypes as C
import os

import numpy as np

from . import _ffi
from ._ffi import (BIAS_CONST, BIAS_LINEAR, BIAS_REFERENCE, FP32, FP64, GRID_DENS, GRID_DZRSD, GRID_MASS,
                   GRID_RVEL, GRID_VPOT, GRID_VX, GRID_VY, GRID_VZ, MODES_INJECTED, MODES_PHILOX, NU_TABLE,
                   NU_UNIFORM, VEL_SPECTRAL, VEL_STENCIL, CrimeGhDesc)

__all__ = ["Params", "GetHI", "Sht", "Batch", "run_batch", "alm_index", "jt_instrument", "CrimeError", "write_ini",
           "write_bbks_pk", "FP32", "FP64",
           "GRID_DENS", "GRID_VPOT", "GRID_RVEL", "GRID_MASS", "GRID_DZRSD", "GRID_VX", "GRID_VY", "GRID_VZ",
           "VEL_STENCIL", "VEL_SPECTRAL", "MODES_PHILOX",
           "MODES_INJECTED", "NU_TABLE", "NU_UNIFORM", "BIAS_REFERENCE", "BIAS_CONST", "BIAS_LINEAR"]


class CrimeError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"{_ffi.err_name(rc)} ({rc}): {msg}")
        self.rc = rc


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def write_ini(path, **keys):
    """Write a parameter file in the reference grammar ("key= value", src/io_gh.c:202-209)."""
    with open(path, "w") as f:
        f.write("# GetHI parameter file\n")
        for k, v in keys.items():
            f.write(f"{k}= {v}\n")
    return path


def write_bbks_pk(path, omega_m, omega_b, h, n_s, kmin=1e-5, kmax=1e2, n=2000):
    """Tabulate A k^ns T_BBKS(k)^2 (Sugiyama shape parameter) in CAMB's two-column format."""
    rc = _ffi.host_lib().gh_write_bbks_pk(os.fsencode(path), omega_m, omega_b, h, n_s, kmin, kmax, n)
    if rc:
        raise OSError(f"could not write {path}")
    return path


class Params:
    """A GetHI run description parsed and set up by the host C library (libgethi_host.so)."""

    def __init__(self, ini_path):
        self._lib = _ffi.host_lib()
        self._p = self._lib.gh_params_new()
        if not self._p:
            raise MemoryError("gh_params_new")
        if self._lib.gh_params_load(self._p, os.fsencode(ini_path)):
            msg = self._lib.gh_params_error(self._p).decode()
            self.close()
            raise ValueError(msg)

    def get(self, name):
        return self._lib.gh_params_get(self._p, name.encode())

    def desc(self):
        """crime_gh_desc pointing into this object's tables (valid while it lives)."""
        d = CrimeGhDesc()
        self._lib.gh_params_to_desc(self._p, C.byref(d))
        return d

    def write_maps(self, maps):
        m = np.ascontiguousarray(maps, dtype=np.float32)
        if self._lib.gh_write_maps(self._p, m.ctypes.data_as(C.POINTER(C.c_float))):
            raise OSError("writing the FITS maps failed")

    def close(self):
        if getattr(self, "_p", None):
            self._lib.gh_params_delete(self._p)
            self._p = None

    def __del__(self):
        self.close()


class GetHI:
    """One device handle of libcrime_gh.so (one HIP stream + resident fields and maps)."""

    def __init__(self, params, **overrides):
        self._lib = _ffi.gpu_lib()
        d = params.desc() if isinstance(params, Params) else params
        keep = []
        for k, v in overrides.items():
            if isinstance(v, np.ndarray):
                arr = np.ascontiguousarray(v, dtype=np.float64)
                keep.append(arr)
                setattr(d, k, _dptr(arr))
            else:
                setattr(d, k, v)
        self.desc = d
        self.n = d.n_grid
        self.n_nu = d.n_nu
        self.n_side = d.n_side
        self.n_sub = d.n_sub
        self.npix = 12 * d.n_side * d.n_side
        self.precision = d.precision
        h = C.c_void_p()
        rc = self._lib.crime_gh_create(C.byref(d), C.byref(h))
        if rc:
            raise CrimeError(rc, self._lib.crime_gh_last_error(None).decode())
        self._h = h
        self._params = params  # keeps the tables alive for the handle's life

    def _check(self, rc):
        if rc:
            raise CrimeError(rc, self._lib.crime_gh_last_error(self._h).decode())

    def _stage(self, fn):
        st = _ffi.CrimeGhStats()
        self._check(fn(self._h, C.byref(st)))
        return st.as_dict()

    def create_d_and_vr_fields(self):
        return self._stage(self._lib.crime_gh_create_d_and_vr_fields)

    def mk_T_maps(self):
        return self._stage(self._lib.crime_gh_mk_T_maps)

    def run(self):
        return self._stage(self._lib.crime_gh_run)

    def maps(self, dtype=np.float32):
        if np.dtype(dtype) == np.float64:
            out = np.empty((self.n_nu, self.npix), np.float64)
            self._check(self._lib.crime_gh_copy_maps_f64(self._h, _dptr(out)))
        else:
            out = np.empty((self.n_nu, self.npix), np.float32)
            self._check(self._lib.crime_gh_copy_maps(self._h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def grid(self, which):
        out = np.empty((self.n, self.n, self.n), np.float64)
        self._check(self._lib.crime_gh_copy_grid(self._h, which, _dptr(out)))
        return out

    def set_seed(self, seed):
        """A new realisation on this handle: the Philox key of the next field stage (crime_gh_set_seed)."""
        self._check(self._lib.crime_gh_set_seed(self._h, int(seed) & 0xFFFFFFFF))

    def grid_range(self, which, cell0, ncell):
        """Cells [cell0, cell0 + ncell) of a grid (x fastest), as FP64 (crime_gh_copy_grid_range)."""
        out = np.empty(ncell, np.float64)
        self._check(self._lib.crime_gh_copy_grid_range(self._h, which, int(cell0), int(ncell), _dptr(out)))
        return out

    def channel_totals(self, cell0=0, ncell=None):
        """FP64 per-channel sums of mass/n_sub x mK factor over every sample of cells [cell0, cell0+ncell),
        and the number of binned samples per channel (crime_gh_channel_totals)."""
        ncell = self.n ** 3 - cell0 if ncell is None else ncell
        tot = np.empty(self.n_nu)
        cnt = np.empty(self.n_nu, np.int64)
        self._check(self._lib.crime_gh_channel_totals(self._h, int(cell0), int(ncell), _dptr(tot),
                                                      cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return tot, cnt

    def screen_check(self):
        """Every sample through the FP32 screen and the FP64 path (crime_gh_screen_check): dict of counts
        (samples, channel_fallbacks, pixel_fallbacks, mismatches, binned) and the two margins."""
        c = np.zeros(5, np.int64)
        eps = np.zeros(2)
        self._check(self._lib.crime_gh_screen_check(self._h, c.ctypes.data_as(C.POINTER(C.c_int64)), _dptr(eps)))
        keys = ("samples", "channel_fallbacks", "pixel_fallbacks", "mismatches", "binned")
        out = {k: int(v) for k, v in zip(keys, c)}
        out.update(eps_nu_mhz=float(eps[0]), eps_pix=float(eps[1]))
        return out

    def reference_maps(self):
        """FP64 maps scattered sample by sample without the cell screen or the LDS aggregation
        (crime_gh_reference_maps): the parity target of the timed map kernel."""
        out = np.empty((self.n_nu, self.npix), np.float64)
        self._check(self._lib.crime_gh_reference_maps(self._h, _dptr(out)))
        return out

    def set_grid(self, which, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.size == self.n ** 3
        self._check(self._lib.crime_gh_set_grid(self._h, which, _dptr(a)))

    def sigma2(self):
        s2, m = C.c_double(), C.c_double()
        self._check(self._lib.crime_gh_get_sigma2(self._h, C.byref(s2), C.byref(m)))
        return s2.value, m.value

    def set_sigma2(self, sigma2, mean=0.0):
        self._check(self._lib.crime_gh_set_sigma2(self._h, sigma2, mean))

    def sample_indices(self, cell0, ncell):
        inu = np.empty((ncell, self.n_sub), np.int32)
        ipix = np.empty((ncell, self.n_sub), np.int64)
        self._check(self._lib.crime_gh_sample_indices(
            self._h, cell0, ncell, inu.ctypes.data_as(C.POINTER(C.c_int32)),
            ipix.ctypes.data_as(C.POINTER(C.c_int64))))
        return inu, ipix

    STAGES = ("fill_fft_z", "fft_y", "fft_x", "pixelize")

    def schedule(self):
        """Kernel variants [fill+kz, ky, kx, pixelize] (crime_gh_get_schedule)."""
        s = (C.c_int * 4)()
        self._check(self._lib.crime_gh_get_schedule(self._h, s))
        return list(s)

    def set_schedule(self, sched):
        s = (C.c_int * 4)(*[int(v) for v in sched])
        self._check(self._lib.crime_gh_set_schedule(self._h, s))

    def autotune(self, reps=3):
        """Time every variant of every stage on this handle and keep the fastest (crime_gh_autotune).
        Returns (schedule, {stage: [ms of variant 0, 1, 2, 3]}); -1 marks a variant the grid or stage
        lacks (variant 3, radix-8 plans, exists for the FFT stages only)."""
        s = (C.c_int * 4)()
        nv = _ffi.N_SCHED
        ms = (C.c_double * (4 * nv))()
        self._check(self._lib.crime_gh_autotune(self._h, int(reps), s, ms))
        table = {st: [ms[nv * i + v] for v in range(nv)] for i, st in enumerate(self.STAGES)}
        return list(s), table

    def autotune_cached(self, path, reps=3):
        """The schedule of this handle's kind from the cache file `path`, or the autotuner's on a miss
        (appended to the file) (crime_gh_autotune_cached).  Returns (schedule, hit)."""
        s = (C.c_int * 4)()
        hit = C.c_int()
        self._check(self._lib.crime_gh_autotune_cached(self._h, os.fsencode(path), int(reps), s, C.byref(hit)))
        return list(s), bool(hit.value)

    def power_spectrum(self, nbins=None, bin_dk=None):
        """Shell-averaged P(k) of the resident Gaussian delta: (k_mean [h/Mpc], P [(Mpc/h)^3], n_modes).
        Default shells are one fundamental mode wide up to the Nyquist k.  Consumes the fields."""
        dk = 2 * np.pi / self.desc.l_box
        bin_dk = dk if bin_dk is None else bin_dk
        nbins = int(np.ceil((self.n // 2) * dk / bin_dk)) if nbins is None else nbins
        k = np.empty(nbins)
        pk = np.empty(nbins)
        cnt = np.empty(nbins, np.int64)
        self._check(self._lib.crime_gh_power_spectrum(self._h, nbins, bin_dk, _dptr(k), _dptr(pk),
                                                      cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return k, pk, cnt

    def device_maps(self):
        p, nb = C.c_void_p(), C.c_size_t()
        self._check(self._lib.crime_gh_device_maps(self._h, C.byref(p), C.byref(nb)))
        return p.value, nb.value

    def mk_psources_maps(self):
        """Point sources (do_psources): Poisson counts per cell, positions, luminosities and the
        point-source maps, on the resident fields.  Returns timings and the number of sources."""
        st = _ffi.CrimeGhPsStats()
        self._check(self._lib.crime_gh_mk_psources_maps(self._h, C.byref(st)))
        return st.as_dict()

    def ps_maps(self):
        """Point-source maps (n_nu, Npix), mK, FP64 copy."""
        out = np.empty((self.n_nu, self.npix), np.float64)
        self._check(self._lib.crime_gh_copy_ps_maps(self._h, _dptr(out)))
        return out

    def ps_pixel_sums(self):
        """The two per-pixel terms (A, B) of the last point-source pass."""
        a, b = np.empty(self.npix), np.empty(self.npix)
        self._check(self._lib.crime_gh_ps_pixel_sums(self._h, _dptr(a), _dptr(b)))
        return a, b

    def device_ps_maps(self):
        p, nb = C.c_void_p(), C.c_size_t()
        self._check(self._lib.crime_gh_device_ps_maps(self._h, C.byref(p), C.byref(nb)))
        return p.value, nb.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.crime_gh_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


class Batch:
    """Every member's field and map stages captured into one hipGraph (crime_gh_batch_*); each
    run() replays it with a single launch.  Members keep their own maps and sigma^2 on the device."""

    def __init__(self, handles, n_streams=4):
        self._lib = _ffi.gpu_lib()
        self.handles = list(handles)
        arr = (C.c_void_p * len(self.handles))(*[h._h for h in self.handles])
        b = C.c_void_p()
        rc = self._lib.crime_gh_batch_create(arr, len(self.handles), n_streams, C.byref(b))
        if rc:
            raise CrimeError(rc, self._lib.crime_gh_batch_last_error(None).decode())
        self._b = b

    def run(self):
        """Replay the graph; returns the wall time of the launch in ms."""
        ms = C.c_double()
        rc = self._lib.crime_gh_batch_run(self._b, C.byref(ms))
        if rc:
            raise CrimeError(rc, self._lib.crime_gh_batch_last_error(self._b).decode())
        return ms.value

    def close(self):
        if getattr(self, "_b", None):
            self._lib.crime_gh_batch_destroy(self._b)
            self._b = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


def run_batch(handles):
    """Run several handles through crime_gh_run_batch; returns one stats dict per handle."""
    lib = _ffi.gpu_lib()
    n = len(handles)
    arr = (C.c_void_p * n)(*[h._h for h in handles])
    st = (_ffi.CrimeGhStats * n)()
    rc = lib.crime_gh_run_batch(arr, n, st)
    if rc:
        raise CrimeError(rc, lib.crime_gh_last_error(None).decode())
    return [s.as_dict() for s in st]


def alm_index(l, m, lmax):
    """he_indexlm (src/healpix_extra.c:36-42): position of a_lm, m >= 0, in a map's alm array."""
    return l + m * lmax - m * (m - 1) // 2 if m > 0 else l


def jt_instrument(nu0, nuf, nside, dish_diameter=0.0, tsys=0.0, time_total=1.0, n_dish=1, fsky=1.0):
    """Per-channel beam FWHM (arcmin) and noise sigma2 (mK^2) of JoinT (src/main_jt.c:27-67, 270-311)
    for channels [nu0, nuf) in MHz: nu = (nu0 + nuf)/2 (src/io_jt.c:231), fwhm = 1030610.55/(nu D)
    (0 when D <= 0), sigma2 = fsky Npix tsys^2 / (3.6e9 dnu t_tot n_dish) (0 when tsys <= 0), tsys in mK,
    t_tot in hours."""
    nu0, nuf = np.asarray(nu0, np.float64), np.asarray(nuf, np.float64)
    nu, dnu = 0.5 * (nu0 + nuf), nuf - nu0
    fwhm = 1030610.55 / (nu * dish_diameter) if dish_diameter > 0 else np.zeros_like(nu)
    npix = 12 * int(nside) ** 2
    sig2 = (fsky * npix * tsys * tsys / (3.6e9 * dnu * time_total * n_dish) if tsys > 0 else np.zeros_like(nu))
    return fwhm, sig2


class Sht:
    """HEALPix plan of libcrime_gh.so (include/crime_sht.h): map2alm and alm2map as he_map2alm and
    he_alm2map (src/healpix_extra.c:44-130) do them, angular power spectra, and JoinT's beam, noise and
    mask stages (src/main_jt.c:211-311), in FP64 on the device."""

    def __init__(self, nside, lmax=None, max_maps=1, device=0):
        self._lib = _ffi.gpu_lib()
        self.nside = int(nside)
        self.lmax = 3 * self.nside - 1 if lmax is None else int(lmax)
        self.max_maps = int(max_maps)
        self.npix = 12 * self.nside ** 2
        self.nalm = (self.lmax + 1) * (self.lmax + 2) // 2
        h = C.c_void_p()
        rc = self._lib.crime_sht_create(self.nside, self.lmax, self.max_maps, device, C.byref(h))
        if rc:
            raise CrimeError(rc, self._lib.crime_sht_last_error(None).decode())
        self._p = h

    def _check(self, rc):
        if rc:
            raise CrimeError(rc, self._lib.crime_sht_last_error(self._p).decode())

    def map2alm(self, maps):
        """maps (nmaps, Npix) RING on the host -> alms (nmaps, nalm) complex, he_indexlm order."""
        m = np.ascontiguousarray(np.atleast_2d(maps), dtype=np.float64)
        assert m.shape[1] == self.npix
        out = np.empty((m.shape[0], self.nalm), np.complex128)
        self._check(self._lib.crime_sht_map2alm(self._p, m.shape[0], _dptr(m), out.ctypes.data_as(_ffi._dp)))
        return out

    def map2alm_device(self, dptr, nmaps, fp64):
        out = np.empty((nmaps, self.nalm), np.complex128)
        self._check(self._lib.crime_sht_map2alm_device(self._p, nmaps, C.c_void_p(dptr), int(fp64),
                                                       out.ctypes.data_as(_ffi._dp)))
        return out

    def anafast(self, maps, cross=True):
        """C_l of host maps (nmaps, Npix): (nmaps, nmaps, lmax+1) if cross else (nmaps, lmax+1)."""
        m = np.ascontiguousarray(np.atleast_2d(maps), dtype=np.float64)
        assert m.shape[1] == self.npix
        n = m.shape[0]
        out = np.empty((n, n, self.lmax + 1) if cross else (n, self.lmax + 1))
        self._check(self._lib.crime_sht_anafast(self._p, n, _dptr(m), int(cross), _dptr(out)))
        return out

    def anafast_device(self, dptr, nmaps, fp64, cross=True, to_host=True):
        """C_l of device maps: (nmaps, nmaps, lmax+1) if cross else (nmaps, lmax+1).  to_host=False keeps
        them in the plan's device buffer and returns (device pointer, bytes)."""
        if not to_host:
            self._check(self._lib.crime_sht_anafast_device(self._p, nmaps, C.c_void_p(dptr), int(fp64), int(cross),
                                                           None))
            p, nb = C.c_void_p(), C.c_size_t()
            self._check(self._lib.crime_sht_device_cls(self._p, C.byref(p), C.byref(nb)))
            return p.value, nb.value
        shape = (nmaps, nmaps, self.lmax + 1) if cross else (nmaps, self.lmax + 1)
        out = np.empty(shape)
        self._check(self._lib.crime_sht_anafast_device(self._p, nmaps, C.c_void_p(dptr), int(fp64), int(cross),
                                                       _dptr(out)))
        return out

    def cls_sets_device(self, dptrs, maps_per_set, fp64, to_host=True):
        """Cross spectra within each set of maps_per_set device maps (set s at dptrs[s]):
        (nsets, maps_per_set, maps_per_set, lmax+1), or (device pointer, bytes) with to_host=False."""
        n = len(dptrs)
        tab = (C.c_void_p * n)(*dptrs)
        if not to_host:
            self._check(self._lib.crime_sht_anafast_sets_device(self._p, n, tab, maps_per_set, int(fp64), None))
            p, nb = C.c_void_p(), C.c_size_t()
            self._check(self._lib.crime_sht_device_cls(self._p, C.byref(p), C.byref(nb)))
            return p.value, nb.value
        out = np.empty((n, maps_per_set, maps_per_set, self.lmax + 1))
        self._check(self._lib.crime_sht_anafast_sets_device(self._p, n, tab, maps_per_set, int(fp64), _dptr(out)))
        return out

    def cls_of_members(self, ghs, to_host=True):
        """Per-member cross spectra of several GetHI handles of this plan's Nside and one precision and
        channel count, in one call."""
        fp64 = {gh.precision == FP64 for gh in ghs}
        nnu = {gh.n_nu for gh in ghs}
        assert len(fp64) == 1 and len(nnu) == 1
        return self.cls_sets_device([gh.device_maps()[0] for gh in ghs], nnu.pop(), fp64.pop(), to_host)

    def cls_pairs_device(self, dptr, nmaps, fp64, pairs):
        """C_l of the listed (a, b) map pairs of nmaps device maps: (npairs, lmax+1)."""
        pr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2))
        out = np.empty((pr.shape[0], self.lmax + 1))
        self._check(self._lib.crime_sht_cls_pairs_device(self._p, nmaps, C.c_void_p(dptr), int(fp64), pr.shape[0],
                                                          pr.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(out)))
        return out

    def cls_of(self, gh, cross=True, channels=None, pairs=None, to_host=True):
        """C_l of a GetHI handle's resident maps: all pairs (cross), autospectra, or the listed pairs."""
        p, _ = gh.device_maps()
        n = gh.n_nu if channels is None else int(channels)
        if pairs is not None:
            return self.cls_pairs_device(p, n, gh.precision == FP64, pairs)
        return self.anafast_device(p, n, gh.precision == FP64, cross, to_host)

    def alm2map(self, alms):
        """alms (nmaps, nalm) complex, he_indexlm order, on the host -> real maps (nmaps, Npix) RING, as
        he_alm2map (src/healpix_extra.c:44-86) computes them."""
        a = np.ascontiguousarray(np.atleast_2d(alms), dtype=np.complex128)
        assert a.shape[1] == self.nalm
        out = np.empty((a.shape[0], self.npix))
        self._check(self._lib.crime_sht_alm2map(self._p, a.shape[0], a.ctypes.data_as(_ffi._dp), _dptr(out)))
        return out

    def alm2map_device(self, d_alms, nmaps, d_maps, fp64):
        """Device alms (nmaps x nalm complex) into device maps (FP32 or FP64)."""
        self._check(self._lib.crime_sht_alm2map_device(self._p, nmaps, C.c_void_p(d_alms), C.c_void_p(d_maps),
                                                       int(fp64)))

    def smooth_device(self, dptr, nmaps, fp64, fwhm_arcmin):
        """JoinT's apply_beam in place: map2alm, Gaussian beam of fwhm_arcmin[t] per map, alm2map."""
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(fwhm_arcmin, np.float64), (nmaps,)))
        self._check(self._lib.crime_sht_smooth_device(self._p, nmaps, C.c_void_p(dptr), int(fp64), _dptr(f)))

    def add_noise_device(self, dptr, nmaps, fp64, sigma2, seed):
        """JoinT's add_noise: += sqrt(-sigma2[t] ln(1-u2)) cos(2 pi u1) per pixel (variance sigma2 / 2)."""
        s2 = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma2, np.float64), (nmaps,)))
        self._check(self._lib.crime_sht_add_noise_device(self._p, nmaps, C.c_void_p(dptr), int(fp64), _dptr(s2),
                                                         int(seed) & 0xFFFFFFFF))

    def mask_device(self, dptr, nmaps, fp64, mask):
        """JoinT's apply_mask: zero every pixel where mask (Npix, nonzero = observed) is 0."""
        mk = np.ascontiguousarray(np.asarray(mask) != 0, dtype=np.uint8)
        assert mk.shape == (self.npix,)
        self._check(self._lib.crime_sht_mask_device(self._p, nmaps, C.c_void_p(dptr), int(fp64),
                                                    mk.ctypes.data_as(C.POINTER(C.c_uint8))))

    def instrument_of(self, gh, fwhm_arcmin=None, sigma2=None, seed=0, mask=None):
        """The JoinT stages in its order (src/main_jt.c:362-382) on a GetHI handle's resident maps: beam if
        fwhm_arcmin is given, noise if sigma2 is given, then the mask if one is given."""
        p, _ = gh.device_maps()
        fp64 = gh.precision == FP64
        if fwhm_arcmin is not None:
            self.smooth_device(p, gh.n_nu, fp64, fwhm_arcmin)
        if sigma2 is not None:
            self.add_noise_device(p, gh.n_nu, fp64, sigma2, seed)
        if mask is not None:
            self.mask_device(p, gh.n_nu, fp64, mask)

    def stats(self):
        st = _ffi.CrimeShtStats()
        self._check(self._lib.crime_sht_get_stats(self._p, C.byref(st)))
        return st.as_dict()

    def close(self):
        if getattr(self, "_p", None):
            self._lib.crime_sht_destroy(self._p)
            self._p = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()
