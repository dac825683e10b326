"""ctypes bindings of the two C ABIs of the GetHI drop-in.

``libcrime_gh.so``   gfx950 kernels behind include/crime_gh.h (the hot path)
``libgethi_host.so`` host C of include/gethi_host.h (parameter file, tables, P(k), FITS)

Both are built in-tree into ``crime_amd/_lib`` by ``crime_amd/csrc/Makefile``.  Loading fails loudly
when a library is missing: there is no CPU fallback for the hot path.
"""
import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")

ABI_VERSION = 2
NZ = 5001
N_SCHED = 4   # CRIME_GH_N_SCHED: variants per FFT stage (the map stage has 3)
FP32, FP64 = 0, 1
MODES_PHILOX, MODES_INJECTED = 0, 1
NU_TABLE, NU_UNIFORM = 0, 1
BIAS_REFERENCE, BIAS_CONST, BIAS_LINEAR = 0, 1, 2
GRID_DENS, GRID_VPOT, GRID_RVEL, GRID_MASS, GRID_DZRSD = 0, 1, 2, 3, 4
GRID_VX, GRID_VY, GRID_VZ = 5, 6, 7
VEL_STENCIL, VEL_SPECTRAL = 0, 1

ERRORS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ENODEV", -5: "ESTATE", -6: "EABI"}

_dp = C.POINTER(C.c_double)


class CrimeGhDesc(C.Structure):
    """Mirror of ``crime_gh_desc`` (include/crime_gh.h)."""

    _fields_ = [
        ("abi_version", C.c_uint32),
        ("precision", C.c_int32),
        ("n_grid", C.c_int32),
        ("n_nu", C.c_int32),
        ("n_side", C.c_int64),
        ("nu_mode", C.c_int32),
        ("rsd", C.c_int32),
        ("nu0", _dp),
        ("nuf", _dp),
        ("nu_min", C.c_double),
        ("nu_max", C.c_double),
        ("l_box", C.c_double),
        ("pos_obs", C.c_double * 3),
        ("omega_b", C.c_double),
        ("hhub", C.c_double),
        ("f0h0", C.c_double),
        ("r2_smooth", C.c_double),
        ("do_smoothing", C.c_int32),
        ("nz", C.c_int32),
        ("z_z2r", _dp),
        ("r_z2r", _dp),
        ("z_r2z", _dp),
        ("r_r2z", _dp),
        ("d_r2z", _dp),
        ("v_r2z", _dp),
        ("idr", C.c_double),
        ("logk", _dp),
        ("pk", _dp),
        ("numk", C.c_int32),
        ("logkmin", C.c_double),
        ("logkmax", C.c_double),
        ("idlogk", C.c_double),
        ("n_s", C.c_double),
        ("bias_model", C.c_int32),
        ("bias_b0", C.c_double),
        ("bias_b1", C.c_double),
        ("seed", C.c_uint32),
        ("n_sub", C.c_int32),
        ("sub_off", _dp),
        ("mode_source", C.c_int32),
        ("unit_modes", _dp),
        ("device", C.c_int32),
        ("velocity_mode", C.c_int32),
        ("deterministic", C.c_int32),
    ]


class CrimeShtStats(C.Structure):
    """Mirror of ``crime_sht_stats``."""

    _fields_ = [("ms_rings", C.c_double), ("ms_legendre", C.c_double), ("ms_alm", C.c_double),
                ("ms_cls", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CrimeGhPsStats(C.Structure):
    """Mirror of ``crime_gh_ps_stats``."""

    _fields_ = [("ms_sources", C.c_double), ("ms_maps", C.c_double), ("n_sources", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CrimeGhStats(C.Structure):
    """Mirror of ``crime_gh_stats``."""

    _fields_ = [
        ("ms_fill_fft_z", C.c_double),
        ("ms_fft_y", C.c_double),
        ("ms_fft_x", C.c_double),
        ("ms_sigma2", C.c_double),
        ("ms_maps_clear", C.c_double),
        ("ms_pixelize", C.c_double),
        ("ms_total", C.c_double),
        ("sigma2_gauss", C.c_double),
        ("mean_gauss", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built: run `make -C crime_amd/csrc` (or __graft_entry__.build()); "
            "the GetHI hot path has no CPU fallback")
    return C.CDLL(path)


_gpu = None
_host = None


def gpu_lib():
    """libcrime_gh.so with prototypes (loads the HIP runtime, touches no device)."""
    global _gpu
    if _gpu is None:
        lib = _load("libcrime_gh.so")
        H = C.c_void_p
        lib.crime_gh_desc_size.restype = C.c_size_t
        lib.crime_gh_desc_size.argtypes = []
        lib.crime_gh_create.restype = C.c_int
        lib.crime_gh_create.argtypes = [C.POINTER(CrimeGhDesc), C.POINTER(H)]
        lib.crime_gh_destroy.restype = None
        lib.crime_gh_destroy.argtypes = [H]
        lib.crime_gh_last_error.restype = C.c_char_p
        lib.crime_gh_last_error.argtypes = [H]
        for fn in ("crime_gh_create_d_and_vr_fields", "crime_gh_mk_T_maps", "crime_gh_run"):
            f = getattr(lib, fn)
            f.restype = C.c_int
            f.argtypes = [H, C.POINTER(CrimeGhStats)]
        lib.crime_gh_copy_maps.restype = C.c_int
        lib.crime_gh_copy_maps.argtypes = [H, C.POINTER(C.c_float)]
        lib.crime_gh_copy_maps_f64.restype = C.c_int
        lib.crime_gh_copy_maps_f64.argtypes = [H, _dp]
        lib.crime_gh_get_sigma2.restype = C.c_int
        lib.crime_gh_get_sigma2.argtypes = [H, _dp, _dp]
        lib.crime_gh_set_sigma2.restype = C.c_int
        lib.crime_gh_set_sigma2.argtypes = [H, C.c_double, C.c_double]
        lib.crime_gh_copy_grid.restype = C.c_int
        lib.crime_gh_copy_grid.argtypes = [H, C.c_int, _dp]
        lib.crime_gh_set_grid.restype = C.c_int
        lib.crime_gh_set_grid.argtypes = [H, C.c_int, _dp]
        lib.crime_gh_sample_indices.restype = C.c_int
        lib.crime_gh_sample_indices.argtypes = [H, C.c_int64, C.c_int64, C.POINTER(C.c_int32),
                                                C.POINTER(C.c_int64)]
        lib.crime_gh_get_schedule.restype = C.c_int
        lib.crime_gh_get_schedule.argtypes = [H, C.POINTER(C.c_int)]
        lib.crime_gh_set_schedule.restype = C.c_int
        lib.crime_gh_set_schedule.argtypes = [H, C.POINTER(C.c_int)]
        lib.crime_gh_autotune_cached.restype = C.c_int
        lib.crime_gh_autotune_cached.argtypes = [H, C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.crime_gh_autotune.restype = C.c_int
        lib.crime_gh_autotune.argtypes = [H, C.c_int, C.POINTER(C.c_int), _dp]
        lib.crime_gh_power_spectrum.restype = C.c_int
        lib.crime_gh_power_spectrum.argtypes = [H, C.c_int, C.c_double, _dp, _dp, C.POINTER(C.c_int64)]
        lib.crime_gh_set_seed.restype = C.c_int
        lib.crime_gh_set_seed.argtypes = [H, C.c_uint32]
        lib.crime_gh_copy_maps_channels.restype = C.c_int
        lib.crime_gh_copy_maps_channels.argtypes = [H, C.c_int, C.c_int, C.POINTER(C.c_float)]
        lib.crime_gh_host_alloc.restype = C.c_int
        lib.crime_gh_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
        lib.crime_gh_host_free.restype = None
        lib.crime_gh_host_free.argtypes = [C.c_void_p]
        lib.crime_gh_copy_grid_range.restype = C.c_int
        lib.crime_gh_copy_grid_range.argtypes = [H, C.c_int, C.c_int64, C.c_int64, _dp]
        lib.crime_gh_channel_totals.restype = C.c_int
        lib.crime_gh_channel_totals.argtypes = [H, C.c_int64, C.c_int64, _dp, C.POINTER(C.c_int64)]
        lib.crime_gh_screen_check.restype = C.c_int
        lib.crime_gh_screen_check.argtypes = [H, C.POINTER(C.c_int64), _dp]
        lib.crime_gh_reference_maps.restype = C.c_int
        lib.crime_gh_reference_maps.argtypes = [H, _dp]
        lib.crime_gh_power_spectrum_noalloc.restype = C.c_int
        lib.crime_gh_power_spectrum_noalloc.argtypes = [H, C.c_int, C.c_double, _dp, _dp, C.POINTER(C.c_int64)]
        lib.crime_gh_device_maps.restype = C.c_int
        lib.crime_gh_device_maps.argtypes = [H, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
        lib.crime_gh_mk_psources_maps.restype = C.c_int
        lib.crime_gh_mk_psources_maps.argtypes = [H, C.POINTER(CrimeGhPsStats)]
        lib.crime_gh_copy_ps_maps.restype = C.c_int
        lib.crime_gh_copy_ps_maps.argtypes = [H, _dp]
        lib.crime_gh_device_ps_maps.restype = C.c_int
        lib.crime_gh_device_ps_maps.argtypes = [H, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
        lib.crime_gh_ps_pixel_sums.restype = C.c_int
        lib.crime_gh_ps_pixel_sums.argtypes = [H, _dp, _dp]
        lib.crime_gh_run_batch.restype = C.c_int
        lib.crime_gh_run_batch.argtypes = [C.POINTER(H), C.c_int, C.POINTER(CrimeGhStats)]
        lib.crime_gh_batch_create.restype = C.c_int
        lib.crime_gh_batch_create.argtypes = [C.POINTER(H), C.c_int, C.c_int, C.POINTER(H)]
        lib.crime_gh_batch_run.restype = C.c_int
        lib.crime_gh_batch_run.argtypes = [H, _dp]
        lib.crime_gh_batch_destroy.restype = None
        lib.crime_gh_batch_destroy.argtypes = [H]
        lib.crime_gh_batch_last_error.restype = C.c_char_p
        lib.crime_gh_batch_last_error.argtypes = [H]
        lib.crime_sht_create.restype = C.c_int
        lib.crime_sht_create.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(H)]
        lib.crime_sht_destroy.restype = None
        lib.crime_sht_destroy.argtypes = [H]
        lib.crime_sht_last_error.restype = C.c_char_p
        lib.crime_sht_last_error.argtypes = [H]
        lib.crime_sht_map2alm.restype = C.c_int
        lib.crime_sht_map2alm.argtypes = [H, C.c_int, _dp, _dp]
        lib.crime_sht_map2alm_device.restype = C.c_int
        lib.crime_sht_map2alm_device.argtypes = [H, C.c_int, C.c_void_p, C.c_int, _dp]
        lib.crime_sht_anafast_device.restype = C.c_int
        lib.crime_sht_anafast_device.argtypes = [H, C.c_int, C.c_void_p, C.c_int, C.c_int, _dp]
        lib.crime_sht_cls_pairs_device.restype = C.c_int
        lib.crime_sht_cls_pairs_device.argtypes = [H, C.c_int, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int32), _dp]
        lib.crime_sht_anafast.restype = C.c_int
        lib.crime_sht_anafast.argtypes = [H, C.c_int, _dp, C.c_int, _dp]
        lib.crime_sht_device_cls.restype = C.c_int
        lib.crime_sht_device_cls.argtypes = [H, C.POINTER(C.c_void_p), C.POINTER(C.c_size_t)]
        lib.crime_sht_anafast_sets_device.restype = C.c_int
        lib.crime_sht_anafast_sets_device.argtypes = [H, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.c_int, _dp]
        lib.crime_sht_alm2map.restype = C.c_int
        lib.crime_sht_alm2map.argtypes = [H, C.c_int, _dp, _dp]
        lib.crime_sht_alm2map_device.restype = C.c_int
        lib.crime_sht_alm2map_device.argtypes = [H, C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        lib.crime_sht_smooth_device.restype = C.c_int
        lib.crime_sht_smooth_device.argtypes = [H, C.c_int, C.c_void_p, C.c_int, _dp]
        lib.crime_sht_add_noise_device.restype = C.c_int
        lib.crime_sht_add_noise_device.argtypes = [H, C.c_int, C.c_void_p, C.c_int, _dp, C.c_uint32]
        lib.crime_sht_mask_device.restype = C.c_int
        lib.crime_sht_mask_device.argtypes = [H, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_uint8)]
        lib.crime_sht_get_stats.restype = C.c_int
        lib.crime_sht_get_stats.argtypes = [H, C.POINTER(CrimeShtStats)]
        lib.crime_mlp_param_count.restype = C.c_int64
        lib.crime_mlp_param_count.argtypes = []
        lib.crime_mlp_create.restype = C.c_int
        lib.crime_mlp_create.argtypes = [C.c_int, C.c_int, C.POINTER(H)]
        lib.crime_mlp_destroy.restype = None
        lib.crime_mlp_destroy.argtypes = [H]
        lib.crime_mlp_last_error.restype = C.c_char_p
        lib.crime_mlp_last_error.argtypes = [H]
        _fp = C.POINTER(C.c_float)
        lib.crime_mlp_set_params.restype = C.c_int
        lib.crime_mlp_set_params.argtypes = [H, _fp]
        lib.crime_mlp_get_params.restype = C.c_int
        lib.crime_mlp_get_params.argtypes = [H, _fp]
        lib.crime_mlp_forward.restype = C.c_int
        lib.crime_mlp_forward.argtypes = [H, C.c_int64, _fp, _fp]
        lib.crime_mlp_forward_mean.restype = C.c_int
        lib.crime_mlp_forward_mean.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int64, _fp, _fp]
        lib.crime_mlp_loss_grad.restype = C.c_int
        lib.crime_mlp_loss_grad.argtypes = [H, C.c_int, _fp, _fp, _fp, _dp, _fp]
        lib.crime_mlp_set_data.restype = C.c_int
        lib.crime_mlp_set_data.argtypes = [H, C.c_int64, _fp, _fp, _fp]
        lib.crime_mlp_set_graph.restype = C.c_int
        lib.crime_mlp_set_graph.argtypes = [H, C.c_int]
        lib.crime_mlp_train.restype = C.c_int
        lib.crime_mlp_train.argtypes = [H, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_float, C.c_float,
                                        C.c_uint32, _fp]
        if lib.crime_gh_desc_size() != C.sizeof(CrimeGhDesc):
            raise ImportError("crime_gh_desc layout mismatch between libcrime_gh.so and crime_amd._ffi")
        _gpu = lib
    return _gpu


def host_lib():
    """libgethi_host.so with prototypes."""
    global _host
    if _host is None:
        lib = _load("libgethi_host.so")
        P = C.c_void_p
        lib.gh_params_new.restype = P
        lib.gh_params_new.argtypes = []
        lib.gh_params_delete.restype = None
        lib.gh_params_delete.argtypes = [P]
        lib.gh_params_load.restype = C.c_int
        lib.gh_params_load.argtypes = [P, C.c_char_p]
        lib.gh_params_read.restype = C.c_int
        lib.gh_params_read.argtypes = [P, C.c_char_p]
        lib.gh_params_setup.restype = C.c_int
        lib.gh_params_setup.argtypes = [P]
        lib.gh_params_error.restype = C.c_char_p
        lib.gh_params_error.argtypes = [P]
        lib.gh_params_get.restype = C.c_double
        lib.gh_params_get.argtypes = [P, C.c_char_p]
        lib.gh_params_to_desc.restype = None
        lib.gh_params_to_desc.argtypes = [P, C.POINTER(CrimeGhDesc)]
        lib.gh_write_maps.restype = C.c_int
        lib.gh_write_maps.argtypes = [P, C.POINTER(C.c_float)]
        lib.gh_write_healpix_fits.restype = C.c_int
        lib.gh_write_healpix_fits.argtypes = [C.c_char_p, C.POINTER(C.c_float), C.c_long]
        lib.gh_write_nutable.restype = C.c_int
        lib.gh_write_nutable.argtypes = [P, C.c_char_p]
        lib.gh_write_bbks_pk.restype = C.c_int
        lib.gh_write_bbks_pk.argtypes = [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_int]
        lib.gh_mt_seed.restype = None
        lib.gh_mt_seed.argtypes = [P, C.c_uint32]
        lib.gh_mt_next.restype = C.c_uint32
        lib.gh_mt_next.argtypes = [P]
        lib.gh_mt_uniform.restype = C.c_double
        lib.gh_mt_uniform.argtypes = [P]
        lib.gh_unit_modes_mt.restype = None
        lib.gh_unit_modes_mt.argtypes = [C.c_int, C.c_uint32, _dp]
        lib.gh_subcell_offsets.restype = C.c_int
        lib.gh_subcell_offsets.argtypes = [C.c_int, C.c_double, C.c_uint32, _dp]
        lib.gh_fraction_hi.restype = C.c_double
        lib.gh_fraction_hi.argtypes = [C.c_double]
        lib.gh_pk_linear0.restype = C.c_double
        lib.gh_pk_linear0.argtypes = [P, C.c_double]
        lib.gh_pk_sigma2_tophat.restype = C.c_double
        lib.gh_pk_sigma2_tophat.argtypes = [P, C.c_double]
        for fn in ("gh_z_of_r", "gh_r_of_z", "gh_dgrowth_of_r", "gh_vgrowth_of_r"):
            f = getattr(lib, fn)
            f.restype = C.c_double
            f.argtypes = [P, C.c_double]
        _host = lib
    return _host


def err_name(rc):
    return ERRORS.get(rc, str(rc))
