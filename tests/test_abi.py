"""The C-ABI libraries load and export what include/*.h declare; failure paths return status codes.

Runs on the CPU: no call here reaches a kernel.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

import crime_amd
from crime_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n.startswith(("crime_gh_", "crime_sht_", "crime_mlp_", "gh_"))))


@pytest.mark.parametrize("header,lib", [("crime_gh.h", "gpu"), ("crime_sht.h", "gpu"), ("crime_mlp.h", "gpu"),
                                        ("gethi_host.h", "host")])
def test_every_declared_symbol_is_exported(header, lib):
    L = _ffi.gpu_lib() if lib == "gpu" else _ffi.host_lib()
    names = declared(header)
    assert len(names) >= 7
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_descriptor_layout():
    assert _ffi.gpu_lib().crime_gh_desc_size() == C.sizeof(_ffi.CrimeGhDesc)


def test_create_rejects_bad_descriptors(cfgdir):
    from conftest import small_params

    lib = _ffi.gpu_lib()
    h = C.c_void_p()
    assert lib.crime_gh_create(None, C.byref(h)) == -1
    par = small_params(cfgdir, "abi")
    d = par.desc()
    d.abi_version = 99
    assert lib.crime_gh_create(C.byref(d), C.byref(h)) == -6
    assert b"abi_version" in lib.crime_gh_last_error(None)
    d = par.desc()
    d.n_grid = 100
    assert lib.crime_gh_create(C.byref(d), C.byref(h)) == -1
    assert b"n_grid=100" in lib.crime_gh_last_error(None)
    d = par.desc()
    d.mode_source = crime_amd.MODES_INJECTED  # without an array
    assert lib.crime_gh_create(C.byref(d), C.byref(h)) == -1
    assert h.value is None


def test_null_handle_calls_fail_cleanly():
    lib = _ffi.gpu_lib()
    assert lib.crime_gh_run(None, None) == -1
    assert lib.crime_gh_mk_T_maps(None, None) == -1
    lib.crime_gh_destroy(None)


def _has_device():
    try:
        hip = C.CDLL("libamdhip64.so")
        n = C.c_int(0)
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def test_no_device_is_reported(cfgdir):
    if _has_device():
        pytest.skip("a device is visible")
    from conftest import small_params

    par = small_params(cfgdir, "nodev")
    with pytest.raises(crime_amd.CrimeError, match="ENODEV"):
        crime_amd.GetHI(par)


def test_cli_errors(tmp_path):
    exe = os.path.join(ROOT, "crime_amd", "_lib", "gethi")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert "Usage" in r.stderr
    bad = tmp_path / "bad.ini"
    bad.write_text("n_grid=64\n")
    r = subprocess.run([exe, str(bad)], capture_output=True, text=True)
    assert r.returncode == 1 and "Error reading file" in r.stderr


def test_sht_null_and_bad_plans():
    lib = _ffi.gpu_lib()
    h = C.c_void_p()
    assert lib.crime_sht_map2alm(None, 1, None, None) == -1
    assert lib.crime_sht_anafast_device(None, 1, None, 1, 1, None) == -1
    assert lib.crime_sht_alm2map(None, 1, None, None) == -1
    assert lib.crime_sht_alm2map_device(None, 1, None, None, 1) == -1
    assert lib.crime_sht_smooth_device(None, 1, None, 1, None) == -1
    assert lib.crime_sht_add_noise_device(None, 1, None, 1, None, 0) == -1
    assert lib.crime_sht_mask_device(None, 1, None, 1, None) == -1
    lib.crime_sht_destroy(None)
    # bad geometry is rejected before any device work
    assert lib.crime_sht_create(0, 10, 1, 0, C.byref(h)) == -1
    assert h.value is None
    assert lib.crime_sht_create(16, 64, 1, 0, C.byref(h)) == -1      # lmax > 4 nside - 1
    assert b"lmax" in lib.crime_sht_last_error(None)
    assert lib.crime_sht_create(16, 47, 0, 0, C.byref(h)) == -1      # no maps
    assert lib.crime_sht_create(2048, 6143, 1, 0, C.byref(h)) == -1  # ring spectra beyond the LDS
    assert b"LDS" in lib.crime_sht_last_error(None)


def test_mlp_abi_without_device():
    lib = _ffi.gpu_lib()
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mlp_oracle

    assert lib.crime_mlp_param_count() == mlp_oracle.layout()[1] == 265729
    h = C.c_void_p()
    assert lib.crime_mlp_create(100, 0, C.byref(h)) == -1        # not a multiple of 64
    assert h.value is None
    assert b"max_batch" in lib.crime_mlp_last_error(None)
    assert lib.crime_mlp_forward(None, 1, None, None) == -1
    assert lib.crime_mlp_forward_mean(None, 1, 1, None, None) == -1
    assert lib.crime_mlp_train(None, 1, 64, 1e-3, 10, 0.9, 0.999, 1e-8, 0, None) == -1
    lib.crime_mlp_destroy(None)
