"""Host C library (libgethi_host.so): parameter file, RNG, cosmology tables, P(k), outputs.

Pins: MT19937 against the C++11 known answer and numpy's independent MT19937 (same init_genrand
seeding); cosmology against Einstein-de Sitter closed forms and scipy quadrature; P(k) normalisation
against the top-hat variance of a pure power law integrated by scipy; FITS files read back by cfitsio.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest
from scipy import integrate

import crime_amd
import crime_amd.configs as cf
from crime_amd import _ffi

H = _ffi.host_lib()


def mt_state(seed):
    st = (C.c_uint32 * 626)()
    H.gh_mt_seed(st, seed)
    return st


# ---------------------------------------------------------------------------------------------
# parameter file (src/io_gh.c:188-296)
# ---------------------------------------------------------------------------------------------
SAMPLE_KEYS = """###File names
prefix_out= {d}/test
pk_filename= {pk}
omega_M= 0.3
omega_L= 0.7
omega_B= 0.049
h= 0.67
w= -1.0
ns= 0.96
sigma_8= 0.8
r_smooth= 2.
frequencies_filename= {nu}
nu_min= 355.
nu_max= 945.
n_nu= 150.
n_side= 512
n_grid= 64
seed= 1001
do_psources= 0
"""


def test_reference_style_parameter_file(tmp_path):
    # the reference sample parameter set with its own P(k) and frequency table (tests/golden)
    gold = os.path.join(os.path.dirname(__file__), "golden")
    ini = tmp_path / "p.ini"
    ini.write_text(SAMPLE_KEYS.format(d=tmp_path, pk=os.path.join(gold, "Pk_CAMB_test.dat"),
                                      nu=os.path.join(gold, "nuTable_ref.txt")))
    par = crime_amd.Params(str(ini))
    assert par.get("n_nu") == 150            # 151 edges (src/io_gh.c:36)
    assert par.get("nu_min") == 355.0 and par.get("nu_max") == 945.0
    assert par.get("n_side") == 512 and par.get("n_grid") == 64 and par.get("seed") == 1001
    assert par.get("n_sub") == 10            # N_SUBPART (src/pixelize.c:25)
    d = par.desc()
    assert d.numk == 807
    assert abs(d.r2_smooth - 4.0) < 1e-15 and d.do_smoothing == 1     # squared (io_gh.c:255-260)
    assert d.nu_mode == crime_amd.NU_TABLE
    # derived box: L = 2 r_max (1 + 2/N) (src/cosmo.c:361), observer at the centre
    assert abs(d.l_box - 2 * par.get("r_max") * (1 + 2.0 / 64)) < 1e-9
    assert all(abs(d.pos_obs[i] - d.l_box / 2) < 1e-12 for i in range(3))
    # sigma_8 of the rescaled table equals the requested one
    assert abs(math.sqrt(H.gh_pk_sigma2_tophat(par._p, 8.0)) - 0.8) < 1e-10


def test_grammar_errors(tmp_path):
    ini = tmp_path / "bad.ini"
    ini.write_text("omega_M=0.3\n")   # sscanf("%s %s") finds one token: fatal in the reference
    with pytest.raises(ValueError, match="Error reading file"):
        crime_amd.Params(str(ini))
    with pytest.raises(ValueError, match="Couldn't open file"):
        crime_amd.Params(str(tmp_path / "missing.ini"))


def test_unknown_key_is_a_warning(tmp_path, capfd):
    p = cf.make(str(tmp_path), "unk", n_grid=16, n_side=8, n_nu=4, nu_lo=1200.0, nu_hi=1300.0, l_box=300.0,
                seed=1, extra={"some_new_key": 3})
    crime_amd.Params(p)
    assert "Unknown parameter some_new_key=" in capfd.readouterr().err


def test_descending_edges_rejected(tmp_path):
    nu = tmp_path / "nu.txt"
    nu.write_text("800\n790\n780\n")
    p = cf.make(str(tmp_path), "desc", n_grid=16, n_side=8, n_nu=4, nu_lo=1200.0, nu_hi=1300.0, l_box=300.0,
                seed=1, extra={"frequencies_filename": str(nu)})
    with pytest.raises(ValueError, match="Frequency bins don't make sense"):
        crime_amd.Params(p)


def test_redshift_beyond_tables_rejected(tmp_path):
    # z_max > 5 leaves the NZ=5001, DZ=0.001 tables ("OMG!", src/cosmo.c:374-375)
    p = cf.make(str(tmp_path), "zhigh", n_grid=16, n_side=8, n_nu=4, nu_lo=200.0, nu_hi=300.0, l_box=300.0,
                seed=1)
    with pytest.raises(ValueError, match="beyond the z<=5 tables"):
        crime_amd.Params(p)


def test_unsupported_grid_rejected(tmp_path):
    p = cf.make(str(tmp_path), "ng", n_grid=100, n_side=8, n_nu=4, nu_lo=1200.0, nu_hi=1300.0, l_box=300.0,
                seed=1)
    with pytest.raises(ValueError, match="n_grid=100"):
        crime_amd.Params(p)


# ---------------------------------------------------------------------------------------------
# MT19937 (src/common.c:133-164)
# ---------------------------------------------------------------------------------------------
def test_mt19937_known_answer():
    st = mt_state(5489)
    for _ in range(9999):
        H.gh_mt_next(st)
    assert H.gh_mt_next(st) == 4123659995  # C++11 [rand.predef]: 10000th output of mt19937


@pytest.mark.parametrize("seed", [1, 1001, 2002, 4095, 2 ** 31 + 7])
def test_mt19937_matches_numpy(seed):
    st = mt_state(seed)
    ours = np.array([H.gh_mt_next(st) for _ in range(2000)], dtype=np.uint32)
    ref = np.random.RandomState(seed).randint(0, 2 ** 32, size=2000, dtype=np.uint32)
    assert np.array_equal(ours, ref)
    st = mt_state(seed)
    u = H.gh_mt_uniform(st)
    assert u == float(ref[0]) / 2 ** 32   # gsl_rng_uniform convention


def test_mt19937_seed_zero_is_4357():
    a, b = mt_state(0), mt_state(4357)
    assert [H.gh_mt_next(a) for _ in range(5)] == [H.gh_mt_next(b) for _ in range(5)]


def test_unit_modes_stream():
    # phase from u1, modulus from u2, kz slowest / kx fastest, no draw for k=0 (fourier.c:260-302)
    n = 8
    out = np.empty(2 * n * n * (n // 2 + 1))
    H.gh_unit_modes_mt(n, 1001, out.ctypes.data_as(C.POINTER(C.c_double)))
    g = out[0::2] + 1j * out[1::2]
    u = np.random.RandomState(1001).randint(0, 2 ** 32, size=2 * g.size, dtype=np.uint32) / 2.0 ** 32
    assert g[0] == 0
    exp = np.sqrt(-np.log(1 - u[1:2 * g.size - 1:2][: g.size - 1])) * np.exp(2j * np.pi * u[0:2 * g.size - 2:2][: g.size - 1])
    assert np.allclose(g[1:], exp, rtol=1e-14, atol=0)


def test_subcell_offsets():
    lcell = 8.1
    out = np.empty(3 * 10)
    n = H.gh_subcell_offsets(0, lcell, 2002, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert n == 10
    u = np.random.RandomState(2002).randint(0, 2 ** 32, size=30, dtype=np.uint32) / 2.0 ** 32
    assert np.array_equal(out, lcell * (u - 0.5))   # x,y,z per sample (src/pixelize.c:159-163)
    out = np.empty(3 * 27)
    assert H.gh_subcell_offsets(3, lcell, 0, out.ctypes.data_as(C.POINTER(C.c_double))) == 27
    o = out.reshape(27, 3)
    assert np.allclose(o.mean(axis=0), 0, atol=1e-15)
    assert np.allclose(np.sort(np.unique(np.round(o[:, 0], 12))), lcell * (np.arange(3) + 0.5) / 3 - lcell / 2)


# ---------------------------------------------------------------------------------------------
# cosmology tables (src/cosmo.c:341-394, src/cosmo_mad.c)
# ---------------------------------------------------------------------------------------------
def _params(tmp_path, name, **kw):
    a = dict(n_grid=16, n_side=8, n_nu=8, nu_lo=400.0, nu_hi=800.0, l_box=9000.0, seed=3)
    a.update(kw)
    return crime_amd.Params(cf.make(str(tmp_path), name, **a))


def _tables(d):
    t = {}
    for k in ("z_z2r", "r_z2r", "z_r2z", "r_r2z", "d_r2z", "v_r2z"):
        t[k] = np.ctypeslib.as_array(getattr(d, k), shape=(_ffi.NZ,)).copy()
    return t


def test_eds_closed_forms(tmp_path):
    # Omega_m = 1: chi = 2 c/H0 (1 - (1+z)^-1/2), D = a, f = 1, D H f/(f0 H0) = a^-1/2
    par = _params(tmp_path, "eds", omega_m=1.0, extra={"omega_L": 0.0, "omega_B": 0.05})
    t = _tables(par.desc())
    z = t["z_z2r"]
    assert np.allclose(z, np.arange(_ffi.NZ) * 0.001, rtol=0, atol=1e-15)
    chi = 2 * 2997.92458 * (1 - 1 / np.sqrt(1 + z))
    assert np.allclose(t["r_z2r"], chi, rtol=1e-12, atol=1e-9)
    a = 1 / (1 + t["z_r2z"])
    assert np.allclose(t["d_r2z"], a, rtol=1e-12)
    assert np.allclose(t["v_r2z"], 1 / np.sqrt(a), rtol=1e-11)
    assert abs(par.get("fgrowth_0") - 1) < 1e-12
    assert abs(par.get("hubble_0") - 1 / 2997.92458) < 1e-18


def test_lcdm_against_scipy(tmp_path):
    om = 0.3
    par = _params(tmp_path, "lcdm", omega_m=om)
    d = par.desc()
    t = _tables(d)
    E = lambda z: math.sqrt(om * (1 + z) ** 3 + 1 - om)
    for i in (1, 100, 777, 2500, 5000):
        chi = 2997.92458 * integrate.quad(lambda z: 1 / E(z), 0, t["z_z2r"][i], epsabs=0, epsrel=1e-13)[0]
        assert abs(t["r_z2r"][i] - chi) < 1e-9 * chi
    # growth: D(a) proportional to E(a) int_0^a da'/(a' E(a'))^3 (Heath 1977)
    Ea = lambda a: math.sqrt(om / a ** 3 + 1 - om)
    Dun = lambda a: Ea(a) * integrate.quad(lambda x: 1 / (x * Ea(x)) ** 3, 0, a, epsabs=0, epsrel=1e-13)[0]
    for i in (10, 1000, 3000, 4999):
        a = 1 / (1 + t["z_r2z"][i])
        assert abs(t["d_r2z"][i] - Dun(a) / Dun(1.0)) < 2e-6
    # the uniform-r table's geometry
    idr = par.get("idr")
    assert abs(idr - 5000 / t["r_z2r"][5000]) < 1e-15 * idr
    assert np.array_equal(t["r_r2z"], np.arange(_ffi.NZ) / idr)
    # z(r) inverts r(z) up to the lerp error of the table
    zz = np.array([H.gh_z_of_r(par._p, r) for r in t["r_z2r"][1:4999:37]])
    assert np.allclose(zz, t["z_z2r"][1:4999:37], atol=1e-7)


def test_table_lookup_edges(tmp_path):
    par = _params(tmp_path, "edges")
    t = _tables(par.desc())
    assert H.gh_z_of_r(par._p, -1.0) == 0 and H.gh_dgrowth_of_r(par._p, 0.0) == 1
    assert H.gh_vgrowth_of_r(par._p, -3.0) == 1
    assert H.gh_z_of_r(par._p, 1e9) == t["z_r2z"][-1]
    assert H.gh_r_of_z(par._p, 7.0) == t["r_z2r"][-1]


# ---------------------------------------------------------------------------------------------
# P(k) (src/cosmo.c:153-170, 232-339)
# ---------------------------------------------------------------------------------------------
def test_pk_power_law_sigma(tmp_path):
    # a pure power law P = k^n tabulated densely: the top-hat variance of the reference's log-k
    # integral must match the textbook k-space integral (scipy), and the pipeline rescales to sigma_8
    n = -1.2
    k = np.logspace(-6, 4, 20001)
    pkf = tmp_path / "pl.dat"
    np.savetxt(pkf, np.c_[k, k ** n], fmt="%.17e")
    p = cf.make(str(tmp_path), "pl", n_grid=16, n_side=8, n_nu=4, nu_lo=1200.0, nu_hi=1300.0, l_box=300.0,
                seed=1, n_s=n, sigma_8=0.9, extra={"pk_filename": str(pkf)})
    par = crime_amd.Params(p)
    W = lambda x: 3 * (math.sin(x) - x * math.cos(x)) / x ** 3 if x > 1e-3 else 1 - x * x / 10
    R = 8.0
    f = lambda kk: kk ** (2 + n) * W(kk * R) ** 2 / (2 * math.pi ** 2)
    edges = np.r_[0, np.arange(1, 40000) * math.pi / R / 4]
    ref = sum(integrate.quad(f, a, b, epsabs=0, epsrel=1e-12, limit=200)[0] for a, b in zip(edges[:-1], edges[1:]))
    s8_in = par.get("sigma8_in")
    assert abs(s8_in ** 2 - ref) < 2e-5 * ref
    assert abs(H.gh_pk_sigma2_tophat(par._p, 8.0) - 0.81) < 1e-10


def test_pk_linear0_branches(tmp_path):
    par = _params(tmp_path, "pkb")
    d = par.desc()
    lk = np.ctypeslib.as_array(d.logk, shape=(d.numk,))
    pk = np.ctypeslib.as_array(d.pk, shape=(d.numk,))
    # inside: linear in log10 k between nodes
    mid = 0.5 * (lk[100] + lk[101])
    assert abs(H.gh_pk_linear0(par._p, mid) - 0.5 * (pk[100] + pk[101])) < 1e-9 * pk[100]
    # below: k^ns; above: k^-3
    assert abs(H.gh_pk_linear0(par._p, lk[0] - 1) - pk[0] * 10 ** (-d.n_s)) < 1e-12 * pk[0]
    assert abs(H.gh_pk_linear0(par._p, lk[-1] + 1) - pk[-1] * 1e-3) < 1e-12 * pk[-1]


def test_hi_fraction():
    assert H.gh_fraction_hi(0.0) == 0.008
    assert abs(H.gh_fraction_hi(1.0) - 0.008 * 2 ** 0.6) < 1e-17


# ---------------------------------------------------------------------------------------------
# outputs (src/io_gh.c:69-131, src/healpix_extra.c:132-164)
# ---------------------------------------------------------------------------------------------
def _cfitsio():
    for p in ("/opt/conda/lib/libcfitsio.so", "libcfitsio.so"):
        try:
            return C.CDLL(p)
        except OSError:
            pass
    return None


def test_fits_roundtrip_with_cfitsio(tmp_path):
    lib = _cfitsio()
    if lib is None:
        pytest.skip("cfitsio not available")
    nside = 4
    m = np.arange(12 * nside * nside, dtype=np.float32) * 0.5 - 3
    fn = str(tmp_path / "m.fits")
    assert H.gh_write_healpix_fits(fn.encode(), m.ctypes.data_as(C.POINTER(C.c_float)), nside) == 0
    assert os.path.getsize(fn) % 2880 == 0
    f = C.c_void_p()
    st = C.c_int(0)
    lib.ffopen(C.byref(f), fn.encode(), 0, C.byref(st))
    assert st.value == 0
    lib.ffmahd(f, 2, None, C.byref(st))
    def key(name):
        buf = C.create_string_buffer(80)
        lib.ffgkys(f, name.encode(), buf, None, C.byref(st))
        return buf.value.decode().strip()
    assert key("PIXTYPE") == "HEALPIX" and key("ORDERING") == "RING" and key("COORDSYS") == "G"
    assert key("TTYPE1") == "T" and key("TUNIT1") == "mK" and key("TFORM1") == "1E"
    ns = C.c_long()
    lib.ffgkyj(f, b"NSIDE", C.byref(ns), None, C.byref(st))
    assert ns.value == nside
    out = np.empty_like(m)
    anynul = C.c_int()
    lib.ffgcve(f, 1, C.c_longlong(1), C.c_longlong(1), C.c_longlong(m.size), C.c_float(0),
               out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(anynul), C.byref(st))
    lib.ffclos(f, C.byref(st))
    assert st.value == 0
    assert np.array_equal(out, m)


def test_write_maps_and_nutable(tmp_path):
    par = _params(tmp_path, "out", n_side=2, n_nu=3, nu_lo=700.0, nu_hi=730.0)
    maps = np.arange(3 * 48, dtype=np.float32).reshape(3, 48)
    par.write_maps(maps)
    for i in range(3):
        assert os.path.exists(tmp_path / f"out_{i + 1:03d}.fits")
    lines = (tmp_path / "out_nuTable.dat").read_text().splitlines()
    assert len(lines) == 3
    f = lines[0].split()
    assert f[0] == "1" and float(f[1]) == 700.0 and float(f[2]) == 710.0
    assert abs(float(f[3]) - (1420.40575177 / 710.0 - 1)) < 1e-6
    assert abs(float(f[4]) - (1420.40575177 / 700.0 - 1)) < 1e-6
