"""Cross-frequency C_l emulator trained on ensembles of this pipeline.

Inputs (Omega_m, sigma_8, n_s, h, b, nu, nu', log10 l) -> output asinh(C_l(nu, nu') / s) of binned angular
spectra of the HI temperature maps, s = 1e-3 of the median auto spectrum: log C_l + log(2/s) wherever
C_l >> s (every auto spectrum), and still defined where cross spectra of neighbouring channels scatter
through zero (a cell wider than a channel splits its fixed mass between them, which anticorrelates the
two).  Network and training run on the GPU through include/crime_mlp.h
(BF16 MFMA, FP32 accumulation, Adam).  The plain pipeline (GetHI maps + map2alm of include/crime_sht.h)
produces the training targets and stays the method the emulator is judged against.

A "mock" here is one GetHI run: Philox Gaussian field of the cosmology's BBKS P(k), lognormal HI with a
constant bias b, RSD, pixelisation into the frequency channels (deterministic fixed-point scatter, so
every mock is bit-reproducible); its C_l are measured per channel pair and averaged in logarithmic l
bins.  Training is deterministic too (fixed-order split-K gradients, Philox minibatches): the whole
pipeline gives the same network on every run.
"""
import ctypes as C
import math
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import CrimeError, GetHI, Params, Sht, _ffi
from . import configs as cf

NIN = 8
# parameter ranges of the cosmologies (Latin hypercube): Omega_m, sigma_8, n_s, h, b
RANGES = np.array([[0.25, 0.35], [0.70, 0.90], [0.92, 1.00], [0.60, 0.80], [0.80, 1.20]])


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """The 8 -> 256 x 5 (SiLU) -> 1 network of libcrime_gh.so (include/crime_mlp.h)."""

    def __init__(self, max_batch=8192, device=0):
        self._lib = _ffi.gpu_lib()
        self.nparam = int(self._lib.crime_mlp_param_count())
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        rc = self._lib.crime_mlp_create(self.max_batch, device, C.byref(h))
        if rc:
            raise CrimeError(rc, self._lib.crime_mlp_last_error(None).decode())
        self._h = h

    def _check(self, rc):
        if rc:
            raise CrimeError(rc, self._lib.crime_mlp_last_error(self._h).decode())

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparam
        self._check(self._lib.crime_mlp_set_params(self._h, _fp(p)))

    def params(self):
        out = np.empty(self.nparam, np.float32)
        self._check(self._lib.crime_mlp_get_params(self._h, _fp(out)))
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NIN)
        y = np.empty(x.shape[0], np.float32)
        self._check(self._lib.crime_mlp_forward(self._h, x.shape[0], _fp(x), _fp(y)))
        return y

    @staticmethod
    def forward_mean(mlps, x):
        """The mean of the networks' outputs at x, the inputs sent to the device once (crime_mlp_forward_mean)."""
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NIN)
        y = np.empty(x.shape[0], np.float32)
        hs = (C.c_void_p * len(mlps))(*[m._h.value for m in mlps])
        mlps[0]._check(mlps[0]._lib.crime_mlp_forward_mean(hs, len(mlps), x.shape[0], _fp(x), _fp(y)))
        return y

    def loss_grad(self, x, t, w=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NIN)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        wp = None if w is None else _fp(w)
        g = np.empty(self.nparam, np.float32)
        loss = C.c_double()
        self._check(self._lib.crime_mlp_loss_grad(self._h, x.shape[0], _fp(x), _fp(t), wp, C.byref(loss), _fp(g)))
        return loss.value, g

    def set_data(self, x, t, w=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, NIN)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        self._check(self._lib.crime_mlp_set_data(self._h, x.shape[0], _fp(x), _fp(t), None if w is None else _fp(w)))

    def set_graph(self, on):
        """Replay training steps as a captured graph (default) or enqueue them step by step."""
        self._check(self._lib.crime_mlp_set_graph(self._h, 1 if on else 0))

    def train(self, steps, batch, lr0, total_steps, seed=0, beta1=0.9, beta2=0.999, eps=1e-8):
        loss = np.empty(steps, np.float32)
        self._check(self._lib.crime_mlp_train(self._h, steps, batch, lr0, total_steps, beta1, beta2, eps, seed,
                                              _fp(loss)))
        return loss

    def close(self):
        if getattr(self, "_h", None):
            self._lib.crime_mlp_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()


def latin_hypercube(n, seed, ranges=RANGES):
    """n points, one per stratum in every dimension (McKay, Beckman & Conover 1979)."""
    rng = np.random.default_rng(seed)
    d = ranges.shape[0]
    u = (np.argsort(rng.random((d, n)), axis=1).T + rng.random((n, d))) / n
    return ranges[:, 0] + u * (ranges[:, 1] - ranges[:, 0])


class MockSpec:
    """Grid, sky and spectrum binning of the emulated mocks."""

    def __init__(self, n_grid=256, n_side=256, n_nu=64, nu_lo=700.0, nu_hi=800.0, channels=(0, 9, 18, 27, 36, 45, 54, 62),
                 l_edges=None):
        self.n_grid, self.n_side, self.n_nu = n_grid, n_side, n_nu
        self.nu_lo, self.nu_hi = nu_lo, nu_hi
        self.lmax = 3 * n_side - 1
        # auto spectra of the listed channels and their cross spectra with the next channel
        self.pairs = [(c, c) for c in channels] + [(c, c + 1) for c in channels]
        lo = 8
        hi = min(n_side, self.lmax)
        self.l_edges = np.unique(np.round(np.geomspace(lo, hi, 17)).astype(int)) if l_edges is None else np.asarray(l_edges)
        nuc = 0.5 * (np.linspace(nu_lo, nu_hi, n_nu + 1)[1:] + np.linspace(nu_lo, nu_hi, n_nu + 1)[:-1])
        self.nu_centres = nuc

    @property
    def nbins(self):
        return len(self.l_edges) - 1

    def bin(self, cl):
        """Mean of C_l over [edge_i, edge_{i+1}) weighted by 2l+1 (per pair)."""
        out = np.empty((cl.shape[0], self.nbins))
        ls = np.arange(cl.shape[1])
        for i in range(self.nbins):
            sel = (ls >= self.l_edges[i]) & (ls < self.l_edges[i + 1])
            w = 2 * ls[sel] + 1.0
            out[:, i] = (cl[:, sel] * w).sum(axis=1) / w.sum()
        return out

    def l_centres(self):
        e = self.l_edges
        return np.sqrt(e[:-1] * (e[1:] - 1.0).clip(1))


def _params_for(cosmo, spec, tmp, name):
    # deterministic scatter: the training targets (and so the trained network) are the same on every run
    om, s8, ns, h, b = cosmo
    ini = cf.make(tmp, name, n_grid=spec.n_grid, n_side=spec.n_side, n_nu=spec.n_nu, nu_lo=spec.nu_lo,
                  nu_hi=spec.nu_hi, l_box=-1.0, seed=1, omega_m=om, sigma_8=s8, n_s=ns, h=h, precision="single",
                  bias_model="const", bias_b0=b, extra=dict(deterministic=1))
    return Params(ini)


def simulate(cosmos, nseeds, spec, seed0, log=None):
    """Binned C_l of every (cosmology, realisation): array (ncosmo, nseeds, npairs, nbins), and the seconds
    per mock (the cosmology's device set-up shared by its realisations, runs and spectra; host parameter
    set-up excluded).  Realisation j of cosmology i uses the Philox seed seed0 + i nseeds + j."""
    ncos = len(cosmos)
    out = np.empty((ncos, nseeds, len(spec.pairs), spec.nbins))
    t_mock = 0.0
    with tempfile.TemporaryDirectory() as tmp, ThreadPoolExecutor(16) as ex:
        pars = list(ex.map(lambda i: _params_for(cosmos[i], spec, tmp, f"c{i:04d}"), range(ncos)))
        with Sht(spec.n_side, spec.lmax, max_maps=spec.n_nu) as sh:
            for i, par in enumerate(pars):
                # one device handle per cosmology; each realisation re-keys the mode fill
                t0 = time.perf_counter()
                with GetHI(par, seed=int(seed0 + i * nseeds)) as gh:
                    for j in range(nseeds):
                        gh.set_seed(seed0 + i * nseeds + j)
                        gh.run()
                        cl = sh.cls_of(gh, pairs=spec.pairs)
                        out[i, j] = spec.bin(cl)
                t_mock += time.perf_counter() - t0
                par.close()
                if log and (i + 1) % max(1, ncos // 8) == 0:
                    log(f"simulated {i + 1}/{ncos} cosmologies")
    return out, t_mock / (ncos * nseeds)


def features(cosmos, spec):
    """Network inputs of every (cosmology, pair, l bin), raw units."""
    ncos = len(cosmos)
    npair, nb = len(spec.pairs), spec.nbins
    x = np.empty((ncos, npair, nb, NIN))
    lc = np.log10(spec.l_centres())
    for p, (a, b) in enumerate(spec.pairs):
        x[:, p, :, 5] = spec.nu_centres[a]
        x[:, p, :, 6] = spec.nu_centres[b]
    x[:, :, :, 7] = lc[None, None, :]
    x[:, :, :, :5] = np.asarray(cosmos)[:, None, None, :]
    return x.reshape(-1, NIN)


def encode(x, dnu):
    """Fixed affine pre-transform of the raw inputs: nu' -> (nu' - nu) / channel width, so that the pair
    type (0 for an auto spectrum, 1 for neighbouring channels) is an exact input rather than a 6% change
    of a frequency."""
    z = np.array(x, dtype=np.float64, copy=True)
    z[:, 6] = (z[:, 6] - z[:, 5]) / dnu
    return z


def fold_standardisation(p, x_mean, x_std, y_mean, y_std):
    """Parameters of the same network acting on encoded raw inputs and returning y in target units:
    W0 <- W0 / x_std, b0 <- b0 - (x_mean / x_std) W0, w5 <- y_std w5, b5 <- y_std b5 + y_mean."""
    q = np.array(p, dtype=np.float64, copy=True)
    wd = 256
    w0 = q[:NIN * wd].reshape(NIN, wd)
    b0 = q[NIN * wd:NIN * wd + wd]
    b0 -= (x_mean / x_std) @ w0
    w0 /= x_std[:, None]
    o5 = q.size - wd - 1
    q[o5:o5 + wd] *= y_std
    q[-1] = q[-1] * y_std + y_mean
    return q.astype(np.float32)


class Emulator:
    """Trained Mlps (an ensemble of networks that differ in their initial weights and batch draws) whose
    parameters have the input and output standardisation folded in; targets y = asinh(C / scale), the
    prediction the networks' mean y.  One network's error on held-out cosmologies depends on where its
    training ends as much as on the data (a change of 1e-8 in the targets can move it several-fold); the
    mean of a few is steadier and smaller."""

    def __init__(self, mlps, x_mean, x_std, y_mean, y_std, scale, dnu):
        self.mlps = list(mlps) if isinstance(mlps, (list, tuple)) else [mlps]
        self.mlp = self.mlps[0]
        self.scale, self.dnu = scale, dnu
        self._z = None   # float32 input buffer, reused: a fresh one costs its page faults on every call
        for mlp in self.mlps:
            mlp.set_params(fold_standardisation(mlp.params(), x_mean, x_std, y_mean, y_std))

    def predict_y(self, x):
        x = np.asarray(x).reshape(-1, NIN)
        if self._z is None or self._z.shape[0] < x.shape[0]:
            self._z = np.empty((x.shape[0], NIN), np.float32)
        z = self._z[:x.shape[0]]
        np.copyto(z, x, casting="unsafe")
        z[:, 6] -= z[:, 5]
        z[:, 6] *= np.float32(1.0 / self.dnu)
        return Mlp.forward_mean(self.mlps, z) if len(self.mlps) > 1 else self.mlps[0].forward(z)

    def close(self):
        for mlp in self.mlps:
            mlp.close()

    def predict(self, x):
        """Binned C_l at raw inputs x (n, 8).  Calls on one Emulator must not overlap (a shared input buffer)."""
        y = self.predict_y(x)
        np.sinh(y, out=y)
        y *= np.float32(self.scale)
        return y


def gaussian_sigma(mean, spec, nseeds):
    """Cosmic-variance error of seed-averaged binned spectra on the full sky (Gaussian fields):
    var C_ab = (C_aa C_bb + C_ab^2) / (N_modes nseeds), N_modes = sum over the bin of (2l + 1); the auto
    spectrum of a cross pair's second channel is taken equal to the first's (neighbouring channels)."""
    nm = np.array([np.sum(2 * np.arange(lo, hi) + 1.0) for lo, hi in zip(spec.l_edges[:-1], spec.l_edges[1:])])
    auto_of = {a: i for i, (a, b) in enumerate(spec.pairs) if a == b}
    sig = np.empty_like(mean)
    for i, (a, b) in enumerate(spec.pairs):
        caa = mean[:, auto_of[a], :]
        sig[:, i, :] = np.sqrt((caa * caa + mean[:, i, :] ** 2) / (nm * nseeds))
    return sig


def train_emulator(cosmos, cls, spec, steps=4000, batch=4096, lr0=2e-3, seed=0, log=None, n_nets=1):
    """Fit asinh(C / scale) of the seed-averaged binned C_l by chi^2 (weights 1/sigma_y^2 from the
    Gaussian cosmic variance of each target), with n_nets networks (seeds seed, seed + 1, ...).  Returns
    (Emulator, per-step losses averaged over the networks)."""
    dnu = (spec.nu_hi - spec.nu_lo) / spec.n_nu
    x = encode(features(cosmos, spec), dnu)
    mean = cls.mean(axis=1)
    autos = [i for i, (a, b) in enumerate(spec.pairs) if a == b]
    scale = 1e-3 * float(np.median(mean[:, autos, :]))
    y = np.arcsinh(mean / scale)
    sig_y = gaussian_sigma(mean, spec, cls.shape[1]) / scale / np.sqrt(1.0 + (mean / scale) ** 2)
    w = 1.0 / sig_y.reshape(-1) ** 2
    w /= w.mean()
    y = y.reshape(-1)
    xm, xs = x.mean(axis=0), x.std(axis=0)
    xs[xs == 0] = 1.0
    ym, ys = y.mean(), y.std()
    mlps, losses = [], []
    for k in range(n_nets):
        mlp = Mlp(max_batch=max(batch, 65536))
        mlp.set_params(init_params(seed + k))
        mlp.set_data((x - xm) / xs, (y - ym) / ys, w)
        losses.append(mlp.train(steps, batch, lr0, steps, seed=seed + k))
        mlps.append(mlp)
        if log:
            log(f"network {k + 1}/{n_nets}: {steps} steps, loss {losses[-1][:10].mean():.3e} -> "
                f"{losses[-1][-50:].mean():.3e}")
    return Emulator(mlps, xm, xs, ym, ys, scale, dnu), np.mean(losses, axis=0)


def init_params(seed):
    """He-uniform weights and zero biases, in the parameter order of include/crime_mlp.h."""
    rng = np.random.default_rng(seed)
    wd, nin = 256, NIN
    parts = [rng.uniform(-1, 1, (nin, wd)) * np.sqrt(6.0 / nin), np.zeros(wd)]
    for _ in range(4):
        parts += [rng.uniform(-1, 1, (wd, wd)) * np.sqrt(6.0 / wd), np.zeros(wd)]
    parts += [rng.uniform(-1, 1, wd) * np.sqrt(3.0 / wd), np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def evaluate(emu, cosmos_train, cls_train, cosmos_test, cls_test, spec):
    """err/sigma of the emulator and of the nearest-training-cosmology lookup on held-out cosmologies.

    truth = mean of the held-out cosmology's nseeds mocks.  sigma = cosmic variance of one realisation:
    the Gaussian sigma of each binned spectrum (gaussian_sigma with one realisation) times a factor per
    (pair, l bin) that calibrates it on the held-out mocks, pooled over all test cosmologies
    (sum of sample variances / sum of Gaussian variances; non-Gaussianity of the lognormal field raises
    it above 1).  The truth itself carries sigma / sqrt(nseeds) of noise; the debiased RMS subtracts
    it in quadrature."""
    nseed = cls_test.shape[1]
    truth = cls_test.mean(axis=1)
    sig_g = gaussian_sigma(truth, spec, 1)
    calib = np.sqrt(cls_test.var(axis=1, ddof=1).sum(axis=0) / (sig_g ** 2).sum(axis=0))
    sigma = sig_g * calib[None]
    pred = emu.predict(features(cosmos_test, spec)).reshape(truth.shape)
    z = (pred - truth) / sigma
    # no-emulator method: the C_l of the nearest training cosmology (standardised parameter distance)
    span = RANGES[:, 1] - RANGES[:, 0]
    d = (((np.asarray(cosmos_test)[:, None, :] - np.asarray(cosmos_train)[None, :, :]) / span) ** 2).sum(axis=2)
    near = cls_train.mean(axis=1)[d.argmin(axis=1)]
    zn = (near - truth) / sigma

    def summary(zz):
        rms = float(np.sqrt(np.mean(zz ** 2)))
        return {"rms": rms, "rms_debiased": float(np.sqrt(max(0.0, rms ** 2 - 1.0 / nseed))),
                "median_abs": float(np.median(np.abs(zz))), "p95_abs": float(np.percentile(np.abs(zz), 95))}

    return {"emulator": summary(z), "nearest_cosmology": summary(zn), "truth_noise_sigma": 1.0 / math.sqrt(nseed),
            "sigma_calibration": {"min": float(calib.min()), "median": float(np.median(calib)),
                                  "max": float(calib.max())}}
