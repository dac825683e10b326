"""Checkers of the P(k) estimator: the estimate from numpy's FFT of a real grid, and the ensemble
expectation of each shell from the input P(k) (lattice enumeration).

Conventions: delta(x) on an N^3 grid of side L, P_est(k) = |sum_x delta(x) e^{-i k.x}|^2 L^3 / N^6,
averaged over the grid modes whose |k| (in units of the fundamental dk = 2 pi / L) falls in
[i, i+1) * bin_dk / dk.  The fill draws |delta_k|^2 with mean P(k)/dk^3 and scales by
(sqrt(2 pi)/L)^3 (src/fourier.c:293-296, 403), so <P_est> = P(k) off the kx = 0 and kx = N/2 planes,
where the real inverse transform keeps the Hermitian part of the modes (half the power).
"""
import numpy as np

import oracle


def pk_numpy(dens, l_box, nbins, bin_dk):
    n = dens.shape[0]
    dk = 2 * np.pi / l_box
    f = np.fft.fftn(dens)
    p = (f.real ** 2 + f.imag ** 2) * (l_box ** 3 / float(n) ** 6)
    a = np.abs(np.fft.fftfreq(n, 1.0 / n)).astype(np.int64)
    m = a[:, None, None] ** 2 + a[None, :, None] ** 2 + a[None, None, :] ** 2
    kk = np.sqrt(m.astype(np.float64))
    b = np.floor(kk / (bin_dk / dk)).astype(np.int64)
    sel = b < nbins
    cnt = np.bincount(b[sel], minlength=nbins)
    ps = np.bincount(b[sel], weights=p[sel], minlength=nbins)
    ks = np.bincount(b[sel], weights=kk[sel], minlength=nbins)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(cnt > 0, ks / cnt * dk, 0.0), np.where(cnt > 0, ps / cnt, 0.0), cnt


def pk_expect(desc, nbins, bin_dk):
    """Mode count and <P_est> per shell, by enumerating every lattice mode with |k| below the last
    shell edge (the edge must not pass the Nyquist k N/2 dk)."""
    n = desc.n_grid
    dk = 2 * np.pi / desc.l_box
    binw = bin_dk / dk
    r = int(np.ceil(nbins * binw))
    assert r <= n // 2, "shells past the Nyquist k need the full grid"
    ax = np.arange(-r, r + 1)
    m = ax[:, None, None] ** 2 + ax[None, :, None] ** 2 + ax[None, None, :] ** 2
    kk = np.sqrt(m.astype(np.float64))
    b = np.floor(kk / binw).astype(np.int64)
    half = (ax == 0) | (np.abs(ax) == n // 2)
    w = np.where(half[None, None, :], 0.5, 1.0) * np.ones_like(kk)
    mu = np.unique(m[(b < nbins) & (m > 0)])
    pm = {}
    for v in mu:
        k2 = float(v) * dk * dk
        p = oracle.pk_linear0(desc, 0.5 * np.log10(k2))
        if desc.do_smoothing:
            p *= np.exp(-desc.r2_smooth * k2)
        pm[int(v)] = p
    lut = np.zeros(int(mu.max()) + 1 if len(mu) else 1)
    for v, p in pm.items():
        lut[v] = p
    sel = (b < nbins)
    e = np.where(m > 0, lut[np.minimum(m, len(lut) - 1)], 0.0) * w
    cnt = np.bincount(b[sel], minlength=nbins)
    es = np.bincount(b[sel], weights=e[sel], minlength=nbins)
    with np.errstate(invalid="ignore", divide="ignore"):
        return cnt, np.where(cnt > 0, es / cnt, 0.0)


def zscores(pk, cnt, expect, first=1):
    """(P_est - <P>) / sigma per shell, sigma = <P> sqrt(2 / count) (the k and -k modes of a real
    field are equal, so a shell holds count/2 independent exponential variates)."""
    sl = slice(first, None)
    return (pk[sl] - expect[sl]) / (expect[sl] * np.sqrt(2.0 / cnt[sl]))


def pk_model_grid(desc):
    """Input P(k) (with the smoothing kernel when it is on) at every FFT-ordered mode of the grid
    (0 at k = 0); small grids only."""
    n = desc.n_grid
    dk = 2 * np.pi / desc.l_box
    a = np.abs(np.fft.fftfreq(n, 1.0 / n)).astype(np.int64)
    m = a[:, None, None] ** 2 + a[None, :, None] ** 2 + a[None, None, :] ** 2
    mu, inv = np.unique(m, return_inverse=True)
    vals = np.zeros(len(mu))
    for j, v in enumerate(mu):
        if v == 0:
            continue
        k2 = float(v) * dk * dk
        vals[j] = oracle.pk_linear0(desc, 0.5 * np.log10(k2))
        if desc.do_smoothing:
            vals[j] *= np.exp(-desc.r2_smooth * k2)
    return vals[inv.reshape(m.shape)]
