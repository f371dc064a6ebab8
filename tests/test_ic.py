"""Host initial-condition generators (CPU).

The band-limited field is pinned three ways: its RMS over interior faces
(the normalisation), a brute-force evaluation of the mode sum at sample
faces against the separable evaluation (a transposed or mis-indexed
partial sum fails it), and its spectrum along axes sampled periodically
(no energy above kmax).  Taylor-Green is checked against its closed form
and the noise against its bound.
"""
import numpy as np
import pytest

from aeroflux_amd import ic


def interior(f, name, nx, ny, nz, g):
    xr, yr, zr = ic.face_ranges(nx, ny, nz, g)[name]
    return f[name][zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop]


@pytest.mark.parametrize("shape", [(16, 12, 10, 1), (24, 16, 20, 2), (20, 20, 20, 1)])
def test_band_limited_rms_and_zero_outside(shape):
    nx, ny, nz, g = shape
    f = ic.band_limited(nx, ny, nz, g, 1.0 / max(nx, ny, nz), seed=3)
    ss = sum(float(np.sum(interior(f, k, nx, ny, nz, g) ** 2)) for k in "UVW")
    cnt = sum(interior(f, k, nx, ny, nz, g).size for k in "UVW")
    assert abs(np.sqrt(ss / cnt) - 1.0) < 1e-13
    for k in "UVW":
        outside = f[k].copy()
        xr, yr, zr = ic.face_ranges(nx, ny, nz, g)[k]
        outside[zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop] = 0
        assert not outside.any()


def test_band_limited_matches_direct_mode_sum():
    nx, ny, nz, g, kmax, seed = 12, 10, 9, 1, 3, 7
    h = 1.0 / 12
    f = ic.band_limited(nx, ny, nz, g, h, seed=seed, kmax=kmax)
    # Unnormalised field by brute force over every mode, at sample faces.
    ks = np.arange(-kmax, kmax + 1)
    nk = len(ks)
    L = np.array([nx * h, ny * h, nz * h])
    rng = np.random.default_rng(0)
    ratios = []
    for ci, name in enumerate("UVW"):
        phase = 2 * np.pi * ic.uniform(seed, nk ** 3, offset=ci * nk ** 3)
        xr, yr, zr = ic.face_ranges(nx, ny, nz, g)[name]
        off = {"U": (0.0, 0.5, 0.5), "V": (0.5, 0.0, 0.5), "W": (0.5, 0.5, 0.0)}[name]
        for _ in range(6):
            i, j, k = rng.integers(len(xr)), rng.integers(len(yr)), rng.integers(len(zr))
            pos = np.array([(i + off[0]) * h, (j + off[1]) * h, (k + off[2]) * h])
            total = 0.0
            m = 0
            for kx in ks:
                for ky in ks:
                    for kz in ks:
                        kk = np.sqrt(kx * kx + ky * ky + kz * kz)
                        if 1 <= kk <= kmax:
                            arg = 2 * np.pi * (kx * pos[0] / L[0] + ky * pos[1] / L[1] +
                                               kz * pos[2] / L[2])
                            total += kk ** (-5.0 / 3.0) * np.cos(arg + phase[m])
                        m += 1
            val = f[name][zr.start + k, yr.start + j, xr.start + i]
            ratios.append(val / total)
    # One common normalisation factor for every sample of every component.
    ratios = np.array(ratios)
    assert np.all(np.abs(ratios / ratios[0] - 1) < 1e-12)


def test_band_limited_spectrum_is_band_limited():
    n, g, kmax = 32, 1, 8
    f = ic.band_limited(n, n, n, g, 1.0 / n, seed=11, kmax=kmax)
    # W at cell centres in x and y (periodic samples); its own axis has
    # n+1 faces, of which the first n are one period.
    w = interior(f, "W", n, n, n, g)[:n]
    spec = np.abs(np.fft.fftn(w))
    k = np.fft.fftfreq(n, 1.0 / n)
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    assert spec[kk > kmax].max() < 1e-9 * spec.max()
    assert spec[(kk >= 1) & (kk <= kmax)].max() > 0.01 * spec.max()
    assert spec[0, 0, 0] < 1e-9 * spec.max()


def test_taylor_green_closed_form_and_noise():
    n, g = 16, 1
    h = 1.0 / n
    f0 = ic.taylor_green(n, n, n, g, h, seed=1, noise=0.0)
    f1 = ic.taylor_green(n, n, n, g, h, seed=1)
    # u(i, j, k) at ((i-g)h, (j-g+1/2)h, (k-g+1/2)h), each axis onto 2 pi.
    i, j, k = 5, 3, 7
    X, Y, Z = 2 * np.pi * (i - g) / n, 2 * np.pi * (j - g + 0.5) / n, 2 * np.pi * (k - g + 0.5) / n
    assert abs(f0["U"][k, j, i] - np.sin(X) * np.cos(Y) * np.cos(Z)) < 1e-15
    Xv, Yv = 2 * np.pi * (i - g + 0.5) / n, 2 * np.pi * (j - g) / n
    assert abs(f0["V"][k, j, i] + np.cos(Xv) * np.sin(Yv) * np.cos(Z)) < 1e-15
    assert not f0["W"].any()
    for name in "UVW":
        d = f1[name] - f0[name]
        assert np.abs(d).max() <= 1e-3
        inner = interior({name: d}, name, n, n, n, g)
        assert np.all(inner != 0) or name != "W"
        assert abs(inner.mean()) < 1e-4
