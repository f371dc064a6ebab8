"""Seeded initial conditions in the padded MAC layout (host, numpy).

No reference counterpart except :func:`seed_vortex` (engine_test.go:302-335);
the reference seeds fields by writing ``Mesh.Data`` views directly.  Each
generator returns ``{"U": u, "V": v, "W": w}`` float64 arrays shaped
``(NZ, NY, NX)`` of the padded U/V/W fields (grid.go:116-123), zero outside
the interior faces; callers upload them with ``Mesh.SetData`` and apply the
velocity BC.  The same arrays feed the CPU oracle, so the trigonometric
library never enters a parity comparison.

Pseudo-random numbers: SplitMix64 (Steele, Lea, Flood 2014), state advanced
by the golden-ratio increment, output mixed by the published two-multiply
finaliser; a uniform double is ``(z >> 11) * 2**-53``.

Interior faces (the faces a sweep writes, stencil.go:54-68):
U: x in [g, g+NX], y in [g, g+NY), z in [g, g+NZ); V and W likewise with the
extra face on their own axis.  Face/centre coordinates are physical MAC
positions: u(i,j,k) sits at ((i-g)H, (j-g+1/2)H, (k-g+1/2)H).
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Outputs offset+1 .. offset+n of SplitMix64 started at `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Doubles in [0, 1) from SplitMix64."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def shapes(nx: int, ny: int, nz: int, g: int) -> dict:
    cx, cy, cz = nx + 2 * g, ny + 2 * g, nz + 2 * g
    return {"U": (cz, cy, cx + 1), "V": (cz, cy + 1, cx), "W": (cz + 1, cy, cx)}


def face_ranges(nx: int, ny: int, nz: int, g: int) -> dict:
    """(x range, y range, z range) of each component's interior faces."""
    return {
        "U": (range(g, g + nx + 1), range(g, g + ny), range(g, g + nz)),
        "V": (range(g, g + nx), range(g, g + ny + 1), range(g, g + nz)),
        "W": (range(g, g + nx), range(g, g + ny), range(g, g + nz + 1)),
    }


def _positions(name: str, nx: int, ny: int, nz: int, g: int, h: float):
    """Physical coordinates (x, y, z) of the interior faces of one component."""
    xr, yr, zr = face_ranges(nx, ny, nz, g)[name]
    off = {"U": (0.0, 0.5, 0.5), "V": (0.5, 0.0, 0.5), "W": (0.5, 0.5, 0.0)}[name]
    x = (np.arange(xr.start, xr.stop) - g + off[0]) * h
    y = (np.arange(yr.start, yr.stop) - g + off[1]) * h
    z = (np.arange(zr.start, zr.stop) - g + off[2]) * h
    return x, y, z


def _add_noise(fields: dict, nx, ny, nz, g, seed: int, amp: float) -> None:
    """U(-amp, amp) on interior faces in (field U,V,W; z; y; x) order."""
    offset = 0
    for name in ("U", "V", "W"):
        xr, yr, zr = face_ranges(nx, ny, nz, g)[name]
        n = len(xr) * len(yr) * len(zr)
        r = uniform(seed, n, offset).reshape(len(zr), len(yr), len(xr))
        offset += n
        fields[name][zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop] += (2.0 * r - 1.0) * amp


def taylor_green(nx: int, ny: int, nz: int, g: int, h: float, seed: int = 1,
                 amp: float = 1.0, noise: float = 1e-3) -> dict:
    """u = A sin X cos Y cos Z, v = -A cos X sin Y cos Z, w = 0, with each
    axis's extent N*H mapped onto [0, 2pi], plus U(-noise, noise)."""
    out = {k: np.zeros(s) for k, s in shapes(nx, ny, nz, g).items()}
    L = (nx * h, ny * h, nz * h)
    for name in ("U", "V"):
        x, y, z = _positions(name, nx, ny, nz, g, h)
        X = 2 * np.pi * x / L[0]
        Y = 2 * np.pi * y / L[1]
        Z = 2 * np.pi * z / L[2]
        if name == "U":
            val = amp * (np.cos(Z)[:, None, None] * np.cos(Y)[None, :, None] * np.sin(X)[None, None, :])
        else:
            val = -amp * (np.cos(Z)[:, None, None] * np.sin(Y)[None, :, None] * np.cos(X)[None, None, :])
        xr, yr, zr = face_ranges(nx, ny, nz, g)[name]
        out[name][zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop] = val
    if noise:
        _add_noise(out, nx, ny, nz, g, seed, noise)
    return out


def band_limited(nx: int, ny: int, nz: int, g: int, h: float, seed: int = 1,
                 kmax: int = 8, slope: float = -5.0 / 3.0, rms: float = 1.0) -> dict:
    """Sum over integer wavevectors 1 <= |k| <= kmax of |k|^slope
    cos(2 pi k.x/L + phi), random phases per (component, k), each
    component evaluated at its own face positions; then scaled so the mean
    square over all interior faces of U, V, W is rms**2.  Evaluated
    separably (three partial sums), O(17 N^3) per component."""
    ks = np.arange(-kmax, kmax + 1)
    KX, KY, KZ = np.meshgrid(ks, ks, ks, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    mask = (kk >= 1) & (kk <= kmax)
    amp = np.zeros_like(kk)
    amp[mask] = kk[mask] ** slope
    n_modes = amp.size
    L = (nx * h, ny * h, nz * h)
    out = {k: np.zeros(s) for k, s in shapes(nx, ny, nz, g).items()}
    for ci, name in enumerate(("U", "V", "W")):
        phase = 2 * np.pi * uniform(seed, n_modes, offset=ci * n_modes).reshape(amp.shape)
        coef = amp * np.exp(1j * phase)  # (kx, ky, kz)
        x, y, z = _positions(name, nx, ny, nz, g, h)
        ex = np.exp(1j * 2 * np.pi * np.outer(ks, x) / L[0])  # (k, x)
        ey = np.exp(1j * 2 * np.pi * np.outer(ks, y) / L[1])
        ez = np.exp(1j * 2 * np.pi * np.outer(ks, z) / L[2])
        t1 = np.tensordot(coef, ez, axes=([2], [0]))        # (kx, ky, z)
        t2 = np.tensordot(t1, ey, axes=([1], [0]))          # (kx, z, y)
        val = np.real(np.tensordot(t2, ex, axes=([0], [0])))  # (z, y, x)
        xr, yr, zr = face_ranges(nx, ny, nz, g)[name]
        out[name][zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop] = val
    ss, cnt = 0.0, 0
    for name in ("U", "V", "W"):
        xr, yr, zr = face_ranges(nx, ny, nz, g)[name]
        blk = out[name][zr.start:zr.stop, yr.start:yr.stop, xr.start:xr.stop]
        ss += float(np.sum(blk * blk))
        cnt += blk.size
    scale = rms / np.sqrt(ss / cnt)
    for name in out:
        out[name] *= scale
    return out


def seed_vortex(nx: int, ny: int, nz: int, g: int, amp: float = 1.0) -> dict:
    """seedVortex, engine_test.go:302-335: index coordinates, no stagger,
    w = 0.25*amp*sin(2 pi (z-g)/NZ) carrying divergence.  The caller applies
    the velocity BC afterwards, as the reference does (:334)."""
    out = {k: np.zeros(s) for k, s in shapes(nx, ny, nz, g).items()}
    fx, fy, fz = 2 * np.pi / nx, 2 * np.pi / ny, 2 * np.pi / nz
    xu = np.arange(g, g + nx + 1) - g
    yu = np.arange(g, g + ny) - g
    out["U"][g:g + nz, g:g + ny, g:g + nx + 1] = (
        amp * np.sin(fx * xu)[None, None, :] * np.cos(fy * yu)[None, :, None]
        * np.ones((nz, 1, 1)))
    xv = np.arange(g, g + nx) - g
    yv = np.arange(g, g + ny + 1) - g
    out["V"][g:g + nz, g:g + ny + 1, g:g + nx] = (
        -amp * np.cos(fx * xv)[None, None, :] * np.sin(fy * yv)[None, :, None]
        * np.ones((nz, 1, 1)))
    zw = np.arange(g, g + nz + 1) - g
    out["W"][g:g + nz + 1, g:g + ny, g:g + nx] = (
        0.25 * amp * np.sin(fz * zw)[:, None, None] * np.ones((1, ny, nx)))
    return out


def shear_layer(nx: int, ny: int, nz: int, g: int) -> dict:
    """The Example's shear layer (example_test.go:33-47): u = -1 below
    y = g + NY/2, +1 above, on U's interior faces; V = W = 0."""
    out = {k: np.zeros(s) for k, s in shapes(nx, ny, nz, g).items()}
    for y in range(g, g + ny):
        val = -1.0 if y < g + ny // 2 else 1.0
        out["U"][g:g + nz, y, g:g + nx + 1] = val
    return out
