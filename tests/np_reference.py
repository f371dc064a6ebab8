"""An independent, vectorised numpy reading of the Go reference's data plane.

Written directly from AeroFlex/stencil.go, project.go and aeroflux.go (not
from the C oracle), in a different structure: whole-array slicing instead of
loops.  Within every reference loop nest the writes are independent of each
other (sweeps write Tmp fields; a colour pass reads only the other colour;
each BC loop reads cells no earlier iteration of the same loop writes), so
the vectorised form performs the same IEEE operations in the same order and
must agree with the oracle bit for bit.  Used only on tiny grids.

Arrays are indexed [z, y, x] over the padded extents.
"""
from __future__ import annotations

import numpy as np


class NpEngine:
    def __init__(self, nx, ny, nz, g, h, dt, nu, rho, iters):
        self.NX, self.NY, self.NZ, self.g, self.H = nx, ny, nz, g, h
        self.DT, self.nu, self.rho, self.iters = dt, nu, rho, iters
        cx, cy, cz = nx + 2 * g, ny + 2 * g, nz + 2 * g
        self.f = {
            "P": np.zeros((cz, cy, cx)), "U": np.zeros((cz, cy, cx + 1)),
            "V": np.zeros((cz, cy + 1, cx)), "W": np.zeros((cz + 1, cy, cx)),
            "Div": np.zeros((cz, cy, cx)), "Tmp0": np.zeros((cz, cy, cx + 1)),
            "Tmp1": np.zeros((cz, cy + 1, cx)), "Tmp2": np.zeros((cz + 1, cy, cx)),
        }

    # ---------------------------------------------------------- stencil.go
    def face_velocity(self, c, X, Y, Z):
        u, v, w = self.f["U"], self.f["V"], self.f["W"]
        if c == 0:
            vx = u[Z, Y, X]
            vy = 0.25 * (v[Z, Y, X - 1] + v[Z, Y, X] + v[Z, Y + 1, X - 1] + v[Z, Y + 1, X])
            vz = 0.25 * (w[Z, Y, X - 1] + w[Z, Y, X] + w[Z + 1, Y, X - 1] + w[Z + 1, Y, X])
        elif c == 1:
            vx = 0.25 * (u[Z, Y - 1, X] + u[Z, Y - 1, X + 1] + u[Z, Y, X] + u[Z, Y, X + 1])
            vy = v[Z, Y, X]
            vz = 0.25 * (w[Z, Y - 1, X] + w[Z, Y, X] + w[Z + 1, Y - 1, X] + w[Z + 1, Y, X])
        else:
            vx = 0.25 * (u[Z - 1, Y, X] + u[Z - 1, Y, X + 1] + u[Z, Y, X] + u[Z, Y, X + 1])
            vy = 0.25 * (v[Z - 1, Y, X] + v[Z - 1, Y + 1, X] + v[Z, Y, X] + v[Z, Y + 1, X])
            vz = w[Z, Y, X]
        return vx, vy, vz

    @staticmethod
    def sample(d, px, py, pz):
        nz, ny, nx = d.shape
        mx, my, mz = float(nx - 1), float(ny - 1), float(nz - 1)
        with np.errstate(invalid="ignore"):
            px = np.where(px < 0, 0.0, np.where(px > mx, mx, px))
            py = np.where(py < 0, 0.0, np.where(py > my, my, py))
            pz = np.where(pz < 0, 0.0, np.where(pz > mz, mz, pz))
        if np.isnan(px).any() or np.isnan(py).any() or np.isnan(pz).any():
            raise FloatingPointError("NaN departure coordinate")
        x0 = px.astype(np.int64)
        y0 = py.astype(np.int64)
        z0 = pz.astype(np.int64)
        x0 = np.where(x0 >= nx - 1, nx - 2, x0)
        y0 = np.where(y0 >= ny - 1, ny - 2, y0)
        z0 = np.where(z0 >= nz - 1, nz - 2, z0)
        tx = px - x0.astype(np.float64)
        ty = py - y0.astype(np.float64)
        tz = pz - z0.astype(np.float64)
        c00 = d[z0, y0, x0] * (1 - tx) + d[z0, y0, x0 + 1] * tx
        c10 = d[z0, y0 + 1, x0] * (1 - tx) + d[z0, y0 + 1, x0 + 1] * tx
        c01 = d[z0 + 1, y0, x0] * (1 - tx) + d[z0 + 1, y0, x0 + 1] * tx
        c11 = d[z0 + 1, y0 + 1, x0] * (1 - tx) + d[z0 + 1, y0 + 1, x0 + 1] * tx
        c0 = c00 * (1 - ty) + c10 * ty
        c1 = c01 * (1 - ty) + c11 * ty
        return c0 * (1 - tz) + c1 * tz

    def sweep(self, ax, dt):
        g = self.g
        invH = 1.0 / self.H
        lam = self.nu * dt * invH * invH
        outs = {}
        for c, (src, dst) in enumerate((("U", "Tmp0"), ("V", "Tmp1"), ("W", "Tmp2"))):
            x1 = g + self.NX + (1 if c == 0 else 0)
            y1 = g + self.NY + (1 if c == 1 else 0)
            z1 = g + self.NZ + (1 if c == 2 else 0)
            Z, Y, X = np.meshgrid(np.arange(g, z1), np.arange(g, y1), np.arange(g, x1),
                                  indexing="ij")
            vx, vy, vz = self.face_velocity(c, X, Y, Z)
            px, py, pz = X.astype(float), Y.astype(float), Z.astype(float)
            if ax == 0:
                px = px - dt * vx * invH
            elif ax == 1:
                py = py - dt * vy * invH
            else:
                pz = pz - dt * vz * invH
            s = self.f[src]
            adv = self.sample(s, px, py, pz)
            lap = np.zeros_like(adv)
            if lam != 0:
                d = [0, 0, 0]
                d[2 - ax] = 1  # array axis order is (z, y, x)
                sm = s[Z - d[0], Y - d[1], X - d[2]]
                sp = s[Z + d[0], Y + d[1], X + d[2]]
                lap = lam * (sm - 2 * s[Z, Y, X] + sp)
            outs[dst] = (Z, Y, X, adv + lap)
        for dst, (Z, Y, X, val) in outs.items():
            self.f[dst][Z, Y, X] = val

    def commit(self):
        f = self.f
        f["U"], f["Tmp0"] = f["Tmp0"], f["U"]
        f["V"], f["Tmp1"] = f["Tmp1"], f["V"]
        f["W"], f["Tmp2"] = f["Tmp2"], f["W"]

    def velocity_bc(self):
        g, NX, NY, NZ = self.g, self.NX, self.NY, self.NZ
        u, v, w = self.f["U"], self.f["V"], self.f["W"]
        y0, y1, z0, z1 = g, g + NY, g, g + NZ
        ux0, ux1 = g, g + NX
        u[z0:z1, y0:y1, ux0] = 0
        u[z0:z1, y0:y1, ux1] = 0
        v[z0:z1, y0, g:g + NX] = 0
        v[z0:z1, y1, g:g + NX] = 0
        w[z0, y0:y1, g:g + NX] = 0
        w[z1, y0:y1, g:g + NX] = 0
        u[z0:z1, y0 - 1, ux0:ux1 + 1] = -u[z0:z1, y0, ux0:ux1 + 1]
        u[z0:z1, y1, ux0:ux1 + 1] = -u[z0:z1, y1 - 1, ux0:ux1 + 1]
        u[z0 - 1, y0 - 1:y1 + 1, ux0:ux1 + 1] = -u[z0, y0 - 1:y1 + 1, ux0:ux1 + 1]
        u[z1, y0 - 1:y1 + 1, ux0:ux1 + 1] = -u[z1 - 1, y0 - 1:y1 + 1, ux0:ux1 + 1]
        v[z0:z1, y0:y1 + 1, g - 1] = -v[z0:z1, y0:y1 + 1, g]
        v[z0:z1, y0:y1 + 1, g + NX] = -v[z0:z1, y0:y1 + 1, g + NX - 1]
        v[z0 - 1, y0:y1 + 1, g - 1:g + NX + 1] = -v[z0, y0:y1 + 1, g - 1:g + NX + 1]
        v[z1, y0:y1 + 1, g - 1:g + NX + 1] = -v[z1 - 1, y0:y1 + 1, g - 1:g + NX + 1]
        w[z0:z1 + 1, y0:y1, g - 1] = -w[z0:z1 + 1, y0:y1, g]
        w[z0:z1 + 1, y0:y1, g + NX] = -w[z0:z1 + 1, y0:y1, g + NX - 1]
        w[z0:z1 + 1, y0 - 1, g - 1:g + NX + 1] = -w[z0:z1 + 1, y0, g - 1:g + NX + 1]
        w[z0:z1 + 1, y1, g - 1:g + NX + 1] = -w[z0:z1 + 1, y1 - 1, g - 1:g + NX + 1]

    # ---------------------------------------------------------- project.go
    def divergence(self):
        g, NX, NY, NZ = self.g, self.NX, self.NY, self.NZ
        u, v, w = self.f["U"], self.f["V"], self.f["W"]
        scale = self.rho / (self.DT * self.H)
        zs, ys, xs = slice(g, g + NZ), slice(g, g + NY), slice(g, g + NX)
        du = u[zs, ys, g + 1:g + NX + 1] - u[zs, ys, xs]
        dv = v[zs, g + 1:g + NY + 1, xs] - v[zs, ys, xs]
        dw = w[g + 1:g + NZ + 1, ys, xs] - w[zs, ys, xs]
        self.f["Div"][zs, ys, xs] = scale * (du + dv + dw)

    def gs_pass(self, col):
        g, NX, NY, NZ = self.g, self.NX, self.NY, self.NZ
        p, dv = self.f["P"], self.f["Div"]
        h2 = self.H * self.H
        Z, Y, X = np.meshgrid(np.arange(g, g + NZ), np.arange(g, g + NY), np.arange(g, g + NX),
                              indexing="ij")
        m = ((X + Y + Z) & 1) == col
        Z, Y, X = Z[m], Y[m], X[m]
        s = (p[Z, Y, X - 1] + p[Z, Y, X + 1] + p[Z, Y - 1, X] + p[Z, Y + 1, X]
             + p[Z - 1, Y, X] + p[Z + 1, Y, X])
        p[Z, Y, X] = (s - h2 * dv[Z, Y, X]) / 6.0

    def pressure_bc(self):
        g, p = self.g, self.f["P"]
        x1, y1, z1 = g + self.NX, g + self.NY, g + self.NZ
        p[g:z1, g:y1, g - 1] = p[g:z1, g:y1, g]
        p[g:z1, g:y1, x1] = p[g:z1, g:y1, x1 - 1]
        p[g:z1, g - 1, g - 1:x1 + 1] = p[g:z1, g, g - 1:x1 + 1]
        p[g:z1, y1, g - 1:x1 + 1] = p[g:z1, y1 - 1, g - 1:x1 + 1]
        p[g - 1, g - 1:y1 + 1, g - 1:x1 + 1] = p[g, g - 1:y1 + 1, g - 1:x1 + 1]
        p[z1, g - 1:y1 + 1, g - 1:x1 + 1] = p[z1 - 1, g - 1:y1 + 1, g - 1:x1 + 1]

    def solve(self):
        self.f["P"][...] = 0
        for _ in range(self.iters):
            self.pressure_bc()
            self.gs_pass(0)
            self.pressure_bc()
            self.gs_pass(1)
        self.pressure_bc()

    def gradient(self):
        g, NX, NY, NZ = self.g, self.NX, self.NY, self.NZ
        p, u, v, w = self.f["P"], self.f["U"], self.f["V"], self.f["W"]
        scale = self.DT / (self.rho * self.H)
        zs, ys, xs = slice(g, g + NZ), slice(g, g + NY), slice(g, g + NX)
        u[zs, ys, g:g + NX + 1] -= scale * (p[zs, ys, g:g + NX + 1] - p[zs, ys, g - 1:g + NX])
        v[zs, g:g + NY + 1, xs] -= scale * (p[zs, g:g + NY + 1, xs] - p[zs, g - 1:g + NY, xs])
        w[g:g + NZ + 1, ys, xs] -= scale * (p[g:g + NZ + 1, ys, xs] - p[g - 1:g + NZ, ys, xs])

    def project(self):
        self.velocity_bc()
        self.divergence()
        self.solve()
        self.gradient()
        self.velocity_bc()

    def step(self):
        half = self.DT * 0.5
        for ax in (0, 1, 2):
            self.velocity_bc()
            self.sweep(ax, half)
            self.commit()
        self.project()
        for ax in (2, 1, 0):
            self.velocity_bc()
            self.sweep(ax, half)
            self.commit()
        self.project()

    def max_divergence(self):
        g, NX, NY, NZ = self.g, self.NX, self.NY, self.NZ
        u, v, w = self.f["U"], self.f["V"], self.f["W"]
        zs, ys, xs = slice(g, g + NZ), slice(g, g + NY), slice(g, g + NX)
        d = (1.0 / self.H) * ((u[zs, ys, g + 1:g + NX + 1] - u[zs, ys, xs])
                              + (v[zs, g + 1:g + NY + 1, xs] - v[zs, ys, xs])
                              + (w[g + 1:g + NZ + 1, ys, xs] - w[zs, ys, xs]))
        a = np.abs(d)
        a = a[~np.isnan(a)]
        return float(max(0.0, a.max())) if a.size else 0.0
