"""Small test-side helpers: grid coordinates (Grid.cc:104-132), the
single-wavenumber test field (test/SingleWavenumber.h:30-59 of the reference)
and index maps of the Scalar/Flux layouts."""
import math

import numpy as np


class G:
    """Grid with zero shift (Grid.cc:64-112)."""

    def __init__(self, nx, ny, ngrid, length, xoff, yoff):
        self.nx, self.ny, self.ngrid = nx, ny, ngrid
        self.length, self.xoff, self.yoff = length, xoff, yoff
        self.dx = length / nx

    def dxl(self, lev):
        return self.dx * (1 << lev)

    def xoffl(self, lev):
        return self.xoff + 0.5 * (0 - 1) * ((1 << lev) - 1) * (self.nx * self.dx)

    def yoffl(self, lev):
        return self.yoff + 0.5 * (0 - 1) * ((1 << lev) - 1) * (self.ny * self.dx)

    def xedge(self, lev, i):
        return self.xoffl(lev) + i * self.dxl(lev)

    def yedge(self, lev, j):
        return self.yoffl(lev) + j * self.dxl(lev)

    def xcenter(self, lev, i):
        return self.xoffl(lev) + (i + 0.5) * self.dxl(lev)

    def ycenter(self, lev, j):
        return self.yoffl(lev) + (j + 0.5) * self.dxl(lev)

    @property
    def ns(self):
        return (self.nx - 1) * (self.ny - 1)

    @property
    def nf(self):
        return 2 * self.nx * self.ny + self.nx + self.ny

    @property
    def nfx(self):
        return (self.nx + 1) * self.ny

    def si(self, lev, i, j):
        return lev * self.ns + (i - 1) * (self.ny - 1) + (j - 1)

    def fxi(self, lev, i, j):
        return lev * self.nf + i * self.ny + j

    def fyi(self, lev, i, j):
        return lev * self.nf + self.nfx + i * (self.ny + 1) + j

    def scalar(self, fn):
        out = np.zeros(self.ngrid * self.ns)
        for lev in range(self.ngrid):
            for i in range(1, self.nx):
                x = self.xedge(lev, i)
                for j in range(1, self.ny):
                    out[self.si(lev, i, j)] = fn(x, self.yedge(lev, j))
        return out

    def single_wavenumber(self, kx, ky):
        x0 = self.xedge(self.ngrid - 1, 0)
        y0 = self.yedge(self.ngrid - 1, 0)
        Lx = self.xedge(self.ngrid - 1, self.nx) - x0
        Ly = self.yedge(self.ngrid - 1, self.ny) - y0
        kxp = kx * math.pi / Lx
        kyp = ky * math.pi / Ly
        return self.scalar(lambda x, y: math.sin(kxp * (x - x0)) * math.sin(kyp * (y - y0)))


def single_wavenumber_2d(kx, ky, nx, ny):
    """InitializeSingleWavenumber on a bare (nx-1)x(ny-1) array (SingleWavenumber.h:44-59)."""
    pi = 4 * math.atan(1.0)
    dx = pi / nx
    dy = pi / ny
    f = np.zeros((nx - 1, ny - 1))
    for i in range(1, nx):
        for j in range(1, ny):
            f[i - 1, j - 1] = math.sin(kx * i * dx) * math.sin(ky * j * dy)
    return f


def bc_ring(nx, ny, left, top, right, bottom):
    """Builds a BC ring (BC.h:22-79) from per-side callables."""
    bc = np.zeros(2 * (nx + ny))
    for i in range(nx + 1):
        bc[ny + i] = top(i)
    for j in range(ny + 1):
        bc[2 * ny + nx - j] = right(j)
    for i in range(1, nx):
        bc[2 * (nx + ny) - i] = bottom(i)
    for j in range(ny + 1):
        bc[j] = left(j)
    return bc


CYLINDER_157 = "body Cylinder\n  circle_n 0 0 0.5 157\nend\n"
CYLINDER_314 = "body Cylinder\n  circle_n 0 0 0.5 314\nend\n"


def plate_geom(npts=100, length=1.0, pitch_amp=0.2, plunge_amp=0.2, freq=0.3, center=(0.0, 0.0)):
    """Flat plate centred at mid-chord, pitching and plunging (PitchPlunge.h:30-58)."""
    cx, cy = center
    return ("body Plate\n  line_n %.17g %.17g %.17g %.17g %d\n  center %.17g %.17g\n"
            "  motion pitchplunge %.17g %.17g 0 %.17g %.17g 0\nend\n"
            % (cx - length / 2, cy, cx + length / 2, cy, npts, cx, cy,
               pitch_amp, freq, plunge_amp, freq))
