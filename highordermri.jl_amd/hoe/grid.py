"""Grid and solid-harmonic basis: host-side utilities of the operator API.

``Grid`` mirrors ``Grid(nX, nY, nZ, Δx, Δy, Δz; exchange_xy, reverse_*)``
(HighOrderMRI.jl src/Grid/Grid.jl:65-96): axis centres at
``(i - (N+1)/2) * Δ`` computed in Float64 with the T-typed spacing, then cast
to T; x varies fastest. ``basisfunc_spha`` mirrors
src/SphericalHarmonics/basisfunction.jl:27-126 (terms 1..36, evaluated in the
precision of the inputs).
"""
from __future__ import annotations

import numpy as np


class Grid:
    def __init__(self, nX: int, nY: int, nZ: int, dx, dy, dz, *, dtype=np.float32,
                 exchange_xy: bool = False, reverse_x: bool = False, reverse_y: bool = False,
                 reverse_z: bool = False):
        T = np.dtype(dtype).type
        self.nX, self.nY, self.nZ = int(nX), int(nY), int(nZ)
        self.dx, self.dy, self.dz = T(dx), T(dy), T(dz)
        self.dtype = np.dtype(dtype)
        i = np.arange(1, self.nX + 1, dtype=np.float64)
        j = np.arange(1, self.nY + 1, dtype=np.float64)
        k = np.arange(1, self.nZ + 1, dtype=np.float64)
        # x fastest: flattened index = ix + nX*(iy + nY*iz)
        X = np.broadcast_to(i[None, None, :], (self.nZ, self.nY, self.nX)).ravel()
        Y = np.broadcast_to(j[None, :, None], (self.nZ, self.nY, self.nX)).ravel()
        Z = np.broadcast_to(k[:, None, None], (self.nZ, self.nY, self.nX)).ravel()
        x = (X - (self.nX + 1) / 2) * np.float64(self.dx)
        y = (Y - (self.nY + 1) / 2) * np.float64(self.dy)
        z = (Z - (self.nZ + 1) / 2) * np.float64(self.dz)
        if exchange_xy:
            x, y = y, x
        if reverse_x:
            x = x[::-1]
        if reverse_y:
            y = y[::-1]
        if reverse_z:
            z = z[::-1]
        self.x = np.ascontiguousarray(x.astype(self.dtype))
        self.y = np.ascontiguousarray(y.astype(self.dtype))
        self.z = np.ascontiguousarray(z.astype(self.dtype))

    @property
    def matrixSize(self):
        return (self.nX, self.nY, self.nZ)

    @property
    def resolution(self):
        return (self.dx, self.dy, self.dz)

    @property
    def n(self) -> int:
        return self.nX * self.nY * self.nZ

    def __repr__(self):
        return (f"Grid [ MatrixSize: {self.nX} x {self.nY} x {self.nZ}, Resolution: "
                f"{self.dx*1e3:.2f} x {self.dy*1e3:.2f} x {self.dz*1e3:.2f} mm³ ]")


def _term(t: int, x, y, z):
    r2 = x * x + y * y + z * z
    if t == 1:
        return np.ones_like(x)
    if t == 2:
        return x
    if t == 3:
        return y
    if t == 4:
        return z
    if t == 5:
        return x * y
    if t == 6:
        return z * y
    if t == 7:
        return 3 * z ** 2 - r2
    if t == 8:
        return x * z
    if t == 9:
        return x ** 2 - y ** 2
    if t == 10:
        return 3 * y * x ** 2 - y ** 3
    if t == 11:
        return x * z * y
    if t == 12:
        return (5 * z ** 2 - r2) * y
    if t == 13:
        return 5 * z ** 3 - 3 * z * r2
    if t == 14:
        return (5 * z ** 2 - r2) * x
    if t == 15:
        return x ** 2 * z - y ** 2 * z
    if t == 16:
        return x ** 3 - 3 * x * y ** 2
    if t == 17:
        return (x * y) * (x ** 2 - y ** 2)
    if t == 18:
        return (y * z) * (3 * x ** 2 - y ** 2)
    if t == 19:
        return (x * y) * (7 * z ** 2 - r2)
    if t == 20:
        return y * (7 * z ** 3 - 3 * z * r2)
    if t == 21:
        return 35 * z ** 4 - 30 * z ** 2 * r2 + 3 * r2 ** 2
    if t == 22:
        return x * (7 * z ** 3 - 3 * z * r2)
    if t == 23:
        return (x ** 2 - y ** 2) * (7 * z ** 2 - r2)
    if t == 24:
        return x * (x ** 2 - 3 * y ** 2) * z
    if t == 25:
        return x ** 2 * (x ** 2 - 3 * y ** 2) - y ** 2 * (3 * x ** 2 - y ** 2)
    if t == 26:
        return y * (5 * x ** 4 - 10 * x ** 2 * y ** 2 + y ** 4)
    if t == 27:
        return x * y * z * (x ** 2 - y ** 2)
    if t == 28:
        return y * (r2 - 9 * z ** 2) * (y ** 2 - 3 * x ** 2)
    if t == 29:
        return x * y * z * (3 * z ** 2 - r2)
    if t == 30:
        return y * (r2 ** 2 - 14 * r2 * z ** 2 + 21 * z ** 4)
    if t == 31:
        return 63 * z ** 5 - 70 * z ** 3 * r2 + 15 * z * r2 ** 2
    if t == 32:
        return x * (r2 ** 2 - 14 * r2 * z ** 2 + 21 * z ** 4)
    if t == 33:
        return z * (3 * z ** 2 - r2) * (x ** 2 - y ** 2)
    if t == 34:
        return x * (r2 - 9 * z ** 2) * (3 * y ** 2 - x ** 2)
    if t == 35:
        return z * (x ** 4 - 6 * x ** 2 * y ** 2 + y ** 4)
    if t == 36:
        return x ** 5 - 10 * x ** 3 * y ** 2 + 5 * x * y ** 4
    raise ValueError("Basis function undefined.")


def basisfunc_spha(x, y, z, terms):
    """Unnormalised real solid harmonics; ``terms`` is an int or a sequence of
    one-based term indices. Returns ``(len(x), len(terms))`` (or a vector for
    an int), in the dtype of ``x``."""
    x, y, z = np.asarray(x), np.asarray(y), np.asarray(z)
    scalar = np.isscalar(terms) or np.ndim(terms) == 0
    tl = [int(terms)] if scalar else [int(t) for t in terms]
    out = np.empty((x.shape[0], len(tl)), dtype=x.dtype)
    for j, t in enumerate(tl):
        out[:, j] = _term(t, x, y, z)
    return out[:, 0] if scalar else out
