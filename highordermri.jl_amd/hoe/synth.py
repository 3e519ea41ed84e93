"""Seeded synthetic workloads for the Explicit operator (numpy only).

Five workloads, indexed 0..4:
  0  2-D 64x64, FOV 220 mm, 8 coils, single-shot Archimedean spiral, 6000 samples at 2 us;
  1  2-D 200x200, FOV 200 mm, 32 coils, spiral R=4, 16000 samples at 1 us, dB0 within +-150 Hz;
  2  2-D 200x200, 32 coils, single-shot EPI R=4 (50 lines x 400 samples at 1 us), third-order
     terms plus concomitant terms at B0 = 7 T with 40 mT/m gradients, off-centre slice;
  3  as 1 (the differentiation workload perturbs it; see ``perturb``);
  4  3-D 96x96x48, FOV 220x220x110 mm, 32 coils, stack of 48 spirals x 8000 samples, dB0 +-200 Hz.

Coefficient rows follow the reference's term order (basisfunction.jl:47-78):
row 1 (order 0, cycles): a random walk clipped to +-0.5 rad; rows 2-4 (order 1,
cycles/m): the k-space trajectory; rows 5-16 (orders 2-3, cycles/m^p): smooth
seeded traces whose phase at the FOV edge is 1-5 % of the first-order phase.
Images are Shepp-Logan magnitudes (Toft's modified parameters) with a smooth
random phase; coil maps are Gaussian profiles on a ring with a linear phase.
"""
from __future__ import annotations

import numpy as np

from .grid import Grid

GAMMA_BAR = 42.577478e6      # Hz/T (1H)

# Modified Shepp-Logan ellipses (Toft 1996): intensity, a, b, x0, y0, phi(deg)
_SL = np.array([
    [1.0, 0.69, 0.92, 0.0, 0.0, 0.0],
    [-0.8, 0.6624, 0.8740, 0.0, -0.0184, 0.0],
    [-0.2, 0.1100, 0.3100, 0.22, 0.0, -18.0],
    [-0.2, 0.1600, 0.4100, -0.22, 0.0, 18.0],
    [0.1, 0.2100, 0.2500, 0.0, 0.35, 0.0],
    [0.1, 0.0460, 0.0460, 0.0, 0.1, 0.0],
    [0.1, 0.0460, 0.0460, 0.0, -0.1, 0.0],
    [0.1, 0.0460, 0.0230, -0.08, -0.605, 0.0],
    [0.1, 0.0230, 0.0230, 0.0, -0.606, 0.0],
    [0.1, 0.0230, 0.0460, 0.06, -0.605, 0.0],
])


def shepp_logan(u, v, w=None):
    """Phantom magnitude at normalised coordinates u, v (and w for 3-D) in [-1, 1]."""
    out = np.zeros_like(u, dtype=np.float64)
    for amp, a, b, x0, y0, phi in _SL:
        t = np.deg2rad(phi)
        xr = (u - x0) * np.cos(t) + (v - y0) * np.sin(t)
        yr = -(u - x0) * np.sin(t) + (v - y0) * np.cos(t)
        r = (xr / a) ** 2 + (yr / b) ** 2
        if w is not None:
            r = r + (w / (0.9 * max(a, b))) ** 2
        out += amp * (r <= 1.0)
    return out


def _smooth_traces(rng, n_rows, n_sam, n_cos=4):
    t = np.linspace(0.0, 1.0, n_sam)
    out = np.zeros((n_rows, n_sam))
    for r in range(n_rows):
        for k in range(1, n_cos + 1):
            out[r] += rng.normal() / k * np.cos(np.pi * k * t + rng.uniform(0, 2 * np.pi))
        out[r] -= out[r, 0] * 0.0
    return out


def spiral(n_sam, kmax, turns, rng=None):
    """Archimedean spiral k(s) in cycles/m with (approximately) uniform arc-length sampling."""
    s = np.arange(n_sam) / max(n_sam - 1, 1)
    theta = 2 * np.pi * turns * np.sqrt(s)
    r = kmax * theta / (2 * np.pi * turns)
    return r * np.cos(theta), r * np.sin(theta)


def coefficient_rows(rng, k1, fov_half, n_sam):
    """16 x n_sam coefficient traces: order 0 random walk, order 1 = k1 (3 rows), orders 2-3 smooth."""
    kspha = np.zeros((16, n_sam))
    walk = np.cumsum(rng.normal(scale=0.02, size=n_sam))
    kspha[0] = np.clip(walk, -0.5, 0.5) / (2 * np.pi)       # rad -> cycles
    kspha[1:4] = k1
    first_phase = np.max(np.abs(k1)) * fov_half + 1e-30
    orders = [2] * 5 + [3] * 7                               # rows 5-9 order 2, rows 10-16 order 3
    sm = _smooth_traces(rng, 12, n_sam)
    for j, p in enumerate(orders):
        frac = rng.uniform(0.01, 0.05)
        tr = sm[j] / (np.max(np.abs(sm[j])) + 1e-30)
        kspha[4 + j] = tr * frac * first_phase / fov_half ** p
    return kspha


def gaussian_coils(grid: Grid, n_cha, rng, sigma, ring_radius, z_spread=0.0):
    x, y, z = grid.x.astype(np.float64), grid.y.astype(np.float64), grid.z.astype(np.float64)
    csm = np.empty((grid.n, n_cha), dtype=np.complex128)
    for c in range(n_cha):
        ang = 2 * np.pi * c / n_cha
        cx, cy = ring_radius * np.cos(ang), ring_radius * np.sin(ang)
        cz = z_spread * (((c * 7) % n_cha) / max(n_cha - 1, 1) - 0.5)
        mag = np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * sigma ** 2))
        ph = rng.uniform(-1, 1) * x / sigma + rng.uniform(-1, 1) * y / sigma
        csm[:, c] = mag * np.exp(1j * ph)
    return csm.astype(np.complex64)


def gaussian_b0(grid: Grid, rng, fmax, n=3):
    x, y, z = grid.x.astype(np.float64), grid.y.astype(np.float64), grid.z.astype(np.float64)
    ext = max(np.ptp(x), np.ptp(y), 1e-9)
    f = np.zeros(grid.n)
    for _ in range(n):
        c = rng.uniform(-0.3, 0.3, size=3) * np.array([np.ptp(x), np.ptp(y), np.ptp(z)])
        s = rng.uniform(0.15, 0.3) * ext
        f += rng.uniform(-1, 1) * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))
    f *= fmax / (np.max(np.abs(f)) + 1e-30)
    return f.astype(np.float32)


def phantom_image(grid: Grid, rng, fov):
    x, y, z = grid.x.astype(np.float64), grid.y.astype(np.float64), grid.z.astype(np.float64)
    u, v = 2 * x / fov[0], -2 * y / fov[1]
    w = None if grid.nZ == 1 else 2 * z / fov[2]
    mag = shepp_logan(v, u, w)    # phantom rows along y
    c = rng.normal(size=5)
    ph = (np.pi / 4) * np.tanh(c[0] * u + c[1] * v + 0.5 * c[2] * u * v + 0.3 * c[3] * u * u + 0.3 * c[4] * v * v)
    return (mag * np.exp(1j * ph)).astype(np.complex64)


def make(config: int, *, n_cha: int | None = None, n_sam: int | None = None, seed: int | None = None):
    """Build workload ``config``; returns a dict of operator inputs (Julia column-major semantics)."""
    rng = np.random.default_rng(1000 + config if seed is None else seed)
    out = {"config": config}
    if config == 0:
        N, fov, nc, ns, dwell, fmax, turns = 64, 0.220, 8, 6000, 2e-6, 50.0, 32
    elif config in (1, 3):
        N, fov, nc, ns, dwell, fmax, turns = 200, 0.200, 32, 16000, 1e-6, 150.0, 25
    elif config == 2:
        N, fov, nc, ns, dwell, fmax, turns = 200, 0.200, 32, 20000, 1e-6, 150.0, None
    elif config == 4:
        N, fov, nc, ns, dwell, fmax, turns = 96, 0.220, 32, 384000, 1e-6, 200.0, 24
    else:
        raise ValueError("config must be 0..4")
    if n_cha is not None:
        nc = n_cha
    if config == 4:
        grid = Grid(96, 96, 48, 0.220 / 96, 0.220 / 96, 0.110 / 48)
        fovv = (0.220, 0.220, 0.110)
    else:
        grid = Grid(N, N, 1, fov / N, fov / N, fov / N)
        fovv = (fov, fov, fov)
    if config == 2:
        # off-centre axial slice so the z^2 concomitant term is active
        grid.z[:] = np.float32(0.04)
    kmax = N / (2 * fov)
    if config == 4:
        n_part, per = 48, 8000
        kx, ky = spiral(per, kmax, turns)
        kz = (np.arange(n_part) - n_part // 2) / fovv[2]
        k1 = np.stack([np.tile(kx, n_part), np.tile(ky, n_part), np.repeat(kz, per)])
        times = np.tile(np.arange(per) * dwell, n_part)
        if n_sam is not None:
            k1, times = k1[:, :n_sam], times[:n_sam]
    elif config == 2:
        n_lines, per = 50, 400
        g_ro = 0.040
        span = GAMMA_BAR * g_ro * per * dwell
        kx_line = np.linspace(-span / 2, span / 2, per)
        kx = np.concatenate([kx_line if (l % 2 == 0) else kx_line[::-1] for l in range(n_lines)])
        ky = np.repeat((np.arange(n_lines) - n_lines / 2) * (4.0 / fov), per)
        k1 = np.stack([kx, ky, np.zeros_like(kx)])
        times = np.arange(n_lines * per) * dwell
        # concomitant (Maxwell) terms of the readout gradient, Gz = 0:
        # phase/cycles = gamma_bar * int (Gx^2+Gy^2)/(2 B0) dt * z^2
        b0 = 7.0
        gx = np.full(n_lines * per, g_ro)
        c_z2 = np.cumsum(GAMMA_BAR * gx ** 2 / (2 * b0) * dwell)
        conc = np.zeros((4, n_lines * per))
        conc[0] = c_z2
        out["concomitant"] = conc.astype(np.float32)
    else:
        kx, ky = spiral(ns, kmax, turns)
        k1 = np.stack([kx, ky, np.zeros_like(kx)])
        times = np.arange(ns) * dwell
        if n_sam is not None:
            k1, times = k1[:, :n_sam], times[:n_sam]
    ns_eff = k1.shape[1]
    kspha = coefficient_rows(rng, k1, fovv[0] / 2, ns_eff)
    out.update(
        grid=grid,
        kspha=kspha.astype(np.float32),
        times=times.astype(np.float32),
        fieldmap=gaussian_b0(grid, rng, fmax),
        csm=gaussian_coils(grid, nc, rng, sigma=fovv[0] / 3, ring_radius=0.6 * fovv[0] / 2,
                           z_spread=(fovv[2] if config == 4 else 0.0)),
        image=phantom_image(grid, rng, fovv),
        dwell=dwell,
        mask=None,
    )
    return out


DESCRIPTIONS = {
    0: "2D 64x64 FOV 220 mm, 8 coils, 6000-sample Archimedean spiral at 2 us, 16 solid-harmonic terms + dB0",
    1: "2D 200x200 FOV 200 mm, 32 coils, 16000-sample spiral R=4 at 1 us, 16 solid-harmonic terms + dB0",
    2: "2D 200x200 FOV 200 mm off-centre slice, 32 coils, single-shot EPI R=4 (50 x 400 samples at 1 us), "
       "16 terms + 4 concomitant terms (7 T, 40 mT/m) + dB0",
    3: "2D 200x200 FOV 200 mm, 32 coils, 16000-sample spiral R=4 at 1 us, 16 terms + dB0; "
       "data from perturbed coefficients and a 3 us delay",
    4: "3D 96x96x48 FOV 220x220x110 mm, 32 coils, stack of 48 spirals x 8000 samples at 1 us, "
       "16 solid-harmonic terms + dB0",
}


def describe(config: int) -> str:
    """One-line description of workload ``config``."""
    return f"workload {config}: {DESCRIPTIONS[config]}"


def perturb(cfg: dict, seed: int = 7, order1: float = 0.01, order23: float = 0.05):
    """Perturbed coefficients for the differentiation workload: seeded noise of
    ``order1`` relative amplitude on rows 2-4 and ``order23`` on rows 5-16."""
    rng = np.random.default_rng(seed)
    k = cfg["kspha"].astype(np.float64).copy()
    ns = k.shape[1]
    sm = _smooth_traces(rng, 16, ns, n_cos=6)
    for r in range(1, 16):
        amp = order1 if r < 4 else order23
        scale = np.max(np.abs(k[r])) + 1e-30
        k[r] += amp * scale * sm[r] / (np.max(np.abs(sm[r])) + 1e-30)
    return k.astype(np.float32)
