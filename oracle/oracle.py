"""CPU oracle of HighOrderMRI.jl's Explicit encoding operator.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; never by the product.

Each function cites the reference code it restates:
  - prep_kspha / prep_concomitant: src/EncodingOperator/prep_kspha.jl:1-56, 78-98
  - masking and ordering: src/EncodingOperator/Explicit/array_operator.jl:109-138
  - FP64 operator: test/test_explicit_gpu.jl:8-35 (explicit_reference_basis/_matrix)
  - CPU Array path (cpu_ref_*): array_operator.jl:241-348
  - deterministic fixture: test/test_HighOrderExplicit.jl:3-85
  - normal weights: explicit_gpu_operator.jl:399-416; reconstruction: recon_HOOp.jl:73-81
  - derivative product B x: array_operator.jl:187-235
  - delay loop: src/Synchronization/FindDelay.jl:152-214, InterpTraj.jl:18-44
  - GIRF prediction: src/GIRF/apply_girf.jl:39-150; trajectory integration: utils/conversion.jl:14-41
The heavy loops live in hoe_oracle.c (liboracle.so, built by oracle/Makefile).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        vp, i64, ci = C.c_void_p, C.c_int64, C.c_int
        L.oracle_basis_f64.argtypes = [vp, vp, vp, i64, vp]
        L.oracle_basis_f64d.argtypes = [vp, vp, vp, i64, vp]
        L.cpu_ref_basis_f32.argtypes = [vp, vp, vp, i64, vp]
        for n in ("oracle_forward_f64", "oracle_adjoint_f64"):
            getattr(L, n).argtypes = [i64, i64, i64, ci, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_grad_f64.argtypes = [i64, i64, i64, ci, vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_forward_dt_f64.argtypes = [i64, i64, i64, ci, vp, vp, vp, vp, vp, vp, vp, vp]
        for n in ("cpu_ref_forward", "cpu_ref_adjoint"):
            getattr(L, n).argtypes = [i64, i64, i64, ci, vp, vp, vp, vp, vp, vp, vp, i64]
        L.oracle_num_threads.restype = ci
        L.oracle_set_threads.argtypes = [ci]
        _lib = L
    return _lib


def _p(a):
    """Address of a contiguous array. The caller must hold a reference to ``a``
    for the duration of the C call (never pass a temporary)."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- model prep

def prep_kspha(kspha, k_nominal, n_term, recon_terms=None):
    """prep_kspha.jl:1-56 (returns a modified copy)."""
    if recon_terms is None:
        recon_terms = "111" if n_term == 9 else "1111"
    expect = 3 if n_term == 9 else 4 if n_term == 16 else 0
    if expect == 0:
        raise ValueError(f"nTerm must be 9 or 16, got {n_term}")
    if len(recon_terms) != expect or any(c not in "01" for c in recon_terms):
        raise ValueError("bad recon_terms")
    t = [c == "1" for c in recon_terms] + ([False] if n_term == 9 else [])
    k = np.array(kspha, copy=True)
    if not t[0]:
        k[0, :] = 0
    if not t[1]:
        k[1:4, :] = k_nominal
    if not t[2]:
        k[4:9, :] = 0
    if not t[3] and n_term == 16:
        k[9:16, :] = 0
    return k


def basis_f64(x, y, z):
    """FP64 basis [n][16] from FP32 coordinates (test_explicit_gpu.jl:8-13)."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, z))
    out = np.empty((x.size, 16), dtype=np.float64)
    lib().oracle_basis_f64(_p(x), _p(y), _p(z), x.size, _p(out))
    return out


def basis_f64d(x, y, z):
    """FP64 basis [n][16] from FP64 coordinates."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    out = np.empty((x.size, 16), dtype=np.float64)
    lib().oracle_basis_f64d(_p(x), _p(y), _p(z), x.size, _p(out))
    return out


def basis_f32(x, y, z):
    """FP32 basis [n][16] in Julia Float32 operation order (basisfunction.jl:47-78)."""
    x, y, z = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, z))
    out = np.empty((x.size, 16), dtype=np.float32)
    lib().cpu_ref_basis_f32(_p(x), _p(y), _p(z), x.size, _p(out))
    return out


@dataclass
class Model:
    """Masked operator inputs; arrays follow the reference's column-major conventions."""
    nv: int
    ns: int
    nc: int
    L: int
    mask_idx: np.ndarray        # [nv] flat grid indices (column-major order)
    n_grid: int
    bf64: np.ndarray            # [nv][L]
    bf32: np.ndarray            # [nv][L]
    k64: np.ndarray             # [ns][L]
    k32: np.ndarray             # [ns][L]
    times: np.ndarray           # [ns] float32
    fieldmap: np.ndarray        # [nv] float32
    csm: np.ndarray             # [nc][nv] complex64


def make_model(grid, kspha, times, fieldmap=None, csm=None, mask=None, concomitant=None, recon_terms=None,
               k_nominal=None, f64=False):
    """array_operator.jl:102-139 (masking, prep_kspha, basisfunc_spha, prep_concomitant).
    ``f64``: the model of a Float64 operator (Grid{Float64}, Float64 coefficients, times and
    maps, ComplexF64 coil maps): inputs kept in FP64 and the basis evaluated from the FP64
    coordinates; otherwise the inputs are Float32, as for the ComplexF32 operators."""
    RT = np.float64 if f64 else np.float32
    kspha = np.asarray(kspha, dtype=RT)
    n_term, ns = kspha.shape
    n_grid = grid.nX * grid.nY * grid.nZ
    m = np.ones(n_grid, dtype=bool) if mask is None else np.asarray(mask).reshape(-1, order="F").astype(bool)
    idx = np.nonzero(m)[0]
    kn = kspha[1:4, :] if k_nominal is None else np.asarray(k_nominal, dtype=RT)
    k = prep_kspha(kspha, kn, n_term, recon_terms)
    x, y, z = grid.x[idx], grid.y[idx], grid.z[idx]
    b64 = (basis_f64d if f64 else basis_f64)(x, y, z)[:, :n_term]
    b32 = basis_f32(x, y, z)[:, :n_term]
    if concomitant is not None:
        conc = np.asarray(concomitant, dtype=RT)
        X, Y, Z = (a.astype(np.float64) for a in (x, y, z))
        c64 = np.stack([Z * Z, X * X + Y * Y, X * Z, Y * Z], axis=1)
        xf, yf, zf = x.astype(np.float32), y.astype(np.float32), z.astype(np.float32)
        c32 = np.stack([zf * zf, xf * xf + yf * yf, xf * zf, yf * zf], axis=1).astype(np.float32)
        b64 = np.concatenate([b64, c64], axis=1)
        b32 = np.concatenate([b32, c32], axis=1)
        k = np.concatenate([k, conc], axis=0)
    fm = np.zeros(n_grid, dtype=RT) if fieldmap is None else \
        np.asarray(fieldmap, dtype=RT).reshape(-1, order="F")
    CT = np.complex128 if f64 else np.complex64
    if csm is None:
        cs = np.ones((1, idx.size), dtype=CT)
    else:
        c = np.asarray(csm, dtype=CT)
        c = c.reshape(n_grid, -1, order="F") if c.ndim != 2 else c
        cs = np.ascontiguousarray(c[idx, :].T)
    L = k.shape[0]
    return Model(nv=idx.size, ns=ns, nc=cs.shape[0], L=L, mask_idx=idx, n_grid=n_grid,
                 bf64=np.ascontiguousarray(b64), bf32=np.ascontiguousarray(b32, dtype=np.float32),
                 k64=np.ascontiguousarray(k.T.astype(np.float64)),
                 k32=np.ascontiguousarray(k.T.astype(np.float32)),
                 times=np.ascontiguousarray(np.asarray(times, dtype=RT).reshape(-1)),
                 fieldmap=np.ascontiguousarray(fm[idx]), csm=cs)


def with_coefficients(m: Model, k_rows) -> Model:
    """Same model with effective coefficients replaced by ``k_rows`` ([L][ns])."""
    k = np.asarray(k_rows, dtype=np.float64)
    return Model(**{**m.__dict__, "k64": np.ascontiguousarray(k.T),
                    "k32": np.ascontiguousarray(k.T.astype(np.float32))})


def gather(m: Model, x_full):
    return np.asarray(x_full).reshape(-1)[m.mask_idx]


def scatter(m: Model, xm, dtype=np.complex128):
    out = np.zeros(m.n_grid, dtype=dtype)
    out[m.mask_idx] = xm
    return out


# ---------------------------------------------------------------- FP64 oracle

def forward_f64(m: Model, x_full):
    """E x in FP64 -> [nc*ns] (samples fastest)."""
    xm = np.ascontiguousarray(gather(m, x_full).astype(np.complex128))
    cs = np.ascontiguousarray(m.csm.astype(np.complex128))
    y = np.empty(m.nc * m.ns, dtype=np.complex128)
    t, fm = _f64(m.times), _f64(m.fieldmap)      # keep the temporaries alive during the call
    lib().oracle_forward_f64(m.nv, m.ns, m.nc, m.L, _p(m.bf64), _p(m.k64), _p(t), _p(fm), _p(cs), _p(xm), _p(y))
    return y


def forward_dt_f64(m: Model, kspha_dt, x_full):
    """Derivative product B x in FP64 -> [nc*ns] (array_operator.jl:187-235). ``kspha_dt`` is
    (nTerm, nSam) as given to the operator (no recon_terms processing); the model must
    carry no concomitant terms (array_operator.jl:139)."""
    kdt = np.ascontiguousarray(np.asarray(kspha_dt, dtype=np.float32).astype(np.float64).T)
    assert kdt.shape == (m.ns, m.L), "kspha_dt must match the model's terms and samples"
    xm = np.ascontiguousarray(gather(m, x_full).astype(np.complex128))
    cs = np.ascontiguousarray(m.csm.astype(np.complex128))
    y = np.empty(m.nc * m.ns, dtype=np.complex128)
    t, fm = _f64(m.times), _f64(m.fieldmap)
    lib().oracle_forward_dt_f64(m.nv, m.ns, m.nc, m.L, _p(m.bf64), _p(m.k64), _p(kdt), _p(t), _p(fm), _p(cs),
                                _p(xm), _p(y))
    return y


def adjoint_f64(m: Model, y):
    ys = np.ascontiguousarray(np.asarray(y).reshape(-1).astype(np.complex128))
    cs = np.ascontiguousarray(m.csm.astype(np.complex128))
    xm = np.empty(m.nv, dtype=np.complex128)
    t, fm = _f64(m.times), _f64(m.fieldmap)
    lib().oracle_adjoint_f64(m.nv, m.ns, m.nc, m.L, _p(m.bf64), _p(m.k64), _p(t), _p(fm), _p(cs), _p(ys), _p(xm))
    return scatter(m, xm)


def expand_weights(m: Model, D):
    if D is None:
        return np.ones(m.ns * m.nc)
    D = np.asarray(D, dtype=np.float64).reshape(-1)
    return np.tile(D, m.nc) if D.size == m.ns else D


def normal_f64(m: Model, x_full, D=None):
    """Aᴴ D A x (D already squared; explicit_gpu_operator.jl:399-416)."""
    return adjoint_f64(m, expand_weights(m, D) * forward_f64(m, x_full))


def grad_f64(m: Model, x_full, y):
    """dL/dk_l(s) for L = ||E x - y||^2, FP64; returns (g [L][ns], loss, residual)."""
    r = forward_f64(m, x_full) - np.asarray(y).reshape(-1)
    xm = np.ascontiguousarray(gather(m, x_full).astype(np.complex128))
    cs = np.ascontiguousarray(m.csm.astype(np.complex128))
    rr = np.ascontiguousarray(r.astype(np.complex128))
    g = np.empty((m.L, m.ns), dtype=np.float64)
    t, fm = _f64(m.times), _f64(m.fieldmap)
    lib().oracle_grad_f64(m.nv, m.ns, m.nc, m.L, _p(m.bf64), _p(m.k64), _p(t), _p(fm), _p(cs), _p(xm), _p(rr),
                          _p(g))
    return g, float(np.vdot(r, r).real), r


def dense_matrix_f64(m: Model):
    """Dense FP64 encoding matrix (nc*ns, n_grid) (test_explicit_gpu.jl:15-35); tiny problems only."""
    ph = m.k64 @ m.bf64.T + np.outer(m.times.astype(np.float64), m.fieldmap.astype(np.float64))
    P = np.exp(2j * np.pi * ph) / np.sqrt(m.nv)
    A = np.zeros((m.nc * m.ns, m.n_grid), dtype=np.complex128)
    for c in range(m.nc):
        A[c * m.ns:(c + 1) * m.ns, m.mask_idx] = P * m.csm[c].astype(np.complex128)[None, :]
    return A


# ---------------------------------------------------------------- reference CPU path

def cpu_ref_forward(m: Model, x_full, nblock=50):
    xm = np.ascontiguousarray(gather(m, x_full).astype(np.complex64))
    y = np.empty(m.nc * m.ns, dtype=np.complex64)
    cs = np.ascontiguousarray(m.csm)
    lib().cpu_ref_forward(m.nv, m.ns, m.nc, m.L, _p(m.bf32), _p(m.k32), _p(m.times), _p(m.fieldmap),
                          _p(cs), _p(xm), _p(y), nblock)
    return y


def cpu_ref_adjoint(m: Model, y, nblock=50):
    ys = np.ascontiguousarray(np.asarray(y).reshape(-1).astype(np.complex64))
    xm = np.empty(m.nv, dtype=np.complex64)
    cs = np.ascontiguousarray(m.csm)
    lib().cpu_ref_adjoint(m.nv, m.ns, m.nc, m.L, _p(m.bf32), _p(m.k32), _p(m.times), _p(m.fieldmap),
                          _p(cs), _p(ys), _p(xm), nblock)
    return scatter(m, xm, np.complex64)


def cpu_ref_normal(m: Model, x_full, D=None, nblock=50):
    y = cpu_ref_forward(m, x_full, nblock)
    return cpu_ref_adjoint(m, (expand_weights(m, D) * y).astype(np.complex64), nblock)


# ---------------------------------------------------------------- solvers

def cgnr(apply_normal, rhs, iters, lam=0.0, rel_tol=0.0):
    """Textbook CG on (N + λI) x = rhs, x0 = 0, FP64 (the CGNR iteration of
    RegularizedLeastSquares invoked at recon_HOOp.jl:80-81)."""
    x = np.zeros_like(rhs, dtype=np.complex128)
    r = rhs.astype(np.complex128).copy()
    p = r.copy()
    rr = np.vdot(r, r).real
    rr0 = rr
    log = [1.0 if rr0 > 0 else 0.0]
    for _ in range(iters):
        if rr0 == 0 or (rel_tol > 0 and np.sqrt(rr / rr0) <= rel_tol):
            log.append(0.0)
            continue
        q = apply_normal(p) + lam * p
        pap = np.vdot(p, q).real
        if pap <= 0:
            break
        a = rr / pap
        x += a * p
        r -= a * q
        rn = np.vdot(r, r).real
        log.append(np.sqrt(rn / rr0))
        p = r + (rn / rr) * p
        rr = rn
    return x, np.array(log)


# ---------------------------------------------------------------- delay model

def catmull_rom_weights(f):
    f2, f3 = f * f, f * f * f
    w = np.array([0.5 * (-f3 + 2 * f2 - f), 0.5 * (3 * f3 - 5 * f2 + 2), 0.5 * (-3 * f3 + 4 * f2 + f),
                  0.5 * (f3 - f2)])
    dw = np.array([0.5 * (-3 * f2 + 4 * f - 1), 0.5 * (9 * f2 - 10 * f), 0.5 * (-9 * f2 + 8 * f + 1),
                   0.5 * (3 * f2 - 2 * f)])
    return w, dw


def delayed(traces, tau, dwell):
    """Traces [L][ns] resampled at s + tau/dwell (uniform Catmull-Rom, clamped ends)."""
    tr = np.asarray(traces, dtype=np.float64)
    if tau == 0:
        return tr.copy()
    u = tau / dwell
    m = int(np.floor(u))
    w, _ = catmull_rom_weights(u - m)
    ns = tr.shape[1]
    out = np.zeros_like(tr)
    s = np.arange(ns)
    for t in range(4):
        j = np.clip(s + m - 1 + t, 0, ns - 1)
        out += w[t] * tr[:, j]
    return out


# ---------------------------------------------------------------- delay estimation

def kspha_dt_stencil(kspha, dt):
    """``conv(kspha, [1/8 1/4 0 -1/4 -1/8]')[3:end-2, :] / dt`` per column of a
    ``(nFieldSample, nTerm)`` array (FindDelay.jl:178-180), FP64."""
    k = np.asarray(kspha, dtype=np.float64)
    ker = np.array([1 / 8, 1 / 4, 0.0, -1 / 4, -1 / 8])
    cols = [np.convolve(k[:, j], ker, mode="full")[2:-2] for j in range(k.shape[1])]
    return np.stack(cols, axis=1) / float(dt)


def interp_traj_time(traj, dt, delTime, datatime):
    """InterpTraj.jl:18-44: column j of ``traj`` ``(nFieldSample, nTerm)`` sampled at
    ``(0:n-1)*dt - delTime``, interpolated at ``datatime`` by Akima's method, zero outside
    the sampled interval. Uses scipy's Akima1DInterpolator, an implementation of the
    same published method (Akima 1970) that is independent of the product's."""
    from scipy.interpolate import Akima1DInterpolator
    tr = np.asarray(traj, dtype=np.float64)
    tt = float(dt) * np.arange(tr.shape[0]) - float(delTime)
    out = Akima1DInterpolator(tt, tr, axis=0)(np.asarray(datatime, dtype=np.float64))
    return np.nan_to_num(out, nan=0.0)


def find_delay_f64(grid, data, kspha, datatime, StartTime, dt, *, JumpFact=3, dtau_min, fieldmap=None,
                   csm=None, recon_terms=None, iter_max=10, lam=0.0, weight=None, max_updates=50):
    """FindDelay.jl:152-214 in FP64: per update, resample kspha and its stencil derivative at
    ``StartTime + τ``, reconstruct x by ``iter_max`` CGNR iterations on the weighted normal,
    ``Δτ = J * real((B x) \\ (y - E x))``, halve J (not below 1) on a sign change, stop at
    ``|Δτ| <= dtau_min``; returns ``(τ - Δτ/J*(J-1), [τ after each update])``."""
    data = np.asarray(data)
    n_sample = data.shape[0]
    y = data.reshape(-1, order="F").astype(np.complex128)
    kdt = kspha_dt_stencil(kspha, dt)
    tau, J, n_it = 0.0, JumpFact, 1
    dtau = dtau_prev = np.inf
    hist = [0.0]
    while abs(dtau) > dtau_min and n_it <= max_updates:
        k_tau = interp_traj_time(kspha, dt, tau + StartTime, datatime)[:n_sample].T.astype(np.float32)
        kdt_tau = interp_traj_time(kdt, dt, tau + StartTime, datatime)[:n_sample].T.astype(np.float32)
        m = make_model(grid, k_tau, datatime, fieldmap, csm, None, recon_terms=recon_terms)
        D = np.ones(n_sample) if weight is None else np.abs(np.asarray(weight, dtype=np.complex128)) ** 2
        rhs = adjoint_f64(m, expand_weights(m, D) * y)
        x, _ = cgnr(lambda p: normal_f64(m, p, D), rhs, iter_max, lam)
        y1 = y - forward_f64(m, x)
        y2 = forward_dt_f64(m, kdt_tau, x)
        dtau = J * float((np.vdot(y2, y1) / np.vdot(y2, y2)).real)
        tau += dtau
        if n_it > 1 and np.sign(dtau_prev) != np.sign(dtau):
            J = max(1, J / 2)
        dtau_prev = dtau
        n_it += 1
        hist.append(tau)
    return tau - dtau / J * (J - 1), hist


# ---------------------------------------------------------------- GIRF prediction

def nextprod235(n):
    """Smallest m >= n whose only prime factors are 2, 3, 5 (``nextprod([2,3,5], n)``)."""
    m = max(int(n), 1)
    while True:
        r = m
        for f in (2, 3, 5):
            while r % f == 0:
                r //= f
        if r == 1:
            return m
        m += 1


def girf_expand_response(Hw, N):
    """apply_girf.jl:72-90 in FP64: time-domain expansion of a DC-centred response
    ``(nFreq, 3, nOut)`` to N frequencies; time sample nFreq//2 (0-based) is left out."""
    Hw = np.asarray(Hw, dtype=np.complex128)
    n_freq = Hw.shape[0]
    hw_t = np.fft.ifft(np.fft.ifftshift(Hw, axes=0), axis=0)
    pad = np.zeros((N,) + Hw.shape[1:], dtype=np.complex128)
    half = n_freq // 2
    pad[:half] = hw_t[:half]
    pad[N - (n_freq - half) + 1:] = hw_t[half + 1:]
    return np.fft.fftshift(np.fft.fft(pad, axis=0), axes=0)


def girf_core_f64(G2, Hw_work, n_point, rbw=1.0):
    """_apply_girf_core (apply_girf.jl:111-150) in FP64 for one (3, nPoint) waveform."""
    N = Hw_work.shape[0]
    gp = np.zeros((3, N), dtype=np.complex128)
    gp[:, :n_point] = G2
    Gf = np.fft.fftshift(np.fft.fft(gp, axis=1), axes=1)
    if rbw < 1.0:
        nz = int(np.round(N * (1.0 - rbw) / 2.0))
        if nz > 0:
            Gf[:, :nz] = 0
            Gf[:, N - nz:] = 0
    Of = np.einsum("fio,if->of", Hw_work, Gf)
    return np.real(np.fft.ifft(np.fft.ifftshift(Of, axes=1), axis=1))[:, :n_point]


def apply_girf_f64(G, Hw, dim_spatial=1, dim_time=2, rbw=1.0):
    """apply_girf (apply_girf.jl:39-109) in FP64; dims are 1-based as in the reference."""
    G = np.asarray(G, dtype=np.float64)
    Hw = np.asarray(Hw, dtype=np.complex128)
    ds, dt_ = dim_spatial - 1, dim_time - 1
    assert G.shape[ds] == 3 and Hw.shape[1] == 3
    n_point = G.shape[dt_]
    N = nextprod235(max(n_point, Hw.shape[0]))
    Hw_work = girf_expand_response(Hw, N) if N > Hw.shape[0] else Hw
    Gm = np.moveaxis(G, (ds, dt_), (0, 1))
    others = Gm.shape[2:]
    G3 = Gm.reshape(3, n_point, -1)
    out = np.stack([girf_core_f64(G3[:, :, i], Hw_work, n_point, rbw) for i in range(G3.shape[2])], axis=2)
    out = out.reshape((Hw.shape[2], n_point) + others)
    return np.moveaxis(out, (0, 1), (ds, dt_))


def grad2traj_f64(x, dt, dim=1):
    """utils/conversion.jl:14-41: zero then the running trapezoid sum (x[j]+x[j+1])*dt/2."""
    x = np.asarray(x, dtype=np.float64)
    a = np.moveaxis(x, dim - 1, 0)
    out = np.zeros_like(a)
    acc = np.zeros(a.shape[1:])
    for j in range(1, a.shape[0]):
        acc = acc + (a[j - 1] + a[j]) * (dt / 2)
        out[j] = acc
    return np.moveaxis(out, 0, dim - 1)


# ---------------------------------------------------------------- reference fixtures

class OracleGrid:
    """The centred Cartesian grid of src/Grid/Grid.jl:63-95, restated for the oracle's
    fixtures: grid points (i, j, k) over 1:nX x 1:nY x 1:nZ with i fastest
    (Grid.jl:78-80), centred at (N + 1)/2 and scaled by the spacing in Float64
    (:81-82), then the optional exchange of x and y and reversals of the flattened
    vectors (:83-94), converted to T last (:95)."""

    def __init__(self, nX, nY, nZ, dx, dy, dz, *, dtype=np.float32, exchange_xy=False, reverse_x=False,
                 reverse_y=False, reverse_z=False):
        T = np.dtype(dtype).type
        self.nX, self.nY, self.nZ = int(nX), int(nY), int(nZ)
        self.dx, self.dy, self.dz = T(dx), T(dy), T(dz)
        self.dtype = np.dtype(dtype)
        I, J, K = np.meshgrid(np.arange(1.0, nX + 1), np.arange(1.0, nY + 1), np.arange(1.0, nZ + 1), indexing="ij")
        coords = []
        for A, n, d in ((I, nX, self.dx), (J, nY, self.dy), (K, nZ, self.dz)):
            coords.append((A.ravel(order="F") - (n + 1) / 2) * np.float64(d))
        x, y, z = coords
        if exchange_xy:
            x, y = y, x
        if reverse_x:
            x = x[::-1]
        if reverse_y:
            y = y[::-1]
        if reverse_z:
            z = z[::-1]
        self.x, self.y, self.z = (np.ascontiguousarray(v.astype(T)) for v in (x, y, z))
        self.matrixSize = (self.nX, self.nY, self.nZ)
        self.resolution = (self.dx, self.dy, self.dz)
        self.n = self.nX * self.nY * self.nZ


def _jrange(a, b, n, T=np.float32):
    return (np.float64(T(a)) + (np.float64(T(b)) - np.float64(T(a))) * np.arange(n) / (n - 1)).astype(T)


def explicit_highorder_test_problem(n_coil=3, grid_cls=OracleGrid):
    """The deterministic fixture of test/test_HighOrderExplicit.jl:3-59 (ranges, sin, cos only)."""
    f32 = np.float32
    n_x, n_y, n_z = 4, 3, 2
    ns, nt = 7, 16
    grid = grid_cls(n_x, n_y, n_z, f32(0.7), f32(0.9), f32(1.2))
    times = _jrange(0.0, 1.3e-3, ns)
    k = np.zeros((nt, ns), dtype=f32)
    k[0] = _jrange(-0.01, 0.015, ns)
    k[1] = _jrange(-0.08, 0.07, ns)
    k[2] = _jrange(0.06, -0.05, ns)
    k[3] = _jrange(-0.03, 0.04, ns)
    ang = _jrange(0.0, np.pi, ns)
    for term in range(5, 10):
        k[term - 1] = f32(4e-4 * (term - 4)) * np.sin(ang).astype(f32)
    for term in range(10, 17):
        k[term - 1] = f32(2e-5 * (term - 9)) * np.cos(ang).astype(f32)
    n = n_x * n_y * n_z
    fieldmap = _jrange(-25.0, 20.0, n)
    nxv = (grid.x / np.max(np.abs(grid.x))).astype(f32)
    nyv = (grid.y / np.max(np.abs(grid.y))).astype(f32)
    nzv = (grid.z / np.max(np.abs(grid.z))).astype(f32)
    csm = np.empty((n, n_coil), dtype=np.complex64)
    for coil in range(1, n_coil + 1):
        angle = f32(2 * np.pi * (coil - 1) / n_coil)
        mag = f32(1) + f32(0.08) * (np.cos(angle).astype(f32) * nxv + np.sin(angle).astype(f32) * nyv)
        ph = f32(0.12 * coil) * (nyv - nzv)
        csm[:, coil - 1] = mag * np.exp(1j * ph.astype(np.float64)).astype(np.complex64)
    mask = np.ones(n, dtype=bool)
    mask[0] = False
    mask[n - 1] = False
    image = (_jrange(0.1, 1.0, n) + 1j * _jrange(-0.2, 0.15, n)).astype(np.complex64)
    return dict(grid=grid, kspha=k, times=times, fieldmap=fieldmap, csm=csm, mask=mask, image=image)


def delay_test_problem(n=16, n_cha=4, n_field=1200, n_adc=1000, dt=2e-6, start=100, tau_samples=3.4,
                       turns=6, seed=5, grid_cls=OracleGrid):
    """A 2-D spiral acquisition whose data are simulated (FP64, noiseless) from field
    coefficients delayed by ``tau_samples * dt``: ``y = E(k(t + StartTime + τ)) x``.
    Coefficients ``(n_field, 16)``: a smooth order-0 drift, the spiral in rows 2-3,
    smooth small orders 2-3. Returns the FindDelay inputs and the planted delay."""
    rng = np.random.default_rng(seed)
    fov = 0.2
    grid = grid_cls(n, n, 1, np.float32(fov / n), np.float32(fov / n), np.float32(fov / n))
    u = np.arange(n_field) / (n_field - 1)
    kmax = n / (2 * fov)
    r = kmax * u
    th = 2 * np.pi * turns * u
    k = np.zeros((n_field, 16))
    k[:, 0] = 0.05 * np.sin(2 * np.pi * 1.3 * u + rng.uniform(0, 6))
    k[:, 1], k[:, 2] = r * np.cos(th), r * np.sin(th)
    for j in range(4, 16):
        order = 2 if j < 9 else 3
        a, b, f = rng.normal(size=3)
        k[:, j] = 0.03 * kmax / (fov / 2) ** (order - 1) * (a * np.sin(2 * np.pi * (1 + abs(f)) * u) + b * u)
    kspha = k.astype(np.float32)
    datatime = (dt * np.arange(n_adc)).astype(np.float32)
    StartTime = np.float32(start * dt)
    tau = tau_samples * dt
    x, y = grid.x.astype(np.float64), grid.y.astype(np.float64)
    csm = np.empty((grid.n, n_cha), np.complex64)
    for c in range(n_cha):
        a = 2 * np.pi * c / n_cha
        csm[:, c] = np.exp(-((x - 0.06 * np.cos(a)) ** 2 + (y - 0.06 * np.sin(a)) ** 2) / (2 * 0.08 ** 2)) * \
            np.exp(1j * (x * np.cos(a) - y * np.sin(a)) * 10)
    fm = (40 * np.exp(-((x - 0.02) ** 2 + y ** 2) / (2 * 0.05 ** 2))).astype(np.float32)
    img = (np.exp(-(x ** 2 + (y / 0.8) ** 2) / (2 * 0.05 ** 2)) * (1 + 0.3 * np.sin(30 * x))).astype(np.complex64)
    k_true = interp_traj_time(kspha, dt, StartTime + tau, datatime).T.astype(np.float32)
    m = make_model(grid, k_true, datatime, fm, csm, None)
    data = forward_f64(m, img).reshape(n_adc, n_cha, order="F").astype(np.complex64)
    return dict(grid=grid, data=data, kspha=kspha, datatime=datatime, StartTime=StartTime, dt=np.float32(dt),
                fieldmap=fm, csm=csm, image=img, tau=tau)


def relerr(a, b):
    a = np.asarray(a).reshape(-1).astype(np.complex128)
    b = np.asarray(b).reshape(-1).astype(np.complex128)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
