"""Field-camera inputs of the 7 T demo data and sampling-density weights.

``load_demo_mat(path)`` reads a MATLAB 5/7 MAT-file of the reference's demo data
(the variables of demo/DemoCommon.jl:13-36: ``gre_csm``, ``gre_b0``, ``gre_mask``,
``kdata``, ``datatime``, ``matrixSize``, ``FOV``, ``k0_adc``, ``dt_Measured``,
``dt_Nominal``, ``ksphaMeasured``, ``startMeasured``, ``ksphaNominal``,
``startNominal``, ``tauNominal``, ``tauMeasured``, ``dt_adc``) with
``scipy.io.loadmat``, a parser of the file's numeric arrays that executes nothing
from it. Scalars are unwrapped.

``prepare_demo(data, trajectory)`` turns them into the inputs of
``HighOrderExplicitOp`` and ``recon_HOOp`` as the demo does:

- coil maps, B0 map and mask: ``circshift`` by (0, -1), ``reverse`` along the
  first dimension, ``permutedims`` (2, 1) (DemoCommon.jl:38-42);
- field coefficients resampled onto the ADC times with the synchronisation
  delay, ``InterpTrajTime(kspha, dt, start + tau * dt_adc, datatime)``
  (DemoCommon.jl:47-49);
- grid ``Grid(nX, nY, nZ, FOV ./ matrixSize...)`` (DemoCommon.jl:54-56);
- square-root density weights of the (kx, ky) rows (DemoCommon.jl:59-60);
- the measured-trajectory data divided by ``exp(2 pi i k0_adc)`` and multiplied
  by ``exp(-2 pi i k_0(t))``, recon terms "0111"; the nominal trajectory takes
  the raw data with recon terms "010" (demo/CompareMeasuredNominal.jl:4-9,
  27, 35-38);
- ``recParams``: L2 regularisation 1e-9, 20 CGNR iterations (DemoCommon.jl:66-71).

``samplingDensity(tr, shape)`` follows the reference's helper
(src/Utils/SampleDensity.jl:36-48): coordinates normalised by ``2 max|tr|``,
ten density-compensation iterations, square root of the density. The reference
computes the density with NFFT.jl's ``sdc`` (not vendored with the reference);
this module uses the published iteration of Pipe and Menon (Magn. Reson. Med.
41:179, 1999), ``w <- w / (C * w)``, where ``C * w`` spreads the weights onto a
2x oversampled periodic grid with a Kaiser-Bessel kernel of half-width 2 grid
cells (Beatty et al., IEEE TMI 24:799, 2005) and interpolates back at the
samples, scaled to Nyquist cells per sample (full Cartesian sampling = 1). Its
weights therefore differ from NFFT.jl's in detail and in overall scale.
"""
from __future__ import annotations

import numpy as np

from ._lib import ArgumentError, DimensionMismatch
from .grid import Grid
from .sync import InterpTrajTime

DEMO_KEYS = ("gre_csm", "gre_b0", "gre_mask", "kdata", "datatime", "matrixSize", "FOV", "k0_adc", "dt_Measured",
             "dt_Nominal", "ksphaMeasured", "startMeasured", "ksphaNominal", "startNominal", "tauNominal",
             "tauMeasured", "dt_adc")


def load_demo_mat(path) -> dict:
    """Variables of a demo MAT-file (MATLAB 5/7 format) as numpy arrays."""
    import scipy.io as sio
    raw = sio.loadmat(str(path), squeeze_me=False, mat_dtype=False)
    missing = [k for k in DEMO_KEYS if k not in raw]
    if missing:
        raise ArgumentError(f"{path}: missing variables {missing}")
    out = {}
    for k in DEMO_KEYS:
        a = np.asarray(raw[k])
        out[k] = a.reshape(()).item() if a.size == 1 else a
    return out


def _demo_orient(a):
    """circshift(a, (0, -1[, 0])), reverse(dims=1), permutedims((2, 1[, 3]))."""
    a = np.roll(a, -1, axis=1)[::-1]
    axes = (1, 0) + tuple(range(2, a.ndim))
    return np.transpose(a, axes)


def _kb_kernel(d, half_width, beta):
    """Kaiser-Bessel window at distance d (grid cells), zero beyond half_width."""
    u = np.clip(1.0 - (d / half_width) ** 2, 0.0, None)
    return np.where(np.abs(d) < half_width, np.i0(beta * np.sqrt(u)), 0.0)


def _torch_gpu():
    """torch when it sees a GPU, else None."""
    try:
        import torch
    except ImportError:
        return None
    return torch if torch.cuda.is_available() else None


def _pipe_menon(k, gshape, iters, hw=2.0, sigma=2.0):
    """Density estimate w (per sample) of normalised coordinates k ((dim, nSam) in
    [-0.5, 0.5]) after ``iters`` iterations w <- w / (C * w) on the periodic grid.
    Trajectories of 4096 samples or more run on the GPU when torch sees one
    (``_pipe_menon_device``, the same FP64 arithmetic)."""
    beta = np.pi * np.sqrt((2 * hw / sigma) ** 2 * (sigma - 0.5) ** 2 - 0.8)
    ndim, ns = k.shape
    tc = _torch_gpu() if ns >= 4096 else None
    if tc is not None:
        return _pipe_menon_device(tc, k, gshape, iters, hw, beta)
    # per dimension: the 5 grid cells around each sample and their kernel values
    idx, val = [], []
    for a in range(ndim):
        u = (k[a] + 0.5) * gshape[a]
        base = np.floor(u).astype(np.int64) - 2
        cells = base[:, None] + np.arange(5)[None, :]
        val.append(_kb_kernel(cells - u[:, None], hw, beta))
        idx.append(np.mod(cells, gshape[a]))
    # flat grid index and separable weight of each (sample, tap)
    flat = np.zeros((ns, 1), dtype=np.int64)
    wt = np.ones((ns, 1))
    stride = 1
    for a in range(ndim):
        flat = (flat[:, :, None] + stride * idx[a][:, None, :]).reshape(ns, -1)
        wt = (wt[:, :, None] * val[a][:, None, :]).reshape(ns, -1)
        stride *= gshape[a]
    w = np.ones(ns)
    for _ in range(int(iters)):
        grid = np.bincount(flat.ravel(), weights=(wt * w[:, None]).ravel(), minlength=stride)
        cw = (wt * grid[flat]).sum(axis=1)
        w = w / np.where(cw > 0, cw, 1.0)
    return w


def _pipe_menon_device(tc, k, gshape, iters, hw, beta):
    """``_pipe_menon`` with torch on the current GPU. The gridding sums run as a
    segmented reduction over the (sample, tap) pairs sorted by grid cell (stable, so each
    cell adds its pairs in the host's order) instead of atomics: the result does not vary
    from run to run."""
    dev = tc.device("cuda", tc.cuda.current_device())
    kd = tc.from_numpy(np.ascontiguousarray(k, dtype=np.float64)).to(dev)
    ndim, ns = kd.shape
    taps = tc.arange(5, device=dev)
    flat = tc.zeros((ns, 1), dtype=tc.int64, device=dev)
    wt = tc.ones((ns, 1), dtype=tc.float64, device=dev)
    stride = 1
    for a in range(ndim):
        u = (kd[a] + 0.5) * gshape[a]
        cells = tc.floor(u).to(tc.int64)[:, None] - 2 + taps[None, :]
        d = cells.to(tc.float64) - u[:, None]
        arg = tc.clamp(1.0 - (d / hw) ** 2, min=0.0)
        val = tc.where(tc.abs(d) < hw, tc.special.i0(beta * tc.sqrt(arg)), tc.zeros_like(d))
        idx = tc.remainder(cells, gshape[a])
        flat = (flat[:, :, None] + stride * idx[:, None, :]).reshape(ns, -1)
        wt = (wt[:, :, None] * val[:, None, :]).reshape(ns, -1)
        stride *= int(gshape[a])
    keys = flat.reshape(-1)
    perm = tc.argsort(keys, stable=True)
    cells_u, counts = tc.unique_consecutive(keys[perm], return_counts=True)
    wt_p = wt.reshape(-1)[perm]
    sample = perm // flat.shape[1]
    w = tc.ones(ns, dtype=tc.float64, device=dev)
    grid = tc.zeros(stride, dtype=tc.float64, device=dev)
    for _ in range(int(iters)):
        grid[cells_u] = tc.segment_reduce(wt_p * w[sample], "sum", lengths=counts)
        cw = (wt * grid[flat]).sum(dim=1)
        w = w / tc.where(cw > 0, cw, tc.ones_like(cw))
    return w.cpu().numpy()


def samplingDensity(tr, shape, *, iters: int = 10, normalize: bool = False):
    """Square-root sampling-density weights of a (dim, nSam) trajectory
    (src/Utils/SampleDensity.jl:36-48); Pipe-Menon iterations on a 2x grid.

    The density is expressed in Nyquist cells per sample: it is divided by the
    density the same iterations give a fully sampled Cartesian lattice, so full
    Cartesian sampling has weight 1 (sparser sampling gets larger weights, up to
    what the kernel's support can resolve). This keeps the weighted normal matrix at O(1) scale, so a small L2
    regularisation (the demo's 1e-9) stays negligible as with the reference's
    weights."""
    tr = np.asarray(tr, dtype=np.float64)
    shape = tuple(int(s) for s in shape)
    if tr.ndim != 2 or tr.shape[0] != len(shape) or tr.shape[0] not in (1, 2, 3):
        raise DimensionMismatch("tr must be (length(shape), nSam)")
    if any(s <= 0 for s in shape):
        raise ArgumentError("shape must have positive dimensions")
    c = 2.0 * np.max(np.abs(tr[np.isfinite(tr)])) if np.isfinite(tr).any() else 0.0
    if not (c > 0.0):
        raise ArgumentError("trajectory must have a nonzero finite value")
    sigma = 2.0
    w = _pipe_menon(tr / c, tuple(int(sigma * s) for s in shape), iters)
    # the same iterations on a fully sampled Cartesian lattice (16 points per dimension)
    ndim, m = tr.shape[0], 16
    lat = (np.stack(np.meshgrid(*([np.arange(m)] * ndim), indexing="ij")).reshape(ndim, -1) - m / 2) / m
    w_full = _pipe_menon(lat, (int(sigma * m),) * ndim, iters)[0]
    weights = np.sqrt(w / w_full)
    if normalize:
        scale = np.linalg.norm(weights) / np.sqrt(weights.size)
        if not (np.isfinite(scale) and scale > 0):
            raise ArgumentError("density weights must have a finite positive RMS")
        weights = weights / scale
    return weights.astype(np.complex64)


def demodulate(kdata, k0_adc, k_order0):
    """``kdata ./ exp(2 pi i k0_adc) .* exp(-2 pi i k_0)`` per sample, over all coils
    (demo/CompareMeasuredNominal.jl:4-6): removes the scanner's eddy-current B0
    compensation and the measured zeroth-order phase from (nSam, nCha) data."""
    kdata = np.asarray(kdata, dtype=np.complex128)
    k0 = np.asarray(k0_adc, dtype=np.float64).reshape(-1)
    kz = np.asarray(k_order0, dtype=np.float64).reshape(-1)
    if kdata.ndim != 2 or kdata.shape[0] != k0.size or k0.size != kz.size:
        raise DimensionMismatch("kdata (nSam, nCha), k0_adc and k_0 must agree in nSam")
    return kdata * np.exp(-2j * np.pi * (k0 + kz))[:, None]


def prepare_demo(data: dict, trajectory: str = "measured", *, dtype=np.float32) -> dict:
    """Operator and reconstruction inputs of the demo for the measured or the
    nominal trajectory (DemoCommon.jl:39-71, CompareMeasuredNominal.jl)."""
    T = np.dtype(dtype).type
    traj = trajectory.lower()
    if traj not in ("measured", "nominal"):
        raise ArgumentError("trajectory must be 'measured' or 'nominal'")
    ms = np.asarray(data["matrixSize"]).reshape(-1).astype(np.int64)
    fov = np.asarray(data["FOV"], dtype=np.float64).reshape(-1)
    nX, nY, nZ = (int(v) for v in ms)
    csm = _demo_orient(np.asarray(data["gre_csm"]))
    b0 = _demo_orient(np.asarray(data["gre_b0"], dtype=np.float64))
    mask = _demo_orient(np.asarray(data["gre_mask"]))
    if csm.shape[:2] != (nX, nY) or b0.shape != (nX, nY) or mask.shape != (nX, nY):
        raise DimensionMismatch("maps do not match matrixSize")
    datatime = np.asarray(data["datatime"], dtype=np.float64).reshape(-1)
    dt_adc = float(data["dt_adc"])
    if traj == "measured":
        kfield = InterpTrajTime(np.asarray(data["ksphaMeasured"], dtype=np.float64), float(data["dt_Measured"]),
                                float(data["startMeasured"]) + float(data["tauMeasured"]) * dt_adc, datatime)
        recon_terms = "0111"
        kdata = demodulate(data["kdata"], data["k0_adc"], kfield[:, 0])
    else:
        kfield = InterpTrajTime(np.asarray(data["ksphaNominal"], dtype=np.float64), float(data["dt_Nominal"]),
                                float(data["startNominal"]) + float(data["tauNominal"]) * dt_adc, datatime)
        recon_terms = "010"
        kdata = np.asarray(data["kdata"], dtype=np.complex128)
    if kdata.shape[0] != datatime.size or kfield.shape[0] != datatime.size:
        raise DimensionMismatch("kdata, datatime and the field coefficients disagree in length")
    d = fov / ms
    grid = Grid(nX, nY, nZ, T(d[0]), T(d[1]), T(d[2]), dtype=dtype)
    weight = samplingDensity(kfield.T[1:3, :], (nX, nY))
    return {
        "grid": grid,
        "kspha": np.ascontiguousarray(kfield.T.astype(T)),
        "times": datatime.astype(T),
        "csm": np.ascontiguousarray(csm.reshape(nX * nY * nZ, -1, order="F").astype(np.complex64)),
        "fieldmap": b0.reshape(-1, order="F").astype(T),
        "mask": mask.reshape(-1, order="F") != 0,
        "data": np.ascontiguousarray(kdata.astype(np.complex64)),
        "weight": weight,
        "recon_terms": recon_terms,
        "reconSize": (nX, nY),
        "reg": 1.0e-9,
        "iterations": 20,
    }


def rotl90(img):
    """Julia's rotl90 (counterclockwise quarter turn) of an (nX, nY) image."""
    return np.rot90(np.asarray(img), 1)
