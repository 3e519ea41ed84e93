"""Field-model inputs upstream of ``kspha``: GIRF prediction on the GPU and
gradient/trajectory conversion.

- ``apply_girf`` / ``GIRFModel`` / ``unpack_girf_channels``
  (src/GIRF/apply_girf.jl:39-173, src/GIRF/GIRF.jl:18-58): the prediction runs in
  libhoe (``hoe_apply_girf``: hipFFT transforms and HIP kernels for the transfer
  product), batched over every waveform of the input.
- ``grad2traj`` / ``traj2grad`` (src/utils/conversion.jl:14-64): trapezoid running
  sum and forward differences, host numpy with Julia's result-type rules.

Dimensions (``dim_spatial``, ``dim_time``, ``dim``) are 1-based, as in the reference,
and arrays carry the reference's shapes (a ``(3, nPoint)`` numpy array is the
reference's ``3 x nPoint`` matrix).
"""
from __future__ import annotations

import ctypes as C
import logging
import warnings
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, check

log = logging.getLogger("hoe.girf")


@dataclass
class GIRFModel:
    """Centred frequency-domain gradient impulse response: ``Hw`` ``(nFreq, 3, nOut)``
    complex, DC centred; ``freqs`` the frequencies in Hz (metadata: ``apply_girf``
    does not resample with it) (GIRF.jl:18-45)."""
    Hw: np.ndarray
    freqs: np.ndarray

    def __post_init__(self):
        self.Hw = np.asarray(self.Hw)
        self.freqs = np.asarray(self.freqs).reshape(-1)
        if not np.iscomplexobj(self.Hw) or self.Hw.ndim != 3:
            raise ArgumentError("Hw must be a complex (nFreq, 3, nOut) array")


def _axis(d, ndim, name):
    if not isinstance(d, (int, np.integer)) or not 1 <= d <= ndim:
        raise ArgumentError(f"{name} must be a dimension in 1:{ndim}")
    return int(d) - 1


def apply_girf(G_DCS_nom, Hw, *, dim_spatial: int = 1, dim_time: int = 2, rbw: float = 1.0, device: int = 0):
    """Predict the realised field channels from nominal DCS gradients (apply_girf.jl:39-109).

    ``G_DCS_nom``: real array with 3 channels along ``dim_spatial`` and the waveform
    along ``dim_time``; other dimensions are independent waveforms. ``Hw``:
    ``(nFreq, 3, nOut)`` complex, DC centred, or a ``GIRFModel``. Returns the input's
    shape with the channel dimension replaced by nOut, in the input's floating type
    (float32 inputs are computed in FP32, float64 in FP64). Circular FFT convolution
    of length ``nextprod([2,3,5], max(nPoint, nFreq))``; ``rbw < 1`` clears the outer
    bins of each waveform's spectrum.
    """
    if isinstance(Hw, GIRFModel):
        Hw = Hw.Hw
    G = np.asarray(G_DCS_nom)
    if not np.issubdtype(G.dtype, np.floating):
        raise ArgumentError("apply_girf needs real floating-point gradients")
    ds = _axis(dim_spatial, G.ndim, "dim_spatial")
    dtm = _axis(dim_time, G.ndim, "dim_time")
    if ds == dtm:
        raise ArgumentError("dim_spatial and dim_time must differ")
    if G.shape[ds] != 3:
        raise ArgumentError("Input gradient must have 3 channels along dim_spatial.")
    Hw = np.asarray(Hw)
    if Hw.ndim != 3:
        raise DimensionMismatch("Hw must be (nFreq, 3, nOut)")
    n_freq, n_in, n_out = Hw.shape
    if n_in != 3:
        raise AssertionError("GIRF model must have 3 input channels (X, Y, Z)")
    if not np.isfinite(rbw):
        raise ArgumentError("rbw must be finite")
    lib = _lib.load()
    n_point = G.shape[dtm]
    n_fft = int(lib.hoe_girf_fft_length(max(n_point, n_freq)))
    if n_fft > n_freq:
        log.info("GIRF Resolution Expansion: Padding Hw from %d to %d points to safely cover gradient length.",
                 n_freq, n_fft)
    f64 = G.dtype == np.float64
    R, Cx = (np.float64, np.complex128) if f64 else (np.float32, np.complex64)
    # C layout (3, nPoint, nOther) channel fastest == numpy (nOther, nPoint, 3)
    Gm = np.moveaxis(G, (ds, dtm), (-1, -2))
    others = Gm.shape[:-2]
    g = np.ascontiguousarray(Gm.reshape(-1, n_point, 3), dtype=R)
    n_other = g.shape[0]
    hw = np.ascontiguousarray(np.transpose(Hw, (2, 1, 0)), dtype=Cx)   # frequency fastest
    out = np.empty((n_other, n_point, n_out), dtype=R)
    fn = lib.hoe_apply_girf_f64 if f64 else lib.hoe_apply_girf
    check(fn(g.ctypes.data, n_point, n_other, hw.ctypes.data, n_freq, n_out, float(rbw), int(device),
             out.ctypes.data, 0))
    res = np.moveaxis(out.reshape(others + (n_point, n_out)), (-1, -2), (ds, dtm))
    return res.astype(G.dtype, copy=False)


def unpack_girf_channels(G_act_all, *, dim_spatial: int = 1):
    """Split predicted channels into ``(linear, zeroth, higher)`` = channels 2:4, 1:1 and
    5:end along ``dim_spatial`` (``higher`` is None for exactly 4 channels); with fewer
    than 4 channels all are taken as linear with a zero zeroth term (apply_girf.jl:153-173)."""
    G = np.asarray(G_act_all)
    ds = _axis(dim_spatial, G.ndim, "dim_spatial")
    n_out = G.shape[ds]
    take = lambda a, b: np.take(G, np.arange(a, b), axis=ds).copy()   # noqa: E731
    if n_out >= 4:
        return take(1, 4), take(0, 1), (take(4, n_out) if n_out > 4 else None)
    warnings.warn("GIRF output has less than 4 channels. Treating all channels as 1st-order gradients.")
    shape = tuple(1 if d == ds else s for d, s in enumerate(G.shape))
    return G, np.zeros(shape, dtype=G.dtype), None


def _scalar_dtype(v):
    """Julia type of a scalar argument: Python int -> Int64, float -> Float64."""
    if isinstance(v, (bool, np.bool_)):
        return np.dtype(np.int64)
    if isinstance(v, (int, np.integer)):
        return np.dtype(np.int64) if isinstance(v, int) else np.asarray(v).dtype
    if isinstance(v, float):
        return np.dtype(np.float64)
    return np.asarray(v).dtype


def _promote(xd, sd):
    """Julia's promotion for real scalars: an integer with a float takes the float."""
    xd, sd = np.dtype(xd), np.dtype(sd)
    if np.issubdtype(xd, np.integer) and np.issubdtype(sd, np.floating):
        return sd
    if np.issubdtype(sd, np.integer) and np.issubdtype(xd, np.floating):
        return xd
    return np.result_type(xd, sd)


def grad2traj(x, dt, *, dim: int = 1):
    """Trapezoid running integral along ``dim`` with a leading zero:
    ``y[1] = 0``, ``y[j+1] = y[j] + (x[j] + x[j+1]) * (dt/2)`` (conversion.jl:14-41).
    The result type follows Julia's arithmetic (``dt/2`` is Float64 for an Int dt)."""
    x = np.asarray(x)
    ax = _axis(dim, x.ndim, "dim")
    sd = _scalar_dtype(dt)
    half_t = np.dtype(np.float64) if np.issubdtype(sd, np.integer) else sd
    R = _promote(x.dtype, half_t)
    a = np.moveaxis(x, ax, 0)
    n = a.shape[0]
    out = np.zeros((max(n, 1),) + a.shape[1:], dtype=R)
    if n > 1:
        trap = (a[:-1] + a[1:]).astype(R) * np.asarray(dt / 2, dtype=half_t).astype(R)
        np.cumsum(trap, axis=0, dtype=R, out=out[1:])
    return np.moveaxis(out, 0, ax)


def traj2grad(x, dt, *, dim: int = 1):
    """Forward differences ``(x[j+1] - x[j]) / dt`` along ``dim`` (one sample shorter,
    conversion.jl:44-64); Julia's division result type (Int / Int is Float64)."""
    x = np.asarray(x)
    ax = _axis(dim, x.ndim, "dim")
    sd = _scalar_dtype(dt)
    R = _promote(x.dtype, sd)
    if np.issubdtype(R, np.integer):
        R = np.dtype(np.float64)
    a = np.moveaxis(x, ax, 0)
    d = (a[1:] - a[:-1]).astype(R) / np.asarray(dt).astype(R)
    return np.moveaxis(d, 0, ax)
