"""Field/data synchronisation over the libhoe operator.

- ``InterpTrajTime``: coefficient resampling onto ADC times with zero outside the
  shifted field interval (src/Synchronization/InterpTraj.jl:18-140).
- ``kspha_derivative``: the 5-tap time derivative of the coefficients
  (FindDelay.jl:178-180).
- ``FindDelay`` / ``FindDelay_multishot``: the model-based delay loop
  (FindDelay.jl:139-214, FindDelay_multishot.jl:65-159). Every product and the
  inner CGNR run on the GPU through one reused handle: ``set_kspha`` replaces
  the coefficients of each trial delay, ``set_kspha_dt`` switches the forward
  to the derivative product ``B x`` for the delay update.

The reference's default interpolant is Interpolations.jl's
``AkimaMonotonicInterpolation`` (Interpolations 0.16.3, not vendored). The
method here is Akima (1970), "A new method of interpolation and smooth curve
fitting based on local procedures", J. ACM 17:589-602: node slopes from the
four neighbouring interval slopes, two extra slopes at each end by linear
extrapolation of the interval slopes, and the mean of the two middle slopes
where both weights vanish; cubic Hermite pieces between nodes.

Density weights: the reference recomputes ``weight = samplingDensity(kspha_τ[2:3,:],
(nX, nY))`` from the resampled first-order rows at every update (FindDelay.jl:183,
FindDelay_multishot.jl:120), and so does the default here, with
``hoe.samplingDensity``. That function runs the published Pipe-Menon iteration on its
own Kaiser-Bessel grid; its parity with NFFT.jl's ``sdc`` (which the reference calls,
not vendored) is unpinned. ``weight`` overrides it: an array of one weight per sample,
or a callable of the resampled ``(nTerm, nSample)`` coefficients.
"""
from __future__ import annotations

import logging

import numpy as np

from ._lib import ArgumentError, DimensionMismatch
from .operator import CGNR, HighOrderExplicitOp, L2Regularization, recon_HOOp

log = logging.getLogger("hoe.sync")

STENCIL = np.array([1 / 8, 1 / 4, 0.0, -1 / 4, -1 / 8])


class AkimaMonotonicInterpolation:
    """Akima (1970) cubic interpolation (the reference's default ``intermode``)."""


class LinearMonotonicInterpolation:
    """Piecewise-linear interpolation."""


def _mode(intermode):
    if intermode is None:
        return "akima"
    if isinstance(intermode, str):
        m = intermode.lower()
    elif isinstance(intermode, type):
        m = intermode.__name__
    else:
        m = type(intermode).__name__
    m = {"akimamonotonicinterpolation": "akima", "linearmonotonicinterpolation": "linear"}.get(m.lower(), m.lower())
    if m not in ("akima", "linear"):
        raise ArgumentError(f"intermode {intermode!r} is not provided (AkimaMonotonicInterpolation, "
                            "LinearMonotonicInterpolation)")
    return m


def akima_tangents(x, y):
    """Node slopes of Akima's method for the columns of ``y`` ((n, m), n >= 3) at increasing ``x``."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = x.size
    if n < 3:
        raise ArgumentError("Akima interpolation needs at least 3 samples")
    d = np.empty((n + 3,) + y.shape[1:])
    d[2:-2] = np.diff(y, axis=0) / np.diff(x).reshape((-1,) + (1,) * (y.ndim - 1))
    d[1] = 2.0 * d[2] - d[3]
    d[0] = 2.0 * d[1] - d[2]
    d[-2] = 2.0 * d[-3] - d[-4]
    d[-1] = 2.0 * d[-2] - d[-3]
    # node i sees interval slopes d[i..i+3] = (d_{i-2}, d_{i-1}, d_i, d_{i+1})
    w_right = np.abs(d[3:] - d[2:-1])      # |d_{i+1} - d_i|
    w_left = np.abs(d[1:-2] - d[:-3])      # |d_{i-1} - d_{i-2}|
    den = w_right + w_left
    mean = 0.5 * (d[1:-2] + d[2:-1])
    with np.errstate(invalid="ignore", divide="ignore"):
        t = (w_right * d[1:-2] + w_left * d[2:-1]) / den
    return np.where(den > 0, t, mean)


def _eval_columns(x, y, t, xq, mode):
    """Cubic Hermite (``t``: node slopes) or linear interpolation of the float64 columns
    ``y`` at ``xq``; zero outside ``[x[0], x[-1]]``. The basis weights are formed once per
    query point and applied to all columns."""
    n = x.size
    inside = (xq >= x[0]) & (xq <= x[-1])
    i = np.clip(np.searchsorted(x, xq, side="right") - 1, 0, n - 2)
    h = x[i + 1] - x[i]
    s = (xq - x[i]) / h
    if mode == "linear":
        w0, w1 = 1.0 - s, s
        w0[~inside] = 0.0
        w1[~inside] = 0.0
        out = np.take(y, i, axis=0)
        out *= w0[:, None]
        tmp = np.take(y, i + 1, axis=0)
        tmp *= w1[:, None]
        out += tmp
        return out
    s2 = s * s
    s3 = s2 * s
    h00 = 2 * s3 - 3 * s2 + 1
    h10 = (s3 - 2 * s2 + s) * h
    h01 = -2 * s3 + 3 * s2
    h11 = (s3 - s2) * h
    for w in (h00, h10, h01, h11):
        w[~inside] = 0.0
    out = np.take(y, i, axis=0)
    out *= h00[:, None]
    tmp = np.take(t, i, axis=0)
    tmp *= h10[:, None]
    out += tmp
    np.take(y, i + 1, axis=0, out=tmp)
    tmp *= h01[:, None]
    out += tmp
    np.take(t, i + 1, axis=0, out=tmp)
    tmp *= h11[:, None]
    out += tmp
    return out


def interp_columns(x, y, xq, intermode="akima"):
    """Interpolate the columns of ``y`` ((n, m)) given at increasing ``x`` onto ``xq``;
    zero outside ``[x[0], x[-1]]`` (``extrapolate(itp, 0)``, InterpTraj.jl:41)."""
    mode = _mode(intermode)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xq = np.asarray(xq, dtype=np.float64).reshape(-1)
    if y.ndim == 1:
        return interp_columns(x, y[:, None], xq, mode)[:, 0]
    if x.size != y.shape[0]:
        raise DimensionMismatch("interpolation nodes and values")
    if not np.all(np.diff(x) > 0):
        raise ArgumentError("interpolation nodes must be strictly increasing")
    t = akima_tangents(x, y) if mode == "akima" else None
    return _eval_columns(x, y, t, xq, mode)


class _Resampler:
    """Field samples ``traj`` ((nFieldSample, nTerm)) at ``j*dt``, resampled at shifted
    times: sample j sits at ``j*dt - delTime`` (InterpTraj.jl:18-44). A shift leaves
    Akima's node slopes unchanged, so they are formed once on the unshifted nodes and
    reused for every ``delTime`` (the delay loops resample the same samples once per update)."""

    def __init__(self, traj, dt, intermode=None):
        self.dtype = np.asarray(traj).dtype
        self.y = np.asarray(traj, dtype=np.float64)
        self.mode = _mode(intermode)
        n = self.y.shape[0]
        if n < 2:
            raise ArgumentError("interpolation needs at least 2 field samples")
        self.x0 = float(dt) * np.arange(n)
        self.t = akima_tangents(self.x0, self.y) if self.mode == "akima" else None
        self.torch = None

    def to_device(self, torch):
        """Evaluate on the current GPU with torch from now on (the same FP64 formulas; the
        result still comes back as a numpy array): the delay loops' 2 x nTerm rows per update
        then cost a few small kernels instead of host passes over every sample."""
        dev = torch.device("cuda", torch.cuda.current_device())
        self.torch, self.dev = torch, dev
        self.x0_d = torch.from_numpy(self.x0).to(dev)
        self.y_d = torch.from_numpy(np.ascontiguousarray(self.y)).to(dev)
        self.t_d = torch.from_numpy(np.ascontiguousarray(self.t)).to(dev) if self.t is not None else None
        self.xq_key, self.xq_d = None, None
        return self

    def __call__(self, delTime, datatime):
        xq = np.asarray(datatime, dtype=np.float64).reshape(-1)
        if self.torch is not None:
            return self._eval_device(float(delTime), xq).astype(self.dtype)
        return _eval_columns(self.x0 - float(delTime), self.y, self.t, xq, self.mode).astype(self.dtype)

    def _eval_device(self, delTime, xq):
        """``_eval_columns`` on the device (float64 result)."""
        tc = self.torch
        if self.xq_key is None or self.xq_key.shape != xq.shape or not np.array_equal(self.xq_key, xq):
            self.xq_key, self.xq_d = xq.copy(), tc.from_numpy(np.ascontiguousarray(xq)).to(self.dev)
        x = self.x0_d - delTime
        q = self.xq_d
        n = x.numel()
        inside = (q >= x[0]) & (q <= x[-1])
        i = tc.clamp(tc.searchsorted(x, q, right=True) - 1, 0, n - 2)
        h = x[i + 1] - x[i]
        s = (q - x[i]) / h
        if self.mode == "linear":
            w0, w1 = 1.0 - s, s
            out = self.y_d[i] * w0[:, None] + self.y_d[i + 1] * w1[:, None]
        else:
            s2 = s * s
            s3 = s2 * s
            h00 = 2 * s3 - 3 * s2 + 1
            h10 = (s3 - 2 * s2 + s) * h
            h01 = -2 * s3 + 3 * s2
            h11 = (s3 - s2) * h
            out = (self.y_d[i] * h00[:, None] + self.t_d[i] * h10[:, None]) + \
                  (self.y_d[i + 1] * h01[:, None] + self.t_d[i + 1] * h11[:, None])
        out = tc.where(inside[:, None], out, tc.zeros_like(out))
        return out.cpu().numpy()


def InterpTrajTime(traj, dt, delTime, datatime=None, *, intermode=None):
    """Resample field coefficients onto ADC times (InterpTraj.jl:18-140).

    2-D: ``traj`` ``(nFieldSample, nTerm)``; sample j sits at ``j*dt - delTime``;
    ``datatime`` defaults to ``dt*(0:nFieldSample-1)``; returns
    ``(len(datatime), nTerm)`` in ``traj``'s dtype. 3-D: ``traj``
    ``(nShot, nFieldSample, nTerm)``, ``delTime`` one offset per shot, ``datatime``
    ``(nShot, nADC)``; returns ``(nShot, nADC, nTerm)``, each shot resampled on its own.
    """
    traj = np.asarray(traj)
    if not np.issubdtype(traj.dtype, np.floating):
        raise ArgumentError("traj must be floating point")
    if traj.ndim == 3:
        n_shot = traj.shape[0]
        dl = np.asarray(delTime).reshape(-1)
        if dl.size != n_shot:
            raise AssertionError("Number of shots in trajectory must match length of delTime vector")
        dtm = np.asarray(datatime)
        if dtm.ndim != 2 or dtm.shape[0] != n_shot:
            raise AssertionError("Number of shots in datatime must match length of delTime vector")
        out = np.empty((n_shot, dtm.shape[1], traj.shape[2]), dtype=traj.dtype)
        for i in range(n_shot):
            out[i] = InterpTrajTime(traj[i], dt, float(dl[i]), dtm[i], intermode=intermode)
        return out
    if traj.ndim != 2:
        raise DimensionMismatch("traj must be (nFieldSample, nTerm) or (nShot, nFieldSample, nTerm)")
    n = traj.shape[0]
    if datatime is None:
        datatime = float(dt) * np.arange(n)
    return _Resampler(traj, dt, intermode)(delTime, datatime)


def kspha_derivative(kspha, dt):
    """d kspha / dt by the 5-tap stencil [1/8, 1/4, 0, -1/4, -1/8] / dt along the field
    samples (axis -2: ``(nFieldSample, nTerm)`` or ``(nShot, nFieldSample, nTerm)``),
    with zeros beyond both ends: ``conv(kspha, kernel)[3:end-2, :] / dt``
    (FindDelay.jl:178-180, FindDelay_multishot.jl:106-111)."""
    k = np.asarray(kspha)
    T = k.dtype.type
    x = k.astype(np.float64)
    pad = [(0, 0)] * x.ndim
    pad[-2] = (2, 2)
    xp = np.pad(x, pad)
    n = x.shape[-2]
    sl = lambda a: xp[(Ellipsis, slice(a, a + n), slice(None))]   # noqa: E731
    # full convolution cropped by 2 on each side: out[i] = sum_j c_j x[i + 2 - j]
    out = sum(STENCIL[j] * sl(4 - j) for j in range(5))
    return (out / float(T(dt))).astype(T)


def density_weight(gridding):
    """The reference's per-update weights: ``samplingDensity`` of the resampled first-order
    rows 2:3 (FindDelay.jl:183, FindDelay_multishot.jl:120)."""
    from .fieldcam import samplingDensity

    def weight(k_tau):
        return samplingDensity(np.asarray(k_tau)[1:3], (gridding.nX, gridding.nY))
    return weight


def _weights(weight, k_tau, n):
    w = weight(k_tau) if callable(weight) else weight
    w = np.asarray(w).reshape(-1)
    if w.size != n:
        raise DimensionMismatch("weight must have one entry per sample")
    return w


def _delay_step(y1, y2):
    """``real(y2 \\ y1)``: the least-squares scalar a minimising ||a y2 - y1||."""
    y1 = np.asarray(y1, dtype=np.complex128).reshape(-1)
    y2 = np.asarray(y2, dtype=np.complex128).reshape(-1)
    return float((np.vdot(y2, y1) / np.vdot(y2, y2)).real)


class _Signals:
    """The measured signal ``yv`` (nSam*nCha, coil-major) and the update's products
    (FindDelay.jl:194-196). With torch and a visible GPU they stay on the operator's
    device: the reconstruction reads the signal there, ``y - A x`` and ``B x`` are formed
    there, and only the scalar step comes back (no host round trip of the signals per
    update). Otherwise they live on the host."""

    def __init__(self, yv):
        self.yv = yv
        self.torch = None
        self.y = None
        try:
            import torch
            if torch.cuda.is_available():
                self.torch = torch
        except ImportError:
            pass

    def data(self, op):
        """The signal for ``recon_HOOp`` on ``op``'s device (or the host array)."""
        if self.torch is None:
            return self.yv
        if self.y is None:
            dev = self.torch.device("cuda", op.device)
            self.y = self.torch.from_numpy(np.ascontiguousarray(self.yv.astype(op._cdtype))).to(dev)
        return self.y

    def step(self, r, x, kdt_tau):
        """``real((B x) \\ (y - A x))`` with ``A`` = ``r`` and ``B`` its derivative model."""
        if self.torch is None:
            y1 = self.yv - r * x
            r.set_kspha_dt(kdt_tau)
            y2 = r * x
            r.set_kspha_dt(None)
            return _delay_step(y1, y2)
        t = self.torch
        y = self.data(r)
        xd = t.from_numpy(np.ascontiguousarray(np.asarray(x).astype(r._cdtype))).to(y.device)
        y1 = (y - r * xd).to(t.complex128)
        r.set_kspha_dt(kdt_tau)
        y2 = (r * xd).to(t.complex128)
        r.set_kspha_dt(None)
        return float((t.vdot(y2, y1) / t.vdot(y2, y2)).real)


class _Loop:
    """The update rule shared by both estimators (FindDelay.jl:196-210)."""

    def __init__(self, JumpFact, Δτ_min, dt, max_updates):
        self.J = JumpFact
        self.Δτ_min = float(Δτ_min)
        self.dt = float(dt)
        self.max_updates = max_updates
        self.τ = 0.0
        self.Δτ = self.Δτ_prev = np.inf
        self.n = 1
        self.history = [0.0]

    def running(self):
        if self.max_updates is not None and self.n > self.max_updates:
            return False
        return abs(self.Δτ) > self.Δτ_min

    def update(self, step):
        self.Δτ = self.J * step
        self.τ += self.Δτ
        log.info("Iteration %d: Δτ = %.5f [dt] | τ = %.5f [dt] | JumpFact = %s", self.n, self.Δτ / self.dt,
                 self.τ / self.dt, self.J)
        if self.n > 1 and np.sign(self.Δτ_prev) != np.sign(self.Δτ):
            self.J = max(1, self.J / 2)
        self.Δτ_prev = self.Δτ
        self.n += 1
        self.history.append(self.τ)

    def result(self, T):
        return T(self.τ - self.Δτ / self.J * (self.J - 1))


def _rec_params(gridding, reg, iter_max, solver):
    return {"reconSize": (gridding.nX, gridding.nY), "reg": reg, "iterations": iter_max, "solver": solver}


def FindDelay(gridding, data, kspha, *args, JumpFact=3, Δτ_min=0.005, intermode=None, fieldmap=None, csm=None,
              recon_terms="111", nBlock=50, arrayType=None, solver=CGNR, reg=None, iter_max=10, verbose=False,
              weight=None, backend="mixed", max_updates=None, return_history=False):
    """Estimate the field/data delay τ (seconds) of one 2-D acquisition (FindDelay.jl:139-214).

    Call as ``FindDelay(gridding, data, kspha, datatime, StartTime, dt; ...)`` or
    ``FindDelay(gridding, data, kspha, StartTime, dt; ...)`` (then ``datatime =
    dt*(0:nSample-1)``). ``data`` is ``(nSample, nCha)``, ``kspha``
    ``(nFieldSample, nTerm)``, ``fieldmap`` ``(nX, nY)`` in Hz, ``csm`` ``(nX, nY, nCha)``.
    Each update reconstructs x with ``iter_max`` CGNR iterations, then takes
    ``Δτ = JumpFact * real((B x) \\ (y - E x))``; ``JumpFact`` halves (not below 1)
    when the update changes sign; the loop stops when ``|Δτ| <= Δτ_min`` and returns
    ``τ - Δτ/J*(J-1)``. ``nBlock`` and ``arrayType`` are accepted and unused.
    Beyond the reference: ``weight`` (default: the reference's density weights, see the
    module docstring), ``backend``,
    ``max_updates`` (None: no limit, as in the reference) and ``return_history``
    (also return the list of τ after each update).
    """
    data = np.asarray(data)
    if len(args) == 3:
        datatime, StartTime, dt = args
    elif len(args) == 2:
        StartTime, dt = args
        datatime = None
    else:
        raise ArgumentError("FindDelay(gridding, data, kspha, [datatime,] StartTime, dt; ...)")
    if data.ndim != 2:
        raise DimensionMismatch("data must be (nSample, nCha)")
    nSample, nCha = data.shape
    T = np.float32 if data.dtype in (np.complex64, np.float32) else np.float64
    if datatime is None:
        datatime = (float(dt) * np.arange(nSample)).astype(T)
    datatime = np.asarray(datatime).reshape(-1)
    if csm is None:
        csm = np.ones((gridding.nX, gridding.nY, 1), dtype=np.complex64)
    csm = np.asarray(csm)
    if data.shape[1] != csm.shape[-1]:
        raise AssertionError("data and csm must have the same number of coil channels")
    if data.shape[0] != datatime.size:
        raise AssertionError("data and datatime must have the same number of spatial points")
    if fieldmap is None:
        fieldmap = np.zeros((gridding.nX, gridding.nY), dtype=np.float32)
    reg = L2Regularization(0.0) if reg is None else reg
    weight = density_weight(gridding) if weight is None else weight
    kspha = np.asarray(kspha)
    kspha_dt = kspha_derivative(kspha, dt)
    rec = _rec_params(gridding, reg, iter_max, solver)
    sig = _Signals(data.reshape(-1, order="F"))
    loop = _Loop(JumpFact, Δτ_min, dt, max_updates)
    if kspha.ndim != 2:
        raise DimensionMismatch("traj must be (nFieldSample, nTerm) or (nShot, nFieldSample, nTerm)")
    # InterpTrajTime of the same samples at every update: node slopes formed once
    res_k, res_dt = _Resampler(kspha, dt, intermode), _Resampler(kspha_dt, dt, intermode)
    if sig.torch is not None:
        res_k.to_device(sig.torch)
        res_dt.to_device(sig.torch)
    op = None
    try:
        while loop.running():
            k_tau = res_k(loop.τ + float(StartTime), datatime)[:nSample].T
            kdt_tau = res_dt(loop.τ + float(StartTime), datatime)[:nSample].T
            w = _weights(weight, k_tau, nSample)
            if op is None:
                op = HighOrderExplicitOp(gridding, k_tau, datatime, fieldmap=fieldmap, csm=csm,
                                         recon_terms=recon_terms, backend=backend, verbose=verbose)
            else:
                op.set_kspha(k_tau)
            x = recon_HOOp(op, sig.data(op), w, rec).reshape(-1, order="F")
            loop.update(sig.step(op, x, kdt_tau))   # real((B x) \ (y - A x))
    finally:
        if op is not None:
            op.close()
    tau = loop.result(T)
    return (tau, loop.history) if return_history else tau


def FindDelay_multishot(gridding, data, kspha, *args, JumpFact=3, Δτ_min=0.005, intermode=None, fieldmap=None,
                        csm=None, recon_terms="111", nBlock=50, arrayType=None, solver=CGNR, reg=None, iter_max=10,
                        verbose=False, weight=None, backend="mixed", max_updates=None, return_history=False):
    """One delay shared by all shots (FindDelay_multishot.jl:65-159).

    Call as ``FindDelay_multishot(gridding, data, kspha, datatime, StartTime, dt; ...)``
    or without ``datatime`` (then each shot uses ``dt*(0:nFieldSample-1)``). ``kspha``
    is ``(nShot, nFieldSample, nTerm)``, ``datatime`` ``(nShot, nADC)``, ``StartTime``
    one offset per shot, ``data`` ``(nADC*nShot, nCha)`` with ADC samples fastest.
    The image is reconstructed with every shot's own times. The residual and the
    derivative product are evaluated, as in the reference's per-shot operators
    (FindDelay_multishot.jl:128-135), with the first shot's ADC times for every shot;
    the two agree when all shots share one time vector.
    """
    data = np.asarray(data)
    kspha = np.asarray(kspha)
    if kspha.ndim != 3:
        raise DimensionMismatch("kspha must be (nShot, nFieldSample, nTerm)")
    n_shot, n_field, n_term = kspha.shape
    if len(args) == 3:
        datatime, StartTime, dt = args
    elif len(args) == 2:
        StartTime, dt = args
        datatime = None
    else:
        raise ArgumentError("FindDelay_multishot(gridding, data, kspha, [datatime,] StartTime, dt; ...)")
    T = np.float32 if data.dtype in (np.complex64, np.float32) else np.float64
    if datatime is None:
        datatime = np.tile((float(dt) * np.arange(n_field)).astype(T), (n_shot, 1))
    datatime = np.asarray(datatime)
    if csm is None:
        csm = np.ones((gridding.nX, gridding.nY, 1), dtype=np.complex64)
    csm = np.asarray(csm)
    if data.shape[1] != csm.shape[-1]:
        raise AssertionError("data and csm must have the same number of coil channels")
    if datatime.shape[0] != n_shot:
        raise AssertionError("kspha and datatime must have the same number of shots")
    n_adc = datatime.shape[1]
    if data.shape[0] != n_adc * n_shot:
        raise DimensionMismatch("data must hold nADC*nShot samples")
    if fieldmap is None:
        fieldmap = np.zeros((gridding.nX, gridding.nY), dtype=np.float32)
    reg = L2Regularization(0.0) if reg is None else reg
    weight = density_weight(gridding) if weight is None else weight
    StartTime = np.asarray(StartTime, dtype=np.float64).reshape(-1)
    kspha_dt = kspha_derivative(kspha, dt)
    rec = _rec_params(gridding, reg, iter_max, solver)
    n_sam = n_adc * n_shot
    t_recon = np.ascontiguousarray(datatime.reshape(-1))            # ADC fastest, then shots
    t_first = np.tile(datatime[0], n_shot)
    same_times = np.array_equal(t_recon, t_first)
    sig = _Signals(data.reshape(-1, order="F"))
    loop = _Loop(JumpFact, Δτ_min, dt, max_updates)
    if StartTime.size != n_shot:
        raise AssertionError("Number of shots in trajectory must match length of delTime vector")
    # InterpTrajTime of each shot's samples at every update: node slopes formed once per shot
    res_k = [_Resampler(kspha[i], dt, intermode) for i in range(n_shot)]
    res_dt = [_Resampler(kspha_dt[i], dt, intermode) for i in range(n_shot)]
    if sig.torch is not None:
        for r in res_k + res_dt:
            r.to_device(sig.torch)
    op = op_res = None
    try:
        while loop.running():
            k_tau = np.stack([res_k[i](loop.τ + StartTime[i], datatime[i]) for i in range(n_shot)])
            kdt_tau = np.stack([res_dt[i](loop.τ + StartTime[i], datatime[i]) for i in range(n_shot)])
            k_r = k_tau.reshape(n_sam, n_term).T          # (nTerm, nADC*nShot), ADC fastest
            kdt_r = kdt_tau.reshape(n_sam, n_term).T
            w = _weights(weight, k_r, n_sam)
            if op is None:
                op = HighOrderExplicitOp(gridding, k_r, t_recon, fieldmap=fieldmap, csm=csm,
                                         recon_terms=recon_terms, backend=backend, verbose=verbose)
                if not same_times:
                    op_res = HighOrderExplicitOp(gridding, k_r, t_first, fieldmap=fieldmap, csm=csm,
                                                 recon_terms=recon_terms, backend=backend, verbose=verbose)
            else:
                op.set_kspha(k_r)
                if op_res is not None:
                    op_res.set_kspha(k_r)
            x = recon_HOOp(op, sig.data(op), w, rec).reshape(-1, order="F")
            loop.update(sig.step(op if op_res is None else op_res, x, kdt_r))
    finally:
        for o in (op, op_res):
            if o is not None:
                o.close()
    tau = loop.result(T)
    return (tau, loop.history) if return_history else tau
