"""Host-side mirror of HighOrderMRI.jl's Explicit operator API over libhoe.

Reference interface (HighOrderMRI.jl):
  - ``HighOrderExplicitOp(grid, kspha, times; implementation, fieldmap, csm,
    mask, concomitant, recon_terms, k_nominal, gpus, backend, ...)``
    (src/EncodingOperator/HighOrderExplicitOp.jl:44-51,
    Explicit/kernel_operator.jl:199-315, explicit_gpu_operator.jl:299-392);
  - ``op * x`` (prod!), ``adjoint(op) * y`` (ctprod!), ``normalOperator(op, D)``
    and ``explicit_gpu_normal(op; weights)`` (explicit_gpu_operator.jl:399-433),
    ``close(op)`` (:435-446);
  - ``recon_HOOp(op, Data, weight, recParams)`` (src/Recon/recon_HOOp.jl:52-88).

Vectors may be numpy arrays (host; copied to and from the GPU) or torch
tensors on the handle's GPU (used in place, no copies). Every application runs
the gfx950 kernels of libhoe; there is no CPU execution path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch, check
from .grid import Grid

try:  # torch is only needed for device-resident vectors
    import torch  # noqa: F401
except Exception:  # pragma: no cover
    torch = None


def _is_torch(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor)


def _f32(a, name):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return a


_NAMES = {np.complex64: "complex64 (ComplexF32)", np.complex128: "complex128 (ComplexF64)"}


def _torch_cdtype(cdtype):
    return torch.complex64 if cdtype == np.complex64 else torch.complex128


class _Vec:
    """A vector argument: host numpy (copied) or device torch (in place), of the operator's
    complex type (complex64, or complex128 for the FP64 backend)."""

    def __init__(self, a, n: int, name: str, cdtype=np.complex64):
        self.dev = _is_torch(a) and a.is_cuda
        self.cdtype = cdtype
        if self.dev:
            if a.dtype != _torch_cdtype(cdtype):
                raise ArgumentError(f"{name} must be {_NAMES[cdtype]}")
            t = a.reshape(-1)
            if not t.is_contiguous():
                t = t.contiguous()
            self.obj = t
            self.ptr = t.data_ptr()
        else:
            if _is_torch(a):
                a = a.detach().numpy()
            h = np.ascontiguousarray(np.asarray(a, dtype=cdtype).reshape(-1))
            self.obj = h
            self.ptr = h.ctypes.data
        size = self.obj.numel() if self.dev else self.obj.size
        if size != n:
            raise DimensionMismatch(f"{name} has length {size}, expected {n}")

    def like(self, n: int):
        if self.dev:
            return torch.empty(n, dtype=_torch_cdtype(self.cdtype), device=self.obj.device)
        return np.empty(n, dtype=self.cdtype)


def _ptr(o):
    return o.data_ptr() if _is_torch(o) else o.ctypes.data


def _out(out, vec: _Vec, n: int, name: str = "out"):
    """The output vector: a new one like ``vec``, or the caller's ``out`` after checking
    that libhoe can write n complex64 values into it in place (same kind as the input,
    host or device, contiguous, exact length)."""
    if out is None:
        return vec.like(n)
    if vec.dev:
        if not (_is_torch(out) and out.is_cuda):
            raise ArgumentError(f"{name} must be a device tensor when the input is one")
        if out.device != vec.obj.device:
            raise ArgumentError(f"{name} must be on the input's device")
        if out.dtype != _torch_cdtype(vec.cdtype):
            raise ArgumentError(f"{name} must be {_NAMES[vec.cdtype]}")
        if not out.is_contiguous():
            raise ArgumentError(f"{name} must be contiguous")
        size = out.numel()
    else:
        if not isinstance(out, np.ndarray):
            raise ArgumentError(f"{name} must be a numpy array when the input is a host array")
        if out.dtype != vec.cdtype:
            raise ArgumentError(f"{name} must be {_NAMES[vec.cdtype]}")
        if not (out.flags.c_contiguous or out.flags.f_contiguous) or not out.flags.writeable:
            raise ArgumentError(f"{name} must be a contiguous, writeable array")
        size = out.size
    if size != n:
        raise DimensionMismatch(f"{name} has length {size}, expected {n}")
    return out


class HighOrderExplicitOp:
    """Explicit higher-order encoding operator of size ``(nSam*nCha, prod(matrixSize))``.

    Parameters mirror the reference constructor. ``backend`` selects the
    libhoe execution path: ``"direct"`` (FP32 VALU contraction), ``"mixed"``
    / ``"mixed_fp16"`` or ``"mixed_bf16"`` (MFMA phase-reuse), or ``"fp64"``
    (the Float64 operator: FP64 model and arithmetic, complex128 vectors; the
    reference's operators with ``Grid{Float64}``). ``dwell`` is the coefficient
    sampling interval used by ``set_coefficients(..., tau)``.
    """

    def __init__(self, grid: Grid, kspha, times, *, implementation: str = "kernel", fieldmap=None,
                 csm=None, mask=None, concomitant=None, recon_terms: str | None = None,
                 k_nominal=None, kspha_dt=None, gpus=None, backend: str = "direct",
                 dwell: float = 0.0, verbose: bool = False):
        if implementation not in ("kernel", "array"):
            raise ArgumentError("implementation must be :array or :kernel")
        if implementation == "array":
            raise ArgumentError("implementation=:array (the reference CPU array path) is not provided; "
                                "use implementation=:kernel")
        if backend not in _lib.BACKENDS:
            raise ArgumentError(f"unknown explicit backend {backend}")
        gpus = [0] if gpus is None else [int(g) for g in gpus]
        if len(gpus) == 0:
            raise ArgumentError("gpus must contain at least one device id")
        if len(set(gpus)) != len(gpus):
            raise ArgumentError(f"gpus contains duplicate device ids: {gpus}")
        if len(gpus) > 1:
            raise ArgumentError("one GPU per operator: sample-sharded multi-GPU execution is not built")
        lib = _lib.load()
        f64 = backend == "fp64"
        RT = np.float64 if f64 else np.float32
        self._cdtype = np.complex128 if f64 else np.complex64
        _real = (lambda a, name: np.ascontiguousarray(np.asarray(a, dtype=np.float64))) if f64 else _f32
        kspha = _real(kspha, "kspha")
        if kspha.ndim != 2:
            raise DimensionMismatch("kspha must be (nTerm, nSam)")
        n_term, n_sam = kspha.shape
        if n_term not in (9, 16):
            raise ArgumentError("kspha requires 9 or 16 terms")
        times = _real(times, "times").reshape(-1)
        if times.size != n_sam:
            raise DimensionMismatch("times and kspha")
        nX, nY, nZ = grid.matrixSize
        n_grid = nX * nY * nZ
        if mask is None:
            mask_u8 = np.ones(n_grid, dtype=np.uint8)
        else:
            m = np.asarray(mask)
            if m.size != n_grid:
                raise DimensionMismatch("mask and grid")
            mask_u8 = np.ascontiguousarray(m.reshape(-1, order="F") if m.ndim > 1 else m,
                                           dtype=np.uint8)
        fm = None
        if fieldmap is not None:
            f = np.asarray(fieldmap, dtype=RT)
            if f.size != n_grid:
                raise DimensionMismatch("fieldmap and mask")
            fm = np.ascontiguousarray(f.reshape(-1, order="F") if f.ndim > 1 else f)
        if csm is None:
            n_cha, cs = 1, None
        else:
            c = np.asarray(csm, dtype=self._cdtype)
            if c.ndim == 1:
                c = c.reshape(-1, 1)
            n_cha = c.shape[-1]
            if c.size != n_grid * n_cha:
                raise DimensionMismatch("csm and grid")
            cs = np.ascontiguousarray(c.reshape(n_grid, n_cha, order="F") if c.ndim > 2 else c)
            cs = np.ascontiguousarray(cs.T)   # [coil][grid]: grid fastest per coil
        kdt = None
        if kspha_dt is not None:
            if f64:
                raise ArgumentError("kspha_dt (the derivative product) is not provided on the FP64 backend")
            kdt = _f32(kspha_dt, "kspha_dt")
            if kdt.shape != kspha.shape:
                raise AssertionError("kspha_dt must have same size as kspha")
            # the derivative operator carries no concomitant terms (array_operator.jl:139)
            concomitant = None
        conc = None
        if concomitant is not None:
            ca = np.asarray(concomitant)
            if np.iscomplexobj(ca) or not np.issubdtype(ca.dtype, np.number):
                raise ArgumentError("concomitant must be a real array or nothing")
            if ca.shape != (4, n_sam):
                raise DimensionMismatch(f"concomitant must have size (4, {n_sam}), got {ca.shape}")
            conc = _real(ca, "concomitant")
        kn = None
        if k_nominal is not None:
            kn = _real(k_nominal, "k_nominal")
            if kn.shape[0] != 3:
                raise AssertionError("k_nominal must have 3 terms (row) for kx, ky, kz")
            if kn.shape != (3, n_sam):
                raise DimensionMismatch("k_nominal must be (3, nSam)")
        # column-major host arrays: Julia (nTerm, nSam) is term-fastest = numpy (nSam, nTerm) C order
        self._keep = [np.ascontiguousarray(grid.x, dtype=RT), np.ascontiguousarray(grid.y, dtype=RT),
                      np.ascontiguousarray(grid.z, dtype=RT), mask_u8, np.ascontiguousarray(kspha.T), times]
        d = _lib.ModelDescF64() if f64 else _lib.ModelDesc()
        d.n_x, d.n_y, d.n_z = nX, nY, nZ
        d.x, d.y, d.z = (self._keep[0].ctypes.data, self._keep[1].ctypes.data, self._keep[2].ctypes.data)
        d.mask = mask_u8.ctypes.data
        d.n_term, d.n_sam = n_term, n_sam
        d.kspha = self._keep[4].ctypes.data
        d.times = times.ctypes.data
        if fm is not None:
            self._keep.append(fm)
            d.fieldmap = fm.ctypes.data
        d.n_cha = n_cha
        if cs is not None:
            self._keep.append(cs)
            d.csm = cs.ctypes.data
        if conc is not None:
            ct = np.ascontiguousarray(conc.T)
            self._keep.append(ct)
            d.concomitant = ct.ctypes.data
        if recon_terms is not None:
            d.recon_terms = recon_terms.encode()
        if kn is not None:
            kt = np.ascontiguousarray(kn.T)
            self._keep.append(kt)
            d.k_nominal = kt.ctypes.data
        if not f64:
            d.dwell = float(dwell)
            d.backend = _lib.BACKENDS[backend]
        d.device = gpus[0]
        h = C.c_void_p()
        check(lib.hoe_create_f64(C.byref(d), C.byref(h)) if f64 else lib.hoe_create(C.byref(d), C.byref(h)))
        self._keep = None
        self._h = h
        self._lib = lib
        self.grid = grid
        self.nSam, self.nCha, self.nGrid = n_sam, n_cha, n_grid
        self.nTerm = n_term + (4 if concomitant is not None else 0)
        self._n_term_in = n_term
        info = _lib.Info()
        check(lib.hoe_get_info(h, C.byref(info)), h)
        self.nVox = int(info.n_vox)
        self._horner_last = int(info.horner)
        self.nrow, self.ncol = n_sam * n_cha, n_grid
        self.device = gpus[0]
        self.backend = backend
        self.symmetric = self.hermitian = False
        self._weights_owner = None
        self.kspha_dt = None
        if kdt is not None:
            self.set_kspha_dt(kdt)

    @property
    def horner(self):
        """Phase paths of the current model: bit 0 the Horner forward, bit 1 the Horner MFMA
        adjoint, bit 2 anchored phases in the MFMA Horner kernels (their fieldmap offsets and,
        for the coefficients last set, their Horner offsets within bounds). A closed operator
        reports its last value."""
        if self._h is None:
            return self._horner_last
        info = _lib.Info()
        check(self._lib.hoe_get_info(self._h, C.byref(info)), self._h)
        self._horner_last = int(info.horner)
        return self._horner_last

    # ---- LinearOperator protocol ----
    @property
    def shape(self):
        return (self.nrow, self.ncol)

    def size(self, dim=None):
        return self.shape if dim is None else self.shape[dim - 1]

    def storage_type(self):
        return "device"

    def _handle(self):
        if self._h is None:
            raise ArgumentError("operator is closed")
        return self._h

    def set_backend(self, backend: str):
        if backend not in _lib.BACKENDS:
            raise ArgumentError(f"unknown explicit backend {backend}")
        check(self._lib.hoe_set_backend(self._handle(), _lib.BACKENDS[backend]), self._h)
        self.backend = backend

    def set_kspha_dt(self, kspha_dt):
        """Set (``(nTerm, nSam)`` array) or clear (``None``) the trajectory derivative:
        while set, ``op * x`` is the derivative product B x (array_operator.jl:187-235)
        and ``op' * y`` stays E^H y, as for the reference's ``kspha_dt`` operator."""
        h = self._handle()
        if kspha_dt is None:
            check(self._lib.hoe_set_kspha_dt(h, None), h)
            self.kspha_dt = None
            return
        kdt = _f32(kspha_dt, "kspha_dt")
        if kdt.shape != (self._n_term_in, self.nSam):
            raise AssertionError("kspha_dt must have same size as kspha")
        kt = np.ascontiguousarray(kdt.T)          # term fastest
        check(self._lib.hoe_set_kspha_dt(h, kt.ctypes.data), h)
        self.kspha_dt = kdt

    def set_kspha(self, kspha, k_nominal=None):
        """Replace ``kspha`` (``(nTerm, nSam)``) and prepare it as the constructor does
        (``recon_terms``, ``k_nominal``; prep_kspha.jl:1-56). Concomitant terms, maps and
        times are kept and the delay is reset to 0: the operator equals one built anew
        with the new coefficients, as FindDelay does for every trial delay
        (FindDelay.jl:184-187)."""
        h = self._handle()
        k = _f32(kspha, "kspha")
        if k.shape != (self._n_term_in, self.nSam):
            raise DimensionMismatch(f"kspha must be ({self._n_term_in}, {self.nSam})")
        kt = np.ascontiguousarray(k.T)
        knp = None
        if k_nominal is not None:
            kn = _f32(k_nominal, "k_nominal")
            if kn.shape != (3, self.nSam):
                raise DimensionMismatch("k_nominal must be (3, nSam)")
            knp = np.ascontiguousarray(kn.T)
        check(self._lib.hoe_set_kspha(h, kt.ctypes.data, None if knp is None else knp.ctypes.data), h)

    def _on_device(self, *vecs):
        """Device vectors must live on the operator's GPU (multi_gpu_operator.jl:81-89)."""
        for v in vecs:
            if v.dev and v.obj.device.index != self.device:
                raise ArgumentError(f"device vectors must be on the operator's GPU (device {self.device})")

    def prod(self, x, out=None):
        """``op * x``: x of length prod(matrixSize) -> signal of length nSam*nCha
        (B x while ``kspha_dt`` is set)."""
        h = self._handle()
        xv = _Vec(x, self.ncol, "x", self._cdtype)
        self._on_device(xv)
        y = _out(out, xv, self.nrow)
        flags = _lib.HOE_DEVICE_PTRS if xv.dev else 0
        if xv.dev:
            torch.cuda.current_stream(xv.obj.device).synchronize()
        check(self._lib.hoe_forward(h, xv.ptr, _ptr(y), flags), h)
        return y

    def ctprod(self, y, out=None):
        """``adjoint(op) * y``."""
        h = self._handle()
        yv = _Vec(y, self.nrow, "y", self._cdtype)
        self._on_device(yv)
        x = _out(out, yv, self.ncol)
        flags = _lib.HOE_DEVICE_PTRS if yv.dev else 0
        if yv.dev:
            torch.cuda.current_stream(yv.obj.device).synchronize()
        check(self._lib.hoe_adjoint(h, yv.ptr, _ptr(x), flags), h)
        return x

    def __mul__(self, x):
        return self.prod(x)

    def __matmul__(self, x):
        return self.prod(x)

    @property
    def H(self):
        return AdjointOp(self)

    def adjoint(self):
        return AdjointOp(self)

    def set_coefficients(self, kspha=None, tau: float = 0.0):
        """Replace the coefficient traces ``(nTerm_eff, nSam)`` and the delay tau (s).
        ``kspha`` may be a host array or a float32 tensor on the operator's GPU (no host
        round trip: the records are rebuilt on the device)."""
        h = self._handle()
        ptr, flags, keep = None, 0, None
        if kspha is not None and _is_torch(kspha) and kspha.is_cuda:
            if kspha.dtype != torch.float32:
                raise ArgumentError("kspha must be float32")
            if tuple(kspha.shape) != (self.nTerm, self.nSam):
                raise DimensionMismatch(f"kspha must be ({self.nTerm}, {self.nSam})")
            if kspha.device.index != self.device:
                raise ArgumentError(f"device arrays must be on the operator's GPU (device {self.device})")
            keep = kspha.t().contiguous()           # term fastest
            torch.cuda.current_stream(keep.device).synchronize()
            ptr, flags = keep.data_ptr(), _lib.HOE_DEVICE_PTRS
        elif kspha is not None:
            k = _f32(kspha.detach().cpu().numpy() if _is_torch(kspha) else kspha, "kspha")
            if k.shape != (self.nTerm, self.nSam):
                raise DimensionMismatch(f"kspha must be ({self.nTerm}, {self.nSam})")
            keep = np.ascontiguousarray(k.T)
            ptr = keep.ctypes.data
        check(self._lib.hoe_set_coefficients(h, ptr, float(tau), flags), h)

    def grad(self, x, y):
        """Gradient of ``||E(h,tau) x - y||^2``: returns (g_h (nTerm, nSam), g_tau, loss).
        With device vectors g_h is a float32 tensor on the operator's GPU (a transposed
        view of the term-fastest buffer libhoe writes), else a numpy array."""
        h = self._handle()
        xv = _Vec(x, self.ncol, "x")
        yv = _Vec(y, self.nrow, "y")
        self._on_device(xv, yv)
        if xv.dev != yv.dev:
            raise ArgumentError("x and y must both be host or both be device arrays")
        flags = _lib.HOE_DEVICE_PTRS if xv.dev else 0
        gt, loss = C.c_double(), C.c_double()
        if xv.dev:
            torch.cuda.current_stream(xv.obj.device).synchronize()
            g = torch.empty((self.nSam, self.nTerm), dtype=torch.float32, device=xv.obj.device)
            check(self._lib.hoe_grad(h, xv.ptr, yv.ptr, g.data_ptr(), C.byref(gt), C.byref(loss), flags), h)
            return g.t(), gt.value, loss.value
        g = np.empty((self.nSam, self.nTerm), dtype=np.float32)
        check(self._lib.hoe_grad(h, xv.ptr, yv.ptr, g.ctypes.data, C.byref(gt), C.byref(loss), flags), h)
        return np.ascontiguousarray(g.T), gt.value, loss.value

    def last_kernel_ms(self, which: int = 2) -> float:
        return float(self._lib.hoe_last_kernel_ms(self._handle(), which))

    def close(self):
        if self._h is not None:
            self._lib.hoe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class AdjointOp:
    def __init__(self, op: HighOrderExplicitOp):
        self.op = op

    @property
    def shape(self):
        return (self.op.ncol, self.op.nrow)

    def __mul__(self, y):
        return self.op.ctprod(y)

    def __matmul__(self, y):
        return self.op.ctprod(y)

    def prod(self, y, out=None):
        return self.op.ctprod(y, out)

    @property
    def H(self):
        return self.op


@dataclass
class WeightingOp:
    """Diagonal weighting ``W`` of raw complex sample weights repeated over coils
    (the LinearOperatorCollection WeightingOp used at recon_HOOp.jl:77)."""
    weights: object
    rep: int = 1


@dataclass
class ComposedOp:
    """``W ∘ op``."""
    W: WeightingOp
    op: HighOrderExplicitOp


def compose(W: WeightingOp, op: HighOrderExplicitOp) -> ComposedOp:
    return ComposedOp(W, op)


def _check_diag(D, nrow, nsam, dtype=np.float32):
    D = np.asarray(D).reshape(-1)
    if D.size == 0:
        raise DimensionMismatch("normal weights")
    if not np.all(np.isfinite(D)):
        raise ArgumentError("normal weights must be finite")
    Dr = np.real(D).astype(np.float64)
    if np.iscomplexobj(D):
        scale = max(1.0, float(np.max(np.abs(D))))
        eps = np.finfo(np.float32).eps
        if float(np.max(np.abs(np.imag(D)))) > 10 * eps * scale:
            raise ArgumentError("normal weights must already be real nonnegative WᴴW")
    if np.min(Dr) < 0:
        raise ArgumentError("normal weights must be nonnegative")
    if D.size not in (nsam, nrow):
        raise DimensionMismatch("normal weights")
    return np.ascontiguousarray(Dr.astype(dtype))


class NormalOp:
    """``Aᴴ D A`` sharing the parent's handle (not reentrant)."""

    _gen = 0

    def __init__(self, op: HighOrderExplicitOp, D=None):
        op._handle()
        self.op = op
        self.D = None if D is None else _check_diag(D, op.nrow, op.nSam,
                                                     np.float64 if op.backend == "fp64" else np.float32)
        NormalOp._gen += 1
        self._id = NormalOp._gen
        self.symmetric, self.hermitian = False, True

    @property
    def shape(self):
        return (self.op.ncol, self.op.ncol)

    def _sync_weights(self, h):
        lib = self.op._lib
        if self.op._weights_owner != self._id:
            setw = lib.hoe_set_weights_f64 if self.op.backend == "fp64" else lib.hoe_set_weights
            if self.D is None:
                check(setw(h, None, 0), h)
            else:
                check(setw(h, self.D.ctypes.data, self.D.size), h)
            self.op._weights_owner = self._id

    def prod(self, x, out=None):
        h = self.op._handle()
        self._sync_weights(h)
        xv = _Vec(x, self.op.ncol, "x", self.op._cdtype)
        self.op._on_device(xv)
        o = _out(out, xv, self.op.ncol)
        flags = _lib.HOE_DEVICE_PTRS if xv.dev else 0
        if xv.dev:
            torch.cuda.current_stream(xv.obj.device).synchronize()
        check(self.op._lib.hoe_normal(h, None, _lib.HOE_CURRENT_WEIGHTS, xv.ptr, _ptr(o), flags), h)
        return o

    def __mul__(self, x):
        return self.prod(x)

    def __matmul__(self, x):
        return self.prod(x)


def normalOperator(op, weights=None):
    """``normalOperator(op, D)``: D is the already-squared real diagonal (length
    nSam or nSam*nCha); ``normalOperator(W ∘ op)`` takes ``|W.weights|²``
    (explicit_gpu_operator.jl:399-416)."""
    if isinstance(op, ComposedOp):
        w = np.asarray(op.W.weights).reshape(-1)
        D = (np.abs(w.astype(np.complex128)) ** 2)
        if D.size * op.W.rep == op.op.nrow:
            D = np.tile(D, op.W.rep)
        return NormalOp(op.op, D)
    if isinstance(weights, WeightingOp):
        weights = weights.weights
    return NormalOp(op, weights)


def explicit_gpu_normal(op: HighOrderExplicitOp, weights=None):
    """Normal ``Aᴴ D A`` with ``D = abs2.(weights)`` from raw weights (explicit_gpu_operator.jl:418-433)."""
    if weights is None:
        return NormalOp(op, None)
    w = np.asarray(weights).reshape(-1).astype(np.complex128)
    return NormalOp(op, np.abs(w) ** 2)


@dataclass
class L2Regularization:
    """``L2Regularization(λ)``: adds λI to the normal equations."""
    λ: float = 0.0


class CGNR:
    """Solver tag for recParams[:solver] (RegularizedLeastSquares.CGNR)."""


def recon_HOOp(HOOp: HighOrderExplicitOp, Data, weight, recParams: dict, *, return_log: bool = False):
    """Weighted CGNR reconstruction (src/Recon/recon_HOOp.jl:52-88).

    Solves ``(Aᴴ WᴴW A + λI) x = Aᴴ WᴴW y`` with x₀ = 0 on the GPU (FP64 inner
    products). ``Data`` is ``(nSam, nCha)``; ``weight`` holds one raw complex
    weight per sample (not squared). ``recParams`` must contain ``solver``,
    ``reg`` and ``reconSize``; ``iterations`` (default 10) and ``relTol``
    (default 0: run every iteration) are optional. Returns the image reshaped
    to ``reconSize`` (column-major) as a numpy array.
    """
    for key in ("solver", "reg", "reconSize"):
        if key not in recParams:
            raise ArgumentError(f"recParams must contain :{key}")
    solver = recParams["solver"]
    if solver is not CGNR and not isinstance(solver, CGNR):
        raise ArgumentError("only the CGNR solver is provided")
    reg = recParams["reg"]
    lam = float(reg.λ if isinstance(reg, L2Regularization) else reg)
    iters = int(recParams.get("iterations", 10))
    rel_tol = float(recParams.get("relTol", 0.0))
    op = HOOp
    h = op._handle()
    data = Data
    if _is_torch(data) and data.is_cuda:
        yv = _Vec(data.t().contiguous().reshape(-1) if data.dim() == 2 else data, op.nrow, "Data", op._cdtype)
    else:
        d = np.asarray(data)
        yv = _Vec(d.reshape(-1, order="F") if d.ndim == 2 else d, op.nrow, "Data", op._cdtype)
    op._on_device(yv)
    w = weight.detach().cpu().numpy() if _is_torch(weight) else np.asarray(weight)
    w = w.reshape(-1)
    if w.size != op.nSam:
        raise DimensionMismatch("weight must have one entry per sample")
    op._weights_owner = None
    x = yv.like(op.ncol)
    log = np.zeros(iters + 1, dtype=np.float64)
    flags = _lib.HOE_DEVICE_PTRS if yv.dev else 0
    if yv.dev:
        torch.cuda.current_stream(yv.obj.device).synchronize()
    if op.backend == "fp64":
        D = np.ascontiguousarray(np.abs(w.astype(np.complex128)) ** 2)
        check(op._lib.hoe_set_weights_f64(h, D.ctypes.data, D.size), h)
        check(op._lib.hoe_cg(h, yv.ptr, None, _lib.HOE_CURRENT_WEIGHTS, iters, lam, rel_tol, _ptr(x),
                             log.ctypes.data, flags), h)
    else:
        D = np.ascontiguousarray((np.abs(w.astype(np.complex128)) ** 2).astype(np.float32))
        check(op._lib.hoe_cg(h, yv.ptr, D.ctypes.data, D.size, iters, lam, rel_tol, _ptr(x), log.ctypes.data,
                             flags), h)
    img = x.detach().cpu().numpy() if _is_torch(x) else x
    img = np.asarray(img).reshape(recParams["reconSize"], order="F")
    return (img, log) if return_log else img
