"""GPU CGNR reconstruction (recon_HOOp) against the oracle.

- test_ReconstructionStability.jl:25-76: 3x3 grid, 37 samples, 2 coils,
  complex random-phase weights, 30 CGNR iterations, lambda = 0, compared
  with a QR least-squares solve of the weighted dense system (3e-4 F32 bar).
- Workload 0 (64x64, 8 coils, 6000-sample spiral): 20 iterations against the
  FP64 textbook CG of the oracle after the same number of iterations. The bar comes
  from the measured amplification of the Krylov recurrence: the oracle CG is run a
  second time on the normal operator perturbed by (I + delta R) N (I + delta R)
  (R a seeded real diagonal, |R| <= 1, delta = 1e-6); the image moves by
  amp * delta, and the GPU image may differ from the FP64 one by at most
  10 * amp * e_op, e_op being the GPU normal operator's measured error against the
  FP64 normal (floor 1e-7).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402

BACKENDS = ["direct", "mixed_fp16", "mixed_bf16"]


def _stability_problem():
    rng = np.random.default_rng(514)
    ns, nc = 37, 2
    data = (rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc))).astype(np.complex64)
    weight = ((0.5 + rng.uniform(size=ns)) * np.exp(1j * rng.uniform(size=ns))).astype(np.complex64)
    grid = hoe.Grid(3, 3, 1, 1.0, 1.0, 1.0)
    kspha = (rng.normal(size=(9, ns)) * 0.01).astype(np.float32)
    for j in range(ns):
        kspha[1, j] = ((j % 3) - 1) / 3
        kspha[2, j] = (((j // 3) % 3) - 1) / 3
        kspha[3, j] = 0
    times = np.linspace(0, 1e-3, ns).astype(np.float32)
    csm = (np.exp(1j * rng.uniform(size=(9, nc))) / np.sqrt(nc)).astype(np.complex64)
    return grid, kspha, times, csm, data, weight


@pytest.mark.parametrize("backend", BACKENDS)
def test_cgnr_matches_qr(backend, gpu_lib):
    grid, kspha, times, csm, data, weight = _stability_problem()
    m = O.make_model(grid, kspha, times, None, csm, None)
    A = O.dense_matrix_f64(m)
    wr = np.tile(weight.astype(np.complex128), 2)
    expected = np.linalg.lstsq(A * wr[:, None], data.reshape(-1, order="F") * wr, rcond=None)[0]
    params = {"solver": hoe.CGNR, "reg": hoe.L2Regularization(0.0), "reconSize": grid.matrixSize,
              "iterations": 30, "relTol": 1e-10}
    with hoe.HighOrderExplicitOp(grid, kspha, times, csm=csm, backend=backend) as op:
        img, log = hoe.recon_HOOp(op, data, weight, params, return_log=True)
    assert img.shape == grid.matrixSize
    assert img.dtype == np.complex64
    err = O.relerr(img.reshape(-1, order="F"), expected)
    print(f"CGNR vs QR {backend}: {err:.2e}, residual log {log[:4]} ... {log[-1]:.2e}")
    assert err < 3e-4
    assert log[0] == 1.0


@pytest.fixture(scope="module")
def workload0_cg():
    c = synth.make(0)
    m = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], None)
    y = O.forward_f64(m, c["image"])
    rng = np.random.default_rng(11)
    y = (y + 0.01 * np.linalg.norm(y) / np.sqrt(y.size) * (rng.normal(size=y.size) + 1j * rng.normal(size=y.size)))
    iters, lam = 20, 1e-3
    rhs = O.adjoint_f64(m, y.astype(np.complex64).astype(np.complex128))
    nf = lambda v: O.normal_f64(m, v)
    x64, log64 = O.cgnr(nf, rhs, iters, lam)
    delta = 1e-6
    s = 1.0 + delta * np.random.default_rng(9).uniform(-1.0, 1.0, rhs.size)
    xp, _ = O.cgnr(lambda v: s * nf(s * v), rhs, iters, lam)
    amp = O.relerr(xp, x64) / delta
    print(f"workload0 CG: FP64 residual {log64[-1]:.3e}, amplification {amp:.2f}", flush=True)
    return c, m, y, iters, lam, x64, log64, amp


@pytest.mark.parametrize("backend", BACKENDS)
def test_cg_workload0_matches_fp64_cg(workload0_cg, backend, gpu_lib):
    c, m, y, iters, lam, x64, log64, amp = workload0_cg
    data = y.astype(np.complex64).reshape(m.nc, m.ns).T          # (nSam, nCha)
    weight = np.ones(m.ns, np.complex64)
    params = {"solver": hoe.CGNR, "reg": hoe.L2Regularization(lam), "reconSize": c["grid"].matrixSize,
              "iterations": iters}
    v = (np.random.default_rng(4).standard_normal(c["grid"].n)
         + 1j * np.random.default_rng(5).standard_normal(c["grid"].n)).astype(np.complex64)
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                 backend=backend) as op:
        img, log = hoe.recon_HOOp(op, data, weight, params, return_log=True)
        e_op = O.relerr(np.asarray(hoe.normalOperator(op) * v), O.normal_f64(m, v))
    err = O.relerr(img.reshape(-1, order="F"), x64)
    bar = 10.0 * amp * max(e_op, 1e-7)
    print(f"workload0 CG {backend}: image vs FP64 CG {err:.2e}, operator error {e_op:.2e}, amplification "
          f"{amp:.2f}, bar {bar:.2e}; residual {log[-1]:.3e} vs {log64[-1]:.3e}")
    assert err <= bar
    assert abs(log[-1] - log64[-1]) <= 0.05 * log64[-1] + 1e-6


@pytest.mark.parametrize("backend", ["direct", "mixed_fp16"])
def test_new_weights_wait_for_queued_work(backend, gpu_lib):
    """A CG queued with HOE_ASYNC keeps the weights it was queued with when the host
    sets new ones right after (the upload waits for the handle's queued work): the
    result equals a synchronous CG with the first weights, bit for bit."""
    import ctypes as C
    import torch
    from hoe import _lib
    c = synth.make(0)
    ns = c["kspha"].shape[1]
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(5)
    y = torch.from_numpy((rng.standard_normal((8, ns)) + 1j * rng.standard_normal((8, ns))).astype(np.complex64)
                         .reshape(-1)).to(dev)
    d1 = np.linspace(0.5, 2.0, ns).astype(np.float32)
    d2 = np.full(ns, 1e-3, dtype=np.float32)
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                 backend=backend) as op:
        lib, h = op._lib, op._handle()
        x_sync = torch.empty(c["grid"].n, dtype=torch.complex64, device=dev)
        x_async = torch.empty_like(x_sync)
        flags = _lib.HOE_DEVICE_PTRS
        _lib.check(lib.hoe_cg(h, y.data_ptr(), d1.ctypes.data, ns, 30, 0.0, 0.0, x_sync.data_ptr(), None, flags), h)
        torch.cuda.synchronize()
        _lib.check(lib.hoe_cg(h, y.data_ptr(), d1.ctypes.data, ns, 30, 0.0, 0.0, x_async.data_ptr(), None,
                              flags | _lib.HOE_ASYNC), h)
        _lib.check(lib.hoe_set_weights(h, d2.ctypes.data, ns), h)
        _lib.check(lib.hoe_sync(h), h)
        torch.cuda.synchronize()
        assert torch.equal(x_sync, x_async)
