"""Gradient and CG checks at the sizes of the differentiation, 2-D and 3-D workloads (GPU).

- Workload 3 in full (200x200, 32 coils, 16 000-sample spiral), delay tau = 3.3 us:
  dL/dh on 64 sample rows spread over the readout against the FP64 oracle gradient of
  exactly the sample rows those depend on. Rows of E are independent, and with a delay
  dL/dh(j) = sum_t w_t dL/dk(j - m + 1 - t) gathers the 4 Catmull-Rom taps, so the
  oracle evaluates the 4 x 64 rows it needs (all voxels and coils).
- Workload 4 in full (96x96x48, 32 coils, 384 000 samples), tau = 0.4 dwell: the same
  row check.
- A 1 024-sample slice of workload 3, tau = 3 us: dL/dtau against central differences
  of the FP64 oracle loss (Richardson-extrapolated over two step sizes), whose delay
  model resamples the traces exactly as the operator does.
- A 1 024-sample slice of workload 1: 30 CGNR iterations against the FP64 textbook CG
  of the oracle after the same number of iterations. The bar comes from the measured
  amplification of the Krylov recurrence: the oracle CG is run a second time on the
  normal operator perturbed by a fixed Hermitian relative perturbation of size delta
  ((I + delta R) N (I + delta R), R a seeded real diagonal with |R| <= 1); the image
  moves by amp * delta. The GPU image may then differ from the FP64 one by at most
  10 * amp * e_op, e_op being the GPU normal operator's own measured error against
  the FP64 normal on the same problem (with a floor of 1e-7 for e_op).

- Workload 1 in full (200x200, 32 coils, 16 000 samples): 30 CGNR iterations of every
  FP32 backend (hoe_cg) against the same textbook CG run in FP64 over the FP64 backend's
  normal operator, whose products match the FP64 oracle at 1e-11 (test_gpu_fp64.py, a
  slice of this workload included): the FP64 oracle cannot run 30 normals of this size
  in seconds. Here the CG's own FP32 vectors dominate the difference: the same CG over
  the exact FP64 normal, with x, r, p and q rounded to complex64 after every update
  (FP64 inner products, as the device CG computes them), lands d32 = 2.5e-3 from the
  FP64 image, while the operators' error alone moves it by less than 1e-3 (FP16, FP32).
  The bar is d32 + 10 * amp * e_op, amp being the diagonal-perturbation amplification
  per unit of measured operator change; the final residuals agree within 10 %, and the
  mixed and FP32 images within 2 * d32.

The y data of the full-size cases comes from the GPU forward with perturbed coefficients
(1 % order 1, 5 % orders 2-3) and a 3 us delay; it is data only, the gradient formula
holds for any y. Bars: 1e-4 relative (the gradient bar of the FP32 operator products) for every backend.
"""
import numpy as np
import pytest

import hoe
import oracle as O
from hoe import synth

pytestmark = pytest.mark.gpu

BACKENDS = ["mixed_fp16", "direct", "mixed_bf16"]


def _op(c, backend, kspha=None):
    return hoe.HighOrderExplicitOp(c["grid"], c["kspha"] if kspha is None else kspha, c["times"],
                                   fieldmap=c["fieldmap"], csm=c["csm"], backend=backend, dwell=c["dwell"])


def _observed(c, backend="mixed_fp16"):
    """y from the perturbed coefficients and a 3 us delay, computed on the GPU."""
    kp = synth.perturb(c)
    with _op(c, backend, kp) as op:
        op.set_coefficients(None, 3e-6)
        return np.asarray(op * c["image"])


def _tap_rows(rows, tau, dwell, ns):
    """Samples s whose effective coefficients depend on trace sample j (the 4 Catmull-Rom
    taps of the delay), with their weights: dL/dh(j) = sum_t w_t dL/dk(s_t)."""
    u = tau / dwell
    m = int(np.floor(u))
    w, _ = O.catmull_rom_weights(u - m)
    taps = {j: [(j - m + 1 - t, w[t]) for t in range(4)] for j in rows}
    for j in rows:
        assert all(0 <= s < ns for s, _ in taps[j]), "pick rows away from the clamped ends"
    return taps


def _oracle_rows_grad(c, y, tau, rows):
    """FP64 dL/dh at trace samples ``rows`` ([L][len(rows)]) from the oracle gradient of the
    sample rows they depend on."""
    ns = c["kspha"].shape[1]
    taps = _tap_rows(rows, tau, c["dwell"], ns)
    need = np.unique([s for j in rows for s, _ in taps[j]])
    m_full = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], None)
    k_eff = O.delayed(m_full.k64.T, tau, c["dwell"])              # [L][ns]
    sub = O.make_model(c["grid"], c["kspha"][:, need], c["times"][need], c["fieldmap"], c["csm"], None)
    sub = O.with_coefficients(sub, k_eff[:, need])
    nc = c["csm"].shape[-1]
    ysub = np.asarray(y).reshape(nc, ns)[:, need].reshape(-1).astype(np.complex128)
    gk, _, _ = O.grad_f64(sub, c["image"], ysub)                    # [L][len(need)]
    pos = {s: i for i, s in enumerate(need)}
    g = np.zeros((gk.shape[0], len(rows)))
    for i, j in enumerate(rows):
        for s, wt in taps[j]:
            g[:, i] += wt * gk[:, pos[s]]
    return g


def _rows(ns, seed):
    rng = np.random.default_rng(seed)
    rows = np.concatenate([np.linspace(16, ns - 16, 32).astype(int), rng.integers(16, ns - 16, 32)])
    return np.unique(rows)


@pytest.fixture(scope="module")
def workload3():
    c = synth.make(3)
    return c, _observed(c)


@pytest.mark.parametrize("backend", BACKENDS)
def test_gradient_rows_workload3_full(workload3, backend, gpu_lib):
    c, y = workload3
    tau = 3.3e-6
    rows = _rows(c["kspha"].shape[1], 31)
    with _op(c, backend) as op:
        op.set_coefficients(None, tau)
        g, gtau, loss = op.grad(c["image"], y)
    g64 = _oracle_rows_grad(c, y, tau, rows)
    e = np.linalg.norm(g[:, rows] - g64) / np.linalg.norm(g64)
    print(f"workload 3 full {backend}: dL/dh on {len(rows)} rows vs FP64 {e:.2e} (tau 3.3 us, loss {loss:.6e})",
          flush=True)
    assert np.isfinite(gtau)
    assert e < 1e-4


@pytest.mark.parametrize("backend", BACKENDS)
def test_dtau_fd_workload3_slice(backend, gpu_lib):
    ns, tau = 1024, 3e-6
    c = synth.make(3, n_sam=ns)
    kp = synth.perturb(c)
    m = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], None)
    mp = O.make_model(c["grid"], kp, c["times"], c["fieldmap"], c["csm"], None)
    y = O.forward_f64(O.with_coefficients(mp, O.delayed(mp.k64.T, tau, c["dwell"])), c["image"])
    y32 = y.astype(np.complex64)
    y128 = y32.astype(np.complex128)
    with _op(c, backend) as op:
        op.set_coefficients(None, tau)
        g, gtau, loss = op.grad(c["image"], y32)

    def loss64(t):
        r = O.forward_f64(O.with_coefficients(m, O.delayed(m.k64.T, t, c["dwell"])), c["image"]) - y128
        return float(np.vdot(r, r).real)

    # tau = 3 dwells is a knot of the Catmull-Rom delay model: the loss is C1 there but its
    # second derivative jumps, so a central difference carries an O(eps) error term;
    # Richardson extrapolation of two step sizes cancels it
    et = 1e-3 * c["dwell"]
    fd1 = (loss64(tau + et) - loss64(tau - et)) / (2 * et)
    fd2 = (loss64(tau + et / 2) - loss64(tau - et / 2)) / et
    fd = 2.0 * fd2 - fd1
    l64 = loss64(tau)
    print(f"workload 3 slice {backend}: dL/dtau {gtau:.8e} vs FD {fd:.8e} ({abs(gtau - fd) / abs(fd):.2e}; "
          f"plain central differences {fd1:.8e}, {fd2:.8e}); loss {loss:.8e} vs {l64:.8e}", flush=True)
    assert abs(gtau - fd) <= 1e-4 * abs(fd)
    assert abs(loss - l64) <= 1e-5 * l64


@pytest.mark.parametrize("backend", ["mixed_fp16", "direct"])
def test_gradient_rows_workload4_full(backend, gpu_lib):
    c = synth.make(4)
    y = _observed(c)
    tau = 0.4 * c["dwell"]
    rows = _rows(c["kspha"].shape[1], 41)
    with _op(c, backend) as op:
        op.set_coefficients(None, tau)
        g, gtau, loss = op.grad(c["image"], y)
    g64 = _oracle_rows_grad(c, y, tau, rows)
    e = np.linalg.norm(g[:, rows] - g64) / np.linalg.norm(g64)
    print(f"workload 4 full {backend}: dL/dh on {len(rows)} rows vs FP64 {e:.2e}", flush=True)
    assert np.isfinite(gtau)
    assert e < 1e-4


def _perturbed_normal(apply_normal, n, delta, seed):
    r = np.random.default_rng(seed).uniform(-1.0, 1.0, n)
    s = 1.0 + delta * r
    return lambda v: s * apply_normal(s * v)


@pytest.fixture(scope="module")
def cg_slice():
    ns, iters = 1024, 30
    c = synth.make(1, n_sam=ns)
    m = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], None)
    y = O.forward_f64(m, c["image"])
    y32 = y.astype(np.complex64)
    rhs = O.adjoint_f64(m, y32.astype(np.complex128))
    nf = lambda v: O.normal_f64(m, v)
    x64, log64 = O.cgnr(nf, rhs, iters)
    delta = 1e-6
    xp, _ = O.cgnr(_perturbed_normal(nf, rhs.size, delta, 9), rhs, iters)
    amp = O.relerr(xp, x64) / delta
    print(f"CG slice: FP64 30-iteration residual {log64[-1]:.3e}, amplification {amp:.2f}", flush=True)
    return c, m, y32, x64, log64, amp


@pytest.mark.parametrize("backend", BACKENDS)
def test_cg30_workload1_slice(cg_slice, backend, gpu_lib):
    c, m, y32, x64, log64, amp = cg_slice
    ns, nc = c["kspha"].shape[1], c["csm"].shape[-1]
    data = y32.reshape(nc, ns).T
    params = {"solver": hoe.CGNR, "reg": hoe.L2Regularization(0.0), "reconSize": c["grid"].matrixSize,
              "iterations": 30}
    rng = np.random.default_rng(4)
    v = (rng.standard_normal(c["grid"].n) + 1j * rng.standard_normal(c["grid"].n)).astype(np.complex64)
    with _op(c, backend) as op:
        img, log = hoe.recon_HOOp(op, data, np.ones(ns, np.complex64), params, return_log=True)
        e_op = O.relerr(np.asarray(hoe.normalOperator(op) * v), O.normal_f64(m, v))
    err = O.relerr(img.reshape(-1, order="F"), x64)
    bar = 10.0 * amp * max(e_op, 1e-7)
    print(f"CG slice {backend}: image vs FP64 CG {err:.2e}, operator error {e_op:.2e}, amplification {amp:.2f}, "
          f"bar {bar:.2e}; residual {log[-1]:.4e} vs {log64[-1]:.4e}", flush=True)
    assert err <= bar
    assert abs(log[-1] - log64[-1]) <= 5e-2 * log64[-1]


def _cg_fp32_vectors(apply_normal, rhs, iters):
    """The textbook CG with its vectors rounded to complex64 after every update and FP64
    inner products (the arithmetic of the device CG), over an FP64 normal."""
    c64 = np.complex64
    x = np.zeros(rhs.size, c64)
    r = rhs.astype(c64)
    p = r.copy()
    rr = np.vdot(r.astype(np.complex128), r).real
    rr0 = rr
    for _ in range(iters):
        q = np.asarray(apply_normal(p.astype(np.complex128))).astype(c64)
        a = rr / np.vdot(p.astype(np.complex128), q).real
        x = (x + c64(a) * p).astype(c64)
        r = (r - c64(a) * q).astype(c64)
        rn = np.vdot(r.astype(np.complex128), r).real
        p = (r + c64(rn / rr) * p).astype(c64)
        rr = rn
    return x.astype(np.complex128), np.sqrt(rr / rr0)


@pytest.fixture(scope="module")
def cg_full():
    """30-iteration FP64 CG over the FP64 backend's normal at workload 1, its
    amplification per unit of operator change, and the FP32-vector CG's deviation."""
    iters = 30
    c = synth.make(1)
    g64 = hoe.Grid(c["grid"].nX, c["grid"].nY, c["grid"].nZ, float(c["grid"].dx), float(c["grid"].dy),
                   float(c["grid"].dz), dtype=np.float64)
    op64 = hoe.HighOrderExplicitOp(g64, c["kspha"].astype(np.float64), c["times"].astype(np.float64),
                                   fieldmap=c["fieldmap"].astype(np.float64), csm=c["csm"].astype(np.complex128),
                                   backend="fp64")
    y32 = np.asarray(op64 * c["image"].astype(np.complex128).reshape(-1)).astype(np.complex64)
    rhs = np.asarray(op64.H * y32.astype(np.complex128))
    n64 = hoe.normalOperator(op64)
    nf = lambda v: np.asarray(n64 * v)
    x64, log64 = O.cgnr(nf, rhs, iters)
    rng = np.random.default_rng(4)
    v = rng.standard_normal(c["grid"].n) + 1j * rng.standard_normal(c["grid"].n)
    nv = nf(v)
    s = 1.0 + 1e-6 * np.random.default_rng(9).uniform(-1.0, 1.0, rhs.size)
    xp, _ = O.cgnr(lambda u: s * nf(s * u), rhs, iters)
    amp = O.relerr(xp, x64) / O.relerr(s * nf(s * v), nv)
    x32, res32 = _cg_fp32_vectors(nf, rhs, iters)
    d32 = O.relerr(x32, x64)
    print(f"workload 1 CG: FP64 residual {log64[-1]:.3e}, amplification per operator change {amp:.1f}, "
          f"FP32-vector CG {d32:.2e} (residual {res32:.3e})", flush=True)
    yield c, y32, x64, log64, amp, d32, v, nv, iters
    op64.close()


def test_cg30_workload1_full(cg_full, gpu_lib):
    c, y32, x64, log64, amp, d32, v, nv, iters = cg_full
    ns, nc = c["kspha"].shape[1], c["csm"].shape[-1]
    data = y32.reshape(nc, ns).T
    params = {"solver": hoe.CGNR, "reg": hoe.L2Regularization(0.0), "reconSize": c["grid"].matrixSize,
              "iterations": iters}
    imgs = {}
    for backend in BACKENDS:
        with _op(c, backend) as op:
            img, log = hoe.recon_HOOp(op, data, np.ones(ns, np.complex64), params, return_log=True)
            e_op = O.relerr(np.asarray(hoe.normalOperator(op) * v.astype(np.complex64)), nv)
        imgs[backend] = img.reshape(-1, order="F")
        err = O.relerr(imgs[backend], x64)
        bar = d32 + 10.0 * amp * max(e_op, 1e-7)
        print(f"workload 1 CG {backend}: image vs FP64 CG {err:.2e}, operator error {e_op:.2e}, bar {bar:.2e}; "
              f"residual {log[-1]:.4e} vs {log64[-1]:.4e}", flush=True)
        assert err <= bar
        assert abs(log[-1] - log64[-1]) <= 0.1 * log64[-1]
    d = O.relerr(imgs["mixed_fp16"], imgs["direct"])
    print(f"workload 1 CG: mixed vs FP32 image {d:.2e}", flush=True)
    assert d <= 2 * d32
