"""GPU gradient of L(h, tau) = ||E(h, tau) x - y||^2 against the FP64 oracle
gradient and central finite differences.

The reference has no gradient; the model derivative it uses for delays is
the tau-sensitivity product of array_operator.jl:187-235 (FindDelay.jl:179-201).
Here tau resamples every coefficient trace at s + tau/dwell with uniform
Catmull-Rom interpolation; t*fieldmap is not delayed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402

BACKENDS = ["direct", "mixed_fp16", "mixed_bf16"]


def _small(seed=6, nc=3):
    rng = np.random.default_rng(seed)
    grid = hoe.Grid(6, 5, 1, np.float32(0.02), np.float32(0.02), np.float32(0.02))
    ns = 40
    s = np.linspace(0, 1, ns)
    kspha = np.zeros((16, ns), np.float32)
    for r in range(16):
        kspha[r] = (rng.normal() * np.cos(3 * s + r) + rng.normal() * s) * (5.0 if r < 4 else 20.0)
    times = (np.arange(ns) * 1e-5).astype(np.float32)
    fm = rng.uniform(-50, 50, grid.n).astype(np.float32)
    csm = (rng.normal(size=(grid.n, nc)) + 1j * rng.normal(size=(grid.n, nc))).astype(np.complex64)
    x = (rng.normal(size=grid.n) + 1j * rng.normal(size=grid.n)).astype(np.complex64)
    return grid, kspha, times, fm, csm, x, 1e-5


def _loss_f64(m, traces, tau, dwell, x, y):
    mm = O.with_coefficients(m, O.delayed(traces, tau, dwell))
    r = O.forward_f64(mm, x) - y
    return float(np.vdot(r, r).real)


# 48 coils: the mixed backends take the FP32 gradient kernel (the MFMA gradient is built
# for one 32-coil group)
@pytest.mark.parametrize("nc", [3, 48])
@pytest.mark.parametrize("backend", BACKENDS)
def test_gradient_small_vs_oracle_and_fd(backend, nc, gpu_lib):
    grid, kspha, times, fm, csm, x, dwell = _small(nc=nc)
    m = O.make_model(grid, kspha, times, fm, csm, None)
    traces = m.k64.T.copy()
    y = O.forward_f64(O.with_coefficients(m, O.delayed(traces * 1.02, 0.4 * dwell, dwell)), x)
    y32 = y.astype(np.complex64)
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, backend=backend, dwell=dwell) as op:
        # tau = 0: dL/dk equals dL/dh
        g, gtau, loss = op.grad(x, y32)
        g64, loss64, _ = O.grad_f64(m, x, y32.astype(np.complex128))
        e = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        print(f"grad {backend}: vs FP64 {e:.2e}, loss {loss:.6e} vs {loss64:.6e}")
        assert e < 1e-4
        assert abs(loss - loss64) < 1e-5 * loss64
        # nonzero delay: gradient through the interpolation and dL/dtau
        tau = 0.3 * dwell
        op.set_coefficients(None, tau)
        g, gtau, loss = op.grad(x, y32)
        y128 = y32.astype(np.complex128)
        rng = np.random.default_rng(3)
        gmax = np.max(np.abs(g))
        for _ in range(16):
            l, s = int(rng.integers(0, 16)), int(rng.integers(0, m.ns))
            eps = 1e-4 * max(1.0, abs(traces[l, s]))
            tp, tm = traces.copy(), traces.copy()
            tp[l, s] += eps
            tm[l, s] -= eps
            fd = (_loss_f64(m, tp, tau, dwell, x, y128) - _loss_f64(m, tm, tau, dwell, x, y128)) / (2 * eps)
            assert abs(fd - g[l, s]) <= 1e-4 * max(abs(fd), 1e-3 * gmax), (l, s, fd, g[l, s])
        et = 1e-3 * dwell
        fdt = (_loss_f64(m, traces, tau + et, dwell, x, y128) - _loss_f64(m, traces, tau - et, dwell, x, y128)) / (2 * et)
        print(f"dL/dtau {backend}: {gtau:.6e} vs FD {fdt:.6e}")
        assert abs(gtau - fdt) <= 1e-4 * abs(fdt)


@pytest.mark.parametrize("backend", BACKENDS)
def test_gradient_workload3_subset(backend, gpu_lib):
    """Workload 3 geometry (200x200, 32 coils) on 512 spiral samples: y from
    perturbed coefficients and a 3 us delay; the gradient at the unperturbed
    coefficients (tau = 0) against the FP64 oracle gradient and central
    differences at 64 seeded coordinates (only the sample row of a
    coordinate changes, so each difference needs one row of the forward)."""
    ns = 512
    c = synth.make(3, n_sam=ns)
    kp = synth.perturb(c)
    m = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], None)
    mp = O.make_model(c["grid"], kp, c["times"], c["fieldmap"], c["csm"], None)
    y = O.forward_f64(O.with_coefficients(mp, O.delayed(mp.k64.T, 3e-6, c["dwell"])), c["image"])
    y32 = y.astype(np.complex64)
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                 backend=backend, dwell=c["dwell"]) as op:
        g, gtau, loss = op.grad(c["image"], y32)
    g64, loss64, r64 = O.grad_f64(m, c["image"], y32.astype(np.complex128))
    e = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    print(f"workload3 subset {backend}: grad vs FP64 {e:.2e}, loss {loss:.6e} vs {loss64:.6e}")
    assert e < 1e-4
    rng = np.random.default_rng(64)
    gmax = np.max(np.abs(g64))
    k0 = m.k64.T
    worst = 0.0
    for _ in range(64):
        l, s = int(rng.integers(0, 16)), int(rng.integers(0, ns))
        row = O.make_model(c["grid"], c["kspha"][:, s:s + 1], c["times"][s:s + 1], c["fieldmap"], c["csm"], None)
        yrow = y32.reshape(m.nc, ns)[:, s].astype(np.complex128)
        eps = 1e-4 * max(1.0, abs(k0[l, s]))
        vals = []
        for sign in (1, -1):
            kk = k0[:, s:s + 1].copy()
            kk[l, 0] += sign * eps
            rr = O.forward_f64(O.with_coefficients(row, kk), c["image"]) - yrow
            vals.append(float(np.vdot(rr, rr).real))
        fd = (vals[0] - vals[1]) / (2 * eps)
        err = abs(fd - g[l, s]) / max(abs(fd), 1e-3 * gmax)
        worst = max(worst, err)
    print(f"workload3 subset {backend}: worst FD relative error over 64 coordinates {worst:.2e}")
    assert worst < 1e-4


@pytest.mark.parametrize("backend", ["mixed_fp16", "direct"])
def test_device_pointer_gradient_io(backend, gpu_lib):
    """dL/dh written to device memory equals the host output bit for bit; coefficient
    traces set from device memory give the same records (hence bit-equal products and
    gradients) as the same traces set from host memory; hoe_set_kspha after a device
    upload keeps the concomitant rows of that upload."""
    import torch
    dev = torch.device("cuda:0")
    ns = 2048
    c = synth.make(2, n_sam=ns)
    kp = synth.perturb(dict(c, kspha=c["kspha"][:, :ns]))
    conc = c["concomitant"][:, :ns]
    kw = dict(fieldmap=c["fieldmap"], csm=c["csm"], concomitant=conc, backend=backend, dwell=c["dwell"])
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"][:, :ns], c["times"][:ns], **kw) as op:
        y = np.asarray(op * c["image"])
        traces = np.concatenate([kp, conc * 1.01], axis=0).astype(np.float32)      # (20, ns)
        op.set_coefficients(traces, 0.7 * c["dwell"])
        g_host, gt_host, l_host = op.grad(c["image"], y)
        y_host = np.asarray(op * c["image"])
        x_d, y_d = torch.from_numpy(c["image"]).to(dev), torch.from_numpy(y).to(dev)
        g_dev, gt_dev, l_dev = op.grad(x_d, y_d)
        assert g_dev.is_cuda and tuple(g_dev.shape) == g_host.shape
        assert np.array_equal(g_dev.cpu().numpy(), g_host)
        assert gt_dev == gt_host and l_dev == l_host
        # the same traces from device memory
        op.set_coefficients(np.zeros_like(traces), 0.0)
        op.set_coefficients(torch.from_numpy(traces).to(dev), 0.7 * c["dwell"])
        assert np.array_equal(np.asarray(op * c["image"]), y_host)
        g2, gt2, l2 = op.grad(c["image"], y)
        assert np.array_equal(g2, g_host) and gt2 == gt_host and l2 == l_host
        # set_kspha keeps the concomitant rows last set (from device memory): equal to a new
        # operator built with those rows
        op.set_kspha(c["kspha"][:, :ns])
        y_k = np.asarray(op * c["image"])
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"][:, :ns], c["times"][:ns],
                                 **dict(kw, concomitant=conc * 1.01)) as ref:
        assert np.array_equal(y_k, np.asarray(ref * c["image"]))
