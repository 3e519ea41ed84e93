"""Contractions run as several passes over tile ranges (GPU).

A contraction whose partial sums would pass the partial-sum budget (512 MiB by
default; the environment variable HOE_PART_BUDGET_MIB overrides it when a backend's
schedules are built) runs as passes over disjoint tile ranges that share one buffer,
each pass reduced before the next one starts. With a 1 MiB budget the 200 x 200,
32-coil, 16 000-sample spiral runs one tile per pass: 63 forward passes of 256 samples
and 157 adjoint passes of 256 voxels. Checks:
- rows of E x and entries of Eᴴ y against the FP64 oracle, at the full-size bars
  (3e-5 relative, the Float32 GPU-operator bar of test/test_explicit_gpu.jl:51);
- E x, Eᴴ y and EᴴE x against the one-pass operator at 2e-6: only the order of the
  FP32 partial sums differs between the two;
- the gradient (its residual comes from the forward passes) and the loss against the
  one-pass operator at 1e-5: the MFMA gradient splits the residual into hi/lo FP16
  fragments, whose rounding moves with the residual's last bits (its bar against
  the FP64 oracle is 1e-4);
- 8 weighted CGNR iterations (hoe_cg) against the one-pass operator at 1e-4.
"""
import numpy as np
import pytest

import hoe
import oracle as O
from hoe import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spiral():
    c = synth.make(1)
    rng = np.random.default_rng(31)
    ns, nc = c["kspha"].shape[1], c["csm"].shape[-1]
    y = ((rng.standard_normal(ns * nc) + 1j * rng.standard_normal(ns * nc)) / np.sqrt(2)).astype(np.complex64)
    return c, c["image"].astype(np.complex64).reshape(-1), y


def _op(c, backend):
    return hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                   backend=backend)


def _products(op, c, x, y):
    ex = np.asarray(op * x)
    ehy = np.asarray(op.H * y)
    nx = np.asarray(hoe.normalOperator(op) * x)
    g, _, loss = op.grad(c["image"], y)
    ns = c["kspha"].shape[1]
    params = {"solver": hoe.CGNR, "reg": hoe.L2Regularization(0.0), "reconSize": c["grid"].matrixSize,
              "iterations": 8, "relTol": 0.0}
    w = (0.6 + 0.4 * np.cos(np.linspace(0, 3, ns))).astype(np.complex64)
    img = hoe.recon_HOOp(op, y.reshape(ns, -1, order="F"), w, params)
    return dict(ex=ex, ehy=ehy, nx=nx, g=np.asarray(g), loss=loss, cg=np.asarray(img).reshape(-1))


@pytest.mark.parametrize("backend", ["mixed_fp16", "direct"])
def test_one_tile_per_pass(spiral, backend, gpu_lib, monkeypatch):
    c, x, y = spiral
    ns, nc, n = c["kspha"].shape[1], c["csm"].shape[-1], c["grid"].n
    with _op(c, backend) as op:
        one = _products(op, c, x, y)
    monkeypatch.setenv("HOE_PART_BUDGET_MIB", "1")
    print(f"{backend}: 1 MiB partial-sum budget", flush=True)
    with _op(c, backend) as op:
        many = _products(op, c, x, y)
    monkeypatch.delenv("HOE_PART_BUDGET_MIB")

    rng = np.random.default_rng(5)
    srow = np.unique(np.concatenate([[0, ns - 1], rng.choice(ns, 62, replace=False)]))
    ms = O.make_model(c["grid"], c["kspha"][:, srow], c["times"][srow], c["fieldmap"], c["csm"], None)
    e_fwd = O.relerr(many["ex"].reshape(nc, ns)[:, srow], O.forward_f64(ms, x).reshape(nc, srow.size))
    vpick = np.unique(np.concatenate([[0, n - 1], rng.choice(n, 62, replace=False)]))
    sub = np.zeros(n, bool)
    sub[vpick] = True
    mv = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], sub)
    e_adj = O.relerr(many["ehy"][vpick], O.adjoint_f64(mv, y)[vpick] * np.sqrt(vpick.size / n))
    d = {k: O.relerr(many[k], one[k]) for k in ("ex", "ehy", "nx", "g", "cg")}
    d_loss = abs(many["loss"] - one["loss"]) / abs(one["loss"])
    print(f"{backend} passes: E x rows vs FP64 {e_fwd:.2e}, Eᴴy entries vs FP64 {e_adj:.2e}; vs one pass "
          + ", ".join(f"{k} {v:.1e}" for k, v in d.items()) + f", loss {d_loss:.1e}", flush=True)
    assert e_fwd < 3e-5 and e_adj < 3e-5
    for k in ("ex", "ehy", "nx"):
        assert d[k] < 2e-6, (k, d[k])
    assert d["g"] < 1e-5 and d_loss < 1e-5
    assert d["cg"] < 1e-4
