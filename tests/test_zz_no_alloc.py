"""The normal operator without allocation (hoe_normal_noalloc) against hoe_normal.

hoe_normal_noalloc runs the same kernels in the same order as hoe_normal with its
workspaces taken from one static device array per GPU, so its output must equal
hoe_normal's bit for bit. Two calls in a row on different operators (the second one
smaller) check that nothing depends on the array's previous contents; an operator
whose workspaces do not fit is rejected with the required size in the message.
Both calls go through the C ABI with device vectors.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import hoe  # noqa: E402
from hoe import _lib, synth  # noqa: E402


def _normal_pair(op, x, weights):
    """(hoe_normal, hoe_normal_noalloc) of x with the stored weights, device vectors."""
    import torch
    lib, h = op._lib, op._handle()
    D = np.ascontiguousarray(weights, dtype=np.float32)
    _lib.check(lib.hoe_set_weights(h, D.ctypes.data, D.size), h)
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    ref = torch.full_like(xd, complex(7.0, -7.0))
    out = torch.full_like(xd, complex(-3.0, 3.0))
    torch.cuda.synchronize()
    _lib.check(lib.hoe_normal(h, None, _lib.HOE_CURRENT_WEIGHTS, xd.data_ptr(), ref.data_ptr(),
                              _lib.HOE_DEVICE_PTRS), h)
    _lib.check(lib.hoe_normal_noalloc(h, xd.data_ptr(), out.data_ptr(), _lib.HOE_DEVICE_PTRS), h)
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def _case(config, backend, n_sam, n_cha, seed):
    cfg = synth.make(config, n_sam=n_sam, n_cha=n_cha, seed=seed)
    op = hoe.HighOrderExplicitOp(cfg["grid"], cfg["kspha"], cfg["times"], fieldmap=cfg["fieldmap"],
                                 csm=cfg["csm"], backend=backend)
    rng = np.random.default_rng(seed)
    w = (0.5 + rng.random(op.nSam)).astype(np.float32)
    return op, cfg["image"], w


@pytest.mark.parametrize("backend", ["mixed_fp16", "direct"])
def test_noalloc_equals_normal_on_two_calls(backend, gpu_lib):
    # first: the 200 x 200, 32-coil, 16 000-sample spiral; second: a 64 x 64, 8-coil,
    # 3 000-sample spiral (smaller model, smaller vectors, fewer segments)
    cases = [(1, 16000, 32, 11), (0, 3000, 8, 12)]
    for config, n_sam, n_cha, seed in cases:
        op, x, w = _case(config, backend, n_sam, n_cha, seed)
        with op:
            ref, out = _normal_pair(op, x, w)
            print(f"noalloc {backend} workload {config} n_sam {n_sam}: max |ref| {np.abs(ref).max():.3e}", flush=True)
            assert np.isfinite(ref).all() and np.abs(ref).max() > 0
            assert np.array_equal(ref, out), np.abs(ref - out).max()


def test_noalloc_rejects_oversized_and_host_vectors(gpu_lib):
    import torch
    # 64 x 64 image, 32 coils, 800 000 samples: the weighted signal and its MFMA
    # fragments alone need more than the static array
    cfg = synth.make(0, n_cha=32, seed=5)
    ns = 800000
    kspha = np.tile(cfg["kspha"], (1, ns // cfg["kspha"].shape[1] + 1))[:, :ns]
    times = (np.arange(ns) * 2e-6).astype(np.float32)
    with hoe.HighOrderExplicitOp(cfg["grid"], kspha, times, csm=cfg["csm"], backend="mixed_fp16") as op:
        lib, h = op._lib, op._handle()
        x = torch.zeros(op.ncol, dtype=torch.complex64, device="cuda:0")
        o = torch.empty_like(x)
        rc = lib.hoe_normal_noalloc(h, x.data_ptr(), o.data_ptr(), _lib.HOE_DEVICE_PTRS)
        assert rc == _lib.HOE_EINVAL
        msg = lib.hoe_last_error(h).decode()
        assert "MiB" in msg and "hoe_normal" in msg, msg
        xh = np.zeros(op.ncol, np.complex64)
        oh = np.zeros(op.ncol, np.complex64)
        rc = lib.hoe_normal_noalloc(h, xh.ctypes.data, oh.ctypes.data, 0)
        assert rc == _lib.HOE_EINVAL
        assert "HOE_DEVICE_PTRS" in lib.hoe_last_error(h).decode()


def test_noalloc_stale_nonfinite_scratch_does_not_leak(gpu_lib):
    """The direct kernels write partial sums only for their coil batches (8 of the 32 padded
    columns for 3 coils); the reductions must read only real coils. A first call with a NaN
    image leaves non-finite partial sums in the static array; a 3-coil operator's call
    afterwards must still equal hoe_normal."""
    import torch
    c = synth.make(1, n_sam=4000, n_cha=32, seed=21)
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], csm=c["csm"], backend="direct") as op:
        lib, h = op._lib, op._handle()
        x = torch.full((op.ncol,), complex(float("nan"), float("nan")), dtype=torch.complex64, device="cuda:0")
        o = torch.empty_like(x)
        _lib.check(lib.hoe_set_weights(h, None, 0), h)
        _lib.check(lib.hoe_normal_noalloc(h, x.data_ptr(), o.data_ptr(), _lib.HOE_DEVICE_PTRS), h)
        torch.cuda.synchronize()
        assert torch.isnan(o).any()
    op, x3, w = _case(0, "direct", 2000, 3, 22)
    with op:
        ref, out = _normal_pair(op, x3, w)
        assert np.isfinite(ref).all() and np.isfinite(out).all()
        assert np.array_equal(ref, out)
