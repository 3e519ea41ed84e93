"""E, Eᴴ and EᴴE at the full sizes of the 2-D and 3-D workloads (GPU).

Workload 1: 200x200 (40 000 voxels), 32 coils, 16 000-sample spiral, 16 terms + dB0.
Workload 4: 96x96x48 (442 368 voxels), 32 coils, 384 000-sample stack of spirals.

The FP64 oracle cannot evaluate a whole product at these sizes in seconds, so the
checks are the ones that do not depend on size (test/test_explicit_gpu.jl:38-92
states them for small shapes):
- rows of E x at 96 sample indices spread over the readout, against the oracle's
  forward on exactly those samples (the full image, all voxels and coils);
- entries of Eᴴ y at 96 voxels, against the oracle's adjoint restricted to those
  voxels over all samples, rescaled from 1/sqrt(96) to 1/sqrt(nVox);
- the adjoint identity <E x, y> = <x, Eᴴ y>;
- the normal operator against Eᴴ(E x) from the separate products.
Bars: 3e-5 relative L2 for the sampled rows and entries (the Float32 GPU-operator
bar of test/test_explicit_gpu.jl:51) and 1e-5 for the identities; the BF16 hi/lo
backend (about 16 significant bits per product) is held to the 1e-4 mixed-precision
target for both.

The device memory an operator holds after the four products (the drop in free device
memory from before its construction) stays under 3 GB at both sizes: partial sums that
would pass 512 MiB per contraction run as passes over tile ranges instead.
"""
import numpy as np
import pytest

import hoe
import oracle as O
from hoe import synth

pytestmark = pytest.mark.gpu

N_PICK = 96


@pytest.fixture(scope="module", params=[1, 4])
def workload(request):
    c = synth.make(request.param)
    rng = np.random.default_rng(100 + request.param)
    n, ns, nc = c["grid"].n, c["kspha"].shape[1], c["csm"].shape[-1]
    y = ((rng.standard_normal(ns * nc) + 1j * rng.standard_normal(ns * nc)) / np.sqrt(2)).astype(np.complex64)
    return request.param, c, c["image"].astype(np.complex64).reshape(-1), y, n, ns, nc


def _picks(n, seed):
    """N_PICK indices: both ends, an even spread and random ones."""
    rng = np.random.default_rng(seed)
    idx = np.concatenate([[0, n - 1], np.linspace(0, n - 1, N_PICK // 2).astype(int),
                          rng.choice(n, N_PICK, replace=False)])
    return np.unique(idx)[:N_PICK]


BARS = {"mixed_fp16": (3e-5, 1e-5), "direct": (3e-5, 1e-5), "mixed_bf16": (1e-4, 1e-4)}


def _free_device_bytes():
    import torch
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info(0)[0]


@pytest.mark.parametrize("backend", ["mixed_fp16", "direct", "mixed_bf16"])
def test_full_size_products(workload, backend, gpu_lib):
    w, c, x, y, n, ns, nc = workload
    print(f"workload {w} {backend}: {n} voxels, {ns} samples, {nc} coils", flush=True)
    mask = c.get("mask")
    free0 = _free_device_bytes()
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                 mask=mask, backend=backend) as op:
        ex = np.asarray(op * x)
        ehy = np.asarray(op.H * y)
        nx = np.asarray(hoe.normalOperator(op) * x)
        ehex = np.asarray(op.H * ex)
        held = free0 - _free_device_bytes()
    print(f"workload {w} {backend}: operator holds {held / 1e9:.2f} GB of device memory", flush=True)
    nv = n if mask is None else int(np.count_nonzero(mask))

    # sampled rows of E x
    srow = _picks(ns, 7)
    ms = O.make_model(c["grid"], c["kspha"][:, srow], c["times"][srow], c["fieldmap"], c["csm"], mask)
    y64 = O.forward_f64(ms, x).reshape(nc, srow.size)
    e_fwd = O.relerr(ex.reshape(nc, ns)[:, srow], y64)

    # sampled entries of Eᴴ y (the oracle's 1/sqrt(nVox) is that of its voxel subset)
    full = np.ones(n, bool) if mask is None else np.asarray(mask).reshape(-1, order="F").astype(bool)
    vpick = np.nonzero(full)[0][_picks(nv, 8)]
    sub = np.zeros(n, bool)
    sub[vpick] = True
    mv = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], sub)
    x64 = O.adjoint_f64(mv, y)[vpick] * np.sqrt(vpick.size / nv)
    e_adj = O.relerr(ehy[vpick], x64)

    # adjoint identity and normal consistency
    lhs = np.vdot(ex.astype(np.complex128), y)
    rhs = np.vdot(x.astype(np.complex128), ehy)
    e_dot = abs(lhs - rhs) / (np.linalg.norm(ex) * np.linalg.norm(y))
    e_norm = O.relerr(nx, ehex)
    print(f"workload {w} {backend}: E x rows vs FP64 {e_fwd:.2e}, Eᴴy entries vs FP64 {e_adj:.2e}, "
          f"adjoint identity {e_dot:.2e}, EᴴE vs Eᴴ(E x) {e_norm:.2e}", flush=True)
    bar, bar_id = BARS[backend]
    assert held < 3e9, held
    assert np.count_nonzero(ehy[~full]) == 0
    assert e_fwd < bar, e_fwd
    assert e_adj < bar, e_adj
    assert e_dot < bar_id, e_dot
    assert e_norm < bar_id, e_norm
