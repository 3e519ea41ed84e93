"""The Float64 operator (backend "fp64", hoe_create_f64) against the FP64 oracle.

The reference runs its operators in Float64 as well and holds them to tighter bars:
- 2e-12 against an independent FP64 matrix for the four shapes of
  test/test_explicit_gpu.jl:38-103 (forward, adjoint, weighted normal);
- 1e-11 with concomitant terms (test/test_Concomitant.jl:262-283: forward, adjoint and
  the normal with a ramp of weights);
- 1e-10 for 30 CGNR iterations against QR least squares with complex random-phase
  sample weights (test/test_ReconstructionStability.jl:25-76).
The same bars hold here, with the oracle built from the same Float64 inputs
(``make_model(..., f64=True)``: FP64 coordinates, coefficients, times, maps). A
256-sample slice of the 200 x 200, 32-coil spiral checks the FP64 kernels at a
realistic size (Horner-run phases, 2 coil batches) against the oracle's FP64 products
at 1e-11. The FP64 handle takes complex128 vectors only and refuses the entry points it
does not provide.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402

TOL64 = 2e-12


def _grid64(g):
    return hoe.Grid(g.nX, g.nY, g.nZ, float(g.dx), float(g.dy), float(g.dz), dtype=np.float64)


def _case(shape, nt, nc, seed):
    rng = np.random.default_rng(seed)
    grid = hoe.Grid(*shape, 0.07, 0.08, 0.09, dtype=np.float64)
    ns = 11
    n = int(np.prod(shape))
    times = np.linspace(0.001, 0.009, ns)
    kspha = rng.normal(size=(nt, ns))
    fieldmap = np.linspace(-37, 49, n)
    csm = rng.normal(size=(n, nc)) + 1j * rng.normal(size=(n, nc))
    mask = np.ones(n, dtype=bool)
    mask[0] = False
    mask[n - 2] = False
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    y = rng.normal(size=ns * nc) + 1j * rng.normal(size=ns * nc)
    w = rng.normal(size=ns) + 1j * rng.normal(size=ns) + 0.7
    return dict(grid=grid, kspha=kspha, times=times, fieldmap=fieldmap, csm=csm, mask=mask, x=x, y=y, w=w)


def _op(p, **kw):
    return hoe.HighOrderExplicitOp(p["grid"], p["kspha"], p["times"], fieldmap=p.get("fieldmap"), csm=p.get("csm"),
                                   mask=p.get("mask"), backend="fp64", **kw)


def _model(p, **kw):
    return O.make_model(p["grid"], p["kspha"], p["times"], p.get("fieldmap"), p.get("csm"), p.get("mask"),
                        f64=True, **kw)


def test_fp64_fixture(gpu_lib):
    p = O.explicit_highorder_test_problem()
    p = dict(p, grid=_grid64(p["grid"]), kspha=p["kspha"].astype(np.float64), times=p["times"].astype(np.float64),
             fieldmap=p["fieldmap"].astype(np.float64), csm=p["csm"].astype(np.complex128))
    m = _model(p)
    A = O.dense_matrix_f64(m)
    x = p["image"].astype(np.complex128)
    with _op(p) as op:
        y = op * x
        assert y.dtype == np.complex128
        e_f = O.relerr(y, A @ x)
        probe = np.linspace(-0.4, 0.5, A.shape[0]) + 1j * np.linspace(0.3, -0.2, A.shape[0])
        e_a = O.relerr(op.H * probe, A.conj().T @ probe)
    print(f"fp64 fixture: forward {e_f:.2e} adjoint {e_a:.2e}")
    assert e_f < TOL64 and e_a < TOL64


@pytest.mark.parametrize("shape,nt,nc", [((5, 4, 1), 9, 3), ((4, 3, 2), 16, 33), ((4, 5, 1), 16, 5),
                                         ((3, 4, 2), 9, 33), ((6, 5, 1), 16, 96), ((3, 3, 3), 16, 130)])
def test_fp64_independent_reference(shape, nt, nc, gpu_lib):
    c = _case(shape, nt, nc, 4321 + nt + nc)
    m = _model(c)
    A = O.dense_matrix_f64(m)
    D = np.tile(np.abs(c["w"]) ** 2, nc)
    with _op(c) as op:
        fwd = op * c["x"]
        bwd = op.H * c["y"]
        e_f, e_a = O.relerr(fwd, A @ c["x"]), O.relerr(bwd, A.conj().T @ c["y"])
        ident = abs(np.vdot(c["y"], fwd) - np.vdot(bwd, c["x"])) / (np.linalg.norm(fwd) * np.linalg.norm(c["y"]))
        expected = A.conj().T @ (D * (A @ c["x"]))
        N = hoe.normalOperator(hoe.compose(hoe.WeightingOp(c["w"], nc), op))
        e_n = O.relerr(N * c["x"], expected)
        e_g = O.relerr(hoe.explicit_gpu_normal(op, c["w"]) * c["x"], expected)
        e_d = O.relerr(hoe.normalOperator(op, D) * c["x"], expected)
        assert np.all(bwd[~c["mask"]] == 0)
    print(f"fp64 {shape} nt {nt} nc {nc}: fwd {e_f:.2e} adj {e_a:.2e} identity {ident:.2e} normal {e_n:.2e}")
    for e in (e_f, e_a, ident, e_n, e_g, e_d):
        assert e < TOL64


@pytest.mark.parametrize("z0", [0.0, 0.03])
def test_fp64_concomitant_and_recon_terms(z0, gpu_lib):
    rng = np.random.default_rng(977)
    grid = hoe.Grid(4, 3, 2, 0.04, 0.03, 0.02, dtype=np.float64)
    grid.z[:] += z0
    ns, nc = 13, 3
    kspha = rng.normal(size=(16, ns))
    k_nominal = rng.normal(size=(3, ns))
    times = np.linspace(0.001, 0.013, ns)
    csm = rng.normal(size=(24, nc)) + 1j * rng.normal(size=(24, nc))
    fieldmap = rng.normal(size=24) * 25
    mask = np.ones(24, dtype=bool)
    mask[0] = mask[12] = False
    conc = rng.normal(size=(4, ns)) * 1e-1
    x = rng.normal(size=24) + 1j * rng.normal(size=24)
    y = rng.normal(size=ns * nc) + 1j * rng.normal(size=ns * nc)
    p = dict(grid=grid, kspha=kspha, times=times, fieldmap=fieldmap, csm=csm, mask=mask)
    m = _model(p, recon_terms="0111", k_nominal=k_nominal, concomitant=conc)
    A = O.dense_matrix_f64(m)
    D = np.linspace(0.2, 1.1, ns * nc)
    with _op(p, recon_terms="0111", k_nominal=k_nominal, concomitant=conc) as op:
        e_f = O.relerr(op * x, A @ x)
        e_a = O.relerr(op.H * y, A.conj().T @ y)
        e_n = O.relerr(hoe.normalOperator(op, D) * x, A.conj().T @ (D * (A @ x)))
    print(f"fp64 concomitant z0={z0}: fwd {e_f:.2e} adj {e_a:.2e} normal {e_n:.2e}")
    assert e_f < 1e-11 and e_a < 1e-11 and e_n < 1e-11


def test_fp64_cgnr_matches_qr(gpu_lib):
    rng = np.random.default_rng(514)
    ns, nc = 37, 2
    data = rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc))
    weight = (0.5 + rng.uniform(size=ns)) * np.exp(1j * rng.uniform(size=ns))
    grid = hoe.Grid(3, 3, 1, 1.0, 1.0, 1.0, dtype=np.float64)
    kspha = rng.normal(size=(9, ns)) * 0.01
    for j in range(ns):
        kspha[1, j] = ((j % 3) - 1) / 3
        kspha[2, j] = (((j // 3) % 3) - 1) / 3
        kspha[3, j] = 0
    times = np.linspace(0, 1e-3, ns)
    csm = np.exp(1j * rng.uniform(size=(9, nc))) / np.sqrt(nc)
    m = O.make_model(grid, kspha, times, None, csm, None, f64=True)
    A = O.dense_matrix_f64(m)
    wr = np.tile(weight, nc)
    expected = np.linalg.lstsq(A * wr[:, None], data.reshape(-1, order="F") * wr, rcond=None)[0]
    params = {"solver": hoe.CGNR, "reg": hoe.L2Regularization(0.0), "reconSize": grid.matrixSize,
              "iterations": 30, "relTol": 1e-14}
    with hoe.HighOrderExplicitOp(grid, kspha, times, csm=csm, backend="fp64") as op:
        img, log = hoe.recon_HOOp(op, data, weight, params, return_log=True)
    assert img.dtype == np.complex128
    err = O.relerr(img.reshape(-1, order="F"), expected)
    print(f"fp64 CGNR vs QR: {err:.2e}, residual log ... {log[-1]:.2e}")
    assert err < 1e-10


def test_fp64_workload1_slice(gpu_lib):
    """256 samples of the 200 x 200, 32-coil spiral in FP64 against the oracle's FP64 products."""
    ns = 256
    c = synth.make(1, n_sam=ns)
    p = dict(grid=_grid64(c["grid"]), kspha=c["kspha"].astype(np.float64), times=c["times"].astype(np.float64),
             fieldmap=c["fieldmap"].astype(np.float64), csm=c["csm"].astype(np.complex128))
    m = _model(p)
    x = c["image"].astype(np.complex128)
    rng = np.random.default_rng(8)
    y = rng.standard_normal(ns * 32) + 1j * rng.standard_normal(ns * 32)
    with _op(p) as op:
        assert op.horner & 1     # the full grid runs the Horner-run phase in the forward
        e_f = O.relerr(op * x, O.forward_f64(m, x))
        e_a = O.relerr(op.H * y, O.adjoint_f64(m, y))
        e_n = O.relerr(hoe.normalOperator(op) * x, O.normal_f64(m, x))
    print(f"fp64 workload-1 slice: fwd {e_f:.2e} adj {e_a:.2e} normal {e_n:.2e}")
    assert e_f < 1e-11 and e_a < 1e-11 and e_n < 1e-11


def test_fp64_device_vectors_and_refusals(gpu_lib):
    import torch
    c = _case((5, 4, 1), 16, 3, 5)
    with _op(c) as op:
        xd = torch.from_numpy(c["x"]).to("cuda:0")
        yd = op * xd
        assert yd.dtype == torch.complex128 and yd.is_cuda
        np.testing.assert_array_equal(yd.cpu().numpy(), op * c["x"])
        with pytest.raises(hoe.ArgumentError):
            op * xd.to(torch.complex64)
        with pytest.raises(hoe.ArgumentError):
            op.grad(c["x"], c["y"])
        with pytest.raises(hoe.ArgumentError):
            op.set_coefficients(None, 0.0)
        with pytest.raises(hoe.ArgumentError):
            op.set_kspha(c["kspha"])
        with pytest.raises(hoe.ArgumentError):
            op.set_backend("mixed_fp16")
    p = O.explicit_highorder_test_problem()
    with hoe.HighOrderExplicitOp(p["grid"], p["kspha"], p["times"], backend="direct") as op32:
        with pytest.raises(hoe.ArgumentError):
            op32.set_backend("fp64")
