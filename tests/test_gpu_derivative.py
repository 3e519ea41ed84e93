"""GPU derivative product B x (the reference's ``kspha_dt`` operator,
explicit_array_derivative, array_operator.jl:187-235) against the FP64 oracle,
and the reference's operator semantics around it:
  - only the forward product changes; ``op' * y`` stays E^H y (array_operator.jl:167-171);
  - concomitant terms are ignored when kspha_dt is given (array_operator.jl:139,
    test_Concomitant.jl:124-129);
  - kspha_dt is used as given (recon_terms processes kspha only).
Every backend is checked; the mixed backends run this product on the FP32 kernels.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402

BACKENDS = ["direct", "mixed_fp16", "mixed_bf16"]


def _problem(seed=31, nz=1):
    rng = np.random.default_rng(seed)
    grid = hoe.Grid(12, 10, nz, np.float32(0.02), np.float32(0.02), np.float32(0.02))
    ns, nc = 300, 3
    s = np.linspace(0, 1, ns)
    kspha = np.zeros((16, ns), np.float32)
    kdt = np.zeros((16, ns), np.float32)
    for r in range(16):
        a, b = rng.normal(size=2)
        kspha[r] = (a * np.cos(5 * s + r) + b * s) * (20.0 if r < 4 else 200.0)
        kdt[r] = (-5 * a * np.sin(5 * s + r) + b) * (20.0 if r < 4 else 200.0) * 1e3
    times = (np.arange(ns) * 2e-6).astype(np.float32)
    fm = rng.uniform(-60, 60, grid.n).astype(np.float32)
    csm = (rng.normal(size=(grid.n, nc)) + 1j * rng.normal(size=(grid.n, nc))).astype(np.complex64)
    x = (rng.normal(size=grid.n) + 1j * rng.normal(size=grid.n)).astype(np.complex64)
    return grid, kspha, kdt, times, fm, csm, x


@pytest.mark.parametrize("backend", BACKENDS)
def test_derivative_product_vs_fp64(backend, gpu_lib):
    for nz in (1, 2):   # 2-D (z terms dropped from the records) and 3-D
        grid, kspha, kdt, times, fm, csm, x = _problem(nz=nz)
        m = O.make_model(grid, kspha, times, fm, csm, None)
        ref = O.forward_dt_f64(m, kdt, x)
        with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, kspha_dt=kdt,
                                     backend=backend) as op:
            bx = op * x
            e = O.relerr(bx, ref)
            print(f"B x {backend} nz={nz}: vs FP64 {e:.2e}")
            assert e < 3e-5
            # the adjoint is the plain E^H (the reference's ctprod)
            y = (np.random.default_rng(2).normal(size=op.nrow) + 0j).astype(np.complex64)
            xa = op.H * y
            assert O.relerr(xa, O.adjoint_f64(m, y)) < 3e-5
            # clearing kspha_dt restores E
            op.set_kspha_dt(None)
            assert O.relerr(op * x, O.forward_f64(m, x)) < 3e-5
            op.set_kspha_dt(kdt)
            assert O.relerr(op * x, ref) < 3e-5


def test_derivative_semantics(gpu_lib):
    grid, kspha, kdt, times, fm, csm, x = _problem(seed=8)
    conc = np.random.default_rng(3).normal(size=(4, kspha.shape[1])).astype(np.float32)
    y = (np.random.default_rng(4).normal(size=kspha.shape[1] * csm.shape[1]) + 1j).astype(np.complex64)
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, kspha_dt=kdt) as d, \
            hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, kspha_dt=kdt,
                                    concomitant=conc) as dc:
        # concomitant terms are ignored with kspha_dt (test_Concomitant.jl:124-129)
        np.testing.assert_array_equal(d * x, dc * x)
        np.testing.assert_array_equal(d.H * y, dc.H * y)
        # only the forward product is defined
        with pytest.raises(hoe.ArgumentError):
            hoe.normalOperator(d) * x
        with pytest.raises(hoe.ArgumentError):
            d.grad(x, y)
        with pytest.raises(AssertionError):
            d.set_kspha_dt(kdt[:, :-1])
    # kspha_dt is not processed by recon_terms; kspha is
    m = O.make_model(grid, kspha, times, fm, csm, None, recon_terms="1010")
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, kspha_dt=kdt,
                                 recon_terms="1010") as op:
        assert O.relerr(op * x, O.forward_dt_f64(m, kdt, x)) < 3e-5


@pytest.mark.parametrize("backend", ["direct", "mixed_fp16"])
def test_derivative_product_workload1_subset(backend, gpu_lib):
    """Workload 1 geometry (200x200, 32 coils, Horner runs) on 256 spiral samples,
    kspha_dt from the 5-tap stencil of FindDelay.jl:179-180."""
    c = synth.make(1, n_sam=256)
    k = c["kspha"].astype(np.float64)
    pad = np.pad(k, ((0, 0), (2, 2)), mode="edge")
    taps = np.array([1 / 8, 1 / 4, 0.0, -1 / 4, -1 / 8])
    kdt = np.stack([np.convolve(pad[r], taps, mode="valid") for r in range(16)]) / c["dwell"]
    kdt = kdt.astype(np.float32)
    m = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], None)
    ref = O.forward_dt_f64(m, kdt, c["image"])
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                 kspha_dt=kdt, backend=backend) as op:
        e = O.relerr(op * c["image"], ref)
    print(f"workload1 subset B x {backend}: vs FP64 {e:.2e}")
    assert e < 3e-5
