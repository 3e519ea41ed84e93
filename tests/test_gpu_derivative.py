"""GPU derivative product B x (the reference's ``kspha_dt`` operator,
explicit_array_derivative, array_operator.jl:187-235) against the FP64 oracle,
and the reference's operator semantics around it:
  - only the forward product changes; ``op' * y`` stays E^H y (array_operator.jl:167-171);
  - concomitant terms are ignored when kspha_dt is given (array_operator.jl:139,
    test_Concomitant.jl:124-129);
  - kspha_dt is used as given (recon_terms processes kspha only).
Every backend is checked, each at the bar of its other products (3e-5; 1e-4 for BF16):
the mixed backends run this product on the matrix cores, their A operand carrying the
factor 2 pi psi scaled by a power of two. Both phase paths are covered: the 12x10 grid
takes the generic phase (runs too short for Horner), the 192x8 and 200x200 ones the
anchored Horner forward.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402

BACKENDS = ["direct", "mixed_fp16", "mixed_bf16"]
TOLS = {"direct": 3e-5, "mixed_fp16": 3e-5, "mixed_bf16": 1e-4}


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
        tol = TOLS[backend]
        with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, kspha_dt=kdt,
                                     backend=backend) as op:
            bx = op * x
            e = O.relerr(bx, ref)
            print(f"B x {backend} nz={nz}: vs FP64 {e:.2e}")
            assert e < tol
            # the adjoint is the plain E^H (the reference's ctprod)
            y = (np.random.default_rng(2).normal(size=op.nrow) + 0j).astype(np.complex64)
            xa = op.H * y
            assert O.relerr(xa, O.adjoint_f64(m, y)) < tol
            # clearing kspha_dt restores E
            op.set_kspha_dt(None)
            assert O.relerr(op * x, O.forward_f64(m, x)) < tol
            op.set_kspha_dt(kdt)
            assert O.relerr(op * x, ref) < tol
            # B x is linear in kspha_dt: the scale of the A operand follows it
            op.set_kspha_dt(kdt * np.float32(2.0 ** -20))
            assert O.relerr(op * x, ref * 2.0 ** -20) < tol


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


def _stencil_dt(kspha, dwell):
    """kspha_dt from the 5-tap stencil of FindDelay.jl:179-180 (edge samples repeated)."""
    k = np.asarray(kspha, np.float64)
    pad = np.pad(k, ((0, 0), (2, 2)), mode="edge")
    taps = np.array([1 / 8, 1 / 4, 0.0, -1 / 4, -1 / 8])
    kdt = np.stack([np.convolve(pad[r], taps, mode="valid") for r in range(k.shape[0])]) / dwell
    return kdt.astype(np.float32)


@pytest.mark.parametrize("backend", BACKENDS)
def test_derivative_product_workload1_subset(backend, gpu_lib):
    """Workload 1 geometry (200x200, 32 coils, Horner runs) on 256 spiral samples,
    kspha_dt from the 5-tap stencil of FindDelay.jl:179-180."""
    c = synth.make(1, n_sam=256)
    kdt = _stencil_dt(c["kspha"], c["dwell"])
    m = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], None)
    ref = O.forward_dt_f64(m, kdt, c["image"])
    with hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                 kspha_dt=kdt, backend=backend) as op:
        assert op.horner & 1, "runs long enough for the Horner forward"
        e = O.relerr(op * c["image"], ref)
    print(f"workload1 subset B x {backend}: vs FP64 {e:.2e}")
    assert e < TOLS[backend]


@pytest.mark.parametrize("case", ["192x8_6c", "disc40_12c", "96x12_40c", "192x8_6c_rough"])
@pytest.mark.parametrize("backend", BACKENDS)
def test_derivative_horner_layouts(case, backend, gpu_lib):
    """The Horner forward's derivative paths and the coil layouts of the MFMA kernel:
    - 192x8, 6 coils: every K unit inside one run (anchored psi), 8-coil packed layout;
    - 40x40 with a disc mask, 12 coils: K units across run ends (psi from the full records
      there), 16-coil packed layout;
    - 96x12, 40 coils: two 32-coil groups per wave;
    - 192x8 with a +-3 kHz uncorrelated fieldmap: anchored phases off (FP64 phase and psi
      on every K step).
    B x against the FP64 derivative operator (array_operator.jl:187-235 restated in the
    oracle), then the plain forward after clearing kspha_dt."""
    rng = np.random.default_rng(41)
    nx, ny, nc = {"192x8_6c": (192, 8, 6), "disc40_12c": (40, 40, 12), "96x12_40c": (96, 12, 40),
                  "192x8_6c_rough": (192, 8, 6)}[case]
    fov = 0.22
    grid = hoe.Grid(nx, ny, 1, np.float32(fov / nx), np.float32(fov / nx), np.float32(fov / nx))
    mask = None
    if case.startswith("disc"):
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        mask = ((i - nx / 2 + 0.5) ** 2 + (j - ny / 2 + 0.5) ** 2 <= (0.45 * nx) ** 2).reshape(nx, ny, 1)
    ns, dwell = 1024, 2e-6
    kx, ky = synth.spiral(ns, nx / (2 * fov), 6)
    kspha = synth.coefficient_rows(rng, np.stack([kx, ky, np.zeros(ns)]), fov / 2, ns).astype(np.float32)
    times = (np.arange(ns) * dwell).astype(np.float32)
    fm = synth.gaussian_b0(grid, rng, 120.0)
    if case.endswith("rough"):
        fm = rng.uniform(-3000.0, 3000.0, size=fm.shape).astype(np.float32)
    csm = synth.gaussian_coils(grid, nc, rng, sigma=fov / 3, ring_radius=0.3 * fov, z_spread=0.0)
    x = synth.phantom_image(grid, rng, (fov, fov * ny / nx, fov / nx))
    kdt = _stencil_dt(kspha, dwell)
    m = O.make_model(grid, kspha, times, fm, csm, mask)
    ref = O.forward_dt_f64(m, kdt, x)
    tol = TOLS[backend]
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, mask=mask, kspha_dt=kdt,
                                 backend=backend) as op:
        assert op.horner & 1, "runs long enough for the Horner forward"
        if backend != "direct":
            assert bool(op.horner & 4) == (not case.endswith("rough"))
        e = O.relerr(op * x, ref)
        op.set_kspha_dt(None)
        e0 = O.relerr(op * x, O.forward_f64(m, x))
    print(f"{case} {backend} (horner {op.horner}): B x {e:.2e}, E x after clearing {e0:.2e}")
    assert e < tol and e0 < tol
