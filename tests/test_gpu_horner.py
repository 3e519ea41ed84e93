"""Phase evaluation by Horner runs against the FP64 oracle, in the geometries that select
each kernel path:

- 3-D 96x12x4: every basis term active (record bucket L = 16), runs of 96 voxels, so
  many 64-voxel waves hold the ends of two runs -- the MFMA adjoint's two-run branch,
  with its per-pair FP64 setup product (the Horner forward and MFMA adjoint both on);
- 2-D 40x40 with a mask: runs of 40 or fewer, so some waves span three runs: in the MFMA
  adjoint those waves take the generic phase while the others stay on Horner (the choice
  is per wave);
- 2-D 192x8: runs of exactly three waves, every wave inside one run (the single-run branch).

Every backend's E x, E^H y and E^H E x are compared with the FP64 operator
(array_operator.jl:241-348 restated in the oracle) at the Float32 bar of
test_explicit_gpu.jl:51 (3e-5; 1e-4 for BF16).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402

TOLS = {"direct": 3e-5, "mixed_fp16": 3e-5, "mixed_bf16": 1e-4}


def _problem(nx, ny, nz, n_sam=1536, n_cha=6, seed=3, mask=None):
    rng = np.random.default_rng(seed)
    fov = (0.22, 0.22 * ny / nx, 0.11 if nz > 1 else 0.22 / nx)
    grid = hoe.Grid(nx, ny, nz, np.float32(fov[0] / nx), np.float32(fov[1] / ny), np.float32(fov[2] / nz))
    kmax = nx / (2 * fov[0])
    parts = 4 if nz > 1 else 1
    per = n_sam // parts
    kx, ky = synth.spiral(per, kmax, 6)
    kz = (np.arange(parts) - parts // 2) / fov[2] if nz > 1 else np.zeros(parts)
    k1 = np.stack([np.tile(kx, parts), np.tile(ky, parts), np.repeat(kz, per)])
    kspha = synth.coefficient_rows(rng, k1, fov[0] / 2, k1.shape[1]).astype(np.float32)
    times = np.tile(np.arange(per) * 2e-6, parts).astype(np.float32)
    fm = synth.gaussian_b0(grid, rng, 120.0)
    csm = synth.gaussian_coils(grid, n_cha, rng, sigma=fov[0] / 3, ring_radius=0.3 * fov[0],
                               z_spread=fov[2] if nz > 1 else 0.0)
    x = synth.phantom_image(grid, rng, fov)
    return grid, kspha, times, fm, csm, x, mask


@pytest.mark.parametrize("shape,mask_kind,expect_adj", [
    ((96, 12, 4), None, True),
    ((40, 40, 1), "disc", True),
    ((192, 8, 1), None, True),
])
@pytest.mark.parametrize("backend", ["direct", "mixed_fp16", "mixed_bf16"])
def test_horner_paths_vs_fp64(shape, mask_kind, expect_adj, backend, gpu_lib):
    nx, ny, nz = shape
    mask = None
    if mask_kind == "disc":
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        mask = ((i - nx / 2 + 0.5) ** 2 + (j - ny / 2 + 0.5) ** 2 <= (0.45 * nx) ** 2).reshape(nx, ny, 1)
    grid, kspha, times, fm, csm, x, mask = _problem(nx, ny, nz, mask=mask)
    m = O.make_model(grid, kspha, times, fm, csm, mask)
    y_ref = O.forward_f64(m, x)
    probe = (np.random.default_rng(1).normal(size=y_ref.size)
             + 1j * np.random.default_rng(2).normal(size=y_ref.size)).astype(np.complex64)
    xa_ref = O.adjoint_f64(m, probe)
    n_ref = O.normal_f64(m, x)
    tol = TOLS[backend]
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, mask=mask, backend=backend) as op:
        assert op.horner & 1, "runs long enough for the Horner forward"
        if backend != "direct":
            assert bool(op.horner & 2) == expect_adj
        e_f = O.relerr(op * x, y_ref)
        e_a = O.relerr(op.H * probe, xa_ref)
        e_n = O.relerr(hoe.normalOperator(op) * x, n_ref)
    print(f"{shape} mask={mask_kind} {backend} (horner {op.horner}): E {e_f:.2e}  E^H {e_a:.2e}  E^H E {e_n:.2e}")
    assert e_f < tol and e_a < tol and e_n < tol


@pytest.mark.parametrize("case", ["terms9_plane", "concomitant_plane", "concomitant_offcentre"])
@pytest.mark.parametrize("backend", ["direct", "mixed_fp16"])
def test_horner_record_buckets_vs_fp64(case, backend, gpu_lib):
    """The other record lengths of the Horner kernels, on a 2-D 192x8 grid (runs of 192
    voxels, every wave in one run):
    - 9 terms on the z = 0 plane: x, y, xy, 3z^2-r^2, x^2-y^2 active, + dB0 -> bucket 6
      (the FP64 setup covers the record's t within its loaded blocks);
    - 16 terms and concomitant fields on z = 0: 9 terms + x^2+y^2, + dB0 -> bucket 12
      (t loaded with the records; one k block fewer than a t column would need);
    - the same on an off-centre plane (z = 0.03): all 15 terms and 4 concomitant
      columns, + dB0 -> bucket 20.
    E x, E^H y and E^H E x against the FP64 operator (array_operator.jl:241-348 and
    prep_concomitant, prep_kspha.jl:78-98, restated in the oracle) at 3e-5."""
    grid, kspha, times, fm, csm, x, _ = _problem(192, 8, 1)
    rng = np.random.default_rng(17)
    conc = None
    if case == "terms9_plane":
        kspha = np.ascontiguousarray(kspha[:9])
    else:
        conc = (rng.normal(size=(4, kspha.shape[1])) * 20.0).astype(np.float32)
        if case == "concomitant_offcentre":
            grid.z[:] = np.float32(0.03)
    m = O.make_model(grid, kspha, times, fm, csm, None, concomitant=conc)
    y_ref = O.forward_f64(m, x)
    probe = (np.random.default_rng(1).normal(size=y_ref.size)
             + 1j * np.random.default_rng(2).normal(size=y_ref.size)).astype(np.complex64)
    xa_ref = O.adjoint_f64(m, probe)
    n_ref = O.normal_f64(m, x)
    kw = {} if conc is None else {"concomitant": conc}
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, backend=backend, **kw) as op:
        assert op.horner & 1, "runs long enough for the Horner forward"
        if backend != "direct":
            assert op.horner & 2, "every wave in one run: Horner MFMA adjoint"
        e_f = O.relerr(op * x, y_ref)
        e_a = O.relerr(op.H * probe, xa_ref)
        e_n = O.relerr(hoe.normalOperator(op) * x, n_ref)
    print(f"{case} {backend} (horner {op.horner}): E {e_f:.2e}  E^H {e_a:.2e}  E^H E {e_n:.2e}")
    assert e_f < 3e-5 and e_a < 3e-5 and e_n < 3e-5


@pytest.mark.parametrize("fm_kind,expect_anchored", [("smooth", True), ("rough_small", True), ("rough_large", False)])
@pytest.mark.parametrize("backend", ["mixed_fp16", "mixed_bf16"])
def test_anchored_phase_bound(fm_kind, expect_anchored, backend, gpu_lib):
    """The MFMA Horner kernels add t * (fieldmap - anchor fieldmap) in FP32 (anchored phases),
    whose rounding grows with its size. hoe_create enables them only while
    max|t| * max|fieldmap offset| <= 8 cycles, and otherwise keeps the FP64 phase of every
    pair. On a 192x8 grid with a 20 ms readout:
    - a smooth map (3 Gaussians, 120 Hz): anchored;
    - an uncorrelated map of +-90 Hz (offsets up to 180 Hz: 3.6 cycles): anchored, at the
      same accuracy;
    - an uncorrelated map of +-3 kHz (120 cycles): FP64 phases.
    Every product against the FP64 operator at 3e-5 (1e-4 for BF16), and the MFMA coefficient
    gradient (its Horner blocks anchored alike) against the FP64 gradient at 1e-4."""
    grid, kspha, times, fm, csm, x, _ = _problem(192, 8, 1)
    times = (np.arange(times.size) * 13e-6).astype(np.float32)   # 20 ms
    rng = np.random.default_rng(29)
    if fm_kind == "rough_small":
        fm = rng.uniform(-90.0, 90.0, size=fm.shape).astype(np.float32)
    elif fm_kind == "rough_large":
        fm = rng.uniform(-3000.0, 3000.0, size=fm.shape).astype(np.float32)
    m = O.make_model(grid, kspha, times, fm, csm, None)
    y_ref = O.forward_f64(m, x)
    probe = (np.random.default_rng(1).normal(size=y_ref.size)
             + 1j * np.random.default_rng(2).normal(size=y_ref.size)).astype(np.complex64)
    xa_ref = O.adjoint_f64(m, probe)
    n_ref = O.normal_f64(m, x)
    tol = TOLS[backend]
    # gradient data: the forward of a perturbed image, so that the residual is not small
    y_g = O.forward_f64(m, 0.8 * x + 0.05).astype(np.complex64)
    g64, _, _ = O.grad_f64(m, x, y_g.astype(np.complex128))
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, backend=backend) as op:
        assert op.horner & 3 == 3, "runs long enough for the Horner kernels"
        assert bool(op.horner & 4) == expect_anchored
        e_f = O.relerr(op * x, y_ref)
        e_a = O.relerr(op.H * probe, xa_ref)
        e_n = O.relerr(hoe.normalOperator(op) * x, n_ref)
        g, _, _ = op.grad(x, y_g)
        e_g = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    print(f"{fm_kind} {backend} (horner {op.horner}): E {e_f:.2e}  E^H {e_a:.2e}  E^H E {e_n:.2e}  grad {e_g:.2e}")
    assert e_f < tol and e_a < tol and e_n < tol
    assert e_g < 1e-4


@pytest.mark.parametrize("kscale,expect_anchored", [(1.0, True), (2.0, True), (8.0, False), (16.0, False)])
@pytest.mark.parametrize("backend", ["mixed_fp16", "mixed_bf16"])
def test_anchored_offset_bound(kscale, expect_anchored, backend, gpu_lib):
    """Trajectories reaching kscale times the grid's Nyquist radius (every coefficient row
    scaled): the Horner offsets P(x_a + u) - P(x_a) of the anchored phases, added in FP32,
    grow with them (about 8 cycles per Nyquist radius in the adjoint's 32-voxel groups). After
    each record build the device bounds them, and above 32 cycles the MFMA kernels take the
    FP64 phase of every pair. On a 192x8 grid: E x, E^H y and E^H E x against the FP64
    operator at 3e-5 (1e-4 for BF16) in both regimes; the decision follows hoe_set_kspha (off
    at 8 times the trajectory, on again at the original)."""
    grid, kspha0, times, fm, csm, x, _ = _problem(192, 8, 1)
    kspha = (kspha0 * np.float32(kscale)).astype(np.float32)
    tol = TOLS[backend]

    def check(op, k):
        m = O.make_model(grid, k, times, fm, csm, None)
        y_ref = O.forward_f64(m, x)
        probe = (np.random.default_rng(1).normal(size=y_ref.size)
                 + 1j * np.random.default_rng(2).normal(size=y_ref.size)).astype(np.complex64)
        e_f = O.relerr(op * x, y_ref)
        e_a = O.relerr(op.H * probe, O.adjoint_f64(m, probe))
        e_n = O.relerr(hoe.normalOperator(op) * x, O.normal_f64(m, x))
        return e_f, e_a, e_n

    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, backend=backend) as op:
        assert op.horner & 3 == 3, "runs long enough for the Horner kernels"
        assert bool(op.horner & 4) == expect_anchored
        e = check(op, kspha)
        print(f"kscale {kscale} {backend} (horner {op.horner}): E {e[0]:.2e}  E^H {e[1]:.2e}  E^H E {e[2]:.2e}")
        assert max(e) < tol
        if kscale == 1.0:
            k8 = (kspha0 * np.float32(8.0)).astype(np.float32)
            op.set_kspha(k8)
            assert not op.horner & 4
            assert max(check(op, k8)) < tol
            op.set_kspha(kspha0)
            assert op.horner & 4
            assert max(check(op, kspha0)) < tol
