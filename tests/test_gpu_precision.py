"""Precision variant (workload 2): 200x200, 32 coils, single-shot EPI R=4 with
third-order terms and concomitant terms on an off-centre plane, so the
accumulated phase reaches hundreds of radians. Errors of E x, E^H y and of a
30-iteration CG image against the FP64 oracle, for every backend, with the
1e-4 bar of the mixed-precision target.

E x is checked on 512 sample rows spread over the readout (rows are
independent), E^H y on 512 voxels (each adjoint entry needs only its own
voxel), the CG image on a reduced EPI problem of the same construction
(64x64, 8 coils) where the FP64 CG finishes in seconds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402

BACKENDS = ["direct", "mixed_fp16", "mixed_bf16"]


def _op(c, backend):
    return hoe.HighOrderExplicitOp(c["grid"], c["kspha"], c["times"], fieldmap=c["fieldmap"], csm=c["csm"],
                                   concomitant=c["concomitant"], backend=backend)


@pytest.fixture(scope="module")
def workload2():
    c = synth.make(2)
    ph = O.make_model(c["grid"], c["kspha"][:, ::97], c["times"][::97], c["fieldmap"], c["csm"], None,
                      concomitant=c["concomitant"][:, ::97])
    maxphase = float(np.max(np.abs(ph.k64 @ ph.bf64.T)))
    return c, maxphase


@pytest.mark.parametrize("backend", BACKENDS)
def test_workload2_forward_adjoint_vs_fp64(backend, workload2, gpu_lib):
    c, maxphase = workload2
    ns = c["kspha"].shape[1]
    nv = c["grid"].n
    rng = np.random.default_rng(22)
    rows = np.sort(rng.choice(ns, 512, replace=False))
    vox = np.sort(rng.choice(nv, 512, replace=False))
    x = c["image"]
    y = (rng.normal(size=ns * c["csm"].shape[1]) + 1j * rng.normal(size=ns * c["csm"].shape[1])).astype(np.complex64)
    with _op(c, backend) as op:
        yg = op * x
        xg = op.H * y
    nc = c["csm"].shape[1]
    # forward rows
    mr = O.make_model(c["grid"], c["kspha"][:, rows], c["times"][rows], c["fieldmap"], c["csm"], None,
                      concomitant=c["concomitant"][:, rows])
    y64 = O.forward_f64(mr, x)
    yg_rows = yg.reshape(nc, ns)[:, rows].reshape(-1)
    ef = O.relerr(yg_rows, y64)
    # adjoint voxels: restrict the model to the sampled voxels, undo its 1/sqrt(nv) normalisation
    mask = np.zeros(nv, bool)
    mask[vox] = True
    mv = O.make_model(c["grid"], c["kspha"], c["times"], c["fieldmap"], c["csm"], mask,
                      concomitant=c["concomitant"])
    x64 = O.adjoint_f64(mv, y)[vox] * np.sqrt(mv.nv) / np.sqrt(nv)
    ea = O.relerr(xg[vox], x64)
    print(f"workload2 {backend}: max |phase| {maxphase:.1f} cycles ({2*np.pi*maxphase:.0f} rad); "
          f"E x err {ef:.2e}, E^H y err {ea:.2e}")
    assert maxphase > 50
    assert ef < 1e-4 and ea < 1e-4


@pytest.mark.parametrize("backend", BACKENDS)
def test_reduced_epi_cg_vs_fp64(backend, gpu_lib):
    rng = np.random.default_rng(5)
    N, nc, n_lines, per = 64, 8, 16, 128
    fov = 0.2
    grid = hoe.Grid(N, N, 1, fov / N, fov / N, fov / N)
    grid.z[:] = np.float32(0.04)
    kmax = N / (2 * fov)
    kx_line = np.linspace(-kmax, kmax, per)
    kx = np.concatenate([kx_line if l % 2 == 0 else kx_line[::-1] for l in range(n_lines)])
    ky = np.repeat((np.arange(n_lines) - n_lines / 2) * (2 * kmax / n_lines), per)
    ns = kx.size
    dwell = 2e-6
    k1 = np.stack([kx, ky, np.zeros_like(kx)])
    kspha = synth.coefficient_rows(rng, k1, fov / 2, ns).astype(np.float32)
    times = (np.arange(ns) * dwell).astype(np.float32)
    conc = np.zeros((4, ns), np.float32)
    conc[0] = np.cumsum(np.full(ns, synth.GAMMA_BAR * 0.04 ** 2 / 14.0 * dwell))
    fm = synth.gaussian_b0(grid, rng, 100.0)
    csm = synth.gaussian_coils(grid, nc, rng, fov / 3, 0.6 * fov / 2)
    img = synth.phantom_image(grid, rng, (fov, fov, fov))
    m = O.make_model(grid, kspha, times, fm, csm, None, concomitant=conc)
    y = O.forward_f64(m, img)
    y = y + 0.01 * np.linalg.norm(y) / np.sqrt(y.size) * (rng.normal(size=y.size) + 1j * rng.normal(size=y.size))
    y32 = y.astype(np.complex64)
    iters, lam = 30, 1e-3
    rhs = O.adjoint_f64(m, y32.astype(np.complex128))
    x64, _ = O.cgnr(lambda v: O.normal_f64(m, v), rhs, iters, lam)
    with hoe.HighOrderExplicitOp(grid, kspha, times, fieldmap=fm, csm=csm, concomitant=conc,
                                 backend=backend) as op:
        data = y32.reshape(nc, ns).T
        params = {"solver": hoe.CGNR, "reg": hoe.L2Regularization(lam), "reconSize": grid.matrixSize,
                  "iterations": iters}
        xg = hoe.recon_HOOp(op, data, np.ones(ns, np.complex64), params)
    e = O.relerr(xg.reshape(-1, order="F"), x64)
    print(f"reduced EPI CG ({iters} it) {backend}: image vs FP64 CG {e:.2e}")
    assert e < 1e-4 if backend != "mixed_bf16" else e < 1e-3
