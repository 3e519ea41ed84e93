"""Pins of the CPU oracle (no GPU).

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign, a transposed operand, a wrong scale or ordering) fails it:
  - known answers of the reference tests (test_Concomitant.jl:92, the
    deterministic fixture of test_HighOrderExplicit.jl:3-128 at its 3e-6 bar);
  - closed forms: a Cartesian first-order trajectory reduces E to a DFT
    (numpy.fft), a pure off-resonance model to a direct exponential sum;
  - invariants: every solid harmonic is harmonic (zero Laplacian), the
    adjoint identity, block-count invariance, zeros outside the mask;
  - brute force: the C oracle against an independently coded dense matrix,
    CGNR against a QR least-squares solve (test_ReconstructionStability.jl),
    the gradient oracle and the derivative product against central finite
    differences of the forward.
"""
import numpy as np
import pytest

import oracle as O
from hoe.grid import Grid, basisfunc_spha


@pytest.mark.parametrize("shape,res,dtype", [((4, 3, 2), (0.7, 0.9, 1.2), np.float32), ((5, 1, 3), (0.01, 0.02, 0.03),
                                               np.float64), ((6, 5, 4), (2e-3, 3e-3, 5e-3), np.float32)])
@pytest.mark.parametrize("opts", [{}, {"exchange_xy": True}, {"reverse_x": True, "reverse_z": True},
                                  {"reverse_y": True}])
def test_host_grid_matches_oracle_grid_and_closed_forms(shape, res, dtype, opts):
    """The host API's Grid against the oracle's restatement of Grid.jl:63-95 (bit for bit)
    and against closed forms: voxel (i, j, k), 0-based, sits at flattened index
    i + nX (j + nY k) with coordinates ((i + 1 - (nX + 1)/2) dx, ...), so z takes the
    values (k + 1 - (nZ + 1)/2) dz, symmetric about 0."""
    nX, nY, nZ = shape
    T = np.dtype(dtype).type
    dx, dy, dz = (T(v) for v in res)
    g = Grid(nX, nY, nZ, dx, dy, dz, dtype=dtype, **opts)
    o = O.OracleGrid(nX, nY, nZ, dx, dy, dz, dtype=dtype, **opts)
    for a in "xyz":
        assert getattr(g, a).dtype == np.dtype(dtype)
        np.testing.assert_array_equal(getattr(g, a), getattr(o, a))
    assert g.matrixSize == o.matrixSize == shape and g.n == o.n == nX * nY * nZ
    i, j, k = np.meshgrid(np.arange(nX), np.arange(nY), np.arange(nZ), indexing="ij")
    idx = (i + nX * (j + nY * k)).ravel()
    cx = np.empty(g.n)
    cy = np.empty(g.n)
    cz = np.empty(g.n)
    cx[idx] = (i.ravel() + 1 - (nX + 1) / 2) * float(dx)
    cy[idx] = (j.ravel() + 1 - (nY + 1) / 2) * float(dy)
    cz[idx] = (k.ravel() + 1 - (nZ + 1) / 2) * float(dz)
    if opts.get("exchange_xy"):
        cx, cy = cy, cx
    if opts.get("reverse_x"):
        cx = cx[::-1]
    if opts.get("reverse_y"):
        cy = cy[::-1]
    if opts.get("reverse_z"):
        cz = cz[::-1]
    for a, c in (("x", cx), ("y", cy), ("z", cz)):
        np.testing.assert_allclose(getattr(g, a).astype(np.float64), c, rtol=0, atol=4 * np.finfo(dtype).eps * max(res) * max(shape))
    assert abs(float(np.sum(g.z.astype(np.float64)))) <= g.n * np.finfo(dtype).eps * float(dz) * nZ
    if nZ == 2 and not opts.get("reverse_z"):
        np.testing.assert_array_equal(np.unique(g.z), [-dz / 2, dz / 2])


def test_basis_known_answer_term17():
    # test_Concomitant.jl:92
    np.testing.assert_array_equal(basisfunc_spha(np.array([1., 2.]), np.array([3., 4.]), np.array([5., 6.]), 17),
                                  [-24., -96.])


def test_oracle_basis_matches_host_polynomials_and_is_harmonic():
    rng = np.random.default_rng(0)
    x, y, z = (rng.uniform(-0.2, 0.2, 50) for _ in range(3))
    b = O.basis_f64d(x, y, z)
    np.testing.assert_allclose(b, basisfunc_spha(x, y, z, range(1, 17)), rtol=1e-13, atol=1e-15)
    # Laplacian by central differences: exact for polynomials of degree <= 3
    h = 1e-2
    lap = -6 * b
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        lap += O.basis_f64d(x + e[0], y + e[1], z + e[2]) + O.basis_f64d(x - e[0], y - e[1], z - e[2])
    lap /= h * h
    assert np.max(np.abs(lap)) < 1e-9
    # and a non-harmonic polynomial (x^2) would fail the same test
    xs = x ** 2
    lapx = ((x + h) ** 2 + (x - h) ** 2 - 2 * xs) / h ** 2
    assert np.min(np.abs(lapx)) > 1.0


def test_fp32_basis_tracks_fp64():
    rng = np.random.default_rng(1)
    x, y, z = (rng.uniform(-0.1, 0.1, 200).astype(np.float32) for _ in range(3))
    b32 = O.basis_f32(x, y, z).astype(np.float64)
    b64 = O.basis_f64(x, y, z)
    scale = np.max(np.abs(b64), axis=0) + 1e-30
    assert np.max(np.abs(b32 - b64) / scale) < 1e-6


def test_dense_matrix_equals_c_oracle():
    p = O.explicit_highorder_test_problem()
    m = O.make_model(p["grid"], p["kspha"], p["times"], p["fieldmap"], p["csm"], p["mask"])
    A = O.dense_matrix_f64(m)
    assert A.shape == (m.ns * m.nc, m.n_grid)
    x = p["image"]
    assert O.relerr(O.forward_f64(m, x), A @ x) < 1e-14
    y = np.random.default_rng(2).normal(size=A.shape[0]) + 0j
    assert O.relerr(O.adjoint_f64(m, y), A.conj().T @ y) < 1e-14


def test_reference_fixture_cpu_path():
    """test_HighOrderExplicit.jl:94-128: blocked CPU operator vs dense matrix at rtol 3e-6,
    block-count invariance, adjoint identity, zeros outside the mask."""
    p = O.explicit_highorder_test_problem()
    m = O.make_model(p["grid"], p["kspha"], p["times"], p["fieldmap"], p["csm"], p["mask"])
    A = O.dense_matrix_f64(m)
    n = m.ns * m.nc
    probe = (np.linspace(-0.4, 0.5, n) + 1j * np.linspace(0.3, -0.2, n)).astype(np.complex64)
    fwd4 = O.cpu_ref_forward(m, p["image"], nblock=4)
    adj4 = O.cpu_ref_adjoint(m, probe, nblock=4)
    assert O.relerr(fwd4, A @ p["image"]) < 3e-6
    assert O.relerr(adj4, A.conj().T @ probe) < 3e-6
    assert O.relerr(O.cpu_ref_forward(m, p["image"], nblock=1), fwd4) < 3e-6
    assert O.relerr(O.cpu_ref_adjoint(m, probe, nblock=1), adj4) < 3e-6
    lhs, rhs = np.vdot(probe, fwd4), np.vdot(adj4, p["image"])
    assert abs(lhs - rhs) / abs(lhs) < 3e-6
    assert np.all(adj4[~p["mask"]] == 0)
    assert np.all(O.adjoint_f64(m, probe)[~p["mask"]] == 0)


def _cartesian_problem(N=8, nc=2):
    dx = np.float32(2.0 ** -4)
    grid = Grid(N, N, 1, dx, dx, dx)
    fov = N * float(dx)
    ms, ns_ = np.meshgrid(np.arange(-N // 2, N // 2), np.arange(-N // 2, N // 2), indexing="ij")
    kx = (ms.ravel() / fov).astype(np.float32)
    ky = (ns_.ravel() / fov).astype(np.float32)
    nsam = kx.size
    kspha = np.zeros((16, nsam), dtype=np.float32)
    kspha[1], kspha[2] = kx, ky
    rng = np.random.default_rng(3)
    csm = (rng.normal(size=(N * N, nc)) + 1j * rng.normal(size=(N * N, nc))).astype(np.complex64)
    x = (rng.normal(size=N * N) + 1j * rng.normal(size=N * N)).astype(np.complex64)
    return grid, kspha, np.zeros(nsam, np.float32), csm, x, ms.ravel(), ns_.ravel()


def test_first_order_cartesian_is_a_dft():
    N, nc = 8, 2
    grid, kspha, times, csm, x, mi, ni = _cartesian_problem(N, nc)
    m = O.make_model(grid, kspha, times, None, csm, None)
    y = O.forward_f64(m, x)
    # y[c, s] = 1/N * sum_{i,j} exp(+2 pi i (m (i - (N+1)/2) + n (j - (N+1)/2)) / N) csm x
    for c in range(nc):
        img = (csm[:, c].astype(np.complex128) * x.astype(np.complex128)).reshape(N, N)      # [j][i]: x fastest
        F = np.fft.ifft2(img) * N * N                                  # sum exp(+2 pi i (n j + m i)/N)
        shift = np.exp(-2j * np.pi * (mi + ni) * (N + 1) / 2 / N)
        # samples: m over x (index i, axis 1), n over y (index j, axis 0); i, j are 0-based here
        expect = F[ni % N, mi % N] * np.exp(2j * np.pi * (mi + ni) / N) * shift / N
        assert O.relerr(y[c * kspha.shape[1]:(c + 1) * kspha.shape[1]], expect) < 1e-12


def test_off_resonance_only_closed_form():
    rng = np.random.default_rng(4)
    grid = Grid(5, 3, 2, np.float32(0.01), np.float32(0.01), np.float32(0.01))
    n = grid.n
    ns = 9
    times = np.linspace(0, 2e-3, ns).astype(np.float32)
    fm = rng.uniform(-100, 100, n).astype(np.float32)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    m = O.make_model(grid, np.zeros((9, ns), np.float32), times, fm, None, None)
    y = O.forward_f64(m, x)
    expect = np.exp(2j * np.pi * np.outer(times.astype(np.float64), fm.astype(np.float64))) @ x / np.sqrt(n)
    assert O.relerr(y, expect) < 1e-13


def test_concomitant_matches_reference_dense_encoding():
    # test_Concomitant.jl:26-50: phase .+= c1 z^2 + c2 (x^2+y^2) + c3 xz + c4 yz
    rng = np.random.default_rng(723)
    grid = Grid(6, 5, 3, np.float32(0.008), np.float32(0.013), np.float32(0.011))
    ns, nc, n = 6, 2, grid.n
    kspha = (rng.normal(size=(16, ns)) * 1e-5).astype(np.float32)
    kspha[:4] *= 1000
    times = (rng.uniform(size=ns) * 1e-3).astype(np.float32)
    fm = (rng.normal(size=n) * 5).astype(np.float32)
    csm = (rng.normal(size=(n, nc)) + 1j * rng.normal(size=(n, nc))).astype(np.complex64)
    conc = (rng.normal(size=(4, ns)) * 1e-1).astype(np.float32)
    mask = np.ones(n, bool)
    mask[0] = mask[-1] = False
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    m = O.make_model(grid, kspha, times, fm, csm, mask, concomitant=conc)
    X, Y, Z = (a.astype(np.float64) for a in (grid.x, grid.y, grid.z))
    bf = O.basis_f64(grid.x, grid.y, grid.z)
    phase = bf @ kspha.astype(np.float64) + np.outer(fm.astype(np.float64), times.astype(np.float64))
    phase += np.outer(Z * Z, conc[0]) + np.outer(X * X + Y * Y, conc[1]) + np.outer(X * Z, conc[2]) + \
        np.outer(Y * Z, conc[3])
    enc = np.zeros((ns * nc, n), dtype=np.complex128)
    for c in range(nc):
        enc[c * ns:(c + 1) * ns] = np.exp(2j * np.pi * phase.T) * csm[:, c][None, :] / np.sqrt(mask.sum())
    enc[:, ~mask] = 0
    assert O.relerr(O.forward_f64(m, x), enc @ x) < 1e-12
    # an absent concomitant equals a zero one (test_Concomitant.jl:117-119)
    m0 = O.make_model(grid, kspha, times, fm, csm, mask)
    mz = O.make_model(grid, kspha, times, fm, csm, mask, concomitant=np.zeros((4, ns), np.float32))
    np.testing.assert_array_equal(O.forward_f64(m0, x), O.forward_f64(mz, x))


def test_prep_kspha_order_selection():
    rng = np.random.default_rng(5)
    k = rng.normal(size=(16, 7)).astype(np.float32)
    kn = rng.normal(size=(3, 7)).astype(np.float32)
    out = O.prep_kspha(k, kn, 16, "0011")
    assert np.all(out[0] == 0)
    np.testing.assert_array_equal(out[1:4], kn)
    np.testing.assert_array_equal(out[4:], k[4:])
    out = O.prep_kspha(k, kn, 16, "1100")
    np.testing.assert_array_equal(out[:4], k[:4])
    assert np.all(out[4:] == 0)
    np.testing.assert_array_equal(O.prep_kspha(k, kn, 16), k)
    k9 = k[:9]
    np.testing.assert_array_equal(O.prep_kspha(k9, kn, 9, "110")[4:9], 0)
    for bad in ("012", "11", "11111"):
        with pytest.raises(ValueError):
            O.prep_kspha(k, kn, 16, bad)


def test_cgnr_matches_qr_least_squares():
    """test_ReconstructionStability.jl:25-76: 3x3 grid, 37 samples, 2 coils, complex weights,
    30 CGNR iterations vs QR of the weighted system (FP64 bar 1e-10)."""
    rng = np.random.default_rng(514)
    ns, nc = 37, 2
    data = rng.normal(size=(ns, nc)) + 1j * rng.normal(size=(ns, nc))
    weight = (0.5 + rng.uniform(size=ns)) * np.exp(1j * rng.uniform(size=ns))
    grid = Grid(3, 3, 1, 1.0, 1.0, 1.0)
    kspha = (rng.normal(size=(9, ns)) * 0.01).astype(np.float32)
    for j in range(ns):
        kspha[1, j] = ((j % 3) - 1) / 3
        kspha[2, j] = (((j // 3) % 3) - 1) / 3
        kspha[3, j] = 0
    times = np.linspace(0, 1e-3, ns).astype(np.float32)
    csm = (np.exp(1j * rng.uniform(size=(9, nc))) / np.sqrt(nc)).astype(np.complex64)
    m = O.make_model(grid, kspha, times, None, csm, None)
    A = O.dense_matrix_f64(m)
    wr = np.tile(weight, nc)
    expected = np.linalg.lstsq(A * wr[:, None], data.reshape(-1, order="F") * wr, rcond=None)[0]
    D = np.abs(weight) ** 2
    rhs = O.adjoint_f64(m, O.expand_weights(m, D) * data.reshape(-1, order="F"))
    x, log = O.cgnr(lambda v: O.normal_f64(m, v, D), rhs, 30)
    assert O.relerr(x, expected) < 1e-10
    assert log[0] == 1.0 and log[-1] < 1e-8


def test_gradient_oracle_matches_finite_differences():
    rng = np.random.default_rng(6)
    grid = Grid(4, 4, 1, np.float32(0.02), np.float32(0.02), np.float32(0.02))
    ns, nc = 12, 3
    kspha = (rng.normal(size=(16, ns)) * 2).astype(np.float32)
    times = np.linspace(0, 1e-3, ns).astype(np.float32)
    fm = rng.uniform(-50, 50, grid.n).astype(np.float32)
    csm = (rng.normal(size=(grid.n, nc)) + 1j * rng.normal(size=(grid.n, nc))).astype(np.complex64)
    x = (rng.normal(size=grid.n) + 1j * rng.normal(size=grid.n)).astype(np.complex64)
    m = O.make_model(grid, kspha, times, fm, csm, None)
    y = O.forward_f64(O.with_coefficients(m, m.k64.T * 1.01), x)
    g, loss, _ = O.grad_f64(m, x, y)
    k0 = m.k64.T.copy()
    for (l, s) in [(0, 0), (1, 3), (2, 7), (5, 11), (12, 5), (15, 2)]:
        eps = 1e-6
        kp, km = k0.copy(), k0.copy()
        kp[l, s] += eps
        km[l, s] -= eps
        lp = np.linalg.norm(O.forward_f64(O.with_coefficients(m, kp), x) - y) ** 2
        lm = np.linalg.norm(O.forward_f64(O.with_coefficients(m, km), x) - y) ** 2
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[l, s]) <= 1e-6 * max(1.0, abs(fd)), (l, s, fd, g[l, s])


def test_derivative_product_is_the_coefficient_derivative_of_the_forward():
    """B x (array_operator.jl:187-235) is d/de E(k + e kdt) x at e = 0: central
    differences of the FP64 forward along kdt pin the 2 pi i factor, the sign,
    the basis contraction and the untouched t * fieldmap term; linearity in kdt."""
    rng = np.random.default_rng(21)
    grid = Grid(5, 4, 1, np.float32(0.02), np.float32(0.02), np.float32(0.02))
    ns, nc = 10, 2
    kspha = (rng.normal(size=(16, ns)) * 3).astype(np.float32)
    kdt = (rng.normal(size=(16, ns)) * 2).astype(np.float32)
    times = np.linspace(0, 2e-3, ns).astype(np.float32)
    fm = rng.uniform(-80, 80, grid.n).astype(np.float32)
    csm = (rng.normal(size=(grid.n, nc)) + 1j * rng.normal(size=(grid.n, nc))).astype(np.complex64)
    x = (rng.normal(size=grid.n) + 1j * rng.normal(size=grid.n)).astype(np.complex64)
    m = O.make_model(grid, kspha, times, fm, csm, None)
    bx = O.forward_dt_f64(m, kdt, x)
    eps = 1e-6
    k0, d = m.k64.T, kdt.astype(np.float64)
    fd = (O.forward_f64(O.with_coefficients(m, k0 + eps * d), x)
          - O.forward_f64(O.with_coefficients(m, k0 - eps * d), x)) / (2 * eps)
    assert np.linalg.norm(bx - fd) <= 1e-6 * np.linalg.norm(fd)
    kdt2 = (rng.normal(size=(16, ns))).astype(np.float32)
    lin = O.forward_dt_f64(m, kdt + 2 * kdt2, x)
    ref = bx + 2 * O.forward_dt_f64(m, kdt2, x)
    assert np.linalg.norm(lin - ref) <= 1e-5 * np.linalg.norm(ref)
    assert np.linalg.norm(O.forward_dt_f64(m, np.zeros_like(kdt), x)) == 0.0


def test_catmull_rom_delay():
    ns = 40
    s = np.arange(ns, dtype=np.float64)
    q = 0.3 * s ** 2 - 2 * s + 1
    for tau_over_dt in (0.25, 0.5, 1.75, -0.4, 3.0):
        out = O.delayed(q[None, :], tau_over_dt * 1e-6, 1e-6)[0]
        u = s + tau_over_dt
        inner = (u >= 1) & (u <= ns - 3)
        np.testing.assert_allclose(out[inner], (0.3 * u ** 2 - 2 * u + 1)[inner], rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(O.delayed(q[None, :], 0.0, 1e-6)[0], q)
    np.testing.assert_allclose(O.delayed(q[None, :], 2e-6, 1e-6)[0][:ns - 3], q[2:ns - 1], atol=1e-12)
    for f in (0.0, 0.3, 0.7):
        w, dw = O.catmull_rom_weights(f)
        assert abs(w.sum() - 1) < 1e-14 and abs(dw.sum()) < 1e-14
        # derivative weights reproduce the slope of a line
        assert abs(dw @ np.array([-1.0, 0.0, 1.0, 2.0]) - 1.0) < 1e-14


def test_golden_fixture():
    """Committed vectors of the deterministic fixture (tests/golden/make_golden.py)."""
    from pathlib import Path
    g = np.load(Path(__file__).parent / "golden" / "explicit_fixture.npz")
    p = O.explicit_highorder_test_problem()
    m = O.make_model(p["grid"], p["kspha"], p["times"], p["fieldmap"], p["csm"], p["mask"])
    np.testing.assert_array_equal(g["image"], p["image"])
    np.testing.assert_array_equal(g["kspha"], p["kspha"])
    assert O.relerr(O.forward_f64(m, p["image"]), g["forward"]) < 1e-14
    assert O.relerr(O.adjoint_f64(m, g["probe"]), g["adjoint"]) < 1e-14
    assert O.relerr(O.normal_f64(m, p["image"], g["weights"]), g["normal"]) < 1e-14
