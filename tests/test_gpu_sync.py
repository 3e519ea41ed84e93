"""Delay estimation on the GPU (FindDelay.jl:139-214, FindDelay_multishot.jl:65-159).

- ``set_kspha`` gives the operator a freshly built one would have (bit-equal products,
  recon_terms applied), the premise of reusing one handle across trial delays;
- FindDelay's iterates follow the FP64 oracle loop on the same inputs and recover the
  planted delay;
- FindDelay_multishot with one shot repeats FindDelay; with two shots it recovers the
  shared planted delay;
- at 200x200, 32 coils, 16000 samples the loop recovers a planted delay (data simulated
  by the FP32 forward, which the operator tests pin against the oracle).
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402
from hoe import synth  # noqa: E402


def test_set_kspha_equals_a_new_operator(gpu_lib):
    p = O.delay_test_problem(n=12, n_cha=3, n_field=300, n_adc=250)
    rng = np.random.default_rng(4)
    k1 = rng.normal(size=(16, 250)).astype(np.float32) * 5
    k2 = rng.normal(size=(16, 250)).astype(np.float32) * 5
    x = (rng.normal(size=p["grid"].n) + 1j * rng.normal(size=p["grid"].n)).astype(np.complex64)
    for backend in ("direct", "mixed_fp16"):
        for rt in ("1111", "1010"):
            kw = dict(fieldmap=p["fieldmap"], csm=p["csm"], recon_terms=rt, backend=backend)
            with hoe.HighOrderExplicitOp(p["grid"], k1, p["datatime"], **kw) as a, \
                    hoe.HighOrderExplicitOp(p["grid"], k2, p["datatime"], **kw) as b:
                a.set_kspha(k2)
                np.testing.assert_array_equal(a * x, b * x)
                y = a * x
                np.testing.assert_array_equal(a.H * y, b.H * y)
            m = O.make_model(p["grid"], k2, p["datatime"], p["fieldmap"], p["csm"], None, recon_terms=rt)
            assert O.relerr(y, O.forward_f64(m, x)) < 3e-5


# unit weights: the oracle loop's default (the delay-update rule is what is compared)
UNIT = lambda k: np.ones(np.asarray(k).shape[1], np.float32)  # noqa: E731


def test_find_delay_default_weights_are_the_density(gpu_lib):
    """Without ``weight`` the estimators recompute the reference's density weights,
    samplingDensity of the resampled rows 2:3, at every update (FindDelay.jl:183,
    FindDelay_multishot.jl:120): the same updates as passing that weight explicitly,
    and not those of unit weights."""
    p = O.delay_test_problem()
    dt = float(p["dt"])
    args = (p["grid"], p["data"], p["kspha"], p["datatime"], p["StartTime"], p["dt"])
    kw = dict(Δτ_min=np.float32(1e-3 * dt), fieldmap=p["fieldmap"], csm=p["csm"], recon_terms="1111",
              iter_max=10, max_updates=4, return_history=True)
    _, h_def = hoe.FindDelay(*args, **kw)
    dens = lambda k: hoe.samplingDensity(np.asarray(k)[1:3], (p["grid"].nX, p["grid"].nY))  # noqa: E731
    _, h_exp = hoe.FindDelay(*args, weight=dens, **kw)
    _, h_unit = hoe.FindDelay(*args, weight=UNIT, **kw)
    assert h_def == h_exp
    assert h_def != h_unit
    _, hm = hoe.FindDelay_multishot(p["grid"], p["data"], p["kspha"][None], p["datatime"][None],
                                    np.array([p["StartTime"]]), p["dt"], **kw)
    np.testing.assert_allclose(np.array(hm) / dt, np.array(h_def) / dt, atol=1e-5)


@pytest.mark.parametrize("backend", ["direct", "mixed_fp16"])
def test_find_delay_follows_the_oracle(backend, gpu_lib):
    """10 inner CGNR iterations: there FP32 and FP64 CG give delay iterates within
    3e-4 dt of each other on this problem. At 20 iterations the CG arithmetic alone
    moves them by 0.01 dt (the FP64 oracle loop run with a complex64 CG reproduces
    the GPU iterates), because the update divides by a residual that CG drives small."""
    p = O.delay_test_problem()
    dt = float(p["dt"])
    args = (p["grid"], p["data"], p["kspha"], p["datatime"], p["StartTime"], p["dt"])
    t0 = time.perf_counter()
    tau, hist = hoe.FindDelay(*args, Δτ_min=np.float32(1e-3 * dt), fieldmap=p["fieldmap"], csm=p["csm"],
                              recon_terms="1111", iter_max=10, backend=backend, max_updates=30,
                              return_history=True, weight=UNIT)
    wall = time.perf_counter() - t0
    tau_o, hist_o = O.find_delay_f64(p["grid"], p["data"], p["kspha"], p["datatime"], float(p["StartTime"]), dt,
                                     dtau_min=1e-3 * dt, fieldmap=p["fieldmap"], csm=p["csm"],
                                     recon_terms="1111", iter_max=10)
    print(f"FindDelay {backend}: τ/dt {tau / dt:.5f} (oracle {tau_o / dt:.5f}, planted {p['tau'] / dt}), "
          f"{len(hist) - 1} updates, {wall * 1e3:.1f} ms")
    assert isinstance(tau, np.float32)
    n = min(len(hist), len(hist_o))
    np.testing.assert_allclose(np.array(hist[:n]) / dt, np.array(hist_o[:n]) / dt, atol=2e-3)
    assert abs(tau - tau_o) < 2e-3 * dt
    assert abs(tau - p["tau"]) < 0.03 * dt      # 10 CG iterations leave a 0.02 dt bias


def test_find_delay_multishot(gpu_lib):
    p = O.delay_test_problem()
    dt = float(p["dt"])
    common = dict(Δτ_min=np.float32(1e-3 * dt), fieldmap=p["fieldmap"], csm=p["csm"], recon_terms="1111",
                  iter_max=20, max_updates=30, return_history=True, weight=UNIT)
    tau1, h1 = hoe.FindDelay(p["grid"], p["data"], p["kspha"], p["datatime"], p["StartTime"], p["dt"], **common)
    tau_m, hm = hoe.FindDelay_multishot(p["grid"], p["data"], p["kspha"][None], p["datatime"][None],
                                        np.array([p["StartTime"]]), p["dt"], **common)
    np.testing.assert_allclose(np.array(hm) / dt, np.array(h1) / dt, atol=1e-5)
    assert abs(tau_m - tau1) < 1e-5 * dt
    # two shots with their own coefficients and start times, one planted delay
    q = O.delay_test_problem(seed=9, start=60)
    data = np.concatenate([p["data"], q["data"]], axis=0)
    tau2, h2 = hoe.FindDelay_multishot(p["grid"], data, np.stack([p["kspha"], q["kspha"]]),
                                       np.stack([p["datatime"], q["datatime"]]),
                                       np.array([p["StartTime"], q["StartTime"]]), p["dt"], **common)
    print(f"FindDelay_multishot: τ/dt {tau2 / dt:.5f} (planted {p['tau'] / dt}), {len(h2) - 1} updates")
    assert abs(tau2 - p["tau"]) < 0.02 * dt


def test_find_delay_at_200x200_32_coils(gpu_lib):
    """Full-size properties on data simulated by the operator with a planted delay. At this
    size (R = 4 spiral, 40000 voxels) the image reconstructed at a wrong delay absorbs most
    of the mismatch, so the alternating updates approach the delay linearly and slowly:
    checked are the direction (every update moves toward the planted delay, none beyond it)
    and the fixed point (with the right StartTime the estimate stays at 0)."""
    c = synth.make(1)
    dt = float(c["dwell"])
    ns = c["kspha"].shape[1]
    # field samples: the workload's coefficients on a longer time grid that starts 40 samples early
    n_field = ns + 120
    kf = np.empty((n_field, 16), np.float64)
    src = c["kspha"].T.astype(np.float64)
    kf[40:40 + ns] = src
    kf[:40] = src[0] + (src[1] - src[0]) * np.arange(-40, 0)[:, None]
    kf[40 + ns:] = src[-1] + (src[-1] - src[-2]) * np.arange(1, 81)[:, None]
    kspha = kf.astype(np.float32)
    start = np.float32(40 * dt)
    planted = 2.6 * dt
    datatime = (dt * np.arange(ns)).astype(np.float32)
    k_true = hoe.InterpTrajTime(kspha, dt, float(start) + planted, datatime).T
    with hoe.HighOrderExplicitOp(c["grid"], k_true, datatime, fieldmap=c["fieldmap"], csm=c["csm"],
                                 backend="direct") as op:
        data = (op * c["image"]).reshape(ns, -1, order="F")
    kw = dict(Δτ_min=np.float32(2e-3 * dt), fieldmap=c["fieldmap"], csm=c["csm"], recon_terms="1111",
              iter_max=10, backend="mixed_fp16", return_history=True, weight=UNIT)
    t0 = time.perf_counter()
    tau, hist = hoe.FindDelay(c["grid"], data, kspha, datatime, start, np.float32(dt), max_updates=12, **kw)
    wall = time.perf_counter() - t0
    h = np.array(hist) / dt
    print(f"FindDelay 200x200x32c, 16000 samples, planted 2.6: τ/dt {np.round(h, 3)}, "
          f"{wall / (len(hist) - 1) * 1e3:.0f} ms per update")
    assert np.all(np.diff(h) > 0) and h[-1] < 2.6 + 0.05
    tau0, hist0 = hoe.FindDelay(c["grid"], data, kspha, datatime, np.float32(float(start) + planted),
                                np.float32(dt), max_updates=12, **kw)
    print(f"  with the right StartTime: τ/dt {np.round(np.array(hist0) / dt, 4)}")
    assert abs(tau0) < 0.05 * dt


@pytest.mark.parametrize("intermode", [None, "LinearMonotonicInterpolation"])
def test_device_resampling_matches_the_host(intermode, gpu_lib):
    """The delay loops resample the field samples on the GPU (torch, FP64) with the host
    resampler's formulas: InterpTrajTime's Akima and linear modes (InterpTraj.jl:18-44,
    restated in hoe/sync.py and pinned there against scipy's Akima) at delays inside and
    beyond the first and last samples (zero outside), to 1e-12 of the largest value."""
    import torch
    from hoe import sync
    rng = np.random.default_rng(12)
    n_field, dt = 900, 1e-6
    s = np.arange(n_field) * dt
    traj = np.stack([np.cos(2e4 * s + r) * (r + 1) + rng.normal(size=n_field) * 1e-2 for r in range(16)],
                    axis=1).astype(np.float32)
    datatime = (dt * np.arange(700) + 40 * dt).astype(np.float32)
    host = sync._Resampler(traj, dt, intermode)
    dev = sync._Resampler(traj, dt, intermode).to_device(torch)
    for tau in (0.0, 3.3 * dt, -41.7 * dt, 250.2 * dt):
        a, b = host(tau, datatime), dev(tau, datatime)
        assert a.dtype == b.dtype == np.float32 and a.shape == b.shape == (700, 16)
        ah = sync._eval_columns(host.x0 - tau, host.y, host.t, datatime.astype(np.float64), host.mode)
        bd = dev._eval_device(tau, datatime.astype(np.float64))
        scale = np.max(np.abs(ah))
        assert bd.dtype == np.float64
        assert np.max(np.abs(ah - bd)) <= 1e-12 * scale, (tau, np.max(np.abs(ah - bd)))
        assert np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))) <= 2 ** -22 * scale


def test_device_density_matches_the_host(gpu_lib):
    """samplingDensity's Pipe-Menon iterations on the GPU (segmented sums over the
    (sample, tap) pairs sorted by grid cell) against the host's bincount form, on a
    16 000-sample spiral for a 200x200 grid: to 1e-12, and the same on a second call."""
    from hoe import fieldcam
    kx, ky = synth.spiral(16000, 200 / (2 * 0.2), 12)
    k = np.stack([kx, ky]) / (2.0 * np.max(np.abs(np.stack([kx, ky]))))
    dev = fieldcam._pipe_menon(k, (400, 400), 10)
    import unittest.mock as um
    with um.patch.object(fieldcam, "_torch_gpu", lambda: None):
        host = fieldcam._pipe_menon(k, (400, 400), 10)
    assert dev.shape == host.shape == (16000,)
    assert np.max(np.abs(dev - host)) <= 1e-12 * np.max(np.abs(host))
    assert np.array_equal(dev, fieldcam._pipe_menon(k, (400, 400), 10))
