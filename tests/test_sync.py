"""Delay-estimation host code without a GPU: the coefficient resampler, the derivative
stencil, the update rule of the delay loop, argument checks, and the FP64 oracle loop.

- Akima resampling (Interpolations.jl's AkimaMonotonicInterpolation is not vendored)
  is pinned against scipy's independent implementation of Akima (1970), and by
  exactness on linear data and at the nodes;
- InterpTrajTime's sign and zero-outside conventions (InterpTraj.jl:18-44) are pinned
  by a linear time function, for which the shifted result is known in closed form;
- the stencil (FindDelay.jl:178-180) against numpy's full convolution cropped by 2;
- the oracle loop (FindDelay.jl:152-214) by recovering a planted delay from noiseless
  data.
"""
import numpy as np
import pytest
from scipy.interpolate import Akima1DInterpolator

import hoe
import oracle as O
from hoe import sync


def test_akima_matches_scipy_and_is_exact_on_lines():
    rng = np.random.default_rng(0)
    x = np.cumsum(rng.uniform(0.5, 1.5, 60))
    y = rng.normal(size=(60, 5)).cumsum(0)
    xq = np.linspace(x[0] - 2, x[-1] + 2, 3001)
    got = sync.interp_columns(x, y, xq)
    ref = np.nan_to_num(Akima1DInterpolator(x, y)(xq), nan=0.0)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(y).max()
    # nodes are reproduced, lines are exact, zero outside
    np.testing.assert_allclose(sync.interp_columns(x, y, x), y, rtol=0, atol=1e-12)
    inside = (xq >= x[0]) & (xq <= x[-1])
    lin = sync.interp_columns(x, (2.5 * x - 1)[:, None], xq)[:, 0]
    np.testing.assert_allclose(lin[inside], 2.5 * xq[inside] - 1, rtol=1e-13)
    assert np.all(lin[~inside] == 0)
    np.testing.assert_allclose(sync.interp_columns(x, y[:, 0], xq, "linear")[inside],
                               np.interp(xq[inside], x, y[:, 0]), rtol=1e-13)
    with pytest.raises(hoe.ArgumentError):
        sync.interp_columns(x[::-1], y, xq)
    with pytest.raises(hoe.ArgumentError):
        sync.interp_columns(x[:2], y[:2], xq)
    with pytest.raises(hoe.ArgumentError):
        sync.interp_columns(x, y, xq, intermode="steffen")


def test_interp_traj_time_conventions():
    dt, n = 2e-6, 400
    t = dt * np.arange(n)
    traj = np.stack([1 + 3e4 * t, -2 + 1e5 * t], axis=1).astype(np.float32)   # linear in time
    datatime = (dt * np.arange(0, 300) + 0.3 * dt).astype(np.float32)
    d = 7.6 * dt
    out = hoe.InterpTrajTime(traj, dt, d, datatime)
    assert out.dtype == np.float32 and out.shape == (300, 2)
    # positive delTime evaluates the field at datatime + delTime
    tq = datatime.astype(np.float64) + d
    np.testing.assert_allclose(out[:, 0], 1 + 3e4 * tq, rtol=2e-6)
    np.testing.assert_allclose(out[:, 1], -2 + 1e5 * tq, rtol=2e-6)
    # beyond the shifted interval the result is zero, not extrapolated
    late = hoe.InterpTrajTime(traj, dt, 150 * dt, datatime)
    assert np.all(late[datatime + 150 * dt > t[-1] + 1e-12] == 0)
    early = hoe.InterpTrajTime(traj, dt, -5 * dt, datatime)
    assert np.all(early[datatime - 5 * dt < -1e-12] == 0)
    # default datatime and the 3-D overload (each shot on its own)
    np.testing.assert_allclose(hoe.InterpTrajTime(traj, dt, 0.0), traj, rtol=1e-6)
    rng = np.random.default_rng(1)
    traj3 = rng.normal(size=(3, n, 4)).cumsum(1).astype(np.float32)
    dts = np.array([0.0, 1.5 * dt, -2.25 * dt], np.float32)
    dtm = np.stack([datatime, datatime + dt, datatime * 1.01]).astype(np.float32)
    out3 = hoe.InterpTrajTime(traj3, dt, dts, dtm)
    for i in range(3):
        np.testing.assert_array_equal(out3[i], hoe.InterpTrajTime(traj3[i], dt, dts[i], dtm[i]))
        np.testing.assert_allclose(out3[i], O.interp_traj_time(traj3[i], dt, dts[i], dtm[i]), atol=2e-5)
    with pytest.raises(AssertionError):
        hoe.InterpTrajTime(traj3, dt, dts[:2], dtm)


def test_kspha_derivative_stencil():
    rng = np.random.default_rng(2)
    k = rng.normal(size=(500, 16)).cumsum(0).astype(np.float32)
    dt = np.float32(1e-6)
    got = hoe.kspha_derivative(k, dt)
    assert got.dtype == np.float32 and got.shape == k.shape
    np.testing.assert_allclose(got, O.kspha_dt_stencil(k, dt), rtol=1e-5, atol=1e-5 * np.abs(got).max())
    # a ramp has derivative slope/dt away from the zero-padded ends
    ramp = (np.arange(100, dtype=np.float32)[:, None] * np.ones((1, 3), np.float32))
    d = hoe.kspha_derivative(ramp, np.float32(0.5))
    np.testing.assert_allclose(d[2:-2], 2.0, rtol=1e-6)
    # the 3-D form applies the stencil per shot
    k3 = np.stack([k, 2 * k])
    d3 = hoe.kspha_derivative(k3, dt)
    np.testing.assert_array_equal(d3[0], got)
    np.testing.assert_allclose(d3[1], 2 * got, rtol=1e-6)


def test_delay_update_rule():
    """FindDelay.jl:196-213: Δτ = J*step, J halves (not below 1) after a sign change,
    the result is τ - Δτ/J*(J-1)."""
    loop = sync._Loop(3, 0.01, 1.0, None)
    for step, exp_tau, exp_J in [(1.0, 3.0, 3), (-0.5, 1.5, 1.5), (0.2, 1.8, 1), (0.004, 1.804, 1)]:
        assert loop.running()
        loop.update(step)
        assert abs(loop.τ - exp_tau) < 1e-12 and loop.J == exp_J
    assert not loop.running()
    assert abs(loop.result(float) - 1.804) < 1e-12
    loop = sync._Loop(3, 0.0, 1.0, 2)        # an update limit ends the loop
    loop.update(0.1)
    loop.update(0.1)
    assert not loop.running()
    assert abs(loop.result(float) - (0.6 - 0.3 / 3 * 2)) < 1e-12
    # y2 \ y1 for vectors: the least-squares scalar
    rng = np.random.default_rng(3)
    y2 = rng.normal(size=50) + 1j * rng.normal(size=50)
    a = 0.7 - 0.2j
    assert abs(sync._delay_step(a * y2, y2) - 0.7) < 1e-12


def test_find_delay_argument_checks():
    p = O.delay_test_problem(n=8, n_cha=2, n_field=200, n_adc=150)
    g, data = p["grid"], p["data"]
    with pytest.raises(AssertionError):         # FindDelay.jl:153
        hoe.FindDelay(g, data, p["kspha"], p["datatime"], p["StartTime"], p["dt"],
                      csm=np.ones((8, 8, 3), np.complex64))
    with pytest.raises(AssertionError):         # FindDelay.jl:154
        hoe.FindDelay(g, data, p["kspha"], p["datatime"][:-1], p["StartTime"], p["dt"],
                      csm=p["csm"].reshape(8, 8, 2, order="F"))
    with pytest.raises(hoe.ArgumentError):
        hoe.FindDelay(g, data, p["kspha"], p["StartTime"])
    with pytest.raises(hoe.DimensionMismatch):
        hoe.FindDelay_multishot(g, data, p["kspha"], p["StartTime"], p["dt"])


def test_oracle_delay_loop_recovers_a_planted_delay():
    p = O.delay_test_problem()
    dt = float(p["dt"])
    tau, hist = O.find_delay_f64(p["grid"], p["data"], p["kspha"], p["datatime"], float(p["StartTime"]), dt,
                                 dtau_min=1e-3 * dt, fieldmap=p["fieldmap"], csm=p["csm"], recon_terms="1111",
                                 iter_max=20)
    print("oracle delay", tau / dt, "planted", p["tau"] / dt, np.round(np.array(hist) / dt, 4))
    assert abs(tau - p["tau"]) < 0.01 * dt
    assert len(hist) <= 8
