"""GIRF prediction on the GPU (hoe_apply_girf: hipFFT + HIP kernels) against the FP64
oracle of apply_girf (apply_girf.jl:39-150): FP32 within 2e-6 of the output norm, FP64
within 1e-12; every branch (response expansion or not, rbw band limit, extra and
permuted dimensions, smooth and non-smooth lengths)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import oracle as O  # noqa: E402
import hoe  # noqa: E402

CASES = [
    # (G shape, dim_spatial, dim_time, nFreq, nOut, rbw)
    ((3, 1000), 1, 2, 1000, 16, 1.0),        # N = nFreq: no expansion
    ((3, 1500), 1, 2, 301, 16, 0.9),         # expansion 301 -> 1500, band limit
    ((4, 3, 700, 2), 2, 3, 700, 4, 1.0),     # extra dims, N = 720 > 700: expansion
    ((1009, 3), 2, 1, 1024, 9, 0.75),        # time first, prime length padded to 1024
    ((3, 64, 5), 1, 2, 61, 16, 1.0),         # odd nFreq, the left-out response sample
]


def _hw(n_freq, n_out, seed):
    rng = np.random.default_rng(seed)
    # a smooth response with a delay-like phase and cross terms
    f = np.fft.fftshift(np.fft.fftfreq(n_freq))
    base = np.exp(-(f / 0.2) ** 2)[:, None, None] * np.exp(-2j * np.pi * f * 3)[:, None, None]
    return (base * (rng.normal(size=(1, 3, n_out)) + 1j * rng.normal(size=(1, 3, n_out)))
            + 0.01 * (rng.normal(size=(n_freq, 3, n_out)) + 1j * rng.normal(size=(n_freq, 3, n_out))))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_apply_girf_vs_fp64(case, gpu_lib):
    shape, ds, dt, n_freq, n_out, rbw = CASES[case]
    G = np.random.default_rng(case).normal(size=shape)
    Hw = _hw(n_freq, n_out, case + 10)
    ref = O.apply_girf_f64(G, Hw, dim_spatial=ds, dim_time=dt, rbw=rbw)
    out64 = hoe.apply_girf(G, Hw, dim_spatial=ds, dim_time=dt, rbw=rbw)
    out32 = hoe.apply_girf(G.astype(np.float32), Hw.astype(np.complex64), dim_spatial=ds, dim_time=dt, rbw=rbw)
    assert out64.shape == ref.shape and out32.dtype == np.float32 and out64.dtype == np.float64
    e64, e32 = O.relerr(out64, ref), O.relerr(out32, ref)
    print(f"apply_girf {shape} nFreq={n_freq} nOut={n_out} rbw={rbw}: FP64 {e64:.1e}, FP32 {e32:.1e}")
    assert e64 < 1e-12
    assert e32 < 2e-6
    # the model overload forwards Hw
    m = hoe.GIRFModel(Hw, np.fft.fftshift(np.fft.fftfreq(n_freq)))
    np.testing.assert_array_equal(hoe.apply_girf(G, m, dim_spatial=ds, dim_time=dt, rbw=rbw), out64)


def test_apply_girf_stack_of_spirals_timing(gpu_lib):
    """48 waveforms x 8000 samples, 16 output channels (a stack-of-spirals readout)."""
    rng = np.random.default_rng(7)
    G = rng.normal(size=(3, 8000, 48)).astype(np.float32)
    Hw = _hw(8000, 16, 3).astype(np.complex64)
    t0 = time.perf_counter()
    hoe.apply_girf(G, Hw)                    # plans are made on the first call
    first_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    out = hoe.apply_girf(G, Hw)
    gpu_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    ref = O.apply_girf_f64(G, Hw)
    cpu_ms = (time.perf_counter() - t0) * 1e3
    e = O.relerr(out, ref)
    print(f"apply_girf 48 x 8000 x 16: GPU {gpu_ms:.1f} ms (host arrays and copies; first call with "
          f"planning {first_ms:.1f} ms), "
          f"numpy FP64 oracle {cpu_ms:.1f} ms, rel err {e:.1e}")
    assert e < 2e-6
