"""Field-model inputs without a GPU: the FP64 GIRF oracle (apply_girf.jl:39-150), the
trajectory conversions (conversion.jl:14-64), channel unpacking and argument checks.

The GIRF oracle is pinned by responses whose output is known in closed form, each in
the branch without response expansion (nFreq == FFT length) and with it (nFreq smaller):
identity, a pure circular delay, a causal FIR filter (direct time-domain circular
convolution, and numpy's linear one where the record does not wrap), a non-causal advance, the time sample the reference's expansion leaves out, and the rbw
band limit on pure tones.
"""
import warnings

import numpy as np
import pytest

import hoe
import oracle as O


def _waves(shape, seed=0):
    return np.random.default_rng(seed).normal(size=shape)


def _hw_from_impulse(h, n_freq, n_out=3):
    """DC-centred response (nFreq, 3, nOut) of the FIR taps h on channel i -> output i."""
    hp = np.zeros(n_freq, complex)
    hp[:len(h)] = h
    H = np.fft.fftshift(np.fft.fft(hp))
    Hw = np.zeros((n_freq, 3, n_out), complex)
    for i in range(3):
        Hw[:, i, i] = H
    return Hw


def test_oracle_identity_and_pure_delay():
    G = _waves((3, 480))
    for n_freq in (480, 37):     # no expansion / expansion to N = 480
        Hw = np.zeros((n_freq, 3, 3), complex)
        for i in range(3):
            Hw[:, i, i] = 1.0
        np.testing.assert_allclose(O.apply_girf_f64(G, Hw), G, atol=1e-12)
    # circular delay by d samples: H[m] = exp(-2 pi i m d / N) (unshifted), N = 500 > nPoint
    G = _waves((3, 400), 1)
    N, d = 500, 7
    H = np.fft.fftshift(np.exp(-2j * np.pi * np.arange(N) * d / N))
    Hw = np.zeros((N, 3, 3), complex)
    for i in range(3):
        Hw[:, i, i] = H
    out = O.apply_girf_f64(G, Hw)
    exp = np.zeros_like(G)
    exp[:, d:] = G[:, :-d]
    np.testing.assert_allclose(out, exp, atol=1e-12)


def _circular(G, h, N):
    """Direct time-domain circular convolution of the zero-padded rows of G with h, length N."""
    n = G.shape[1]
    gp = np.zeros((G.shape[0], N))
    gp[:, :n] = G
    out = np.zeros_like(gp)
    for k, hk in enumerate(h):
        out += hk * np.roll(gp, k, axis=1)
    return out[:, :n]


def test_oracle_fir_and_expansion_conventions():
    G = _waves((3, 400), 2)
    h = np.random.default_rng(3).normal(size=9)
    for n_freq, N in ((500, 500), (61, 400)):   # exact taps (no wrap), taps expanded from 61 (wraps)
        out = O.apply_girf_f64(G, _hw_from_impulse(h, n_freq))
        np.testing.assert_allclose(out, _circular(G, h, N), atol=1e-11)
        if N >= 400 + len(h) - 1:
            np.testing.assert_allclose(out, np.stack([np.convolve(G[i], h)[:400] for i in range(3)]), atol=1e-11)
    # a two-sample advance sits at the end of the impulse response; the expansion keeps it there
    for n_freq, N in ((500, 500), (61, 400)):
        hp = np.zeros(n_freq)
        hp[n_freq - 2] = 1.0
        Hw = np.zeros((n_freq, 3, 3), complex)
        for i in range(3):
            Hw[:, i, i] = np.fft.fftshift(np.fft.fft(hp))
        out = O.apply_girf_f64(G, Hw)
        gp = np.zeros((3, N))
        gp[:, :400] = G
        np.testing.assert_allclose(out, np.roll(gp, -2, axis=1)[:, :400], atol=1e-11)
    # the time sample nFreq//2 (0-based): a plain delay without expansion, dropped with it
    n_freq = 61
    hp = np.zeros(n_freq)
    hp[n_freq // 2] = 1.0
    Hw = np.zeros((n_freq, 3, 3), complex)
    for i in range(3):
        Hw[:, i, i] = np.fft.fftshift(np.fft.fft(hp))
    np.testing.assert_allclose(O.apply_girf_f64(G, Hw), 0.0, atol=1e-12)
    G61 = _waves((3, 60), 4)
    out = O.apply_girf_f64(G61, Hw)   # nPoint 60 < nFreq 61 -> N = 64: expansion
    np.testing.assert_allclose(out, 0.0, atol=1e-12)
    hp2 = np.zeros(64)
    hp2[32] = 1.0
    Hw2 = np.zeros((64, 3, 3), complex)
    for i in range(3):
        Hw2[:, i, i] = np.fft.fftshift(np.fft.fft(hp2))
    out = O.apply_girf_f64(G61, Hw2)           # N = 64 = nFreq: circular delay by 32
    exp = np.zeros_like(G61)
    exp[:, 32:] = G61[:, :28]
    exp[:, :28] = G61[:, 32:]              # wrapped: the record is only 64 long
    np.testing.assert_allclose(out, exp, atol=1e-12)


def test_oracle_rbw_band_limit_on_tones():
    N = 480
    t = np.arange(N)
    Hw = np.zeros((N, 3, 3), complex)
    for i in range(3):
        Hw[:, i, i] = 1.0
    keep = np.cos(2 * np.pi * 50 * t / N)
    drop = np.cos(2 * np.pi * 150 * t / N)
    G = np.stack([keep, drop, keep + drop])
    out = O.apply_girf_f64(G, Hw, rbw=0.5)      # round(480*0.25) = 120 bins cleared at each end
    np.testing.assert_allclose(out, np.stack([keep, 0 * keep, keep]), atol=1e-12)
    np.testing.assert_allclose(O.apply_girf_f64(G, Hw, rbw=1.0), G, atol=1e-12)


def test_oracle_dims_and_transfer_matrix():
    rng = np.random.default_rng(5)
    G = rng.normal(size=(4, 3, 120, 2))
    Hw = rng.normal(size=(125, 3, 5)) + 1j * rng.normal(size=(125, 3, 5))
    out = O.apply_girf_f64(G, Hw, dim_spatial=2, dim_time=3)
    assert out.shape == (4, 5, 120, 2)
    N = O.nextprod235(125)
    for a in range(4):
        for b in range(2):
            np.testing.assert_allclose(out[a, :, :, b], O.girf_core_f64(G[a, :, :, b], Hw, 120), atol=1e-12)
    # the transfer matrix sums the input channels: linearity in Hw over inputs
    H1 = Hw.copy()
    H1[:, 1:, :] = 0
    H2 = Hw - H1
    np.testing.assert_allclose(O.apply_girf_f64(G, Hw, dim_spatial=2, dim_time=3),
                               O.apply_girf_f64(G, H1, dim_spatial=2, dim_time=3)
                               + O.apply_girf_f64(G, H2, dim_spatial=2, dim_time=3), atol=1e-12)
    assert N == 125


def test_nextprod_and_library_agree():
    lib = hoe.load()
    for n in (1, 2, 7, 11, 97, 1000, 1009, 16001, 384000, 384001):
        assert lib.hoe_girf_fft_length(n) == O.nextprod235(n)


def test_grad2traj_traj2grad():
    dt = 1e-5
    t = np.arange(200) * dt
    g = 3.0 + 2e4 * t                       # linear: the trapezoid rule is exact
    k = hoe.grad2traj(g, dt)
    np.testing.assert_allclose(k, 3.0 * t + 1e4 * t * t, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(k, O.grad2traj_f64(g, dt), rtol=1e-14)
    np.testing.assert_allclose(hoe.traj2grad(k, dt), 0.5 * (g[1:] + g[:-1]), rtol=1e-10)
    # dimension argument and shapes
    G = np.random.default_rng(6).normal(size=(3, 50, 4))
    k2 = hoe.grad2traj(G, dt, dim=2)
    assert k2.shape == G.shape and np.all(k2[:, 0, :] == 0)
    np.testing.assert_allclose(k2, O.grad2traj_f64(G, dt, dim=2), rtol=1e-12)
    assert hoe.traj2grad(G, dt, dim=3).shape == (3, 50, 3)
    # Julia result types: Float32 with a Float64 dwell promotes, with a Float32 dwell stays
    g32 = g.astype(np.float32)
    assert hoe.grad2traj(g32, dt).dtype == np.float64
    assert hoe.grad2traj(g32, np.float32(dt)).dtype == np.float32
    assert hoe.traj2grad(np.arange(5), 2).dtype == np.float64
    assert hoe.traj2grad(g32, 2).dtype == np.float32
    assert hoe.grad2traj(np.arange(5), 1).dtype == np.float64
    with pytest.raises(hoe.ArgumentError):
        hoe.grad2traj(g, dt, dim=2)


def test_unpack_girf_channels():
    G = np.arange(6 * 10, dtype=np.float32).reshape(6, 10)
    lin, k0, ho = hoe.unpack_girf_channels(G)
    np.testing.assert_array_equal(lin, G[1:4])
    np.testing.assert_array_equal(k0, G[0:1])
    np.testing.assert_array_equal(ho, G[4:6])
    lin, k0, ho = hoe.unpack_girf_channels(G[:4].T, dim_spatial=2)
    np.testing.assert_array_equal(lin, G[1:4].T)
    assert ho is None
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        lin, k0, ho = hoe.unpack_girf_channels(G[:3])
        assert len(w) == 1
    np.testing.assert_array_equal(lin, G[:3])
    assert k0.shape == (1, 10) and np.all(k0 == 0) and ho is None


def test_apply_girf_argument_checks():
    with pytest.raises(hoe.ArgumentError):            # apply_girf.jl:62
        hoe.apply_girf(np.zeros((2, 10)), np.zeros((10, 3, 4), complex))
    with pytest.raises(AssertionError):               # apply_girf.jl:66
        hoe.apply_girf(np.zeros((3, 10)), np.zeros((10, 2, 4), complex))
    with pytest.raises(hoe.ArgumentError):
        hoe.apply_girf(np.zeros((3, 10), np.int32), np.zeros((10, 3, 4), complex))
    with pytest.raises(hoe.ArgumentError):
        hoe.apply_girf(np.zeros((3, 10)), np.zeros((10, 3, 4), complex), dim_time=1)
    with pytest.raises(hoe.ArgumentError):
        hoe.GIRFModel(np.zeros((10, 3, 4)), np.zeros(10))
    m = hoe.GIRFModel(np.zeros((10, 3, 4), complex), np.zeros((1, 10)))
    assert m.freqs.shape == (10,)
