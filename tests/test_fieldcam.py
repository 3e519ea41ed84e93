"""hoe.fieldcam on the CPU: density weights, data demodulation, and the demo-data
preparation of demo/DemoCommon.jl checked index by index against the reference's
.mat file (skipped where the reference's demo data is absent)."""
from pathlib import Path

import numpy as np
import pytest

import hoe
from hoe import fieldcam

DEMO = Path("/root/reference/demo")


def _lattice(n, step=1):
    g = (np.arange(n) - n / 2) / n
    kx, ky = np.meshgrid(g, g[::step], indexing="ij")
    return np.stack([kx.ravel(), ky.ravel()])


def test_density_full_cartesian_is_one():
    w = np.abs(fieldcam.samplingDensity(_lattice(32), (32, 32))) ** 2
    assert np.allclose(w, 1.0, rtol=1e-6)


def test_density_scale_free_and_denser_centre_gets_less_weight():
    rng = np.random.default_rng(3)
    # radial-like sampling: density falls with radius
    r = np.sqrt(rng.uniform(0, 1, 4000)) ** 2 * 0.5
    th = rng.uniform(0, 2 * np.pi, 4000)
    tr = np.stack([r * np.cos(th), r * np.sin(th)])
    w1 = np.abs(fieldcam.samplingDensity(tr, (32, 32)))
    w2 = np.abs(fieldcam.samplingDensity(tr * 250.0, (32, 32)))
    assert np.allclose(w1, w2, rtol=1e-5)
    inner, outer = w1[r < 0.1].mean(), w1[r > 0.4].mean()
    assert outer > 2 * inner


def test_density_sparser_lattice_gets_more_weight():
    w = np.abs(fieldcam.samplingDensity(_lattice(32, step=2), (32, 32))) ** 2
    assert np.all(w > 1.1)


def test_density_errors():
    with pytest.raises(hoe.DimensionMismatch):
        fieldcam.samplingDensity(np.zeros((3, 10)), (8, 8))
    with pytest.raises(hoe.ArgumentError):
        fieldcam.samplingDensity(np.zeros((2, 10)), (8, 8))
    with pytest.raises(hoe.ArgumentError):
        fieldcam.samplingDensity(np.ones((2, 10)), (8, 0))


def test_demodulate():
    rng = np.random.default_rng(1)
    kd = rng.standard_normal((50, 3)) + 1j * rng.standard_normal((50, 3))
    k0, kz = rng.uniform(-1, 1, 50), rng.uniform(-1, 1, 50)
    ref = kd / np.exp(2j * np.pi * k0)[:, None] * np.exp(-2j * np.pi * kz)[:, None]
    assert np.allclose(fieldcam.demodulate(kd, k0, kz), ref, rtol=1e-12, atol=0)
    with pytest.raises(hoe.DimensionMismatch):
        fieldcam.demodulate(kd, k0[:-1], kz)


@pytest.mark.parametrize("stem", ["7T_2D_Spiral_1p0_200_r4", "7T_2D_EPI_1p0_200_r4"])
def test_prepare_demo_against_the_mat_file(stem):
    path = DEMO / f"{stem}.mat"
    if not path.exists():
        pytest.skip("the reference's demo data is not present")
    raw = fieldcam.load_demo_mat(path)
    p = fieldcam.prepare_demo(raw, "measured")
    nX, nY = (int(v) for v in np.asarray(raw["matrixSize"]).reshape(-1)[:2])
    # permutedims(reverse(circshift(a, (0, -1)), dims=1), [2, 1]) in 1-based Julia indexing:
    # b[i, j] = a[nX + 1 - j, mod1(i + 1, nY)]; flattened x fastest.
    rng = np.random.default_rng(0)
    for _ in range(200):
        i, j, c = rng.integers(nX), rng.integers(nY), rng.integers(32)
        src = (nX - 1 - j, (i + 1) % nY)
        assert p["csm"][i + nX * j, c] == raw["gre_csm"][src + (c,)]
        assert p["fieldmap"][i + nX * j] == np.float32(raw["gre_b0"][src])
        assert p["mask"][i + nX * j] == bool(raw["gre_mask"][src])
    # coefficients at the ADC times with the delay: a field sample that falls on an ADC time
    kf = np.asarray(raw["ksphaMeasured"], dtype=np.float64)
    t = np.asarray(raw["datatime"]).reshape(-1)
    t_field = float(raw["dt_Measured"]) * np.arange(kf.shape[0]) - (
        float(raw["startMeasured"]) + float(raw["tauMeasured"]) * float(raw["dt_adc"]))
    j = kf.shape[0] // 2
    s = int(np.argmin(np.abs(t - t_field[j])))
    if abs(t[s] - t_field[j]) < 1e-12:
        assert np.allclose(p["kspha"][:, s], kf[j], rtol=1e-6, atol=1e-6)
    assert p["kspha"].shape == (16, t.size) and p["data"].shape == (t.size, 32)
    assert p["recon_terms"] == "0111" and p["reconSize"] == (nX, nY)
    assert p["grid"].matrixSize == (nX, nY, 1)
    nom = fieldcam.prepare_demo(raw, "nominal")
    assert nom["kspha"].shape == (9, t.size) and nom["recon_terms"] == "010"
    assert np.array_equal(nom["data"], np.asarray(raw["kdata"]).astype(np.complex64))


def test_committed_reference_values_of_the_7t_demo():
    """tests/golden/7t_reference.npz (committed; the large operator inputs are generated):
    the reference's own rendered reconstructions of its two 7 T demo datasets
    (demo/result/*_{Measured,Nominal}.png on the 200 x 200 grid) and the delays stored with
    the data (tauMeasured / tauNominal, in ADC dwells), which the GPU real-data tests
    compare against."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).parent / "golden"))
    import fixture7t
    r = fixture7t.reference()
    taus = {("spiral", "measured"): -0.568, ("spiral", "nominal"): 5.008,
            ("epi", "measured"): -0.408, ("epi", "nominal"): 1.675}
    for (name, traj), tau in taus.items():
        assert abs(float(r[f"{name}_tau_{traj}"]) - tau) < 1e-3
        img = r[f"{name}_golden_{traj}"].astype(np.float32)
        assert img.shape == (200, 200)
        assert img.min() >= 0.0 and img.max() <= 1.0 and img.std() > 0.1
    assert float(r["spiral_dt_adc"]) == 1e-6 and abs(float(r["epi_dt_adc"]) - 1.4e-6) < 1e-12
    # the four images are distinct reconstructions
    ims = [r[k].astype(np.float64).ravel() for k in r.files if "golden" in k]
    for i in range(len(ims)):
        for j in range(i + 1, len(ims)):
            assert fixture7t.corr(ims[i], ims[j]) < 0.99
