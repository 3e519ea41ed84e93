"""Reader of the 7 T fixtures written by make_7t_fixture.py (shared by the real-data
tests and bench.py). load() returns None when the large per-dataset fixture has not been
generated; reference() reads the small committed one (the reference's rendered images
and stored delays)."""
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def reference():
    """The committed reference-held values: rendered reconstructions and stored delays."""
    return np.load(HERE / "7t_reference.npz")


def load(name, trajectory):
    """Operator inputs, data, weights and the reference's rendered images of one
    dataset ("spiral" or "epi") and trajectory ("measured" or "nominal")."""
    import hoe
    from hoe import fieldcam
    path = HERE / f"7t_{name}.npz"
    if not path.exists():
        return None
    z = np.load(path)
    r = reference()
    gm = r[f"{name}_golden_measured"].astype(np.float32)
    gn = r[f"{name}_golden_nominal"].astype(np.float32)
    nX, nY, nZ = (int(v) for v in z["matrix"])
    res = z["resolution"]
    grid = hoe.Grid(nX, nY, nZ, res[0], res[1], res[2])
    mask = z["mask"]
    csm = np.zeros((grid.n, z["csm_masked"].shape[1]), dtype=np.complex64)
    csm[mask] = z["csm_masked"]
    if trajectory == "measured":
        kspha = z["kspha_measured"]
        data = fieldcam.demodulate(z["kdata"], z["k0_adc"], kspha[0]).astype(np.complex64)
        weight, terms, golden, other = z["weight_measured"], "0111", gm, gn
    else:
        kspha = z["kspha_nominal"]
        data = z["kdata"]
        weight, terms, golden, other = z["weight_nominal"], "010", gn, gm
    return dict(grid=grid, kspha=kspha, times=z["times"], csm=csm, fieldmap=z["fieldmap"], mask=mask,
                data=data, weight=weight, recon_terms=terms, golden=golden, golden_other=other)


def rendered(img, n):
    """plt_image's gray levels (src/plot/plt_image.jl:36-48) of |img| after rotl90:
    minimum to the 99.9th percentile, clipped to [0, 1]."""
    from hoe import fieldcam
    a = np.abs(fieldcam.rotl90(np.asarray(img).reshape(n, n, order="F")))
    lo, hi = a.min(), np.quantile(a, 0.999)
    return np.clip((a - lo) / (hi - lo), 0.0, 1.0)


def corr(a, b):
    """Pearson correlation of two images."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    a, b = a - a.mean(), b - b.mean()
    return float(a @ b / np.sqrt((a @ a) * (b @ b)))
