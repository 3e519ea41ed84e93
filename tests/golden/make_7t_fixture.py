"""Writes tests/golden/7t_spiral.npz and tests/golden/7t_epi.npz from the reference's
7 T demo data (demo/7T_2D_{Spiral,EPI}_1p0_200_r4.mat), and tests/golden/7t_reference.npz
from its own reconstructions of them (demo/result/*_{Measured,Nominal}.png) and the
delays stored with the data, for the real-data GPU tests.

The arrays are the operator inputs prepared by hoe.fieldcam.prepare_demo (coil maps
stored on the masked voxels only: they are zero elsewhere), the raw k-space data with
its k0_adc, both field coefficient sets at the ADC times and as measured (with their
start times, sampling intervals and the delays stored with the data), their density
weights. These two files are large (about 16 MB each) and are generated, not committed.
7t_reference.npz is small and committed: per dataset the two rendered reconstructions
of the reference reduced to the 200 x 200 image grid (the plot area of each PNG below the
title, box-averaged onto 200 x 200 pixels, gray levels in [0, 1], stored as float16) and
the stored delays tauMeasured / tauNominal with the ADC dwell. Run from the repository
root:

    python tests/golden/make_7t_fixture.py [REFERENCE_ROOT]
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "highordermri.jl_amd"))
from hoe import fieldcam  # noqa: E402

DATASETS = {"spiral": "7T_2D_Spiral_1p0_200_r4", "epi": "7T_2D_EPI_1p0_200_r4"}


def rendered_image(png, n):
    """Plot area of a plt_image PNG (rows below the title band) box-averaged to n x n."""
    from PIL import Image
    im = Image.open(png).convert("L")
    a = np.asarray(im, dtype=np.float64)
    rows = np.where((a < 250).sum(axis=1) > 0.5 * a.shape[1])[0]
    box = (0, int(rows[0]), a.shape[1], int(rows[-1]) + 1)
    small = im.crop(box).resize((n, n), Image.BOX)
    return (np.asarray(small, dtype=np.float32) / 255.0)


def main(reference_root):
    demo = Path(reference_root) / "demo"
    ref = {}
    for name, stem in DATASETS.items():
        raw = fieldcam.load_demo_mat(demo / f"{stem}.mat")
        meas = fieldcam.prepare_demo(raw, "measured")
        nom = fieldcam.prepare_demo(raw, "nominal")
        g = meas["grid"]
        mask = meas["mask"]
        out = dict(
            matrix=np.array(g.matrixSize, dtype=np.int64),
            resolution=np.array(g.resolution, dtype=np.float32),
            mask=mask, fieldmap=meas["fieldmap"], csm_masked=meas["csm"][mask], times=meas["times"],
            kdata=np.asarray(raw["kdata"]).astype(np.complex64),
            k0_adc=np.asarray(raw["k0_adc"], dtype=np.float64).reshape(-1),
            kspha_measured=meas["kspha"], kspha_nominal=nom["kspha"],
            weight_measured=meas["weight"], weight_nominal=nom["weight"],
            tau_measured=float(raw["tauMeasured"]), tau_nominal=float(raw["tauNominal"]),
            dt_adc=float(raw["dt_adc"]),
            field_measured=np.asarray(raw["ksphaMeasured"], dtype=np.float32),
            field_nominal=np.asarray(raw["ksphaNominal"], dtype=np.float32),
            start_measured=float(raw["startMeasured"]), start_nominal=float(raw["startNominal"]),
            dt_measured=float(raw["dt_Measured"]), dt_nominal=float(raw["dt_Nominal"]),
        )
        path = Path(__file__).parent / f"7t_{name}.npz"
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({path.stat().st_size / 1e6:.1f} MB)")
        for traj in ("Measured", "Nominal"):
            ref[f"{name}_golden_{traj.lower()}"] = rendered_image(demo / "result" / f"{stem}_{traj}.png",
                                                                  g.nX).astype(np.float16)
        ref[f"{name}_tau_measured"] = float(raw["tauMeasured"])
        ref[f"{name}_tau_nominal"] = float(raw["tauNominal"])
        ref[f"{name}_dt_adc"] = float(raw["dt_adc"])
    path = Path(__file__).parent / "7t_reference.npz"
    np.savez_compressed(path, **ref)
    print(f"wrote {path} ({path.stat().st_size / 1e3:.0f} kB)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
