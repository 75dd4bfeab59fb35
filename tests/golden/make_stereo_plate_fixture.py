"""Writes tests/golden/stereo_plate.npz from data files the reference holds.

Inputs (read as data, paths inside the Traditional-DIC tree):
  case/stereo_DIC/plate_center_load/cam1/00_L.bmp, cam1/04_L.bmp,
  cam2/00_R.bmp, cam2/04_R.bmp         8-bit speckle images, 1440 x 1080
  ROI.bmp                               region of interest (read_mask: L > 0)
  tests/data/f0b/stereo_plate/fields/{reference_disparity,left_temporal,
  deformed_disparity}.npz               per-POI golden fields (97200 rows)

Output: the four images as uint8 (the reference loads them with
read_gray = float32 / max, workflows/subset_2d.py:21-27), the ROI as 0/1
uint8 (read_mask, workflows/subset_2d.py:30-31) and, per field,
the rows with valid == 1 (x, y, u, v, correlation) plus the row count.
The fields were produced with config/subset_2d.yaml: radius 41, quintic
B-spline, first-order IC-GN, ZNSSD, 50 iterations, threshold 1e-6, spacing 3.

Run:  python tests/golden/make_stereo_plate_fixture.py /root/reference
"""
import os
import sys

import numpy as np
from PIL import Image

FIELDS = {
    "reference_disparity": ("00_L", "00_R"),
    "left_temporal": ("00_L", "04_L"),
    "deformed_disparity": ("00_L", "04_R"),
}
IMAGES = {"00_L": "cam1/00_L.bmp", "04_L": "cam1/04_L.bmp", "00_R": "cam2/00_R.bmp", "04_R": "cam2/04_R.bmp"}


def main(ref_root: str):
    out = {}
    case = os.path.join(ref_root, "case", "stereo_DIC", "plate_center_load")
    for key, rel in IMAGES.items():
        im = Image.open(os.path.join(case, rel))
        u8 = np.asarray(im.convert("L"), dtype=np.uint8)
        f = np.asarray(im.convert("F"), dtype=np.float32)
        assert np.array_equal(f, u8.astype(np.float32)), "palette is not the identity grey ramp"
        out["img_" + key] = u8
    out["roi"] = (np.asarray(Image.open(os.path.join(case, "ROI.bmp")).convert("L")) > 0).astype(np.uint8)
    fdir = os.path.join(ref_root, "tests", "data", "f0b", "stereo_plate", "fields")
    for name, (a, b) in FIELDS.items():
        z = np.load(os.path.join(fdir, name + ".npz"), allow_pickle=False)
        valid = z["valid"].astype(bool)
        rows = np.stack([z["x"][valid], z["y"][valid], z["u"][valid], z["v"][valid], z["correlation"][valid]],
                        axis=1)
        out["field_" + name] = rows
        out["rows_" + name] = np.array([z["x"].size, int(valid.sum())], np.int64)
        out["pair_" + name] = np.array([a, b])
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stereo_plate.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
