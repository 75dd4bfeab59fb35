"""Writes tests/golden/subset_ring.npz from data files the reference holds.

Inputs (read as data, paths inside the Traditional-DIC tree):
  case/mono_DIC/ring/001.bmp, 002.bmp   8-bit speckle images, 1280 x 1280 (reference, deformed)
  case/mono_DIC/ring/003.bmp            region of interest (read_mask: L > 0)
  tests/data/f0/subset_ring/summary.json  the golden run's displacement row count and the
                                          statistics of its stats.json (workflows/subset_2d.py:94-113)

Output: the two images as uint8 (loaded with read_gray = float32 / max,
workflows/subset_2d.py:21-27), the ROI as 0/1 uint8, and the summary JSON
text.  The golden run used config/subset_2d.yaml (radius 41, spacing 3,
first-order IC-GN, ZNSSD, 50 iterations, threshold 1e-6, SIFT seeding).

Run:  python tests/golden/make_ring_fixture.py /root/reference
"""
import json
import os
import sys

import numpy as np
from PIL import Image


def main(ref_root: str):
    case = os.path.join(ref_root, "case", "mono_DIC", "ring")
    out = {}
    for key, name in (("reference", "001.bmp"), ("deformed", "002.bmp")):
        im = Image.open(os.path.join(case, name))
        u8 = np.asarray(im.convert("L"), dtype=np.uint8)
        f = np.asarray(im.convert("F"), dtype=np.float32)
        assert np.array_equal(f, u8.astype(np.float32)), "palette is not the identity grey ramp"
        out["img_" + key] = u8
    out["roi"] = (np.asarray(Image.open(os.path.join(case, "003.bmp")).convert("L")) > 0).astype(np.uint8)
    with open(os.path.join(ref_root, "tests", "data", "f0", "subset_ring", "summary.json")) as f:
        summary = json.load(f)
    out["summary_json"] = np.array(json.dumps(summary, sort_keys=True))
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "subset_ring.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
