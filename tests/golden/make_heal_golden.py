"""Writes the heal-selection fixtures under tests/golden/ (run in the build container only):
the input images of the six heal cases of Test/testResynth.py:331-366 and, from their
expected outputs Test/reference_out_images/heal*-*.ppm (PPM or PGM, the test harness's
GIMP export of the healed image, flattened onto white when it has alpha), the pixels of
the selection's rectangle and of an 8-pixel ring around it.

The ring holds untouched pixels: the script asserts that GIMP's flattening, restated in
tests/pins.py:gimp_flatten, reproduces them from the input image (and the whole image
outside the selection), which pins that formula.

Usage:  python tests/golden/make_heal_golden.py   (needs /root/reference)
"""
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
REF = "/root/reference"
import pins  # noqa: E402

SELECT1 = (100, 90, 100, 50)   # Test/testResynth.py:81
RING = 8
# name, image, (samplingRadius, direction, order): Test/testResynth.py:331-366
HEAL_CASES = [
    ("heal", "ufo-input", (50, 1, 1)),
    ("healaroundrandom", "ufo-input", (50, 0, 0)),
    ("healincludedalpha", "ufo-input-w-alpha", (50, 1, 1)),
    ("healalpha", "apollo11_w_alpha", (50, 1, 1)),
    ("healgray", "wander", (50, 1, 1)),
    ("healalphagray", "ufo-input-w-alpha-gray", (50, 0, 1)),
]


def read_pnm(path):
    toks = []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            toks += line.split()
    kind, w, h = toks[0], int(toks[1]), int(toks[2])
    a = np.array(toks[4:], dtype=np.int32).astype(np.uint8)
    assert kind in ("P2", "P3")
    return a.reshape(h, w, 3 if kind == "P3" else 1)


def main():
    imgdir = os.path.join(HERE, "images")
    out = []
    for name, image, (radius, direction, order) in HEAL_CASES:
        shutil.copyfile(os.path.join(REF, "Test/in_images/%s.png" % image), os.path.join(imgdir, image + ".png"))
        px = pins.load_png_raw(image)
        ref = read_pnm(os.path.join(REF, "Test/reference_out_images/%s-%s.ppm" % (name, image)))
        h, w = px.shape[:2]
        x0, y0, sw, sh = SELECT1
        x1, y1 = min(w, x0 + sw), min(h, y0 + sh)
        outside = np.ones((h, w), bool)
        outside[y0:y1, x0:x1] = False
        assert np.array_equal(pins.gimp_flatten(px)[outside], ref[outside]), name
        rx0, ry0, rx1, ry1 = max(0, x0 - RING), max(0, y0 - RING), min(w, x1 + RING), min(h, y1 + RING)
        np.save(os.path.join(HERE, "expect_%s-%s.npy" % (name, image)), np.ascontiguousarray(ref[ry0:ry1, rx0:rx1]))
        out.append(dict(name="%s-%s" % (name, image), image=image, select=[x0, y0, x1, y1],
                        region=[rx0, ry0, rx1, ry1], radius=radius, direction=direction, order=order))
    with open(os.path.join(HERE, "heal_pins.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(out), "heal cases")


if __name__ == "__main__":
    main()
