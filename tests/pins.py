"""Loads the reference's regression cases committed under tests/golden/ (inputs and
expected target pixels; written by tests/golden/make_golden.py)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_png(name):
    from PIL import Image
    return np.array(Image.open(os.path.join(GOLDEN, "images", name + ".png")).convert("RGB"))


def cases():
    with open(os.path.join(GOLDEN, "pins.json")) as f:
        return {c["name"]: c for c in json.load(f)}


def build(case):
    """Returns (params dict, (colors, maps), target pixmap, corpus pixmap, rect, expected).
    Pixmaps are uint8 (H, W, depth) in the engine's internal layout [mask, colours, maps]."""
    img = load_png(case["image"])
    h, w, _ = img.shape
    x0, y0, x1, y1 = case["rect"]
    mask = np.zeros((h, w), np.uint8)
    mask[y0:y1, x0:x1] = 0xFF
    if case["corpus"] == "inverse":
        cimg, cmask = img, (~mask).astype(np.uint8)
    elif case["corpus"] == "same":
        cimg, cmask = img, mask
    else:
        cimg = load_png(case["corpus"])
        cmask = np.full(cimg.shape[:2], 0xFF, np.uint8)
    if case["maps"]:
        tp = np.concatenate([mask[..., None], img, img], axis=2)
        cp = np.concatenate([cmask[..., None], cimg, cimg], axis=2)
        fmt = (3, 3)
    else:
        tp = np.concatenate([mask[..., None], img], axis=2)
        cp = np.concatenate([cmask[..., None], cimg], axis=2)
        fmt = (3, 0)
    expected = np.load(os.path.join(GOLDEN, "expect_%s.npy" % case["name"]))
    return case["params"], fmt, np.ascontiguousarray(tp), np.ascontiguousarray(cp), (x0, y0, x1, y1), expected
