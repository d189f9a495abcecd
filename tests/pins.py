"""Loads the reference's regression cases committed under tests/golden/ (inputs and
expected target pixels; written by tests/golden/make_golden.py)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_png(name):
    from PIL import Image
    return np.array(Image.open(os.path.join(GOLDEN, "images", name + ".png")).convert("RGB"))


def load_png_raw(name):
    """uint8 (H, W, C) with the file's own channels: 1 gray, 2 gray + alpha, 3 RGB, 4 RGBA."""
    from PIL import Image
    a = np.array(Image.open(os.path.join(GOLDEN, "images", name + ".png")))
    return a[..., None] if a.ndim == 2 else a


def gimp_flatten(px):
    """GIMP's flattening of a layer with alpha onto a white background, as the regression
    harness exports its results (Test/testResynth.py:163-166): each colour channel c with
    alpha a becomes 255 + ((c - 255) * a + 128 + (((c - 255) * a + 128) >> 8)) >> 8 (the
    8-bit blend with its rounding; >> on negatives is an arithmetic shift).  Pinned by the
    untouched pixels of the heal cases (tests/golden/make_heal_golden.py)."""
    if px.shape[2] not in (2, 4):
        return px
    c = px[..., :-1].astype(np.int64)
    a = px[..., -1:].astype(np.int64)
    t = (c - 255) * a + 0x80
    return ((((t >> 8) + t) >> 8) + 255).astype(np.uint8)


def heal_cases():
    with open(os.path.join(GOLDEN, "heal_pins.json")) as f:
        return {c["name"]: c for c in json.load(f)}


def heal_input(case):
    """(pixels, selection, expected region, region box, selection box) of a heal case."""
    px = load_png_raw(case["image"])
    h, w = px.shape[:2]
    x0, y0, x1, y1 = case["select"]
    sel = np.zeros((h, w), np.uint8)
    sel[y0:y1, x0:x1] = 0xFF
    expected = np.load(os.path.join(GOLDEN, "expect_%s.npy" % case["name"]))
    return px, sel, expected, case["region"], case["select"]


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
