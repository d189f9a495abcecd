"""Writes the golden fixtures under tests/golden/ (run in the build container only).

Sources of truth, none of them this project's code:
  * GLib 2.72.4 (system libglib-2.0.so.0) via ctypes: g_rand_new_with_seed /
    g_rand_int / g_rand_int_range and g_array_sort with the engine's comparators
    (lib/engineTypes.h:51-141 restated as Python callbacks), and the whole target
    ordering of lib/orderTarget.h:269-343 (types 0-8) driven by those calls and by the
    C library's atan2 (lib/brushfire.h:63-148).
  * glibc via ctypes: srand / rand (the PRNG of the stand-alone library build,
    lib/glibProxy.c:30-49).
  * The reference's own regression data: input images Test/in_images/*.png and the
    expected outputs Test/reference_out_images/*.ppm of the direct-engine cases of
    Test/testResynth.py:216-328 and the map-style case :262-267.  Only the pixels the
    engine may change (the target region) are kept, as uint8 arrays.

Usage:  python tests/golden/make_golden.py   (needs /root/reference)
"""
import ctypes
import hashlib
import json
import math
import os
import shutil
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def glib():
    g = ctypes.CDLL("libglib-2.0.so.0")
    g.g_rand_new_with_seed.restype = ctypes.c_void_p
    g.g_rand_new_with_seed.argtypes = [ctypes.c_uint32]
    g.g_rand_int.restype = ctypes.c_uint32
    g.g_rand_int.argtypes = [ctypes.c_void_p]
    g.g_rand_int_range.restype = ctypes.c_int32
    g.g_rand_int_range.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
    g.g_rand_free.argtypes = [ctypes.c_void_p]
    g.g_array_sized_new.restype = ctypes.c_void_p
    g.g_array_sized_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint, ctypes.c_uint]
    g.g_array_append_vals.restype = ctypes.c_void_p
    g.g_array_append_vals.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint]
    g.g_array_free.argtypes = [ctypes.c_void_p, ctypes.c_int]
    return g


CMP = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)


class GArray(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("len", ctypes.c_uint)]


def g_array_sort(g, arr, elem_words, key_fn):
    """Sort an int32 array of shape (n, elem_words) with g_array_sort and a comparator
    that returns -1 if key(a) < key(b) else 1 (never 0), as the engine's do."""
    n = arr.shape[0]
    ga = g.g_array_sized_new(0, 1, 4 * elem_words, n)
    a = np.ascontiguousarray(arr, dtype=np.int32)
    g.g_array_append_vals(ga, a.ctypes.data, n)

    def cmp(pa, pb):
        x = ctypes.cast(pa, ctypes.POINTER(ctypes.c_int32 * elem_words))[0]
        y = ctypes.cast(pb, ctypes.POINTER(ctypes.c_int32 * elem_words))[0]
        return key_fn(x, y)

    cb = CMP(cmp)
    g.g_array_sort.argtypes = [ctypes.c_void_p, CMP]
    g.g_array_sort(ga, cb)
    st = ctypes.cast(ga, ctypes.POINTER(GArray))[0]
    out = np.ctypeslib.as_array(ctypes.cast(st.data, ctypes.POINTER(ctypes.c_int32)), shape=(n * elem_words,))
    out = out.reshape(n, elem_words).copy()
    g.g_array_free(ga, 1)
    return out


def less_cart(a, b):
    return -1 if a[1] * a[1] + a[0] * a[0] < b[1] * b[1] + b[0] * b[0] else 1


def less_h(a, b):
    return -1 if a[0] * a[0] < b[0] * b[0] else 1


def more_h(a, b):
    return -1 if less_h(a, b) == 1 else 1


def less_v(a, b):
    return -1 if a[1] * a[1] < b[1] * b[1] else 1


def more_v(a, b):
    return -1 if less_v(a, b) == 1 else 1


def _prop(word):
    return struct.unpack("<f", struct.pack("<i", int(word)))[0]


def less_inward(a, b):
    # lib/engineTypes.h:86-94: to_sort_result(a->proportionToCenter < b->proportionToCenter);
    # a NaN ratio compares false
    return -1 if _prop(a[2]) < _prop(b[2]) else 1


def more_inward(a, b):
    return -1 if less_inward(a, b) == 1 else 1


def libm():
    m = ctypes.CDLL("libm.so.6")
    m.atan2.restype = ctypes.c_double
    m.atan2.argtypes = [ctypes.c_double, ctypes.c_double]
    return m


def order_points(g, m, match_type, pts):
    """orderTargetPoints (lib/orderTarget.h:269-343) of a row-major (n, 2) int32 point list
    with a fresh GRand of the engine's seed (lib/engine.c:643), using GLib's own
    g_array_sort and g_rand_int_range and the C library's atan2; returns the ordered list."""
    pts = np.array(pts, dtype=np.int64)
    n = len(pts)
    r = g.g_rand_new_with_seed(1198472)
    if match_type in (0, 1):
        # lib/orderTarget.h:36-47
        for i in range(n):
            j = g.g_rand_int_range(r, 0, n)
            pts[[i, j]] = pts[[j, i]]
        g.g_rand_free(r)
        return pts.astype(np.int32)
    # get_center (lib/engineTypes.h:187-226): bounds start at {G_MAXINT, G_MAXINT, 0, 0}
    ulx, uly = min(2 ** 31 - 1, pts[:, 0].min()), min(2 ** 31 - 1, pts[:, 1].min())
    lrx, lry = max(0, pts[:, 0].max()), max(0, pts[:, 1].max())
    cx, cy = (lrx - ulx) // 2 + ulx, (lry - uly) // 2 + uly
    off = pts - np.array([cx, cy])
    if match_type in (2, 5, 8):
        # lib/brushfire.h:63-148: grad = (guint)(atan2((gfloat)y, (gfloat)x) * 200 / G_PI + 200)
        def grad(x, y):
            return int(m.atan2(float(np.float32(y)), float(np.float32(x))) * 200 / math.pi + 200)
        maxc = [0] * 401
        for x, y in off:
            c = int(x * x + y * y) & 0xFFFFFFFF
            maxc[grad(x, y)] = max(maxc[grad(x, y)], c)
        elems = np.zeros((n, 3), np.int32)
        elems[:, :2] = off
        with np.errstate(invalid="ignore", divide="ignore"):
            for i, (x, y) in enumerate(off):
                prop = np.float32(np.float32(int(y * y + x * x)) / np.float32(maxc[grad(x, y)]))
                elems[i, 2] = np.array([prop], np.float32).view(np.int32)[0]
        # type 8 (lib/orderTarget.h:209-262) sorts with moreInward and then overwrites its
        # interleaved copy with that sorted array
        srt = g_array_sort(g, elems, 3, less_inward if match_type == 5 else more_inward)[:, :2]
    else:
        cmp = {3: more_h, 4: more_v, 6: less_h, 7: less_v}[match_type]
        srt = g_array_sort(g, off.astype(np.int32), 2, cmp)
    out = srt.astype(np.int64) + np.array([cx, cy])
    # randomizeBandsTargetPoints (lib/orderTarget.h:58-79), halfBand = (gint)(len * 0.1)
    last = n - 1
    half = int(n * 0.1)
    for i in range(n):
        start, end = max(i - half, 0), min(i + half, last)
        j = start + g.g_rand_int_range(r, 0, end - start)
        out[[i, j]] = out[[j, i]]
    g.g_rand_free(r)
    return out.astype(np.int32)


def order_shapes():
    """Target point sets (row-major, as prepareTargetPoints lists them) for the ordering
    vectors: convex and concave holes, ties, and holes whose centre pixel has no other
    target pixel on its ray (ratio 0/0 = NaN, lib/brushfire.h:117-128)."""
    shapes = {}
    yy, xx = np.mgrid[0:60, 0:80]
    shapes["ellipse"] = (((xx - 41) / 23.0) ** 2 + ((yy - 29) / 13.0) ** 2 <= 1.0)
    rect = np.zeros((60, 80), bool)
    rect[9:26, 5:28] = True
    shapes["rect"] = rect
    line = np.zeros((60, 80), bool)
    line[3:41, 7] = True
    shapes["vline"] = line
    # the bounding-box centre is (29, 30); nothing lies to its right on row 30
    disc = ((xx - 30) ** 2 + (yy - 30) ** 2 <= 144)
    disc[30, 30:] = False
    shapes["disc_gap"] = disc
    big = np.zeros((120, 160), bool)
    yy3, xx3 = np.mgrid[0:120, 0:160]
    big |= ((xx3 - 70) ** 2 + (yy3 - 60) ** 2 <= 44 ** 2)
    big[60, 70:] = False          # centre (69, 60) after the cut, with an empty ray to its right
    shapes["disc_gap_big"] = big
    cross = np.zeros((60, 80), bool)
    cross[5:56, 40] = True
    cross[30, 12:41] = True
    cross[29, 40:69] = True       # bounding box x 12..68, centre (40, 30): ray 200 is empty
    shapes["cross"] = cross
    one = np.zeros((60, 80), bool)
    one[17, 33] = True
    shapes["pixel"] = one
    two = np.zeros((60, 80), bool)
    two[5:15, 5:15] = True
    two[40:52, 60:75] = True
    shapes["two_blobs"] = two
    hline = np.zeros((60, 80), bool)
    hline[22, 10:70] = True
    shapes["hline"] = hline
    rng = np.random.RandomState(11)
    blob = np.zeros((120, 160), bool)
    for _ in range(6):
        cx0, cy0, a, b = rng.uniform(20, 140), rng.uniform(20, 100), rng.uniform(6, 30), rng.uniform(6, 30)
        yy2, xx2 = np.mgrid[0:120, 0:160]
        blob |= ((xx2 - cx0) / a) ** 2 + ((yy2 - cy0) / b) ** 2 <= 1.0
    shapes["blobs"] = blob
    out = {}
    for name, msk in shapes.items():
        ys, xs = np.nonzero(msk)
        out[name] = np.stack([xs, ys], axis=1).astype(np.int32)   # np.nonzero is row-major
    return out


def order_vectors():
    g = glib()
    m = libm()
    out = {}
    for name, pts in order_shapes().items():
        entry = {"points": pts.tolist(), "types": {}}
        for t in range(9):
            res = order_points(g, m, t, pts)
            if len(pts) <= 1000:
                entry["types"][str(t)] = res.tolist()
            entry["types"][str(t) + "_sha256"] = hashlib.sha256(res.astype("<i4").tobytes()).hexdigest()
        out[name] = entry
    return out


def offsets_rowmajor(w, h):
    return np.array([(x, y) for y in range(-h + 1, h) for x in range(-w + 1, w)], dtype=np.int32).reshape(-1, 2)


def prng_vectors():
    g = glib()
    out = {}
    r = g.g_rand_new_with_seed(1198472)
    out["grand_int_seed1198472"] = [g.g_rand_int(r) for _ in range(1500)]
    g.g_rand_free(r)
    ranges = [(0, 1), (0, 2), (0, 3), (0, 7), (0, 100), (0, 4096), (0, 61440), (0, 65536), (0, 1 << 20),
              (0, 3686400), (0, (1 << 31) - 1), (-5, 9), (-1500000000, 1500000000), (-2147483648, 2147483647),
              (0, 5000), (0, 97), (0, 1 << 24), (0, 12345678), (-1073741824, 1073741825)]
    seq = []
    r = g.g_rand_new_with_seed(1198472)
    for i in range(1200):
        b, e = ranges[i % len(ranges)]
        seq.append((b, e, g.g_rand_int_range(r, b, e)))
    g.g_rand_free(r)
    out["grand_int_range_mixed"] = seq
    for seed in (1, 5, 4294967295):
        r = g.g_rand_new_with_seed(seed)
        out["grand_int_seed%d" % seed] = [g.g_rand_int(r) for _ in range(700)]
        g.g_rand_free(r)
    c = ctypes.CDLL("libc.so.6")
    c.srand.argtypes = [ctypes.c_uint]
    c.srand(1198472)
    out["libc_rand_seed1198472"] = [c.rand() for _ in range(1500)]
    c.srand(7)
    out["libc_rand_seed7"] = [c.rand() for _ in range(400)]
    return out


def sort_vectors():
    g = glib()
    out = {}
    for (w, h) in [(1, 1), (2, 1), (3, 3), (7, 5), (20, 13), (64, 64), (31, 47), (128, 96)]:
        s = g_array_sort(g, offsets_rowmajor(w, h), 2, less_cart)
        key = "offsets_%dx%d" % (w, h)
        out[key + "_sha256"] = hashlib.sha256(s.astype("<i4").tobytes()).hexdigest()
        if s.shape[0] <= 400:
            out[key] = s.tolist()
    # directional comparators on a point set with many ties (lib/engineTypes.h:113-141)
    rng = np.random.RandomState(3)
    pts = rng.randint(-6, 7, size=(300, 2)).astype(np.int32)
    out["points_300"] = pts.tolist()
    out["points_300_lessHorizontal"] = g_array_sort(g, pts, 2, less_h).tolist()
    out["points_300_moreHorizontal"] = g_array_sort(g, pts, 2, more_h).tolist()
    out["points_300_lessVertical"] = g_array_sort(g, pts, 2, less_v).tolist()
    out["points_300_moreVertical"] = g_array_sort(g, pts, 2, more_v).tolist()
    return out


def read_ppm(path):
    toks = []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                continue
            toks += line.split()
    assert toks[0] == "P3"
    w, h = int(toks[1]), int(toks[2])
    return np.array(toks[4:], dtype=np.int32).astype(np.uint8).reshape(h, w, 3)


# Direct-engine regression cases of Test/testResynth.py (plug_in_resynthesizer arguments:
# h_tile, v_tile, use_border, corpus, in_map, out_map, map_weight, autism, neighbours, trys)
SELECT1 = (100, 90, 100, 50)   # Test/testResynth.py:81
SELECT2 = (90, 175, 135, 100)  # Test/testResynth.py:82
CASES = [
    # name, image, selection, corpus kind, params (tileH, tileV, type, mapWeight, sens, patch, probes)
    ("resynthfull-zap-texture", "zap-texture", SELECT1, "inverse", (0, 0, 1, 0.0, 0.117, 16, 500)),
    ("resynthfull-ufo-input", "ufo-input", SELECT1, "inverse", (0, 0, 1, 0.0, 0.117, 16, 500)),
    ("resynthfull-brick", "brick", SELECT1, "inverse", (0, 0, 1, 0.0, 0.117, 16, 500)),
    ("resynthfull-donkey_original", "donkey_original", SELECT2, "inverse", (0, 0, 1, 0.0, 0.117, 16, 500)),
    ("resynth-ufo-input", "ufo-input", SELECT1, "same", (0, 0, 1, 0.0, 0.117, 16, 500)),
    ("resynthtwoimages-ufo-input", "ufo-input", SELECT1, "grass-input-alpha", (0, 0, 1, 0.0, 0.117, 16, 500)),
    ("resynthtileable-ufo-input", "ufo-input", SELECT1, "grass-input-alpha", (1, 1, 0, 0.0, 0.117, 16, 500)),
    ("mapstyleangel-angel_target", "angel_target", None, "angel_texture",
     (1, 1, 1, math.acos((10.0 / 100) * 2 - 1) / (2 * 3.14), 0.117, 9, 200)),
]


def pin_cases():
    from PIL import Image
    imgdir = os.path.join(HERE, "images")
    os.makedirs(imgdir, exist_ok=True)
    index = []
    for name, image, sel, corpus, params in CASES:
        src = os.path.join(REF, "Test/in_images/%s.png" % image)
        shutil.copyfile(src, os.path.join(imgdir, image + ".png"))
        if corpus not in ("inverse", "same"):
            shutil.copyfile(os.path.join(REF, "Test/in_images/%s.png" % corpus), os.path.join(imgdir, corpus + ".png"))
        img = np.array(Image.open(src).convert("RGB"))
        ref = read_ppm(os.path.join(REF, "Test/reference_out_images/%s.ppm" % name))
        h, w, _ = img.shape
        if sel is None:
            y0, y1, x0, x1 = 0, h, 0, w
        else:
            x0, y0, sw, sh = sel
            x1, y1 = min(w, x0 + sw), min(h, y0 + sh)
        outside = np.ones((h, w), bool)
        outside[y0:y1, x0:x1] = False
        assert np.array_equal(img[outside], ref[outside]), name
        np.save(os.path.join(HERE, "expect_%s.npy" % name), np.ascontiguousarray(ref[y0:y1, x0:x1]))
        index.append(dict(name=name, image=image, rect=[x0, y0, x1, y1], corpus=corpus,
                          params=dict(tileH=params[0], tileV=params[1], matchContextType=params[2],
                                      mapWeight=params[3], sensitivity=params[4], patchSize=params[5],
                                      maxProbeCount=params[6]),
                          maps=name.startswith("mapstyle")))
    return index


def main():
    data = {"prng": prng_vectors(), "sort": sort_vectors(), "order": order_vectors()}
    with open(os.path.join(HERE, "vectors.json"), "w") as f:
        json.dump(data, f)
    with open(os.path.join(HERE, "pins.json"), "w") as f:
        json.dump(pin_cases(), f, indent=1)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
