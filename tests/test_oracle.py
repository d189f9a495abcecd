"""Pins the CPU oracle against sources of truth that are not this project's code:
GLib / glibc golden vectors, the reference's own regression images and known
answers (src/testSynth.c), closed forms and brute force."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle_py as O
import pins

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def vectors():
    with open(os.path.join(GOLDEN, "vectors.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- PRNG

def test_grand_ints_match_glib(vectors):
    for seed in (1198472, 1, 5, 4294967295):
        want = vectors["prng"]["grand_int_seed%d" % seed]
        assert O.grand_ints(seed, len(want)) == want


def test_grand_int_range_matches_glib(vectors):
    st = ctypes.create_string_buffer(624 * 4 + 16)
    O.lib().orc_grand_seed(st, 1198472)
    for b, e, want in vectors["prng"]["grand_int_range_mixed"]:
        assert O.lib().orc_grand_int_range(st, b, e) == want, (b, e)


def test_libc_rand_matches_glibc(vectors):
    assert O.libc_rands(1198472, 1500) == vectors["prng"]["libc_rand_seed1198472"]
    assert O.libc_rands(7, 400) == vectors["prng"]["libc_rand_seed7"]


def test_libc_int_range_formula():
    # lib/glibProxy.c:37-49 on the glibc stream: lo + rand() / (RAND_MAX/n + 1)
    raw = O.libc_rands(1198472, 50)
    dists = [1, 2, 3, 1000, 61440] * 10
    got = O.libc_ranges(1198472, dists)
    want = [r // (2147483647 // d + 1) for r, d in zip(raw, dists)]
    assert got == want
    assert O.libc_ranges(1198472, [0]) == [0]


# ---------------------------------------------------------------- sort

def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("wh", [(1, 1), (2, 1), (3, 3), (7, 5), (20, 13), (64, 64), (31, 47), (128, 96)])
def test_sorted_offsets_match_g_array_sort(vectors, wh):
    w, h = wh
    got = O.sorted_offsets(w, h)
    key = "offsets_%dx%d" % wh
    assert _sha(got) == vectors["sort"][key + "_sha256"]
    if key in vectors["sort"]:
        assert got.tolist() == vectors["sort"][key]
    assert tuple(got[0]) == (0, 0)


def test_sorted_offsets_closed_form(vectors):
    # The never-zero comparator under a merge sort yields (d^2 ascending, row-major index descending).
    for (w, h) in [(3, 3), (7, 5), (20, 13), (64, 64), (31, 47), (128, 96)]:
        pts = np.array([(x, y) for y in range(-h + 1, h) for x in range(-w + 1, w)], dtype=np.int32)
        idx = np.arange(len(pts))
        order = np.lexsort((-idx, pts[:, 0] ** 2 + pts[:, 1] ** 2))
        assert _sha(pts[order]) == vectors["sort"]["offsets_%dx%d_sha256" % (w, h)]


CMP = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)


def _merge_sort(points, key_cmp):
    a = np.ascontiguousarray(points, dtype=np.int32).copy()

    def cmp(pa, pb):
        x = ctypes.cast(pa, ctypes.POINTER(ctypes.c_int32 * 2))[0]
        y = ctypes.cast(pb, ctypes.POINTER(ctypes.c_int32 * 2))[0]
        return key_cmp(x, y)
    cb = CMP(cmp)
    O.lib().orc_merge_sort.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, CMP]
    O.lib().orc_merge_sort(a.ctypes.data, len(a), 8, cb)
    return a


def test_merge_sort_directional_ties(vectors):
    pts = np.array(vectors["sort"]["points_300"], dtype=np.int32)
    less_h = lambda a, b: -1 if a[0] * a[0] < b[0] * b[0] else 1
    more_h = lambda a, b: -1 if less_h(a, b) == 1 else 1
    assert _merge_sort(pts, less_h).tolist() == vectors["sort"]["points_300_lessHorizontal"]
    assert _merge_sort(pts, more_h).tolist() == vectors["sort"]["points_300_moreHorizontal"]
    less_v = lambda a, b: -1 if a[1] * a[1] < b[1] * b[1] else 1
    more_v = lambda a, b: -1 if less_v(a, b) == 1 else 1
    assert _merge_sort(pts, less_v).tolist() == vectors["sort"]["points_300_lessVertical"]
    assert _merge_sort(pts, more_v).tolist() == vectors["sort"]["points_300_moreVertical"]


def _ordered(mtype, pts):
    err, out, _ = O.order_targets(mtype, np.array(pts, dtype=np.int32))
    assert err == 0
    return np.asarray(out, dtype="<i4").reshape(-1, 2)


@pytest.mark.parametrize("mtype", range(9))
def test_target_ordering_matches_glib(vectors, mtype):
    """orderTargetPoints (lib/orderTarget.h:269-343) for every match-context type, on hole
    shapes that include ties, a target whose centre lies outside it, and targets whose
    centre pixel has no other pixel on its ray (ratio 0/0 = NaN, lib/brushfire.h:117-128):
    the oracle equals the whole ordering computed by GLib's g_array_sort and
    g_rand_int_range and the C library's atan2 (tests/golden/make_golden.py)."""
    import hashlib
    for name, e in vectors["order"].items():
        got = _ordered(mtype, e["points"])
        if str(mtype) in e["types"]:
            assert got.tolist() == e["types"][str(mtype)], name
        assert hashlib.sha256(got.tobytes()).hexdigest() == e["types"][str(mtype) + "_sha256"], name


def test_target_ordering_pins_catch_plausible_mistakes(vectors):
    """Each ordering family differs from its plausible confusions on the pinned shapes, so
    the vectors above would catch a swapped comparator or a dropped band shuffle."""
    e = vectors["order"]["ellipse"]
    t = {k: e["types"][k] for k in map(str, range(9))}
    assert t["2"] != t["5"] and t["3"] != t["6"] and t["4"] != t["7"] and t["3"] != t["4"]
    assert t["2"] == t["8"]          # the squeeze's result is the inward order (lib/orderTarget.h:209-262)
    assert t["0"] == t["1"]
    assert sorted(map(tuple, t["5"])) == sorted(map(tuple, e["points"]))


# ---------------------------------------------------------------- tables and schedule

def test_metric_table_closed_form():
    T, Tm = O.quantize(0.117, 0.5)
    # T[256+i] = trunc(ln((i/s)^2+1)/ln((256/s)^2+1) * 65535), s = float(0.117f*256)
    assert [int(T[256 + d]) for d in (1, 2, 10, 30, 255)] == [16, 67, 1608, 10576, 65417]
    assert [int(T[256 - d]) for d in (1, 2, 10, 30, 255)] == [16, 67, 1608, 10576, 65417]
    assert T[256] == 0 and T[0] == 65535
    assert np.all(np.diff(T[256:].astype(np.int64)) > 0)
    assert Tm[0] == 131072 and Tm[256] == 0 and Tm[256 + 3] == 18
    T0, Tm0 = O.quantize(0.117, 0.0)
    assert np.all(Tm0 == 0) and np.array_equal(T0, T)


def test_metric_table_against_independent_float_model():
    # Independent evaluation in numpy with the same float narrowing of the parameter.
    for sens in (0.117, 0.05, 0.3):
        T, _ = O.quantize(sens, 0.5)
        s = np.float64(np.float32(np.float32(sens) * np.float32(256.0)))
        i = np.arange(-256, 256, dtype=np.float64)
        v = np.log((i / s) * (i / s) + 1.0) / np.log((256.0 / s) * (256.0 / s) + 1.0) * 65535.0
        assert np.array_equal(T, v.astype(np.uint16))


def test_repetition_schedule():
    assert O.repetition(4096) == [4096, 4096, 3072, 2304, 1728, 1296]
    assert sum(O.repetition(4096)) == 16592
    assert O.repetition(1) == [1, 1, 0, 0, 0, 0]


# ---------------------------------------------------------------- scorer (brute force)

def _brute(corpus, ix, T, Tm, offs, pix, cands):
    """Full patch sums of every candidate, then the first strict minimum (sequential
    strict-less update with non-negative terms) and the first perfect match."""
    H, W, _ = corpus.shape
    sums = []
    for cx, cy in cands:
        s = 0
        for i, (ox, oy) in enumerate(offs):
            qx, qy = cx + ox, cy + oy
            if qx < 0 or qy < 0 or qx >= W or qy >= H or corpus[qy, qx, 0] != 255:
                s += 65535 * ix.img_match_bpp + int(Tm[0]) * ix.map_match_bpp
                continue
            if i:
                for j in range(1, ix.colorEndBip):
                    s += int(T[256 + int(pix[i][j]) - int(corpus[qy, qx, j])])
            for j in range(ix.map_start_bip, ix.map_end_bip):
                s += int(Tm[256 + int(pix[i][j]) - int(corpus[qy, qx, j])])
        sums.append(s)
    best, win, stop = None, -1, len(cands)
    for k, s in enumerate(sums):
        if best is None or s < best:
            best, win = s, k
            if s == 0:
                stop = k + 1
                break
    return win, best, stop


@pytest.mark.parametrize("maps", [False, True])
def test_best_of_matches_brute_force(maps):
    rng = np.random.RandomState(11 + maps)
    H, W = 23, 31
    if maps:
        ix = O.prepare_indices(3, 2, False, False, True)
    else:
        ix = O.format_indices(0)[1]
    depth = ix.total_bpp
    corpus = rng.randint(0, 256, size=(H, W, depth)).astype(np.uint8)
    corpus[..., 0] = np.where(rng.rand(H, W) < 0.85, 255, rng.randint(0, 255, size=(H, W)))
    T, Tm = O.quantize(0.117, 0.5)
    for trial in range(20):
        n = rng.randint(1, 40)
        offs = rng.randint(-4, 5, size=(n, 2)).astype(np.int32)
        offs[0] = 0
        pix = rng.randint(0, 256, size=(n, 8)).astype(np.uint8)
        cands = np.stack([rng.randint(-3, W + 3, 60), rng.randint(-3, H + 3, 60)], 1).astype(np.int32)
        if trial % 5 == 0:   # plant an exact copy so that a perfect match exists
            cx, cy = 10, 9
            for i, (ox, oy) in enumerate(offs):
                qx, qy = cx + ox, cy + oy
                if 0 <= qx < W and 0 <= qy < H:
                    corpus[qy, qx, 0] = 255
                    pix[i, :depth] = corpus[qy, qx]
            cands[30] = (cx, cy)
        got = O.best_of(corpus, ix, T, Tm, offs, pix, cands)
        assert got == _brute(corpus, ix, T, Tm, offs, pix, cands)


# ---------------------------------------------------------------- src/testSynth.c known answers

def _hex(img, w, bpp):
    return [" ".join("%02x" % b for b in row[: w * bpp]) for row in img]


def test_testsynth_known_answers():
    """src/testSynth.c:163-212, built with the stand-alone PRNG (lib/glibProxy.c)."""
    L = O.PRNG_LIBC
    P = O.default_params()
    image = np.array([[0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0],
                      [0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0],
                      [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 8, 0, 0]], np.uint8)
    mask = np.array([[0, 0, 0], [0, 0xFF, 0], [0, 0, 0]], np.uint8)
    r, out = O.image_synth(image, mask, 1, P, prng=L, width=3)
    assert r.err == 0
    assert _hex(out, 3, 4)[1] == "00 00 00 00 00 00 00 01 00 00 00 00"  # (1,1) -> 0,0,0 alpha 1 kept
    mask2 = np.array([[0, 0xFF, 0]], np.uint8)
    r, out = O.image_synth(np.array([[128, 128, 128, 0xFF, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0]], np.uint8), mask2, 1, P,
                           prng=L, width=3)
    assert (r.err, _hex(out, 3, 4)) == (0, ["80 80 80 ff 80 80 80 01 00 00 00 00"])
    rgb = np.array([[128, 128, 128, 1, 1, 1, 2, 2, 2, 5, 5], [64, 64, 64, 4, 4, 4, 3, 3, 3, 7, 7]], np.uint8)
    r, out = O.image_synth(rgb, np.array([[0, 0, 0], [0, 0xFF, 0]], np.uint8), 0, P, prng=L, width=3)
    assert (r.err, _hex(out, 3, 3)) == (0, ["80 80 80 01 01 01 02 02 02", "40 40 40 01 01 01 03 03 03"])
    r, out = O.image_synth(np.array([[128, 0xFF, 64, 1, 1, 0, 0, 0]], np.uint8), mask2, 3, P, prng=L, width=3)
    assert (r.err, _hex(out, 3, 2)) == (0, ["80 ff 80 01 01 00"])
    gray = np.array([[128, 64, 1, 0, 0]], np.uint8)
    r, out = O.image_synth(gray, mask2, 2, None, prng=L, width=3)
    assert (r.err, _hex(out, 3, 1)) == (0, ["80 01 01"])
    # error cases (:200-212)
    r, _ = O.image_synth(np.array([[128, 128, 128, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0]], np.uint8), mask2, 1, P,
                         prng=L, width=3)
    assert r.err == 6
    r, out = O.image_synth(gray, mask2, 2, O.default_params(patchSize=65), prng=L, width=3)
    assert r.err == 3 and out.tolist() == gray.tolist()
    r, _ = O.image_synth(gray, np.array([[0]], np.uint8), 2, O.default_params(patchSize=10), prng=L, width=3)
    assert r.err == 2
    r, _ = O.image_synth(gray, np.array([[0, 0, 0]], np.uint8), 2, O.default_params(patchSize=10), prng=L, width=3)
    assert r.err == 5
    r, _ = O.image_synth(gray, mask2, 7, P, prng=L, width=3)
    assert r.err == 1


def test_match_context_type_range_error():
    img = np.full((8, 8, 3), 100, np.uint8)
    mask = np.zeros((8, 8), np.uint8)
    mask[3:5, 3:5] = 255
    r, out = O.image_synth(img, mask, 0, O.default_params(matchContextType=9))
    assert r.err == 4
    for t in range(9):
        r, _ = O.image_synth(img, mask, 0, O.default_params(matchContextType=t))
        assert r.err == 0


# ---------------------------------------------------------------- reference regression images

FAST_PINS = ["resynthfull-zap-texture", "resynthfull-ufo-input", "resynthfull-brick", "resynth-ufo-input",
             "resynthtwoimages-ufo-input", "resynthtileable-ufo-input", "resynthfull-donkey_original"]


def _run_pin(name):
    c = pins.cases()[name]
    params, (colors, maps), tp, cp, rect, expected = pins.build(c)
    ix = O.prepare_indices(colors, maps, False, False, maps > 0)
    r, out = O.engine(params, ix, tp, cp, prng=O.PRNG_GRAND)
    x0, y0, x1, y1 = rect
    return r, out[y0:y1, x0:x1, 1:4], expected


@pytest.mark.parametrize("name", FAST_PINS)
def test_reference_regression_image(name):
    """Test/testResynth.py direct-engine cases: the serial GRand engine must reproduce
    Test/reference_out_images/<name>.ppm byte for byte."""
    r, got, expected = _run_pin(name)
    assert r.err == 0
    assert np.array_equal(got, expected)


@pytest.mark.slow
def test_reference_regression_image_mapstyle():
    r, got, expected = _run_pin("mapstyleangel-angel_target")
    assert r.err == 0
    assert np.array_equal(got, expected)


# ---------------------------------------------------------------- heal selection (plug-in level)

HEAL = sorted(pins.heal_cases())


def _heal_region(name, **kw):
    import heal_oracle as H
    c = pins.heal_cases()[name]
    px, sel, expected, (rx0, ry0, rx1, ry1), _ = pins.heal_input(c)
    r, out = H.heal(px, sel, c["radius"], c["direction"], c["order"], prng=O.PRNG_GRAND, **kw)
    return r, pins.gimp_flatten(out)[ry0:ry1, rx0:rx1], expected


@pytest.mark.parametrize("name", HEAL)
def test_heal_regression_image(name):
    """Test/testResynth.py:331-366: the heal-selection script's corpus (selection grown by
    GIMP's disc, less the selection, cropped by direction), the plug-in's adaptation and the
    serial GRand engine reproduce Test/reference_out_images/<name>.ppm byte for byte over
    the selection and an 8-pixel ring around it (after GIMP's flattening onto white)."""
    r, got, expected = _heal_region(name)
    assert r.err == 0
    assert np.array_equal(got, expected)


def test_heal_pins_catch_plausible_mistakes():
    """The heal pins reject a grow disc measured at the column centres (the all-around
    crop keeps the disc's corners) and inwards orders encoded without their offset."""
    import heal_oracle as H
    _, got, expected = _heal_region("healaroundrandom-ufo-input", inner_edge=False)
    assert not np.array_equal(got, expected)
    c = pins.heal_cases()["heal-ufo-input"]
    px, sel, expected, (rx0, ry0, rx1, ry1), _ = pins.heal_input(c)
    corpus, csel, ub = H.heal_corpus(px, sel, 50, 1, 1)
    assert ub == 3
    ix, tp, cp = H.plugin_maps(px, sel, corpus, csel)
    params = dict(tileH=0, tileV=0, matchContextType=1, mapWeight=0.0, sensitivity=0.117, patchSize=16,
                  maxProbeCount=500)
    r, out = O.engine(params, ix, tp, cp, prng=O.PRNG_GRAND)
    assert not np.array_equal(out[ry0:ry1, rx0:rx1, 1:4], expected)


def test_heal_grow_disc_properties():
    """The grow disc is symmetric, spans r rows at its centre column, and a single selected
    pixel grows into exactly the disc's pixels."""
    import heal_oracle as H
    for r in (1, 5, 50):
        c = H.grow_disc(r)
        assert c == c[::-1] and c[r] == r and max(c) == r
        sel = np.zeros((2 * r + 3, 2 * r + 3), np.uint8)
        sel[r + 1, r + 1] = 255
        g = H.grow(sel, r)
        assert int((g > 0).sum()) == sum(2 * v + 1 for v in c)


# ---------------------------------------------------------------- invariants

def _texture(h, w, seed, c=3):
    rng = np.random.RandomState(seed)
    y, x = np.mgrid[0:h, 0:w]
    img = np.zeros((h, w, c))
    for ch in range(c):
        for _ in range(3):
            th, per, ph = rng.uniform(0, np.pi), rng.uniform(6, 40), rng.uniform(0, 2 * np.pi)
            img[..., ch] += 40 * np.sin(2 * np.pi * (x * np.cos(th) + y * np.sin(th)) / per + ph)
    img += rng.randint(-16, 17, size=img.shape) + 128
    return np.clip(img, 0, 255).astype(np.uint8)


def test_engine_invariants_and_threaded_mode():
    img = _texture(48, 56, 3)
    mask = np.zeros((48, 56), np.uint8)
    mask[14:34, 18:40] = 255
    for threads in (1, 8):
        r, out = O.image_synth(img, mask, 0, O.default_params(), threads=threads, want_maps=True)
        assert r.err == 0
        sel = mask != 0
        assert np.array_equal(out[~sel], img[~sel])
        src = r.sourceOf[sel]
        assert (src >= 0).all()
        # every synthesized pixel carries the colour of its recorded source, outside the hole
        assert np.array_equal(out[sel], img[src[:, 1], src[:, 0]])
        assert (mask[src[:, 1], src[:, 0]] == 0).all()
        assert r.passes >= 2 and r.visits == sum(r.passSize)
        assert np.isfinite(r.meanError) and r.meanError > 0


def test_cancel_leaves_image_unchanged():
    img = _texture(32, 32, 5)
    mask = np.zeros((32, 32), np.uint8)
    mask[8:20, 8:20] = 255
    r, out = O.image_synth(img, mask, 0, O.default_params(), cancel=1)
    assert r.err == 0 and np.array_equal(out, img)
    # the cancel poll happens at the first target of every pass, before any synthesis
    assert r.visits == 0 and r.progress == [int(4095 / sum(O.repetition(144)) * 100)]


def test_progress_percent_can_exceed_100_on_small_targets():
    img = _texture(64, 64, 6)
    mask = np.zeros((64, 64), np.uint8)
    mask[16:48, 16:48] = 255     # 1024 targets: estimated 4148 pixels, +4095 per pass
    r, _ = O.image_synth(img, mask, 0, O.default_params())
    assert r.progress[0] == 98 and all(p > 100 for p in r.progress[1:])
    assert len(r.progress) == r.passes
