"""GPU parity of libresynth_mi355x.so against the CPU oracle and the reference's own
regression data.  Integer work: every comparison is bit-exact, except the batched mode,
which is judged by mean best-match error (within 2% of the serial engine's)."""
import numpy as np
import pytest

import oracle_py as O
import pins
import resynthesizer_amd as R
from resynthesizer_amd import workloads as W

pytestmark = pytest.mark.gpu


def _oparams(p):
    return dict(tileH=p.isMakeSeamlesslyTileableHorizontally, tileV=p.isMakeSeamlesslyTileableVertically,
                matchContextType=p.matchContextType, mapWeight=p.mapWeight, sensitivity=p.sensitivityToOutliers,
                patchSize=p.patchSize, maxProbeCount=p.maxProbeCount)


def test_device_is_gfx950():
    rc, arch, cus = R.device_info()
    assert rc == 0 and arch.startswith("gfx950") and cus >= 200


# ---------------------------------------------------------------- sorted offsets

def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()


@pytest.mark.parametrize("wh", [(1, 1), (2, 1), (3, 3), (7, 5), (20, 13), (64, 64), (31, 47), (128, 96)])
def test_device_sorted_offsets_match_g_array_sort(wh):
    """The device-built offset table (quadrant sort + expansion) equals GLib's g_array_sort
    order of the reference (lib/engine.c:465-497), pinned by the sha256 golden vectors."""
    import json
    import os
    v = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "vectors.json")))["sort"]
    got = R.sorted_offsets(*wh)
    assert _sha(got) == v["offsets_%dx%d_sha256" % wh]


@pytest.mark.parametrize("wh", [(1, 300), (300, 1), (257, 190), (190, 257), (1024, 768), (1023, 1001)])
def test_device_sorted_offsets_closed_form(wh):
    """Larger and degenerate offset ranges against the closed form of that order: x^2 + y^2
    ascending, ties in descending row-major index."""
    w, h = wh
    ys, xs = np.mgrid[-h + 1:h, -w + 1:w]
    pts = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.int64)
    order = np.lexsort((-np.arange(len(pts)), pts[:, 0] ** 2 + pts[:, 1] ** 2))
    assert np.array_equal(R.sorted_offsets(w, h), pts[order])


@pytest.mark.parametrize("whd", [(1, 1, 0), (7, 5, 0), (7, 5, 1), (7, 5, 2), (7, 5, 8), (33, 17, 100), (33, 17, 1300),
                                 (33, 17, 1279), (5, 300, 100), (300, 5, 2500), (257, 190, 4097), (1023, 1001, 40000)])
def test_device_sorted_offsets_prefix_within_a_distance(whd):
    """The table cut at x^2 + y^2 <= maxD2, as an engine run builds it when no scan can reach
    farther, is exactly the whole table's prefix of offsets within that distance (cuts
    narrower than the range in x or y, between two distances, and at or past the farthest
    offset, which is the whole table)."""
    w, h, d2 = whd
    full = R.sorted_offsets(w, h)
    dist = full[:, 0].astype(np.int64) ** 2 + full[:, 1].astype(np.int64) ** 2
    n = int((dist <= d2).sum())
    assert np.all(dist[:n] <= d2) and np.all(dist[n:] > d2)
    got = R.sorted_offsets(w, h, max_d2=d2)
    assert got.shape == (n, 2) and np.array_equal(got, full[:n])


# ---------------------------------------------------------------- frozen-state scorer

@pytest.mark.parametrize("layout", ["rgb", "gray", "rgb_maps3", "gray_map1", "rgba"])
def test_frozen_scorer_bit_exact(layout):
    """computeBestFit over ordered candidate lists (lib/synthesize.h:266-400): winner,
    best sum and the stop position equal the oracle's on every problem."""
    rng = np.random.RandomState(hash(layout) % 1000)
    ix = {"rgb": O.prepare_indices(3, 0, 0, 0, 0), "gray": O.prepare_indices(1, 0, 0, 0, 0),
          "rgb_maps3": O.prepare_indices(3, 3, 0, 0, 1), "gray_map1": O.prepare_indices(1, 1, 0, 0, 1),
          "rgba": O.prepare_indices(3, 0, 1, 1, 0)}[layout]
    rix = R.prepare_indices(ix.img_match_bpp, ix.map_match_bpp, ix.isAlphaTarget, ix.isAlphaSource,
                            ix.map_match_bpp > 0)
    H, W_ = 37, 53
    depth = ix.total_bpp
    corpus = rng.randint(0, 256, size=(H, W_, depth)).astype(np.uint8)
    corpus[..., 0] = np.where(rng.rand(H, W_) < 0.8, 255, rng.randint(0, 255, size=(H, W_)))
    T, Tm = O.quantize(0.117, 0.5 if ix.map_match_bpp else 0.0)
    P, M = 96, 240
    nbc = rng.randint(1, 65, P).astype(np.uint32)
    nbo = rng.randint(-5, 6, size=(P, 64, 2)).astype(np.int32)
    nbo[:, 0] = 0
    nbp = rng.randint(0, 256, size=(P, 64, 8)).astype(np.uint8)
    cc = rng.randint(0, M + 1, P).astype(np.uint32)
    cands = np.stack([rng.randint(-4, W_ + 4, (P, M)), rng.randint(-4, H + 4, (P, M))], 2).astype(np.int32)
    for p in range(0, P, 4):   # plant exact copies so perfect matches and early stops occur
        cx, cy = rng.randint(5, W_ - 5), rng.randint(5, H - 5)
        for i in range(nbc[p]):
            qx, qy = cx + nbo[p, i, 0], cy + nbo[p, i, 1]
            if 0 <= qx < W_ and 0 <= qy < H:
                corpus[qy, qx, 0] = 255
                nbp[p, i, :depth] = corpus[qy, qx]
        if cc[p] > 0:
            cands[p, rng.randint(0, cc[p])] = (cx, cy)
    win, best, stop = R.score_frozen(corpus, rix, T, Tm, nbc, nbo, nbp, cc, cands)
    for p in range(P):
        w, b, s = O.best_of(corpus, ix, T, Tm, nbo[p, : nbc[p]], nbp[p, : nbc[p]], cands[p, : cc[p]])
        assert (win[p], best[p], stop[p]) == (w, b, s), p


# ---------------------------------------------------------------- serial mode: exact

def _hex(img, w, bpp):
    return [" ".join("%02x" % b for b in row[: w * bpp]) for row in img]


def test_serial_testsynth_known_answers():
    """src/testSynth.c:163-212 through imageSynth on the GPU (stand-alone PRNG)."""
    S, L = R.MODE_SERIAL, R.PRNG_LIBC
    P = R.default_params()
    image = np.array([[0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0],
                      [0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0],
                      [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 8, 0, 0]], np.uint8)
    mask = np.array([[0, 0, 0], [0, 0xFF, 0], [0, 0, 0]], np.uint8)
    r, out = R.image_synth(image, mask, R.T_RGBA, P, mode=S, prng=L, width=3)
    assert r.err == 0 and _hex(out, 3, 4)[1] == "00 00 00 00 00 00 00 01 00 00 00 00"
    mask2 = np.array([[0, 0xFF, 0]], np.uint8)
    r, out = R.image_synth(np.array([[128, 128, 128, 0xFF, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0]], np.uint8), mask2,
                           R.T_RGBA, P, mode=S, prng=L, width=3)
    assert (r.err, _hex(out, 3, 4)) == (0, ["80 80 80 ff 80 80 80 01 00 00 00 00"])
    rgb = np.array([[128, 128, 128, 1, 1, 1, 2, 2, 2, 5, 5], [64, 64, 64, 4, 4, 4, 3, 3, 3, 7, 7]], np.uint8)
    r, out = R.image_synth(rgb, np.array([[0, 0, 0], [0, 0xFF, 0]], np.uint8), R.T_RGB, P, mode=S, prng=L, width=3)
    assert (r.err, _hex(out, 3, 3)) == (0, ["80 80 80 01 01 01 02 02 02", "40 40 40 01 01 01 03 03 03"])
    r, out = R.image_synth(np.array([[128, 0xFF, 64, 1, 1, 0, 0, 0]], np.uint8), mask2, R.T_GrayA, P, mode=S,
                           prng=L, width=3)
    assert (r.err, _hex(out, 3, 2)) == (0, ["80 ff 80 01 01 00"])
    gray = np.array([[128, 64, 1, 0, 0]], np.uint8)
    r, out = R.image_synth(gray, mask2, R.T_Gray, None, mode=S, prng=L, width=3)
    assert (r.err, _hex(out, 3, 1)) == (0, ["80 01 01"])
    r, _ = R.image_synth(np.array([[128, 128, 128, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0]], np.uint8), mask2, R.T_RGBA, P,
                         mode=S, prng=L, width=3)
    assert r.err == R.IMAGE_SYNTH_ERROR_EMPTY_CORPUS
    r, out = R.image_synth(gray, mask2, R.T_Gray, R.default_params(patchSize=65), mode=S, width=3)
    assert r.err == R.IMAGE_SYNTH_ERROR_PATCH_SIZE_EXCEEDED and out.tolist() == gray.tolist()
    r, _ = R.image_synth(gray, np.array([[0]], np.uint8), R.T_Gray, R.default_params(patchSize=10), mode=S, width=3)
    assert r.err == R.IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH
    r, _ = R.image_synth(gray, np.array([[0, 0, 0]], np.uint8), R.T_Gray, R.default_params(patchSize=10), mode=S,
                         width=3)
    assert r.err == R.IMAGE_SYNTH_ERROR_EMPTY_TARGET
    r, _ = R.image_synth(gray, mask2, 7, P, mode=S, width=3)
    assert r.err == R.IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT


@pytest.mark.parametrize("name", ["resynthfull-zap-texture", "resynthfull-ufo-input", "resynthfull-brick",
                                  "resynth-ufo-input",
                                  "resynthtwoimages-ufo-input", "resynthtileable-ufo-input",
                                  "resynthfull-donkey_original", "mapstyleangel-angel_target"])
def test_serial_reproduces_reference_regression_image(name):
    """The GPU serial mode reproduces Test/reference_out_images/<name>.ppm byte for byte."""
    c = pins.cases()[name]
    params, (colors, maps), tp, cp, rect, expected = pins.build(c)
    ix = R.prepare_indices(colors, maps, False, False, maps > 0)
    r, out = R.engine(R.params_from_dict(params), ix, tp, cp, mode=R.MODE_SERIAL, prng=R.PRNG_GRAND)
    x0, y0, x1, y1 = rect
    assert r.err == 0
    assert np.array_equal(out[y0:y1, x0:x1, 1:4], expected)
    # nothing outside the target changes
    keep = np.ones(out.shape[:2], bool)
    keep[y0:y1, x0:x1] = False
    assert np.array_equal(out[keep], tp[keep])


CASES = [
    # fmt, patch, probes, type, tileH, tileV, prng, size
    (R.T_RGB, 30, 200, 1, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGBA, 30, 200, 1, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_Gray, 10, 50, 1, 0, 0, R.PRNG_LIBC, (33, 47)),
    (R.T_GrayA, 64, 200, 1, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 1, 30, 1, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 2, 0, 1, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 9, 7, 0, 1, 1, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 16, 100, 0, 1, 0, R.PRNG_LIBC, (40, 52)),
    (R.T_RGB, 25, 64, 2, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 25, 65, 3, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 25, 63, 4, 0, 0, R.PRNG_LIBC, (40, 52)),
    (R.T_RGB, 25, 128, 5, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 25, 129, 6, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 25, 200, 7, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 25, 200, 8, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 30, 200, 1, 0, 0, R.PRNG_GRAND, (130, 150)),
    (R.T_RGB, 20, 255, 1, 0, 0, R.PRNG_LIBC, (40, 52)),
    (R.T_RGB, 20, 256, 1, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGB, 20, 257, 1, 0, 0, R.PRNG_GRAND, (40, 52)),
    (R.T_RGBA, 12, 600, 1, 0, 0, R.PRNG_LIBC, (40, 52)),
]


def _case_inputs(fmt, size, seed):
    h, w = size
    bpp = {R.T_RGB: 3, R.T_RGBA: 4, R.T_Gray: 1, R.T_GrayA: 2}[fmt]
    img = W.grating_noise(h, w, bpp, seed)
    if fmt in (R.T_RGBA, R.T_GrayA):     # some fully transparent and partly transparent pixels
        rng = np.random.RandomState(seed)
        a = rng.choice([0, 0, 7, 255, 255, 255, 255, 255], size=(h, w)).astype(np.uint8)
        img[..., bpp - 1] = a
    mask = np.zeros((h, w), np.uint8)
    mask[h // 4: h // 4 + h // 2, w // 3: w // 3 + w // 2] = 255
    mask[h // 4, w // 3] = 128                 # a partially selected pixel
    return img, mask


@pytest.mark.parametrize("case", range(len(CASES)))
def test_serial_matches_oracle(case):
    fmt, patch, probes, mtype, th, tv, prng, size = CASES[case]
    img, mask = _case_inputs(fmt, size, 100 + case)
    p = R.default_params(patchSize=patch, maxProbeCount=probes, matchContextType=mtype, tileH=th, tileV=tv)
    rg, og = R.image_synth(img, mask, fmt, p, mode=R.MODE_SERIAL, prng=prng)
    rc, oc = O.image_synth(img, mask, fmt, _oparams(p), prng=prng)
    assert rg.err == rc.err == 0
    assert rg.passes == rc.passes and rg.betters == rc.betters
    assert rg.candidates == rc.scorerCalls and rg.visits == rc.visits
    assert rg.progress == rc.progress
    assert np.array_equal(og, oc)


def test_serial_deep_holes_and_small_corpus():
    """Long pass-0 scans: a hole leaving only a 3-pixel frame of context, and an engine()
    call whose corpus is smaller than the target (the offset table then covers less than
    the target image)."""
    img = W.grating_noise(90, 110, 3, 21)
    mask = np.full((90, 110), 255, np.uint8)
    mask[:3, :] = 0
    mask[-3:, :] = 0
    mask[:, :3] = 0
    mask[:, -3:] = 0
    p = R.default_params(maxProbeCount=60)
    rg, og = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
    rc, oc = O.image_synth(img, mask, 0, _oparams(p))
    assert rg.err == rc.err == 0 and np.array_equal(og, oc)
    tgt = np.concatenate([np.zeros((70, 80, 1), np.uint8), W.grating_noise(70, 80, 3, 22)], axis=2)
    tgt[20:50, 25:60, 0] = 255
    cor = np.concatenate([np.full((31, 43, 1), 255, np.uint8), W.grating_noise(31, 43, 3, 23)], axis=2)
    ix = R.prepare_indices(3, 0, False, False, False)
    p = R.default_params(patchSize=20, maxProbeCount=50)
    rg, og = R.engine(p, ix, tgt, cor, mode=R.MODE_SERIAL)
    rc, oc = O.engine(_oparams(p), O.prepare_indices(3, 0, 0, 0, 0), tgt, cor)
    assert rg.err == rc.err == 0 and np.array_equal(og, oc)


def test_serial_stream_refill_path(monkeypatch):
    """A tiny device PRNG buffer forces visits to abort and restart after refills; the
    output must not change."""
    img, mask = _case_inputs(R.T_RGB, (40, 52), 5)
    p = R.default_params(maxProbeCount=300)
    r0, o0 = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
    monkeypatch.setenv("RESYNTH_SERIAL_STREAM_CAP", "1100")
    r1, o1 = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
    monkeypatch.delenv("RESYNTH_SERIAL_STREAM_CAP")
    rc, oc = O.image_synth(img, mask, 0, _oparams(p))
    assert r1.launches > r0.launches
    assert np.array_equal(o1, oc) and np.array_equal(o0, oc)


def test_serial_stream_grow_path(monkeypatch):
    """A visit that needs more raw PRNG outputs than the whole device buffer holds (1 500
    probes against a 1 024-word buffer) grows the buffer and resumes at the same visit;
    the output must still equal the oracle's byte for byte."""
    img, mask = _case_inputs(R.T_RGB, (36, 44), 9)
    p = R.default_params(patchSize=12, maxProbeCount=1500)
    monkeypatch.setenv("RESYNTH_SERIAL_STREAM_CAP", "1024")
    r1, o1 = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
    monkeypatch.delenv("RESYNTH_SERIAL_STREAM_CAP")
    rc, oc = O.image_synth(img, mask, 0, _oparams(p))
    assert r1.err == rc.err == 0
    assert r1.passes == rc.passes and r1.betters == rc.betters and r1.visits == rc.visits
    assert np.array_equal(o1, oc)


def test_serial_imagesynth2_and_cancel():
    img, mask = _case_inputs(R.T_RGB, (48, 60), 7)
    corpus = np.zeros_like(mask)
    corpus[:, :20] = 255                     # corpus: left band only (lib/imageSynth.c:140-217)
    p = R.default_params(patchSize=20, maxProbeCount=80)
    rg, og = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL, mask2=corpus)
    rc, oc = O.image_synth(img, mask, 0, _oparams(p), mask2=corpus)
    assert rg.err == rc.err == 0 and np.array_equal(og, oc)
    small = np.full((10, 20), 255, np.uint8)   # mask2 of other dimensions fills sequentially
    rg, og = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL, mask2=small)
    rc, oc = O.image_synth(img, mask, 0, _oparams(p), mask2=small)
    assert rg.err == rc.err == 0 and np.array_equal(og, oc)
    rg, og = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL, cancel=1)
    rc, oc = O.image_synth(img, mask, 0, _oparams(p), cancel=1)
    assert rg.err == rc.err == 0 and np.array_equal(og, img) and rg.progress == rc.progress


def test_device_entry_matches_host_entry():
    import torch
    img, mask = W.config1(3)
    img = img[:128, :128].copy()
    mask = mask[:128, :128].copy()
    mask[:] = 0
    mask[40:80, 50:90] = 255
    p = R.default_params()
    rh, oh = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
    d_img = torch.from_numpy(img).cuda()
    d_mask = torch.from_numpy(mask).cuda()
    torch.cuda.synchronize()
    rd = R.image_synth_device(d_img.data_ptr(), 128, 128, 128 * 3, d_mask.data_ptr(), 128, R.T_RGB, p,
                              mode=R.MODE_SERIAL)
    assert rd.err == rh.err == 0
    assert np.array_equal(d_img.cpu().numpy(), oh)


# ---------------------------------------------------------------- batched mode

def test_batched_invariants_and_quality_config1():
    """Batched passes on config 1 images: untouched context, every target colour copied
    from a corpus pixel, and a mean best-match error close to the serial engine's.  The
    serial engine's own error varies by about 4% between PRNG seeds at this size, so both
    sides are averaged over 6 images x 2 seeds."""
    p = R.default_params()
    ser, bat = [], []
    for img_seed in range(1, 7):
        img, mask = W.config1(img_seed)
        sel = mask != 0
        corpus_colours = {tuple(c) for c in img[~sel].reshape(-1, 3)}
        for s in (0, 11):
            r, _ = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL, seed=s)
            assert r.err == 0
            ser.append(r.meanError)
        for s in (1, 2):
            r, out = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_BATCHED, seed=s)
            assert r.err == 0
            assert np.array_equal(out[~sel], img[~sel])
            assert all(tuple(c) in corpus_colours for c in out[sel].reshape(-1, 3)[::5])
            bat.append(r.meanError)
    ratio = np.mean(bat) / np.mean(ser)
    print("config 1 mean error: batched %.0f serial %.0f ratio %.4f" % (np.mean(bat), np.mean(ser), ratio))
    assert ratio <= 1.05, (bat, ser)


@pytest.mark.parametrize("mtype,th,tv,bound", [(0, 1, 1, 1.6), (2, 0, 0, 1.2), (5, 0, 0, 1.2), (8, 1, 0, 1.2),
                                               (3, 0, 1, 1.2)])
def test_batched_orderings_and_tiling(mtype, th, tv, bound):
    """Batched passes under the target orderings and seamless tiling that the heal,
    uncrop and render-texture callers select (lib/orderTarget.h:269-343,
    lib/synthesize.h:81-113): context untouched, corpus colours only, and a mean error
    close to the serial engine's under the same parameters.  At 256x256 one run's error
    varies between PRNG seeds by 6-14% with context and by about 50% without it (type 0),
    so each side averages 3 images x 4 seeds and the bound leaves room for that spread
    (measured ratios: 1.01-1.05 for types 2, 3, 8 and 1.06-1.11 for type 5)."""
    p = R.default_params(matchContextType=mtype, tileH=th, tileV=tv)
    ser, bat = [], []
    for img_seed in (1, 2, 3):
        img, mask = W.config1(img_seed)
        sel = mask != 0
        corpus_colours = {tuple(c) for c in img[~sel].reshape(-1, 3)}
        for s in range(4):
            r, _ = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL, seed=s)
            assert r.err == 0
            ser.append(r.meanError)
            r, out = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_BATCHED, seed=s)
            assert r.err == 0
            assert np.array_equal(out[~sel], img[~sel])
            assert all(tuple(c) in corpus_colours for c in out[sel].reshape(-1, 3)[::3])
            bat.append(r.meanError)
    ratio = np.mean(bat) / np.mean(ser)
    print("type %d tile %d%d mean error: batched %.0f serial %.0f ratio %.4f"
          % (mtype, th, tv, np.mean(bat), np.mean(ser), ratio))
    assert ratio <= bound, (bat, ser)


def test_batched_config2_768():
    """Config 2's generators at 768x768 (about 75k targets, chunks up to the 2.3k cap):
    context pixels stay untouched, every target takes a corpus colour, and the mean
    best-match error stays close to the serial engine's (2 PRNG seeds each side)."""
    img, mask = W.config2(size=768)
    sel = mask != 0
    p = R.default_params()
    corpus_colours = {tuple(c) for c in img[~sel][:, :3]}
    ser = []
    for s in (0, 11):
        r, _ = R.image_synth(img, mask, R.T_RGBA, p, mode=R.MODE_SERIAL, seed=s)
        assert r.err == 0
        ser.append(r.meanError)
    bat = []
    for s in (1, 2):
        r, out = R.image_synth(img, mask, R.T_RGBA, p, mode=R.MODE_BATCHED, seed=s)
        assert r.err == 0
        assert r.stats["windowLaunches"] == 0, r.stats
        assert r.stats["windowWaitsExpired"] == 0
        assert np.array_equal(out[~sel], img[~sel])
        assert np.array_equal(out[sel][:, 3], img[sel][:, 3])
        assert all(tuple(c) in corpus_colours for c in out[sel][::7, :3])
        bat.append(r.meanError)
    ratio = np.mean(bat) / np.mean(ser)
    print("768 mean error: batched %.0f serial %.0f ratio %.4f" % (np.mean(bat), np.mean(ser), ratio))
    assert ratio <= 1.05, (bat, ser)


# ---------------------------------------------------------------- map-guided transfer (engine)

def test_serial_map_transfer_matches_oracle():
    """Config 4's map-guided transfer through engine() at reduced size (corpus 48x48,
    whole 72x72 target canvas with no context): the serial mode equals the oracle byte for
    byte, the map term included."""
    tp, cp = W.config4(seed=4, corpus_size=48, target_size=72)
    p = R.default_params(mapWeight=0.5)
    ix = R.prepare_indices(3, 1, False, False, True)
    assert (ix.colorEndBip, ix.map_start_bip, ix.map_end_bip, ix.total_bpp) == (4, 4, 5, 5)
    rg, og = R.engine(p, ix, tp, cp, mode=R.MODE_SERIAL)
    rc, oc = O.engine(_oparams(p), O.prepare_indices(3, 1, False, False, True), tp, cp)
    assert rg.err == rc.err == 0
    assert rg.passes == rc.passes and rg.betters == rc.betters
    assert rg.nTargets == 72 * 72
    assert np.array_equal(og, oc)
    assert np.array_equal(og[..., 4], tp[..., 4])     # maps are matched, never written


def test_batched_map_transfer_config4():
    """Config 4 at full size (768x768 corpus, 1536x1536 target canvas): every target pixel
    gets a corpus colour, the target map is untouched, and the batched error stays close to
    the serial engine's on a 96x96 crop of the same problem."""
    tp, cp = W.config4()
    p = R.default_params(mapWeight=0.5)
    ix = R.prepare_indices(3, 1, False, False, True)
    r, out = R.engine(p, ix, tp, cp, mode=R.MODE_BATCHED, seed=1)
    assert r.err == 0 and r.nTargets == 1536 * 1536
    assert r.stats["windowWaitsExpired"] == 0
    assert np.array_equal(out[..., 4], tp[..., 4])
    corpus_colours = {tuple(c) for c in cp[..., 1:4].reshape(-1, 3)}
    assert all(tuple(c) in corpus_colours for c in out[..., 1:4].reshape(-1, 3)[::997])
    assert np.isfinite(r.meanError) and r.meanError > 0
    ts, cs = W.config4(corpus_size=96, target_size=96)
    ser = [R.engine(p, ix, ts, cs, mode=R.MODE_SERIAL, seed=s)[0].meanError for s in (0, 11)]
    bat = [R.engine(p, ix, ts, cs, mode=R.MODE_BATCHED, seed=s)[0].meanError for s in (1, 2)]
    ratio = np.mean(bat) / np.mean(ser)
    print("config 4 crop mean error: batched %.0f serial %.0f ratio %.4f" % (np.mean(bat), np.mean(ser), ratio))
    assert ratio <= 1.08, (bat, ser)


FORMAT_MATRIX = [
    # colour channels, map channels, alpha in target, alpha in corpus, patch, probes, mapWeight
    (3, 3, True, True, 30, 200, 0.5),       # RGBA with RGB maps: total_bpp 8, the widest pixel
    (3, 3, False, False, 20, 100, 0.3),
    (1, 1, True, True, 16, 64, 0.5),
    (1, 3, False, True, 30, 120, 0.8),      # grey image, RGB maps, alpha only in the corpus
    (3, 1, True, False, 12, 50, 0.5),       # alpha only in the target
    (1, 0, True, True, 10, 60, 0.5),        # grey+alpha, no map
]


def _engine_pixmaps(colors, maps, alpha_t, alpha_s, seed, tsize=(40, 52), csize=(36, 44)):
    """Target and corpus pixmaps in the engine layout [mask, colours, alpha?, maps]
    (lib/imageFormat.c:128-150): the target selection holds a hole, a partially selected
    pixel and unselected context; the corpus has an unselected band, partially selected
    pixels (not corpus) and, where the format has alpha, fully transparent pixels."""
    alpha = alpha_t or alpha_s
    depth = 1 + colors + int(alpha) + maps
    rng = np.random.RandomState(seed)

    def build(h, w, s):
        p = np.zeros((h, w, depth), np.uint8)
        p[..., 1:1 + colors] = W.grating_noise(h, w, colors, s)
        if alpha:
            p[..., 1 + colors] = rng.choice([0, 0, 9, 255, 255, 255, 255, 255], size=(h, w))
        if maps:
            p[..., 1 + colors + int(alpha):] = W.grating_noise(h, w, maps, s + 1)
        return p

    th, tw = tsize
    tp = build(th, tw, seed)
    tp[th // 4: th // 4 + th // 2, tw // 3: tw // 3 + tw // 2, 0] = 255
    tp[th // 5, tw // 2, 0] = 77
    ch, cw = csize
    cp = build(ch, cw, seed + 10)
    cp[..., 0] = 255
    cp[: ch // 6, :, 0] = 0
    cp[ch // 2, : cw // 3, 0] = 100
    return tp, cp


@pytest.mark.parametrize("case", range(len(FORMAT_MATRIX)))
def test_engine_format_matrix(case):
    """engine() over the pixel formats the plugin hands it (lib/imageFormat.c:117-205):
    grey or RGB colours, none/one/three map channels, alpha in the target, the corpus, both
    or neither.  The serial mode equals the oracle byte for byte; the batched mode writes
    only colour channels of target pixels, and only colours of corpus pixels."""
    colors, maps, at, as_, patch, probes, mw = FORMAT_MATRIX[case]
    tp, cp = _engine_pixmaps(colors, maps, at, as_, 500 + case)
    p = R.default_params(patchSize=patch, maxProbeCount=probes, mapWeight=mw)
    ix = R.prepare_indices(colors, maps, at, as_, maps > 0)
    assert ix.total_bpp == tp.shape[2]
    print("format case", case, FORMAT_MATRIX[case], "total_bpp", ix.total_bpp, flush=True)
    rg, og = R.engine(p, ix, tp, cp, mode=R.MODE_SERIAL)
    rc, oc = O.engine(_oparams(p), O.prepare_indices(colors, maps, at, as_, maps > 0), tp, cp)
    assert rg.err == rc.err == 0
    assert rg.nTargets == rc.nTargets > 0
    assert rg.passes == rc.passes and rg.betters == rc.betters
    assert rg.candidates == rc.scorerCalls
    assert rg.pixelWordBytes == (4 if colors + maps <= 3 else 8)
    assert np.array_equal(og, oc)

    rb, ob = R.engine(p, ix, tp, cp, mode=R.MODE_BATCHED, seed=3)
    assert rb.err == 0 and rb.nTargets == rg.nTargets
    target = tp[..., 0] != 0
    assert np.array_equal(ob[~target], tp[~target])
    keep = [0] + list(range(1 + colors, tp.shape[2]))      # mask, alpha, maps are never written
    assert np.array_equal(ob[..., keep], tp[..., keep])
    # Random probes come from opaque selected corpus pixels, but heuristic-1 candidates are
    # checked against the selection only (lib/engine.c:505-517), so a transparent selected
    # corpus pixel can lend its colour too.
    corpus = cp[..., 0] == 255
    colours = {tuple(c) for c in cp[corpus][:, 1:1 + colors]}
    assert all(tuple(c) in colours for c in ob[target][:, 1:1 + colors])
    assert rb.meanError <= 1.3 * rg.meanError, (rb.meanError, rg.meanError)


# ---------------------------------------------------------------- many requests in one launch

def _uneven_requests(n=10, seed=77):
    rng = np.random.RandomState(seed)
    imgs, masks, fmts, params = [], [], [], []
    for k in range(n):
        w, h = (int(v) for v in rng.randint(48, 321, size=2))
        fmt = [R.T_RGB, R.T_RGBA, R.T_Gray][k % 3]
        bpp = {R.T_RGB: 3, R.T_RGBA: 4, R.T_Gray: 1}[fmt]
        img = W.grating_noise(h, w, bpp, seed + k)
        if fmt == R.T_RGBA:
            img[..., 3] = 255
        m = W.ellipse_mask(h, w, int(rng.randint(1, 4)), seed + k, axes=(0.08 * min(w, h), 0.25 * min(w, h)))
        imgs.append(img)
        masks.append((m * 255).astype(np.uint8))
        fmts.append(fmt)
        params.append(R.default_params(patchSize=int(rng.choice([10, 30, 50]))))
    return imgs, masks, fmts, params


def test_batch_of_uneven_requests():
    """Requests of uneven sizes, formats and patch sizes in one multi-request launch: each
    request keeps its context, takes corpus colours, reports the errors imageSynth would
    (mask mismatch, bad format, empty target), and its mean match error stays close to the
    same request run alone in batched mode."""
    imgs, masks, fmts, params = _uneven_requests()
    # three bad requests at the end
    imgs += [imgs[0], imgs[1], imgs[2]]
    masks += [masks[0][:-1], masks[1], np.zeros_like(masks[2])]
    fmts += [fmts[0], 9, fmts[2]]
    params += [params[0], params[1], params[2]]
    n = len(imgs) - 3
    seeds = (1, 2, 3)
    multi = np.zeros((len(seeds), n))
    for si, seed in enumerate(seeds):
        rc, res, st, outs = R.image_synth_batch(imgs, masks, fmts, params, seed=seed)
        assert rc == 0
        assert [r["error"] for r in res[n:]] == [R.IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH,
                                               R.IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT,
                                               R.IMAGE_SYNTH_ERROR_EMPTY_TARGET]
        assert st["requests"] == n and st["waitsExpired"] == 0 and st["graphLaunches"] == 1
        for i in range(n):
            r = res[i]
            assert r["error"] == 0 and r["passes"] >= 1
            sel = masks[i] != 0
            assert r["nTargets"] == int(sel.sum())
            out, img = outs[i], imgs[i]
            assert np.array_equal(out[~sel], img[~sel])
            ncol = 1 if fmts[i] == R.T_Gray else 3
            corpus = {tuple(c) for c in img[~sel].reshape(-1, img.shape[-1])[:, :ncol]}
            assert all(tuple(c) in corpus for c in out[sel].reshape(-1, img.shape[-1])[::3, :ncol])
            multi[si, i] = r["meanError"]
    # the same requests one by one, over the same probe seeds (tiny requests: the error of
    # one run spreads by +-15% from seed to seed)
    single = np.array([[R.image_synth(imgs[i], masks[i], fmts[i], params[i], mode=R.MODE_BATCHED, seed=s)[0].meanError
                        for i in range(n)] for s in seeds])
    ratios = multi.mean(0) / single.mean(0)
    print("batch vs single mean error ratios", np.round(ratios, 3))
    assert np.mean(ratios) <= 1.05, ratios


def test_batch_of_rgb_requests_with_odd_sizes():
    """An all-RGB batch (the multi-request kernel's three-colour scorer) whose images have odd
    pixel counts large enough for the multi-block sorts of the offset tables: every request
    succeeds, context pixels stay untouched and targets take corpus colours."""
    sizes = [(1001, 769), (517, 1023), (403, 401)]
    imgs, masks = [], []
    for k, (h, w) in enumerate(sizes):
        img = W.grating_noise(h, w, 3, 700 + k)
        m = W.ellipse_mask(h, w, 2, 700 + k, axes=(0.08 * min(w, h), 0.2 * min(w, h)))
        imgs.append(img)
        masks.append((m * 255).astype(np.uint8))
    params = [R.default_params(patchSize=p) for p in (10, 30, 50)]
    print("odd RGB batch", sizes, flush=True)
    rc, res, st, outs = R.image_synth_batch(imgs, masks, [R.T_RGB] * 3, params)
    assert rc == 0
    for img, m, r, out in zip(imgs, masks, res, outs):
        assert r["error"] == 0 and r["nTargets"] == int((m != 0).sum())
        sel = m != 0
        assert np.array_equal(out[~sel], img[~sel])
        colours = {tuple(c) for c in img[~sel].reshape(-1, 3)}
        assert all(tuple(c) in colours for c in out[sel].reshape(-1, 3)[::11])


def test_batch_graph_is_kept_and_updated_across_calls():
    """The multi-request graph is instantiated once per kernel shape: later calls with other
    requests (count, sizes, buffers) update its nodes in place.  Every replay, with eager
    single-request runs and synchronisations between them, keeps each request's context,
    takes corpus colours and keeps the error level of the first call."""
    imgs, masks, fmts, params = _uneven_requests(n=6, seed=77)
    rgb = [i for i, f in enumerate(fmts) if f == R.T_RGB]
    first = None
    for call, subset in enumerate([rgb, rgb[:3], rgb[1:], rgb]):
        rc, res, st, outs = R.image_synth_batch([imgs[i] for i in subset], [masks[i] for i in subset],
                                                [fmts[i] for i in subset], [params[i] for i in subset],
                                                config=(4, 64), seed=5)
        assert rc == 0 and st["graphLaunches"] == 1
        assert st["graphInstantiations"] == (1 if call == 0 else 0) or call == 0, st
        for k, i in enumerate(subset):
            sel = masks[i] != 0
            assert res[k]["error"] == 0
            assert np.array_equal(outs[k][~sel], imgs[i][~sel])
        errs = {i: res[k]["meanError"] for k, i in enumerate(subset)}
        if first is None:
            first = errs
        else:
            ratio = np.mean([errs[i] / first[i] for i in subset])
            assert 0.8 < ratio < 1.25, (call, ratio)
        # eager work between replays
        R.image_synth(imgs[rgb[0]], masks[rgb[0]], R.T_RGB, params[rgb[0]], seed=call + 1)
    # another shape instantiates its own graph
    rc, res, st, _ = R.image_synth_batch([imgs[i] for i in rgb], [masks[i] for i in rgb], [fmts[i] for i in rgb],
                                         [params[i] for i in rgb], config=(2, 32), seed=5)
    assert rc == 0 and st["graphInstantiations"] == 1 and st["config"] == (2, 32)


@pytest.mark.parametrize("lanes", [16, 32])
def test_batch_segmented_lanes_keep_quality(lanes):
    """Multi-request launches with 4 or 2 targets per wavefront at a time (16 or 32 lanes per
    target) keep every request's invariants and the error of one target per wavefront."""
    imgs, masks, fmts, params = _uneven_requests()
    ratios = []
    seeds = (1, 2, 3)
    base = [R.image_synth_batch(imgs, masks, fmts, params, config=(4, 64), seed=s)[1] for s in seeds]
    seg = [R.image_synth_batch(imgs, masks, fmts, params, config=(4, lanes), seed=s) for s in seeds]
    for rc, res, st, outs in seg:
        assert rc == 0 and st["config"] == (4, lanes)
        for i, (img, m, out) in enumerate(zip(imgs, masks, outs)):
            assert res[i]["error"] == 0
            assert np.array_equal(out[m == 0], img[m == 0])
    for i in range(len(imgs)):
        b = np.mean([base[k][i]["meanError"] for k in range(len(seeds))])
        g = np.mean([seg[k][1][i]["meanError"] for k in range(len(seeds))])
        ratios.append(g / b)
    # tiny requests: one run's error spreads by +-15% from seed to seed
    print("lanes", lanes, "error ratios to 64 lanes", np.round(ratios, 3))
    assert np.mean(ratios) <= 1.07


def test_batch_device_buffers_and_autotuner():
    """The device-buffer form of the batch entry and the autotuner: the tuner times all 9
    kernel shapes (waves per block x lanes per target), leaves its sample unchanged, and
    later batches of the sample's class use the shape it chose."""
    import torch
    imgs, masks, fmts, params = _uneven_requests(n=4, seed=91)
    d_imgs = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in imgs]
    d_masks = [torch.from_numpy(m).cuda() for m in masks]
    dev_img = [(t.data_ptr(), a.shape[1], a.shape[0], a.strides[0]) for t, a in zip(d_imgs, imgs)]
    dev_mask = [(t.data_ptr(), m.shape[1], m.shape[0], m.strides[0]) for t, m in zip(d_masks, masks)]
    rc, best, ms = R.autotune_batch(dev_img, dev_mask, fmts, params, device=True)
    assert rc == 0 and best[0] in (2, 4, 8) and best[1] in (16, 32, 64)
    assert (ms > 0).all(), ms
    torch.cuda.synchronize()
    for t, a in zip(d_imgs, imgs):
        assert np.array_equal(t.cpu().numpy(), a)
    rc, res, st, _ = R.image_synth_batch(dev_img, dev_mask, fmts, params, device=True)
    assert rc == 0 and st["config"] == best
    torch.cuda.synchronize()
    for t, a, m, r in zip(d_imgs, imgs, masks, res):
        assert r["error"] == 0
        out = t.cpu().numpy()
        sel = m != 0
        assert np.array_equal(out[~sel], a[~sel])
        assert not np.array_equal(out[sel], a[sel])


def test_batched_window_mode_config2_768(monkeypatch):
    """RESYNTH_WINDOW=1: the later passes and the rest of pass 0 run as persistent
    sliding-window launches; no wait hits its spin limit and the result keeps the batched
    invariants and error level."""
    monkeypatch.setenv("RESYNTH_WINDOW", "1")
    img, mask = W.config2(size=768)
    sel = mask != 0
    corpus_colours = {tuple(c) for c in img[~sel][:, :3]}
    r, out = R.image_synth(img, mask, R.T_RGBA, R.default_params(), mode=R.MODE_BATCHED, seed=1)
    assert r.err == 0
    assert r.stats["windowLaunches"] >= 2 and r.stats["windowWaitsExpired"] == 0, r.stats
    assert np.array_equal(out[~sel], img[~sel])
    assert all(tuple(c) in corpus_colours for c in out[sel][::7, :3])
    monkeypatch.setenv("RESYNTH_WINDOW", "0")
    r0, _ = R.image_synth(img, mask, R.T_RGBA, R.default_params(), mode=R.MODE_BATCHED, seed=1)
    assert r0.stats["windowLaunches"] == 0
    assert r.meanError <= 1.08 * r0.meanError, (r.meanError, r0.meanError)


# ---------------------------------------------------------------- ordering on the device

def _order_vectors():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors.json")) as f:
        return json.load(f)["order"]


@pytest.mark.parametrize("mtype", range(2, 9))
def test_device_target_ordering_matches_glib(mtype):
    """The engine's device sort of types 2..8 plus the host's banded shuffle equal the whole
    ordering computed by GLib's g_array_sort / g_rand_int_range and the C library's atan2
    (tests/golden/make_golden.py), on shapes with ties, an off-target centre and a centre
    pixel whose ratio is 0/0 (the NaN case of lib/brushfire.h:117-128)."""
    import hashlib
    for name, e in _order_vectors().items():
        print("ordering type", mtype, "shape", name, flush=True)
        err, got = R.order_targets(mtype, np.array(e["points"], np.int32))
        assert err == 0
        got = np.asarray(got, "<i4").reshape(-1, 2)
        if str(mtype) in e["types"]:
            assert got.tolist() == e["types"][str(mtype)], name
        assert hashlib.sha256(got.tobytes()).hexdigest() == e["types"][str(mtype) + "_sha256"], name


def test_serial_engine_uses_device_ordering_on_a_nan_shape():
    """A whole serial engine run with the inward order (type 2) on a hole whose centre pixel
    has a NaN ratio equals the oracle byte for byte."""
    img = W.grating_noise(60, 80, 3, 31)
    yy, xx = np.mgrid[0:60, 0:80]
    mask = np.where((xx - 30) ** 2 + (yy - 30) ** 2 <= 144, 255, 0).astype(np.uint8)
    mask[30, 30:] = 0
    for mtype in (2, 5, 8):
        p = R.default_params(matchContextType=mtype, patchSize=20, maxProbeCount=60)
        rg, og = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
        rc, oc = O.image_synth(img, mask, 0, _oparams(p))
        assert rg.err == rc.err == 0 and np.array_equal(og, oc), mtype


# ---------------------------------------------------------------- the reference's C entry points

def _reference_api(tmp_path):
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "reference_api")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "c", "reference_api.c"), "-L", os.path.dirname(R.LIB_PATH),
                    "-lresynth_mi355x", "-Wl,-rpath," + os.path.dirname(R.LIB_PATH), "-o", exe], check=True)
    return exe


def _run_c(exe, args, mode=None, prng=None):
    import os
    import subprocess
    env = dict(os.environ)
    env.pop("RESYNTH_MODE", None)
    env.pop("RESYNTH_PRNG", None)
    if mode:
        env["RESYNTH_MODE"] = mode
    if prng:
        env["RESYNTH_PRNG"] = prng
    print("reference_api", args[0], mode, prng, flush=True)
    r = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    return r


@pytest.mark.parametrize("mode", ["serial", None])
def test_reference_entry_points_from_c_known_answers(tmp_path, mode):
    """src/testSynth.c:163-212's known answers through the reference-signature imageSynth,
    called from C, in the serial mode and in the default (batched) mode.  testSynth.c is
    built with the stand-alone PRNG (glibc rand(), lib/glibProxy.c:30-49), and its Gray case
    is a tie that the PRNG stream decides, so the library runs with RESYNTH_PRNG=libc."""
    prng = "libc"
    exe = _reference_api(tmp_path)
    r = _run_c(exe, ["known"], mode, prng)
    assert r.returncode == 0 and "0 failure(s)" in r.stdout


def test_reference_entry_points_from_c_engine_and_imagesynth2(tmp_path):
    """engine() from C with reference Maps reproduces Test/reference_out_images/
    resynthfull-brick.ppm in the serial mode; imageSynth2 from C equals the oracle; the
    default mode leaves everything outside the target as it was."""
    c = pins.cases()["resynthfull-brick"]
    params, (colors, maps), tp, cp, rect, expected = pins.build(c)
    th, tw, depth = tp.shape
    tp.tofile(tmp_path / "t.raw")
    cp.tofile(tmp_path / "c.raw")
    exe = _reference_api(tmp_path)
    args = ["engine", tw, th, tmp_path / "t.raw", cp.shape[1], cp.shape[0], tmp_path / "c.raw", colors, maps,
            params["tileH"], params["tileV"], params["matchContextType"], repr(params["mapWeight"]),
            repr(params["sensitivity"]), params["patchSize"], params["maxProbeCount"], tmp_path / "o.raw"]
    assert _run_c(exe, args, "serial").returncode == 0
    out = np.fromfile(tmp_path / "o.raw", np.uint8).reshape(th, tw, depth)
    x0, y0, x1, y1 = rect
    assert np.array_equal(out[y0:y1, x0:x1, 1:4], expected)
    assert _run_c(exe, args, None).returncode == 0
    out = np.fromfile(tmp_path / "o.raw", np.uint8).reshape(th, tw, depth)
    keep = np.ones((th, tw), bool)
    keep[y0:y1, x0:x1] = False
    assert np.array_equal(out[keep], tp[keep]) and np.array_equal(out[..., 0], tp[..., 0])
    # imageSynth2 with a corpus band
    img, mask = _case_inputs(R.T_RGB, (48, 60), 7)
    corpus = np.zeros_like(mask)
    corpus[:, :20] = 255
    img.tofile(tmp_path / "i.raw")
    mask.tofile(tmp_path / "m.raw")
    corpus.tofile(tmp_path / "m2.raw")
    r = _run_c(exe, ["synth2", R.T_RGB, 60, 48, tmp_path / "i.raw", tmp_path / "m.raw", tmp_path / "m2.raw",
                     tmp_path / "o2.raw"], "serial")
    assert r.returncode == 0
    rc, oc = O.image_synth(img, mask, 0, _oparams(R.default_params()), mask2=corpus)
    assert rc.err == 0 and np.array_equal(np.fromfile(tmp_path / "o2.raw", np.uint8).reshape(oc.shape), oc)


# ---------------------------------------------------------------- failures return codes

def test_out_of_device_memory_returns_an_error_and_recovers():
    """An engine call that cannot get its device memory returns RESYNTH_ERROR_DEVICE and
    leaves the caller's image as it was; the next call with memory succeeds and is exact."""
    img, mask = _case_inputs(R.T_RGB, (64, 80), 3)
    p = R.default_params(maxProbeCount=50)
    try:
        R.set_device_memory_limit(1 << 20)
        r, out = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
        assert r.err == R.RESYNTH_ERROR_DEVICE and np.array_equal(out, img)
    finally:
        R.set_device_memory_limit(0)
    rg, og = R.image_synth(img, mask, R.T_RGB, p, mode=R.MODE_SERIAL)
    rc, oc = O.image_synth(img, mask, 0, _oparams(p))
    assert rg.err == rc.err == 0 and np.array_equal(og, oc)


def test_batch_rejects_inconsistent_requests_and_honours_options():
    """A request whose row pitch is shorter than its rows fails alone with
    RESYNTH_ERROR_INVALID_ARGUMENT (and the autotuner skips it); maxPasses caps every
    request's passes."""
    imgs, masks = [], []
    for k in range(3):
        im, m = _case_inputs(R.T_RGB, (50 + 7 * k, 60 + 5 * k), 40 + k)
        imgs.append(im)
        masks.append(m)
    reqs, keep = R._requests(imgs, masks, [R.T_RGB] * 3, [R.default_params(maxProbeCount=40)] * 3, False)
    reqs[1].image.rowBytes = 10
    st = R.ResynthBatchStats()
    opt = R.options(max_passes=1)
    import ctypes
    assert R.lib().resynth_image_synth_batch(reqs, 3, 0, ctypes.byref(opt), None, ctypes.byref(st)) == 0
    assert reqs[1].error == R.RESYNTH_ERROR_INVALID_ARGUMENT
    assert reqs[0].error == 0 and reqs[2].error == 0 and reqs[0].passes == 1 and reqs[2].passes == 1
    best = R.ResynthKernelConfig()
    assert R.lib().resynth_autotune_batch(reqs, 3, 0, ctypes.byref(best), None) == 0


# ---------------------------------------------------------------- segmented visits

@pytest.mark.parametrize("fmt,probes,mtype", [(R.T_RGB, 200, 1), (R.T_RGBA, 30, 2), (R.T_Gray, 7, 1)])
def test_segmented_visits_give_each_target_the_wave_kernels_result(fmt, probes, mtype):
    """With one target per launch (batch 1: deterministic), wavefronts split into 16-, 32- or
    64-lane segments give every target exactly what one wavefront per target gives: the same
    image, the same candidates scored and the same betters."""
    img, mask = _case_inputs(fmt, (40, 52), 60 + probes)
    p = R.default_params(maxProbeCount=probes, matchContextType=mtype, patchSize=25)
    ref = None
    for lpt in (1, 16, 32, 64):
        print("lanes per target", lpt, flush=True)
        r, out = R.image_synth(img, mask, fmt, p, batch=1, lanes_per_target=lpt)
        assert r.err == 0
        got = (out, r.candidates, list(r.betters), r.passes, r.meanError)
        if ref is None:
            ref = got
        else:
            assert np.array_equal(got[0], ref[0]) and got[1:] == ref[1:], lpt


def test_pass_cap_keeps_the_per_pass_counts():
    """A run capped at two passes reads its last pass's counters together with the run's
    statistics instead of after the pass.  With one target per launch (deterministic) its two
    passes are those of a three-pass run, whose second pass is read the ordinary way: the
    same targets per pass, the same betters."""
    img, mask = _case_inputs(R.T_RGB, (40, 52), 77)
    p = R.default_params(maxProbeCount=30)
    r2, _ = R.image_synth(img, mask, R.T_RGB, p, batch=1, max_passes=2)
    r3, _ = R.image_synth(img, mask, R.T_RGB, p, batch=1, max_passes=3)
    assert r2.err == 0 and r3.err == 0 and r2.passes == 2 and r3.passes >= 2
    assert list(r2.passSize)[:2] == list(r3.passSize)[:2]
    assert list(r2.betters)[:2] == list(r3.betters)[:2] and r2.betters[1] > 0


@pytest.mark.parametrize("lpt", [16, 32, 64])
def test_segmented_visits_keep_batched_quality(lpt):
    """Concurrent segmented visits keep the batched mode's invariants and its mean error
    (within the seed-to-seed spread of the wavefront kernel on a 256 x 256 heal)."""
    img, mask = W.config1()
    p = R.default_params()
    e_wave = np.mean([R.image_synth(img, mask, R.T_RGB, p, seed=s, lanes_per_target=1)[0].meanError
                      for s in range(1, 5)])
    errs = []
    for s in range(1, 5):
        r, out = R.image_synth(img, mask, R.T_RGB, p, seed=s, lanes_per_target=lpt)
        assert r.err == 0 and np.array_equal(out[mask == 0], img[mask == 0])
        errs.append(r.meanError)
    assert np.mean(errs) <= 1.04 * e_wave, (np.mean(errs), e_wave)
