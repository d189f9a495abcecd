"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The shipped library never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PRNG_GRAND = 0
PRNG_LIBC = 1


class OImageBuffer(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_uint), ("height", ctypes.c_uint),
                ("rowBytes", ctypes.c_size_t)]


class OParams(ctypes.Structure):
    _fields_ = [("tileH", ctypes.c_int), ("tileV", ctypes.c_int), ("matchContextType", ctypes.c_int),
                ("mapWeight", ctypes.c_double), ("sensitivity", ctypes.c_double),
                ("patchSize", ctypes.c_uint), ("maxProbeCount", ctypes.c_uint)]


class OFormatIndices(ctypes.Structure):
    _fields_ = [("colorEndBip", ctypes.c_uint8), ("alpha_bip", ctypes.c_uint8),
                ("map_start_bip", ctypes.c_uint8), ("map_end_bip", ctypes.c_uint8),
                ("img_match_bpp", ctypes.c_uint8), ("map_match_bpp", ctypes.c_uint8),
                ("total_bpp", ctypes.c_uint8), ("isAlphaTarget", ctypes.c_int),
                ("isAlphaSource", ctypes.c_int)]


class OPixmap(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint), ("height", ctypes.c_uint), ("depth", ctypes.c_uint),
                ("data", ctypes.c_void_p)]


class OOptions(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int), ("prng", ctypes.c_int), ("batch", ctypes.c_int),
                ("conflictRadius", ctypes.c_int), ("rampPermille", ctypes.c_int),
                ("seed", ctypes.c_uint), ("minChunk", ctypes.c_int)]


class OStats(ctypes.Structure):
    _fields_ = [("nTargets", ctypes.c_uint), ("nCorpus", ctypes.c_uint), ("nOffsets", ctypes.c_uint),
                ("passes", ctypes.c_uint), ("passSize", ctypes.c_uint * 6), ("betters", ctypes.c_uint * 6),
                ("scorerCalls", ctypes.c_ulonglong), ("visits", ctypes.c_ulonglong),
                ("progressCalls", ctypes.c_uint), ("lastPercent", ctypes.c_int),
                ("lastBest", ctypes.c_void_p), ("sourceOf", ctypes.c_void_p),
                ("meanError", ctypes.c_double)]


PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_void_p)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_grand_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.orc_grand_int.argtypes = [ctypes.c_void_p]
        L.orc_grand_int.restype = ctypes.c_uint32
        L.orc_grand_int_range.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        L.orc_grand_int_range.restype = ctypes.c_int32
        L.orc_libc_srand.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.orc_libc_rand.argtypes = [ctypes.c_void_p]
        L.orc_libc_rand.restype = ctypes.c_int32
        L.orc_libc_int_range.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_libc_int_range.restype = ctypes.c_uint32
        L.orc_sorted_offsets.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_quantize.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_repetition.argtypes = [ctypes.c_uint, ctypes.c_void_p]
        L.orc_best_of.argtypes = [ctypes.POINTER(OPixmap), ctypes.POINTER(OFormatIndices), ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_uint, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32),
                                  ctypes.POINTER(ctypes.c_int)]
        L.orc_best_of.restype = ctypes.c_int
        L.orc_order_targets.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_uint, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_engine.argtypes = [OParams, ctypes.POINTER(OFormatIndices), ctypes.POINTER(OPixmap),
                                    ctypes.POINTER(OPixmap), PROGRESS, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(OOptions),
                                    ctypes.POINTER(OStats)]
        for name in ("oracle_imageSynth",):
            getattr(L, name).argtypes = [ctypes.POINTER(OImageBuffer), ctypes.POINTER(OImageBuffer), ctypes.c_int,
                                         ctypes.POINTER(OParams), PROGRESS, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(OOptions),
                                         ctypes.POINTER(OStats)]
        L.oracle_imageSynth2.argtypes = [ctypes.POINTER(OImageBuffer), ctypes.POINTER(OImageBuffer),
                                         ctypes.POINTER(OImageBuffer), ctypes.c_int, ctypes.POINTER(OParams),
                                         PROGRESS, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(OOptions), ctypes.POINTER(OStats)]
        L.oracle_count_pixelels.argtypes = [ctypes.c_int]
        L.oracle_count_pixelels.restype = ctypes.c_uint
        L.oracle_format_indices.argtypes = [ctypes.POINTER(OFormatIndices), ctypes.c_int]
        L.oracle_prepare_indices.argtypes = [ctypes.POINTER(OFormatIndices), ctypes.c_uint, ctypes.c_uint,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def default_params(**kw):
    p = dict(tileH=0, tileV=0, matchContextType=1, mapWeight=0.5, sensitivity=0.117, patchSize=30,
             maxProbeCount=200)
    p.update(kw)
    return p


def _params(p):
    return OParams(**p)


# ---------------------------------------------------------------- PRNG helpers

def grand_ints(seed, n):
    st = ctypes.create_string_buffer(624 * 4 + 16)
    lib().orc_grand_seed(st, seed)
    return [lib().orc_grand_int(st) for _ in range(n)]


def grand_ranges(seed, dists):
    st = ctypes.create_string_buffer(624 * 4 + 16)
    lib().orc_grand_seed(st, seed)
    return [lib().orc_grand_int_range(st, 0, d) for d in dists]


def libc_rands(seed, n):
    st = ctypes.create_string_buffer(31 * 4 + 16)
    lib().orc_libc_srand(st, seed)
    return [lib().orc_libc_rand(st) for _ in range(n)]


def libc_ranges(seed, dists):
    st = ctypes.create_string_buffer(31 * 4 + 16)
    lib().orc_libc_srand(st, seed)
    return [lib().orc_libc_int_range(st, 0, d) for d in dists]


def sorted_offsets(w, h):
    out = np.empty(((2 * w - 1) * (2 * h - 1), 2), dtype=np.int32)
    lib().orc_sorted_offsets(w, h, out.ctypes.data)
    return out


def quantize(sensitivity=0.117, mapWeight=0.5):
    T = np.zeros(512, dtype=np.uint16)
    Tm = np.zeros(512, dtype=np.uint32)
    lib().orc_quantize(sensitivity, mapWeight, T.ctypes.data, Tm.ctypes.data)
    return T, Tm


def repetition(n):
    out = (ctypes.c_uint * 6)()
    lib().orc_repetition(n, out)
    return list(out)


def order_targets(match_type, points, prng=PRNG_GRAND):
    pts = np.ascontiguousarray(points, dtype=np.int32).copy()
    used = ctypes.c_uint64(0)
    err = lib().orc_order_targets(match_type, pts.ctypes.data, len(pts), prng, ctypes.byref(used))
    return err, pts, used.value


def format_indices(fmt):
    ix = OFormatIndices()
    err = lib().oracle_format_indices(ctypes.byref(ix), fmt)
    return err, ix


def prepare_indices(colors, maps, alpha_t, alpha_s, is_map):
    ix = OFormatIndices()
    lib().oracle_prepare_indices(ctypes.byref(ix), colors, maps, int(alpha_t), int(alpha_s), int(is_map))
    return ix


def best_of(corpus_pixmap, ix, T, Tm, offs, pix, cands):
    """corpus_pixmap: uint8 (H, W, depth); offs int32 (n,2); pix uint8 (n,8); cands int32 (m,2)."""
    cp = np.ascontiguousarray(corpus_pixmap, dtype=np.uint8)
    pm = OPixmap(cp.shape[1], cp.shape[0], cp.shape[2], cp.ctypes.data)
    offs = np.ascontiguousarray(offs, dtype=np.int32)
    pix = np.ascontiguousarray(pix, dtype=np.uint8)
    cands = np.ascontiguousarray(cands, dtype=np.int32)
    T = np.ascontiguousarray(T, dtype=np.uint16)
    Tm = np.ascontiguousarray(Tm, dtype=np.uint32)
    best = ctypes.c_uint32(0)
    stop = ctypes.c_int(0)
    win = lib().orc_best_of(ctypes.byref(pm), ctypes.byref(ix), T.ctypes.data, Tm.ctypes.data, len(offs),
                            offs.ctypes.data, pix.ctypes.data, len(cands), cands.ctypes.data,
                            ctypes.byref(best), ctypes.byref(stop))
    return win, best.value, stop.value


# ---------------------------------------------------------------- engine entries

class Result:
    pass


def _stats(shape_hw, want_maps):
    st = OStats()
    keep = {}
    if want_maps:
        h, w = shape_hw
        keep["lastBest"] = np.zeros((h, w), dtype=np.uint32)
        keep["sourceOf"] = np.zeros((h, w, 2), dtype=np.int32)
        st.lastBest = keep["lastBest"].ctypes.data
        st.sourceOf = keep["sourceOf"].ctypes.data
    return st, keep


def _result(err, st, keep, progress_log):
    r = Result()
    r.err = err
    r.nTargets, r.nCorpus, r.nOffsets = st.nTargets, st.nCorpus, st.nOffsets
    r.passes = st.passes
    r.passSize = list(st.passSize)[: st.passes]
    r.betters = list(st.betters)[: st.passes]
    r.scorerCalls = st.scorerCalls
    r.visits = st.visits
    r.meanError = st.meanError
    r.progress = progress_log
    r.lastBest = keep.get("lastBest")
    r.sourceOf = keep.get("sourceOf")
    return r


def _buffer(arr2d_bytes, width, height, row_bytes):
    return OImageBuffer(arr2d_bytes.ctypes.data, width, height, row_bytes)


def image_synth(image, mask, fmt, params=None, threads=1, prng=PRNG_GRAND, cancel=0, want_maps=False,
                mask2=None, width=None, mask_width=None, mask2_width=None, batch=0, radius=0, ramp=0, seed=0,
                min_chunk=0):
    """image: uint8 (H, W*bpp + pad) row buffer or (H, W, C); mask: uint8 (H, W[+pad]).
    Returns (result, image_out) with image_out the same shape as image."""
    img = np.array(image, dtype=np.uint8, copy=True, order="C")
    bpp = int(lib().oracle_count_pixelels(fmt)) or 1
    if img.ndim == 3:
        h, w = img.shape[0], img.shape[1]
        row = img.shape[1] * img.shape[2]
    else:
        h, row = img.shape
        w = width if width is not None else row // bpp
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    mh = m.shape[0]
    mw = mask_width if mask_width is not None else m.shape[1]
    ib = OImageBuffer(img.ctypes.data, w, h, row)
    mb = OImageBuffer(m.ctypes.data, mw, mh, m.shape[1])
    st, keep = _stats((h, w), want_maps)
    log = []
    cb = PROGRESS(lambda pct, ctx: log.append(pct))
    cf = ctypes.c_int(cancel)
    opt = OOptions(threads, prng, batch, radius, ramp, seed, min_chunk)
    pp = ctypes.byref(_params(params)) if params is not None else None
    if mask2 is None:
        err = lib().oracle_imageSynth(ctypes.byref(ib), ctypes.byref(mb), fmt, pp, cb, None, ctypes.byref(cf),
                                      ctypes.byref(opt), ctypes.byref(st))
    else:
        m2 = np.ascontiguousarray(mask2, dtype=np.uint8)
        mb2 = OImageBuffer(m2.ctypes.data, mask2_width if mask2_width is not None else m2.shape[1],
                           m2.shape[0], m2.shape[1])
        err = lib().oracle_imageSynth2(ctypes.byref(ib), ctypes.byref(mb), ctypes.byref(mb2), fmt, pp, cb, None,
                                       ctypes.byref(cf), ctypes.byref(opt), ctypes.byref(st))
    return _result(err, st, keep, log), img


def engine(params, ix, target_pixmap, corpus_pixmap, threads=1, prng=PRNG_GRAND, cancel=0, want_maps=False,
           batch=0, radius=0, ramp=0, seed=0, min_chunk=0):
    """target/corpus pixmaps: uint8 (H, W, depth) interleaved [mask, colours, alpha?, maps].
    Returns (result, target_out)."""
    t = np.array(target_pixmap, dtype=np.uint8, copy=True, order="C")
    c = np.ascontiguousarray(corpus_pixmap, dtype=np.uint8)
    tp = OPixmap(t.shape[1], t.shape[0], t.shape[2], t.ctypes.data)
    cp = OPixmap(c.shape[1], c.shape[0], c.shape[2], c.ctypes.data)
    st, keep = _stats(t.shape[:2], want_maps)
    log = []
    cb = PROGRESS(lambda pct, ctx: log.append(pct))
    cf = ctypes.c_int(cancel)
    opt = OOptions(threads, prng, batch, radius, ramp, seed, min_chunk)
    err = lib().oracle_engine(_params(params), ctypes.byref(ix), ctypes.byref(tp), ctypes.byref(cp), cb, None,
                              ctypes.byref(cf), ctypes.byref(opt), ctypes.byref(st))
    return _result(err, st, keep, log), t
