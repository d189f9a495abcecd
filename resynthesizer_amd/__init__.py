"""Python host interface of libresynth_mi355x.so (the MI355X resynthesizer engine).

Mirrors the reference library's entry points (lib/imageSynth.h:32-52, lib/engine.h:4-12):
same argument meaning, error codes and in-place semantics, on numpy arrays.  Every call
runs on the GPU through the C ABI declared in include/resynth_mi355x.h; there is no CPU
path, and importing this package fails when the HIP library has not been built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libresynth_mi355x.so")

# TImageFormat (lib/imageFormat.h:36-42)
T_RGB, T_RGBA, T_Gray, T_GrayA = 0, 1, 2, 3
# TImageSynthError (lib/engineParams.h:13-28)
IMAGE_SYNTH_SUCCESS = 0
IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT = 1
IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH = 2
IMAGE_SYNTH_ERROR_PATCH_SIZE_EXCEEDED = 3
IMAGE_SYNTH_ERROR_MATCH_CONTEXT_TYPE_RANGE = 4
IMAGE_SYNTH_ERROR_EMPTY_TARGET = 5
IMAGE_SYNTH_ERROR_EMPTY_CORPUS = 6
# Return codes beyond TImageSynthError (include/resynth_mi355x.h)
RESYNTH_ERROR_INVALID_ARGUMENT = -1
RESYNTH_ERROR_DEVICE = 100
RESYNTH_ERROR_TOO_LARGE = 101
RESYNTH_ERROR_PLUGIN = 102
# limits of resynth_image_synth_device_noalloc
NOALLOC_MAX_PIXELS = 1 << 20
NOALLOC_SCRATCH_BYTES = 192 << 20

MODE_SERIAL, MODE_BATCHED = 0, 1
PRNG_GRAND, PRNG_LIBC = 0, 1


class ImageBuffer(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_uint), ("height", ctypes.c_uint),
                ("rowBytes", ctypes.c_size_t)]


class TImageSynthParameters(ctypes.Structure):
    _fields_ = [("isMakeSeamlesslyTileableHorizontally", ctypes.c_int),
                ("isMakeSeamlesslyTileableVertically", ctypes.c_int),
                ("matchContextType", ctypes.c_int), ("mapWeight", ctypes.c_double),
                ("sensitivityToOutliers", ctypes.c_double), ("patchSize", ctypes.c_uint),
                ("maxProbeCount", ctypes.c_uint)]


class TFormatIndices(ctypes.Structure):
    _fields_ = [("colorEndBip", ctypes.c_uint8), ("alpha_bip", ctypes.c_uint8),
                ("map_start_bip", ctypes.c_uint8), ("map_end_bip", ctypes.c_uint8),
                ("img_match_bpp", ctypes.c_uint8), ("map_match_bpp", ctypes.c_uint8),
                ("total_bpp", ctypes.c_uint8), ("isAlphaTarget", ctypes.c_int), ("isAlphaSource", ctypes.c_int)]


class GArray(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("len", ctypes.c_uint)]


class Map(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint), ("height", ctypes.c_uint), ("depth", ctypes.c_uint),
                ("data", ctypes.POINTER(GArray))]


class ResynthOptions(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("prng", ctypes.c_int), ("seed", ctypes.c_uint), ("batch", ctypes.c_uint),
                ("timing", ctypes.c_int), ("model", ctypes.c_void_p), ("learnedProbes", ctypes.c_uint),
                ("learnedPasses", ctypes.c_uint), ("learnedClusters", ctypes.c_uint),
                ("learnedProbesLater", ctypes.c_uint), ("learnedRefresh", ctypes.c_uint),
                ("maxPasses", ctypes.c_uint), ("rampPermille", ctypes.c_uint), ("lanesPerTarget", ctypes.c_uint),
                ("firstChunk", ctypes.c_uint)]


class ResynthStats(ctypes.Structure):
    _fields_ = [("nTargets", ctypes.c_uint), ("nCorpus", ctypes.c_uint), ("nOffsets", ctypes.c_uint),
                ("passes", ctypes.c_uint), ("passSize", ctypes.c_uint * 6), ("betters", ctypes.c_uint * 6),
                ("visits", ctypes.c_ulonglong), ("candidates", ctypes.c_ulonglong),
                ("patchTerms", ctypes.c_ulonglong), ("meanError", ctypes.c_double), ("launches", ctypes.c_uint),
                ("synthKernelMs", ctypes.c_double), ("prepMs", ctypes.c_double), ("synthMs", ctypes.c_double),
                ("totalMs", ctypes.c_double), ("pixelWordBytes", ctypes.c_uint),
                ("windowLaunches", ctypes.c_uint), ("windowWaitsExpired", ctypes.c_uint),
                ("waveLaunches", ctypes.c_uint), ("waveKernelMs", ctypes.c_double),
                ("wavePatchTerms", ctypes.c_ulonglong), ("waveScanned", ctypes.c_ulonglong),
                ("waveStaged", ctypes.c_ulonglong), ("waveVisits", ctypes.c_ulonglong),
                ("learnedLaunches", ctypes.c_uint),
                ("learnedKeys", ctypes.c_uint), ("learnedClusters", ctypes.c_uint),
                ("learnedIndexMs", ctypes.c_double), ("learnedProposeMs", ctypes.c_double),
                ("learnedKeysScored", ctypes.c_ulonglong), ("scratchBytes", ctypes.c_ulonglong)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["passSize"] = list(self.passSize)[: self.passes]
        d["betters"] = list(self.betters)[: self.passes]
        return d


class ResynthDrawable(ctypes.Structure):
    _fields_ = [("pixels", ctypes.c_void_p), ("width", ctypes.c_uint), ("height", ctypes.c_uint),
                ("rowBytes", ctypes.c_size_t), ("colorChannels", ctypes.c_uint), ("hasAlpha", ctypes.c_int),
                ("selection", ctypes.c_void_p), ("selectionRowBytes", ctypes.c_size_t)]


class TGimpAdapterParameters(ctypes.Structure):
    _fields_ = [("h_tile", ctypes.c_int), ("v_tile", ctypes.c_int), ("use_border", ctypes.c_int),
                ("corpus_id", ctypes.c_int), ("input_map_id", ctypes.c_int), ("output_map_id", ctypes.c_int),
                ("map_weight", ctypes.c_double), ("autism", ctypes.c_double), ("neighbours", ctypes.c_int),
                ("trys", ctypes.c_int)]


class ResynthRequest(ctypes.Structure):
    _fields_ = [("image", ImageBuffer), ("mask", ImageBuffer), ("format", ctypes.c_int),
                ("params", TImageSynthParameters), ("error", ctypes.c_int), ("nTargets", ctypes.c_uint),
                ("passes", ctypes.c_uint), ("meanError", ctypes.c_double), ("candidates", ctypes.c_ulonglong)]


class ResynthKernelConfig(ctypes.Structure):
    _fields_ = [("wavesPerBlock", ctypes.c_int), ("firstGroup", ctypes.c_int), ("laterGroup", ctypes.c_int),
                ("lanesPerTarget", ctypes.c_int)]


class ResynthBatchStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint), ("targets", ctypes.c_ulonglong), ("config", ResynthKernelConfig),
                ("candidates", ctypes.c_ulonglong), ("patchTerms", ctypes.c_ulonglong), ("prepMs", ctypes.c_double),
                ("synthKernelMs", ctypes.c_double), ("totalMs", ctypes.c_double),
                ("graphLaunches", ctypes.c_uint), ("waitsExpired", ctypes.c_uint),
                ("graphInstantiations", ctypes.c_uint)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        c = self.config
        d["config"] = (c.wavesPerBlock, c.lanesPerTarget)
        return d


PROGRESS = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_void_p)

# Every function declared in include/resynth_mi355x.h, with its ctypes signature.
_P = ctypes.POINTER
SIGNATURES = {
    "imageSynth": (ctypes.c_int, [_P(ImageBuffer), _P(ImageBuffer), ctypes.c_int, _P(TImageSynthParameters),
                                  PROGRESS, ctypes.c_void_p, _P(ctypes.c_int)]),
    "imageSynth2": (ctypes.c_int, [_P(ImageBuffer), _P(ImageBuffer), _P(ImageBuffer), ctypes.c_int,
                                   _P(TImageSynthParameters), PROGRESS, ctypes.c_void_p, _P(ctypes.c_int)]),
    "engine": (ctypes.c_int, [TImageSynthParameters, _P(TFormatIndices), _P(Map), _P(Map), PROGRESS,
                              ctypes.c_void_p, _P(ctypes.c_int)]),
    "setDefaultParams": (None, [_P(TImageSynthParameters)]),
    "countPixelelsPerPixelForFormat": (ctypes.c_uint, [ctypes.c_int]),
    "prepareImageFormatIndicesFromFormatType": (ctypes.c_int, [_P(TFormatIndices), ctypes.c_int]),
    "prepareImageFormatIndices": (None, [_P(TFormatIndices), ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]),
    "prepareDefaultFormatIndices": (None, [_P(TFormatIndices)]),
    "free_map": (None, [_P(Map)]),
    "new_pixmap": (None, [_P(Map), ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]),
    "new_bytemap": (None, [_P(Map), ctypes.c_uint, ctypes.c_uint]),
    "new_intmap": (None, [_P(Map), ctypes.c_uint, ctypes.c_uint]),
    "new_coordmap": (None, [_P(Map), ctypes.c_uint, ctypes.c_uint]),
    "set_bytemap": (None, [_P(Map), ctypes.c_ubyte]),
    "invert_bytemap": (None, [_P(Map)]),
    "interleave_mask": (None, [_P(Map), _P(Map)]),
    "resynth_set_default_mode": (None, [ctypes.c_int]),
    "resynth_set_device_memory_limit": (None, [ctypes.c_size_t]),
    "resynth_default_options": (None, [_P(ResynthOptions)]),
    "imageSynthEx": (ctypes.c_int, [_P(ImageBuffer), _P(ImageBuffer), _P(ImageBuffer), ctypes.c_int,
                                    _P(TImageSynthParameters), PROGRESS, ctypes.c_void_p, _P(ctypes.c_int),
                                    _P(ResynthOptions), _P(ResynthStats)]),
    "engineEx": (ctypes.c_int, [TImageSynthParameters, _P(TFormatIndices), _P(Map), _P(Map), PROGRESS,
                                ctypes.c_void_p, _P(ctypes.c_int), _P(ResynthOptions), _P(ResynthStats)]),
    "resynth_image_synth_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_size_t,
                                                  ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                                  ctypes.c_int, _P(TImageSynthParameters), PROGRESS,
                                                  ctypes.c_void_p, _P(ctypes.c_int), _P(ResynthOptions),
                                                  _P(ResynthStats)]),
    "resynth_image_synth_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint,
                                                          ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                                          ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                                          _P(TImageSynthParameters), PROGRESS, ctypes.c_void_p,
                                                          _P(ctypes.c_int), _P(ResynthOptions), _P(ResynthStats)]),
    "resynth_image_synth_batch": (ctypes.c_int, [_P(ResynthRequest), ctypes.c_uint, ctypes.c_int,
                                                 _P(ResynthOptions), _P(ResynthKernelConfig),
                                                 _P(ResynthBatchStats)]),
    "resynth_autotune_batch": (ctypes.c_int, [_P(ResynthRequest), ctypes.c_uint, ctypes.c_int,
                                              _P(ResynthKernelConfig), ctypes.c_void_p]),
    "resynth_struct_sizes": (ctypes.c_uint, [ctypes.c_void_p, ctypes.c_uint]),
    "resynth_model_param_count": (ctypes.c_uint, []),
    "resynth_model_create": (ctypes.c_void_p, [ctypes.c_uint]),
    "resynth_model_free": (None, [ctypes.c_void_p]),
    "resynth_model_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_model_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_model_embed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p]),
    "resynth_model_infonce": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
                                             ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_model_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
                                           ctypes.c_uint, ctypes.c_uint, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_uint, ctypes.c_void_p]),
    "resynth_model_log_pairs": (ctypes.c_int, [ctypes.c_void_p, _P(ImageBuffer), _P(ImageBuffer), ctypes.c_int,
                                               _P(TImageSynthParameters), ctypes.c_uint, ctypes.c_uint]),
    "resynth_model_pair_count": (ctypes.c_uint, [ctypes.c_void_p]),
    "resynth_model_clear_pairs": (None, [ctypes.c_void_p]),
    "resynth_model_fit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_uint, ctypes.c_void_p]),
    "resynth_gemm_bf16": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_sorted_offsets": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "resynth_sorted_offsets_within": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p,
                                                     ctypes.POINTER(ctypes.c_ulonglong)]),
    "resynth_learn_route": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_uint, ctypes.c_uint, ctypes.c_int, ctypes.c_void_p]),
    "resynth_learn_fine_top16": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_uint,
                                                ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_score_frozen": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, _P(TFormatIndices),
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_prng_sample": (None, [ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p]),
    "resynth_prng_ranges": (None, [ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_metric_tables": (None, [ctypes.c_double, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]),
    "resynth_order_targets": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint]),
    "resynth_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, _P(ctypes.c_int)]),
    "resynth_version": (ctypes.c_char_p, []),
    "resynth_plugin_run": (ctypes.c_int, [_P(TGimpAdapterParameters), _P(ResynthDrawable), _P(ResynthDrawable),
                                          _P(ResynthDrawable), _P(ResynthDrawable), PROGRESS, ctypes.c_void_p,
                                          _P(ctypes.c_int), _P(ctypes.c_char_p)]),
    "resynth_heal_selection": (ctypes.c_int, [_P(ResynthDrawable), ctypes.c_int, ctypes.c_int, ctypes.c_int, PROGRESS,
                                              ctypes.c_void_p, _P(ctypes.c_int), _P(ctypes.c_char_p)]),
}

_lib = None


def lib():
    """The loaded HIP library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("resynthesizer_amd: %s is not built (run __graft_entry__.build() or "
                               "python -m resynthesizer_amd.build); there is no CPU fallback" % LIB_PATH)
        # One HIP runtime per process: torch ships its own libamdhip64.so.7 under the same
        # soname, and torch fails to initialize if another copy was loaded first.  Loading
        # torch first makes this library bind to torch's copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def default_params(**kw):
    p = TImageSynthParameters()
    lib().setDefaultParams(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, _PARAM_ALIASES.get(k, k), v)
    return p


_PARAM_ALIASES = {"tileH": "isMakeSeamlesslyTileableHorizontally", "tileV": "isMakeSeamlesslyTileableVertically",
                  "sensitivity": "sensitivityToOutliers"}


def params_from_dict(d):
    return default_params(**d)


def options(mode=MODE_BATCHED, prng=PRNG_GRAND, seed=0, batch=0, timing=False, model=None, learned_probes=0,
            learned_passes=0, learned_clusters=0, learned_probes_later=0, learned_refresh=0, max_passes=0,
            ramp_permille=0, lanes_per_target=0, first_chunk=0):
    return ResynthOptions(mode, prng, seed, batch, int(bool(timing)), model.h if model is not None else None,
                          learned_probes, learned_passes, learned_clusters, learned_probes_later, learned_refresh,
                          max_passes, ramp_permille, lanes_per_target, first_chunk)


def format_indices(fmt):
    ix = TFormatIndices()
    err = lib().prepareImageFormatIndicesFromFormatType(ctypes.byref(ix), fmt)
    return err, ix


def prepare_indices(colors, maps, alpha_t, alpha_s, is_map):
    ix = TFormatIndices()
    lib().prepareImageFormatIndices(ctypes.byref(ix), colors, maps, int(alpha_t), int(alpha_s), int(is_map))
    return ix


class Result:
    def __init__(self, err, stats, progress):
        self.err = err
        self.stats = stats.as_dict()
        self.progress = progress

    def __getattr__(self, k):
        try:
            return self.__dict__["stats"][k]
        except KeyError:
            raise AttributeError(k)


def image_synth(image, mask, fmt, params=None, mode=MODE_BATCHED, prng=PRNG_GRAND, cancel=0, mask2=None,
                width=None, mask_width=None, mask2_width=None, seed=0, batch=0, timing=False, model=None,
                learned_probes=0, learned_passes=0, learned_clusters=0, learned_probes_later=0, learned_refresh=0,
                max_passes=0, ramp_permille=0, lanes_per_target=0, first_chunk=0):
    """imageSynth / imageSynth2 (lib/imageSynth.c:62-217) on the GPU.

    image: uint8 (H, W, C) or a padded row buffer (H, rowBytes) with `width` given;
    mask (and mask2): uint8 (H, W) or padded rows with the width given.
    Returns (Result, image_out)."""
    img = np.array(image, dtype=np.uint8, copy=True, order="C")
    bpp = int(lib().countPixelelsPerPixelForFormat(fmt)) or 1
    if img.ndim == 3:
        h, w = img.shape[0], img.shape[1]
        row = img.shape[1] * img.shape[2]
    else:
        h, row = img.shape
        w = width if width is not None else row // bpp
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    ib = ImageBuffer(img.ctypes.data, w, h, row)
    mb = ImageBuffer(m.ctypes.data, mask_width if mask_width is not None else m.shape[1], m.shape[0], m.shape[1])
    mb2p = None
    if mask2 is not None:
        m2 = np.ascontiguousarray(mask2, dtype=np.uint8)
        mb2 = ImageBuffer(m2.ctypes.data, mask2_width if mask2_width is not None else m2.shape[1], m2.shape[0],
                          m2.shape[1])
        mb2p = ctypes.byref(mb2)
    log = []
    cb = PROGRESS(lambda pct, ctx: log.append(pct))
    cf = ctypes.c_int(cancel)
    st = ResynthStats()
    opt = options(mode, prng, seed, batch, timing, model, learned_probes, learned_passes, learned_clusters,
                  learned_probes_later, learned_refresh, max_passes, ramp_permille, lanes_per_target, first_chunk)
    pp = ctypes.byref(params) if params is not None else None
    err = lib().imageSynthEx(ctypes.byref(ib), ctypes.byref(mb), mb2p, fmt, pp, cb, None, ctypes.byref(cf),
                             ctypes.byref(opt), ctypes.byref(st))
    return Result(err, st, log), img


def engine(params, ix, target_pixmap, corpus_pixmap, mode=MODE_BATCHED, prng=PRNG_GRAND, cancel=0, seed=0,
           batch=0, timing=False, max_passes=0, ramp_permille=0):
    """engine() (lib/engine.c:539-690) on engine-layout pixmaps uint8 (H, W, depth).
    Returns (Result, target_out)."""
    t = np.array(target_pixmap, dtype=np.uint8, copy=True, order="C")
    c = np.ascontiguousarray(corpus_pixmap, dtype=np.uint8)
    ta = GArray(t.ctypes.data, 0)
    ca = GArray(c.ctypes.data, 0)
    tm = Map(t.shape[1], t.shape[0], t.shape[2], ctypes.pointer(ta))
    cm = Map(c.shape[1], c.shape[0], c.shape[2], ctypes.pointer(ca))
    log = []
    cb = PROGRESS(lambda pct, ctx: log.append(pct))
    cf = ctypes.c_int(cancel)
    st = ResynthStats()
    opt = options(mode, prng, seed, batch, timing, max_passes=max_passes, ramp_permille=ramp_permille)
    err = lib().engineEx(params, ctypes.byref(ix), ctypes.byref(tm), ctypes.byref(cm), cb, None, ctypes.byref(cf),
                         ctypes.byref(opt), ctypes.byref(st))
    return Result(err, st, log), t


def image_synth_device(d_image, width, height, row_bytes, d_mask, mask_row_bytes, fmt, params=None,
                       mode=MODE_BATCHED, prng=PRNG_GRAND, d_mask2=None, mask2_row_bytes=0, seed=0, batch=0,
                       timing=False, cancel=0, model=None, learned_probes=0, learned_passes=0,
                       learned_clusters=0, learned_probes_later=0, learned_refresh=0, max_passes=0,
                       ramp_permille=0, lanes_per_target=0, first_chunk=0):
    """imageSynth on device-resident buffers (integer device addresses, e.g. torch data_ptr()).
    Synthesizes d_image in place.  Returns a Result."""
    log = []
    cb = PROGRESS(lambda pct, ctx: log.append(pct))
    cf = ctypes.c_int(cancel)
    st = ResynthStats()
    opt = options(mode, prng, seed, batch, timing, model, learned_probes, learned_passes, learned_clusters,
                  learned_probes_later, learned_refresh, max_passes, ramp_permille, lanes_per_target, first_chunk)
    pp = ctypes.byref(params) if params is not None else None
    err = lib().resynth_image_synth_device(d_image, width, height, row_bytes, d_mask, mask_row_bytes, d_mask2,
                                           mask2_row_bytes, fmt, pp, cb, None, ctypes.byref(cf), ctypes.byref(opt),
                                           ctypes.byref(st))
    return Result(err, st, log)


def _requests(images, masks, fmts, params, device):
    n = len(images)
    reqs = (ResynthRequest * n)()
    keep = []
    for i in range(n):
        if device:
            img, m = images[i], masks[i]     # (data_ptr, width, height, rowBytes) tuples
            reqs[i].image = ImageBuffer(*img)
            reqs[i].mask = ImageBuffer(*m)
        else:
            img = np.ascontiguousarray(images[i], dtype=np.uint8)
            m = np.ascontiguousarray(masks[i], dtype=np.uint8)
            keep += [img, m]
            reqs[i].image = ImageBuffer(img.ctypes.data, img.shape[1], img.shape[0], img.strides[0])
            reqs[i].mask = ImageBuffer(m.ctypes.data, m.shape[1], m.shape[0], m.strides[0])
        reqs[i].format = fmts[i]
        reqs[i].params = params[i] if params[i] is not None else default_params()
    return reqs, keep


def image_synth_batch(images, masks, fmts, params, device=False, config=None, seed=0, batch=0):
    """Many imageSynth requests in one multi-request launch.  Host mode: images are uint8
    (H, W, C) arrays (copied, the copies synthesized); device mode: images and masks are
    (data_ptr, width, height, rowBytes) tuples synthesized in place.  config: None or
    (wavesPerBlock, lanesPerTarget).  Returns (rc, per-request dicts, stats dict,
    output arrays in host mode)."""
    if not device:
        images = [np.array(a, dtype=np.uint8, copy=True, order="C") for a in images]
    reqs, keep = _requests(images, masks, fmts, params, device)
    cfg = ResynthKernelConfig(config[0], 2, 4, config[1]) if config else None
    st = ResynthBatchStats()
    opt = options(MODE_BATCHED, PRNG_GRAND, seed, batch, False)
    rc = lib().resynth_image_synth_batch(reqs, len(images), int(device), ctypes.byref(opt),
                                         ctypes.byref(cfg) if cfg else None, ctypes.byref(st))
    out = [dict(error=r.error, nTargets=r.nTargets, passes=r.passes, meanError=r.meanError,
                candidates=r.candidates) for r in reqs]
    if not device:
        out_imgs = keep[0::2]
        return rc, out, st.as_dict(), out_imgs
    return rc, out, st.as_dict(), None


def autotune_batch(images, masks, fmts, params, device=False):
    """Runs the autotuner on a sample of requests.  Returns (rc, best (waves per block, lanes
    per target), the 9 median timings in ms, waves {2, 4, 8} x lanes {16, 32, 64})."""
    reqs, keep = _requests(images, masks, fmts, params, device)
    best = ResynthKernelConfig()
    ms = np.zeros(9, np.float64)
    rc = lib().resynth_autotune_batch(reqs, len(images), int(device), ctypes.byref(best), ms.ctypes.data)
    return rc, (best.wavesPerBlock, best.lanesPerTarget), ms


class Model:
    """The learned-proposal network on the device (resynth_model_*)."""

    def __init__(self, seed=1):
        self.h = lib().resynth_model_create(seed)
        if not self.h:
            raise RuntimeError("resynth_model_create failed (no device)")
        self.n = lib().resynth_model_param_count()

    def close(self):
        if self.h:
            lib().resynth_model_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.n, np.float32)
        assert lib().resynth_model_get_params(self.h, out.ctypes.data) == 0
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n
        assert lib().resynth_model_set_params(self.h, p.ctypes.data) == 0

    def embed(self, x_bits):
        x = np.ascontiguousarray(x_bits, dtype=np.uint16)
        out = np.zeros((x.shape[0], 64), np.float32)
        assert lib().resynth_model_embed(self.h, x.ctypes.data, x.shape[0], out.ctypes.data) == 0
        return out

    def infonce(self, xq_bits, xk_bits, tau):
        xq = np.ascontiguousarray(xq_bits, dtype=np.uint16)
        xk = np.ascontiguousarray(xk_bits, dtype=np.uint16)
        loss = np.zeros(1, np.float32)
        grad = np.zeros(self.n, np.float32)
        rc = lib().resynth_model_infonce(self.h, xq.ctypes.data, xk.ctypes.data, xq.shape[0], tau, loss.ctypes.data,
                                         grad.ctypes.data)
        if rc:
            raise ValueError("resynth_model_infonce rejected its arguments")
        return float(loss[0]), grad

    def train(self, xq_bits, xk_bits, steps, batch=256, lr=1e-3, tau=0.1, seed=1):
        xq = np.ascontiguousarray(xq_bits, dtype=np.uint16)
        xk = np.ascontiguousarray(xk_bits, dtype=np.uint16)
        losses = np.zeros(steps, np.float32)
        rc = lib().resynth_model_train(self.h, xq.ctypes.data, xk.ctypes.data, xq.shape[0], steps, batch, lr, tau,
                                       seed, losses.ctypes.data)
        if rc:
            raise ValueError("resynth_model_train rejected its arguments")
        return losses


def _model_log_pairs(self, image, mask, fmt, params=None, max_pairs=4096, seed=1):
    """Runs the plain batched engine on a copy of image and logs up to max_pairs training
    pairs.  Returns (error code, synthesized copy)."""
    img = np.array(image, dtype=np.uint8, copy=True, order="C")
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    ib = ImageBuffer(img.ctypes.data, img.shape[1], img.shape[0], img.strides[0])
    mb = ImageBuffer(m.ctypes.data, m.shape[1], m.shape[0], m.strides[0])
    p = params if params is not None else default_params()
    err = lib().resynth_model_log_pairs(self.h, ctypes.byref(ib), ctypes.byref(mb), fmt, ctypes.byref(p), max_pairs,
                                        seed)
    return err, img


def _model_fit(self, steps, batch=1024, lr=1e-3, tau=0.1, seed=1):
    losses = np.zeros(steps, np.float32)
    rc = lib().resynth_model_fit(self.h, steps, batch, lr, tau, seed, losses.ctypes.data)
    if rc:
        raise ValueError("resynth_model_fit failed (no pairs or bad batch)")
    return losses


Model.log_pairs = _model_log_pairs
Model.fit = _model_fit
Model.pair_count = lambda self: lib().resynth_model_pair_count(self.h)
Model.clear_pairs = lambda self: lib().resynth_model_clear_pairs(self.h)


def gemm_bf16(a_bits, b_bits, ta=False, tb=False):
    """MFMA GEMM on host arrays of bf16 bit patterns: op(A) op(B) in fp32."""
    A = np.ascontiguousarray(a_bits, dtype=np.uint16)
    B = np.ascontiguousarray(b_bits, dtype=np.uint16)
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    C = np.zeros((M, N), np.float32)
    rc = lib().resynth_gemm_bf16(int(ta), int(tb), M, N, K, A.ctypes.data, B.ctypes.data, C.ctypes.data)
    if rc:
        raise ValueError("resynth_gemm_bf16 rejected its arguments")
    return C


def sorted_offsets(w, h, max_d2=None):
    """The engine's device-built sorted offset table for a w x h offset range, (N, 2) int32;
    with max_d2, its prefix of offsets with x^2 + y^2 <= max_d2."""
    out = np.zeros(((2 * w - 1) * (2 * h - 1), 2), np.int32)
    if max_d2 is None:
        if lib().resynth_sorted_offsets(w, h, out.ctypes.data):
            raise ValueError("resynth_sorted_offsets rejected its arguments")
        return out
    n = ctypes.c_ulonglong(0)
    if lib().resynth_sorted_offsets_within(w, h, int(max_d2), out.ctypes.data, ctypes.byref(n)):
        raise ValueError("resynth_sorted_offsets_within rejected its arguments")
    return out[:n.value]


def route(queries_bits, supers, sub_centroids_bits, S, Q, P):
    """The learned index's query routing on host arrays: for each query, the ids s*Q + c of
    the P best sub-centroids of its super-cluster s, best first.  Returns int32 [nq x P]."""
    q = np.ascontiguousarray(queries_bits, dtype=np.uint16)
    sup = np.ascontiguousarray(supers, dtype=np.int32)
    sub = np.ascontiguousarray(sub_centroids_bits, dtype=np.uint16)
    out = np.zeros((q.shape[0], P), np.int32)
    rc = lib().resynth_learn_route(q.ctypes.data, q.shape[0], sup.ctypes.data, sub.ctypes.data, S, Q, P,
                                   out.ctypes.data)
    if rc:
        raise ValueError("resynth_learn_route rejected its arguments")
    return out


def fine_top16(queries_bits, clusters, keys_bits, start, key_pts):
    """The learned index's fine search on host arrays: queries and keys are bf16 bit patterns
    [n x 64], clusters int32 [nq x P], start uint32 [nClusters + 1], key_pts int32 [nKeys x 2].
    Returns int32 [nq x 16 x 2]: 2 * key_pts of each query's top-16 keys, best first."""
    q = np.ascontiguousarray(queries_bits, dtype=np.uint16)
    cl = np.ascontiguousarray(clusters, dtype=np.int32)
    if cl.ndim == 1:
        cl = cl[:, None]
    k = np.ascontiguousarray(keys_bits, dtype=np.uint16)
    st = np.ascontiguousarray(start, dtype=np.uint32)
    pts = np.ascontiguousarray(key_pts, dtype=np.int32)
    out = np.zeros((q.shape[0], 16, 2), np.int32)
    rc = lib().resynth_learn_fine_top16(q.ctypes.data, q.shape[0], cl.ctypes.data, cl.shape[1], k.ctypes.data,
                                        k.shape[0], st.ctypes.data, len(st) - 1, pts.ctypes.data, out.ctypes.data)
    if rc:
        raise ValueError("resynth_learn_fine_top16 rejected its arguments")
    return out


def struct_sizes():
    """sizeof of the extension structs as the library was compiled, in header order."""
    out = np.zeros(16, np.uint64)
    n = lib().resynth_struct_sizes(out.ctypes.data, 16)
    return [int(v) for v in out[:n]]


def score_frozen(corpus_pixmap, ix, T, Tm, nb_counts, nb_offsets, nb_pixels, cand_counts, cands):
    """GPU frozen-state scorer.  corpus_pixmap uint8 (H, W, depth); nb_offsets int32 (P, 64, 2);
    nb_pixels uint8 (P, 64, 8); cands int32 (P, M, 2).  Returns (winner, bestSum, stopped)."""
    cp = np.ascontiguousarray(corpus_pixmap, dtype=np.uint8)
    P = len(nb_counts)
    nbc = np.ascontiguousarray(nb_counts, dtype=np.uint32)
    nbo = np.ascontiguousarray(nb_offsets, dtype=np.int32)
    nbp = np.ascontiguousarray(nb_pixels, dtype=np.uint8)
    cc = np.ascontiguousarray(cand_counts, dtype=np.uint32)
    cd = np.ascontiguousarray(cands, dtype=np.int32)
    T = np.ascontiguousarray(T, dtype=np.uint16)
    Tm = np.ascontiguousarray(Tm, dtype=np.uint32)
    win = np.zeros(P, np.int32)
    best = np.zeros(P, np.uint32)
    stop = np.zeros(P, np.int32)
    rc = lib().resynth_score_frozen(cp.ctypes.data, cp.shape[1], cp.shape[0], ctypes.byref(ix), T.ctypes.data,
                                    Tm.ctypes.data, P, nbc.ctypes.data, nbo.ctypes.data, nbp.ctypes.data,
                                    cc.ctypes.data, cd.shape[1], cd.ctypes.data, win.ctypes.data, best.ctypes.data,
                                    stop.ctypes.data)
    if rc != 0:
        raise ValueError("resynth_score_frozen rejected its arguments")
    return win, best, stop


def prng_sample(kind, seed, n):
    out = np.zeros(n, np.uint32)
    lib().resynth_prng_sample(kind, seed, n, out.ctypes.data)
    return out


def prng_ranges(kind, seed, dists):
    d = np.ascontiguousarray(dists, dtype=np.uint32)
    out = np.zeros(len(d), np.uint32)
    lib().resynth_prng_ranges(kind, seed, len(d), d.ctypes.data, out.ctypes.data)
    return out


def metric_tables(sensitivity=0.117, map_weight=0.5):
    T = np.zeros(512, np.uint16)
    Tm = np.zeros(512, np.uint32)
    lib().resynth_metric_tables(sensitivity, map_weight, T.ctypes.data, Tm.ctypes.data)
    return T, Tm


def order_targets(match_type, points, prng=PRNG_GRAND):
    pts = np.ascontiguousarray(points, dtype=np.int32).copy()
    err = lib().resynth_order_targets(match_type, prng, pts.ctypes.data, len(pts))
    return err, pts


def set_device_memory_limit(nbytes):
    """Caps the device memory of the library's engine buffers (0 = no cap)."""
    lib().resynth_set_device_memory_limit(int(nbytes))


def device_info():
    name = ctypes.create_string_buffer(64)
    cus = ctypes.c_int(0)
    rc = lib().resynth_device_info(name, 64, ctypes.byref(cus))
    return rc, name.value.decode(), cus.value


# ---------------------------------------------------------------- GIMP plug-in shim, heal selection

def _drawable(pixels, selection=None):
    """ResynthDrawable over uint8 (H, W, C) pixels, C = 1 gray, 2 gray + alpha, 3 RGB, 4 RGBA,
    and an optional uint8 (H, W) selection.  Returns (struct, keep-alive list)."""
    if pixels.ndim == 2:
        pixels = pixels[..., None]
    h, w, c = pixels.shape
    if c not in (1, 2, 3, 4):
        raise ValueError("drawables have 1 to 4 channels")
    d = ResynthDrawable(pixels.ctypes.data, w, h, pixels.strides[0], 3 if c >= 3 else 1, int(c in (2, 4)), None, 0)
    keep = [pixels]
    if selection is not None:
        m = np.ascontiguousarray(selection, dtype=np.uint8)
        if m.shape != (h, w):
            raise ValueError("the selection must cover the drawable")
        d.selection, d.selectionRowBytes = m.ctypes.data, m.strides[0]
        keep.append(m)
    return d, keep


def plugin_run(image, selection=None, corpus=None, corpus_selection=None, input_map=None, output_map=None,
               h_tile=0, v_tile=0, use_border=1, map_weight=0.5, autism=0.117, neighbours=30, trys=200, cancel=0):
    """The engine plug-in (resynth_plugin_run) on uint8 (H, W, C) drawables.  corpus None: the
    image itself.  Returns (error, message, image_out, progress)."""
    img = np.array(image, dtype=np.uint8, copy=True, order="C")
    d, keep = _drawable(img, selection)
    cp = None
    if corpus is not None:
        c, k = _drawable(np.ascontiguousarray(corpus, dtype=np.uint8), corpus_selection)
        cp, keep = ctypes.byref(c), keep + k
    maps = []
    for mimg in (input_map, output_map):
        if mimg is None:
            maps.append(None)
        else:
            md, k = _drawable(np.ascontiguousarray(mimg, dtype=np.uint8))
            maps.append(ctypes.byref(md))
            keep += k
    p = TGimpAdapterParameters(h_tile, v_tile, use_border, -1, -1, -1, map_weight, autism, neighbours, trys)
    log = []
    cb = PROGRESS(lambda pct, ctx: log.append(pct))
    cf = ctypes.c_int(cancel)
    msg = ctypes.c_char_p()
    err = lib().resynth_plugin_run(ctypes.byref(p), ctypes.byref(d), cp, maps[0], maps[1], cb, None,
                                   ctypes.byref(cf), ctypes.byref(msg))
    return err, (msg.value.decode() if msg.value else None), img.reshape(np.shape(image)), log


def heal_selection(image, selection, sampling_radius=50, direction=0, order=0, cancel=0):
    """Heal selection (resynth_heal_selection) on a uint8 (H, W, C) drawable and its uint8
    (H, W) selection.  Returns (error, message, image_out, progress)."""
    img = np.array(image, dtype=np.uint8, copy=True, order="C")
    d, keep = _drawable(img, selection)
    log = []
    cb = PROGRESS(lambda pct, ctx: log.append(pct))
    cf = ctypes.c_int(cancel)
    msg = ctypes.c_char_p()
    err = lib().resynth_heal_selection(ctypes.byref(d), sampling_radius, direction, order, cb, None, ctypes.byref(cf),
                                       ctypes.byref(msg))
    return err, (msg.value.decode() if msg.value else None), img.reshape(np.shape(image)), log
