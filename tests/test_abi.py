"""CPU checks of libresynth_mi355x.so: it loads, exports every function declared in
include/resynth_mi355x.h, keeps the reference's struct layouts, and its host-side pieces
(PRNG streams, metric tables, target ordering) match the golden vectors and the oracle.
No GPU work is launched here."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_py as O
import resynthesizer_amd as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "for", "while", "return", "sizeof"):
                names.append(name)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert "imageSynth" in names and "engine" in names and len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = R.lib()
    for n in names:
        assert hasattr(L, n)
    assert set(R.SIGNATURES) == set(names)


def test_struct_layouts_match_reference():
    # lib/imageBuffer.h, lib/engineParams.h, lib/imageFormatIndicies.h, lib/map.h (x86-64 LP64)
    assert ctypes.sizeof(R.ImageBuffer) == 24 and R.ImageBuffer.rowBytes.offset == 16
    assert ctypes.sizeof(R.TImageSynthParameters) == 40
    assert R.TImageSynthParameters.mapWeight.offset == 16 and R.TImageSynthParameters.patchSize.offset == 32
    assert ctypes.sizeof(R.TFormatIndices) == 16 and R.TFormatIndices.isAlphaTarget.offset == 8
    assert ctypes.sizeof(R.Map) == 24 and R.Map.data.offset == 16


def test_defaults_and_format_indices_match_oracle():
    p = R.default_params()
    assert (p.isMakeSeamlesslyTileableHorizontally, p.isMakeSeamlesslyTileableVertically, p.matchContextType,
            p.mapWeight, p.sensitivityToOutliers, p.patchSize, p.maxProbeCount) == (0, 0, 1, 0.5, 0.117, 30, 200)
    for fmt in range(4):
        e1, a = R.format_indices(fmt)
        e2, b = O.format_indices(fmt)
        assert e1 == e2 == 0
        assert bytes(a) == bytes(b)
        assert R.lib().countPixelelsPerPixelForFormat(fmt) == O.lib().oracle_count_pixelels(fmt)
    assert R.format_indices(4)[0] == 1
    for args in [(3, 3, 0, 0, 1), (3, 1, 1, 0, 1), (1, 1, 0, 1, 1), (3, 0, 1, 1, 0)]:
        assert bytes(R.prepare_indices(*args)) == bytes(O.prepare_indices(*args))
    ix = R.TFormatIndices()
    R.lib().prepareDefaultFormatIndices(ctypes.byref(ix))
    assert (ix.colorEndBip, ix.alpha_bip, ix.map_start_bip, ix.total_bpp, ix.isAlphaTarget) == (4, 4, 5, 5, 1)


def test_map_helpers():
    L = R.lib()
    m = R.Map()
    L.new_pixmap(ctypes.byref(m), 5, 3, 4)
    assert (m.width, m.height, m.depth) == (5, 3, 4)
    k = R.Map()
    L.new_bytemap(ctypes.byref(k), 5, 3)
    L.set_bytemap(ctypes.byref(k), 0x0F)
    L.invert_bytemap(ctypes.byref(k))
    L.interleave_mask(ctypes.byref(m), ctypes.byref(k))
    px = np.ctypeslib.as_array(ctypes.cast(m.data.contents.data, ctypes.POINTER(ctypes.c_uint8)), shape=(60,))
    assert list(px[::4]) == [0xF0] * 15 and px[1:4].tolist() == [0, 0, 0]
    L.free_map(ctypes.byref(k))
    L.free_map(ctypes.byref(m))
    assert not m.data


@pytest.fixture(scope="module")
def vectors():
    with open(os.path.join(ROOT, "tests", "golden", "vectors.json")) as f:
        return json.load(f)


def test_host_prng_matches_glib_and_glibc(vectors):
    for seed in (1198472, 1, 5, 4294967295):
        want = vectors["prng"]["grand_int_seed%d" % seed]
        assert R.prng_sample(R.PRNG_GRAND, seed, len(want)).tolist() == want
    assert R.prng_sample(R.PRNG_LIBC, 1198472, 1500).tolist() == vectors["prng"]["libc_rand_seed1198472"]
    assert R.prng_sample(R.PRNG_LIBC, 7, 400).tolist() == vectors["prng"]["libc_rand_seed7"]
    seq = [(b, e, v) for b, e, v in vectors["prng"]["grand_int_range_mixed"] if b == 0]
    # the engine only draws int_range(0, n); replay the mixed sequence's b == 0 prefix
    k = 0
    while k < len(vectors["prng"]["grand_int_range_mixed"]) and vectors["prng"]["grand_int_range_mixed"][k][0] == 0:
        k += 1
    pref = vectors["prng"]["grand_int_range_mixed"][:k]
    got = R.prng_ranges(R.PRNG_GRAND, 1198472, [e for _, e, _ in pref])
    assert got.tolist() == [v for _, _, v in pref]
    assert len(seq) > 0


def test_host_prng_ranges_match_oracle():
    rng = np.random.RandomState(1)
    dists = np.concatenate([[0, 1, 2, 3], rng.randint(1, 1 << 31, 500), rng.randint(1, 5000, 500)]).astype(np.uint32)
    assert R.prng_ranges(R.PRNG_GRAND, 1198472, dists).tolist() == O.grand_ranges(1198472, dists.tolist())
    assert R.prng_ranges(R.PRNG_LIBC, 1198472, dists).tolist() == O.libc_ranges(1198472, dists.tolist())


@pytest.mark.parametrize("sens,mw", [(0.117, 0.5), (0.05, 0.0), (0.3, 1.0), (0.117, 0.3977857), (1.0, 0.1)])
def test_metric_tables_match_oracle(sens, mw):
    T1, Tm1 = R.metric_tables(sens, mw)
    T2, Tm2 = O.quantize(sens, mw)
    assert np.array_equal(T1, T2) and np.array_equal(Tm1, Tm2)


@pytest.mark.parametrize("prng", [O.PRNG_GRAND, O.PRNG_LIBC])
@pytest.mark.parametrize("mtype", [0, 1])
def test_random_target_order_matches_oracle(mtype, prng):
    """Types 0 and 1 are a host shuffle (lib/orderTarget.h:36-47); types 2..8 sort on the
    device and are checked against GLib's own ordering by the GPU tests."""
    h, w = 40, 57
    y, x = np.mgrid[0:h, 0:w]
    sel = ((x - 30) / 17.0) ** 2 + ((y - 19) / 11.0) ** 2 < 1
    sel[2, 3] = True
    pts = np.stack([x[sel], y[sel]], 1).astype(np.int32)
    e1, a = R.order_targets(mtype, pts, prng)
    e2, b, _ = O.order_targets(mtype, pts, prng)
    assert e1 == e2 == 0 and np.array_equal(a, b)
    for single in ([[5, 5]], [[0, 0], [0, 1]]):
        e1, a = R.order_targets(mtype, np.array(single), prng)
        e2, b, _ = O.order_targets(mtype, np.array(single), prng)
        assert e1 == e2 == 0 and np.array_equal(a, b)
    assert R.order_targets(9, pts, prng)[0] == R.IMAGE_SYNTH_ERROR_MATCH_CONTEXT_TYPE_RANGE


def _has_device():
    import torch
    return torch.cuda.is_available()


def test_entry_points_return_an_error_without_a_device():
    """With no usable device every entry returns RESYNTH_ERROR_DEVICE instead of ending the
    process (the reference's callers expect an error code, lib/engineParams.h:13-28)."""
    if _has_device():
        pytest.skip("a device is present")
    img = np.zeros((8, 8, 3), np.uint8)
    mask = np.zeros((8, 8), np.uint8)
    mask[2:5, 2:5] = 255
    r, out = R.image_synth(img, mask, R.T_RGB, R.default_params())
    assert r.err == R.RESYNTH_ERROR_DEVICE and np.array_equal(out, img)
    tp = np.concatenate([mask[..., None], img], axis=2)
    r, _ = R.engine(R.default_params(), R.prepare_indices(3, 0, False, False, False), tp, tp)
    assert r.err == R.RESYNTH_ERROR_DEVICE
    assert R.order_targets(3, np.array([[1, 2], [3, 4]]))[0] == R.RESYNTH_ERROR_DEVICE
    assert R.lib().resynth_sorted_offsets(3, 3, np.zeros(50, np.int32).ctypes.data) == R.RESYNTH_ERROR_DEVICE


def _build_reference_api(tmp_path):
    exe = str(tmp_path / "reference_api")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c", "reference_api.c"), "-L", os.path.dirname(R.LIB_PATH),
                    "-lresynth_mi355x", "-Wl,-rpath," + os.path.dirname(R.LIB_PATH), "-o", exe], check=True)
    return exe


def test_reference_signature_caller_compiles_and_links(tmp_path):
    """A C program written against include/resynth_mi355x.h with the reference's names and
    signatures (as src/testSynth.c calls imageSynth) compiles as C99 and links to the
    library; without a device its calls return RESYNTH_ERROR_DEVICE."""
    exe = _build_reference_api(tmp_path)
    if not _has_device():
        r = subprocess.run([exe, "known"], capture_output=True, text=True)
        assert r.returncode == 1 and "(100)" in r.stdout


def test_extension_struct_layouts_match_library():
    """The ctypes mirrors of the extension structs have the library's sizes and offsets."""
    got = R.struct_sizes()
    want = [ctypes.sizeof(R.ResynthOptions), ctypes.sizeof(R.ResynthStats), ctypes.sizeof(R.ResynthRequest),
            ctypes.sizeof(R.ResynthKernelConfig), ctypes.sizeof(R.ResynthBatchStats),
            R.ResynthRequest.params.offset, R.ResynthRequest.error.offset, R.ResynthBatchStats.prepMs.offset]
    assert got == want


def test_plugin_shim_checks_and_messages():
    """resynth_plugin_run / resynth_heal_selection make the engine plug-in's and the heal
    script's checks before any device work, with the reference's messages
    (src/resynthesizer/resynthesizer.c:357-421, PluginScripts/plugin-heal-selection.py:41-43)."""
    rgb = np.zeros((20, 30, 3), np.uint8)
    gray = np.zeros((20, 30, 1), np.uint8)
    sel = np.zeros((20, 30), np.uint8)
    sel[5:9, 5:9] = 255
    err, msg, _, _ = R.plugin_run(rgb, sel, corpus=gray)
    assert err == R.RESYNTH_ERROR_PLUGIN
    assert msg == "The input texture and output image must have the same number of color channels."
    err, msg, _, _ = R.plugin_run(rgb, sel, input_map=gray, output_map=rgb)
    assert (err, msg) == (R.RESYNTH_ERROR_PLUGIN, "The input and output maps must have the same mode")
    err, msg, _, _ = R.plugin_run(rgb, sel, input_map=rgb[:10], output_map=rgb)
    assert (err, msg) == (R.RESYNTH_ERROR_PLUGIN, "The input map should be the same size as the input texture image")
    err, msg, _, _ = R.plugin_run(rgb, sel, input_map=rgb, output_map=rgb[:10])
    assert (err, msg) == (R.RESYNTH_ERROR_PLUGIN, "The output map should be the same size as the output image")
    err, msg, _, _ = R.heal_selection(rgb, np.zeros_like(sel))
    assert (err, msg) == (R.RESYNTH_ERROR_PLUGIN, "You must first select a region to heal.")
    assert R.heal_selection(rgb, sel, direction=3)[0] == R.RESYNTH_ERROR_INVALID_ARGUMENT
    assert R.heal_selection(rgb, sel, order=-1)[0] == R.RESYNTH_ERROR_INVALID_ARGUMENT
    if not _has_device():
        # past the checks the engine needs the device
        assert R.plugin_run(rgb, sel)[0] == R.RESYNTH_ERROR_DEVICE
        assert R.heal_selection(rgb, sel)[0] == R.RESYNTH_ERROR_DEVICE
