"""The GIMP plug-in shim (resynth_plugin_run) and heal selection (resynth_heal_selection)
on the GPU.  In the serial mode both reproduce the reference's regression images byte for
byte: the six heal cases of Test/testResynth.py:331-366 (through the heal script's corpus
construction) and direct-engine cases of :216-328 driven through the plug-in's own
adaptation (selection masks, a second corpus image, maps)."""
import contextlib

import numpy as np
import pytest

import pins
import resynthesizer_amd as R

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def default_mode(mode):
    R.lib().resynth_set_default_mode(mode)
    try:
        yield
    finally:
        R.lib().resynth_set_default_mode(-1)


@pytest.mark.parametrize("name", sorted(pins.heal_cases()))
def test_heal_selection_reproduces_reference_image(name):
    c = pins.heal_cases()[name]
    px, sel, expected, (rx0, ry0, rx1, ry1), _ = pins.heal_input(c)
    with default_mode(R.MODE_SERIAL):
        err, msg, out, progress = R.heal_selection(px, sel, c["radius"], c["direction"], c["order"])
    assert (err, msg) == (0, None)
    assert progress and progress == sorted(progress)   # lib/progress.c's percentages, on this thread
    assert np.array_equal(pins.gimp_flatten(out)[ry0:ry1, rx0:rx1], expected)
    # only colour channels of selected pixels change; alpha never does
    keep = sel == 0
    assert np.array_equal(out[keep], px[keep])
    if px.shape[2] in (2, 4):
        assert np.array_equal(out[..., -1], px[..., -1])


@pytest.mark.parametrize("name", ["heal-ufo-input", "healaroundrandom-ufo-input", "healalpha-apollo11_w_alpha",
                                  "healalphagray-ufo-input-w-alpha-gray"])
def test_heal_selection_batched(name):
    """The default (batched) mode: context and alpha untouched, every healed pixel takes the
    colour of a pixel of the band around the selection (the corpus)."""
    c = pins.heal_cases()[name]
    px, sel, _, _, _ = pins.heal_input(c)
    err, msg, out, _ = R.heal_selection(px, sel, c["radius"], c["direction"], c["order"])
    assert (err, msg) == (0, None)
    keep = sel == 0
    assert np.array_equal(out[keep], px[keep])
    nc = 3 if px.shape[2] >= 3 else 1
    x0, y0, x1, y1 = c["select"]
    band = np.zeros_like(keep)
    band[max(0, y0 - c["radius"]):y1 + c["radius"], max(0, x0 - c["radius"]):x1 + c["radius"]] = True
    band &= keep
    if px.shape[2] in (2, 4):
        band &= px[..., -1] > 0   # transparent pixels are not in the corpus
    code = lambda a: sum(a[..., k].astype(np.int64) << (8 * k) for k in range(nc))
    assert np.isin(code(out[~keep]), np.unique(code(px[band]))).all()


@pytest.mark.parametrize("name", ["resynthfull-ufo-input", "resynthtwoimages-ufo-input", "resynthtileable-ufo-input",
                                  "mapstyleangel-angel_target"])
def test_plugin_run_reproduces_reference_image(name):
    """Direct-engine regression cases through the plug-in shim: the target is the drawable
    with its selection (none: the whole drawable), the corpus the same image with the
    inverted selection or another image, the maps the images themselves."""
    c = pins.cases()[name]
    p = c["params"]
    img = pins.load_png(c["image"])
    h, w, _ = img.shape
    x0, y0, x1, y1 = c["rect"]
    sel = np.zeros((h, w), np.uint8)
    sel[y0:y1, x0:x1] = 255
    kw = dict(h_tile=p["tileH"], v_tile=p["tileV"], use_border=p["matchContextType"], map_weight=p["mapWeight"],
              autism=p["sensitivity"], neighbours=p["patchSize"], trys=p["maxProbeCount"])
    if c["corpus"] == "inverse":
        kw.update(corpus=img, corpus_selection=255 - sel)
    else:
        kw.update(corpus=pins.load_png(c["corpus"]))
    if c["maps"]:
        kw.update(input_map=kw["corpus"], output_map=img)
        sel = None   # the whole canvas is the target
    with default_mode(R.MODE_SERIAL):
        err, msg, out, _ = R.plugin_run(img, sel, **kw)
    assert (err, msg) == (0, None)
    expected = np.load(pins.os.path.join(pins.GOLDEN, "expect_%s.npy" % name))
    assert np.array_equal(out[y0:y1, x0:x1], expected)


def test_plugin_run_reports_empty_corpus():
    img = np.full((24, 24, 4), 90, np.uint8)
    img[..., 3] = 0   # every corpus pixel transparent
    sel = np.zeros((24, 24), np.uint8)
    sel[8:12, 8:12] = 255
    err, msg, out, _ = R.plugin_run(img, sel, corpus=img)
    assert err == R.IMAGE_SYNTH_ERROR_EMPTY_CORPUS
    assert msg == "The texture source is empty. Does any selection include non-transparent pixels?"
    assert np.array_equal(out, img)
