"""CPU restatement of the GIMP plug-in's adaptation (src/resynthesizer/resynthesizer.c:
286-546, src/resynthesizer/adaptGimp.h:178-322) and of the heal-selection script's
preprocessing (PluginScripts/plugin-heal-selection.py:36-152), feeding the oracle's
engine().

Test infrastructure only (imported by tests/); the shipped library never imports it.

Pinned by the reference's own regression images: the six heal cases of
Test/testResynth.py:331-366 (Test/reference_out_images/heal*.ppm) are reproduced byte for
byte (tests/test_oracle.py), which fixes GIMP's selection-grow disc (a plausible variant,
heights taken at the column centre instead of its inner edge, fails the all-around case),
the frisket crop per direction and the order/direction encoding of use_border.
"""
import math

import numpy as np

import oracle_py as O

MASK_TOTALLY_SELECTED = 0xFF


def grow_disc(r, inner_edge=True):
    """Half-heights of GIMP's selection-grow disc: column i - r spans rows -c[i]..c[i].
    inner_edge=False takes the height at the column centre (a variant the pins reject)."""
    out = []
    for i in range(2 * r + 1):
        if inner_edge:
            t = (i - r) - 0.5 if i > r else ((r - i) - 0.5 if i < r else 0.0)
        else:
            t = float(abs(i - r))
        out.append(int(math.floor(math.sqrt(r * r - t * t) + 0.5)))
    return out


def grow(sel, r, inner_edge=True):
    """Grey-level dilation of a uint8 (H, W) selection by the disc (a brute-force sweep over
    the disc's columns)."""
    h, w = sel.shape
    out = np.zeros_like(sel)
    for i, c in enumerate(grow_disc(r, inner_edge)):
        dx = i - r
        sh = np.zeros_like(sel)
        if dx >= 0:
            sh[:, :w - dx] = sel[:, dx:]
        else:
            sh[:, -dx:] = sel[:, :w + dx]
        # vertical maximum over rows y - c .. y + c
        pad = np.pad(sh, ((c, c), (0, 0)))
        win = np.zeros_like(sh)
        for k in range(2 * c + 1):
            win = np.maximum(win, pad[k:k + h])
        out = np.maximum(out, win)
    return out


def _bounds(m):
    ys, xs = np.nonzero(m)
    if len(xs) == 0:
        return None
    return int(xs.min()), int(ys.min()), int(xs.max()) + 1, int(ys.max()) + 1


def heal_corpus(pixels, selection, radius, direction, order, inner_edge=True):
    """The heal script's corpus (plugin-heal-selection.py:48-118) and use_border (:120-133):
    returns (corpus pixels, corpus selection, use_border)."""
    h, w = selection.shape
    grown = grow(selection, radius, inner_edge)
    frisket = np.maximum(grown.astype(np.int32) - selection, 0).astype(np.uint8)
    tx0, ty0, tx1, ty1 = _bounds(selection)
    fb = _bounds(frisket) or (0, 0, w, h)
    fx0, fy0, fx1, fy1 = fb
    if direction == 0:
        nw, nh, nx, ny = fx1 - fx0, fy1 - fy0, fx0, fy0
    elif direction == 1:
        nw, nh, nx, ny = fx1 - fx0, ty1 - ty0, fx0, ty0
    else:
        nw, nh, nx, ny = tx1 - tx0, fy1 - fy0, tx0, fy0
    nw, nh = min(w - nx, nw), min(h - ny, nh)
    use_border = 1 if order == 0 else (direction + 2 if order == 1 else direction + 5)
    return pixels[ny:ny + nh, nx:nx + nw], frisket[ny:ny + nh, nx:nx + nw], use_border


def _pixmap(pixels, selection, total_bpp, map_pixels=None, map_start=0):
    """fetch_image_mask_map (adaptGimp.h:257-322): [mask, colour + alpha, map channels]."""
    h, w, c = pixels.shape
    pm = np.zeros((h, w, total_bpp), np.uint8)
    sel = selection if selection is not None and selection.any() else None
    pm[..., 0] = sel if sel is not None else MASK_TOTALLY_SELECTED
    pm[..., 1:1 + c] = pixels
    if map_pixels is not None:
        mc = 3 if map_pixels.shape[2] >= 3 else 1   # the map's alpha is dropped
        pm[..., map_start:map_start + mc] = map_pixels[..., :mc]
    return pm


def plugin_maps(pixels, selection, corpus, corpus_selection, input_map=None, output_map=None):
    """Format indices and the two engine pixmaps the plug-in builds (resynthesizer.c:448-498)
    from uint8 (H, W, C) drawables (C = 1, 2, 3, 4: gray, gray + alpha, RGB, RGBA)."""
    colors = 3 if pixels.shape[2] >= 3 else 1
    alpha_t, alpha_c = pixels.shape[2] in (2, 4), corpus.shape[2] in (2, 4)
    with_map = input_map is not None and output_map is not None
    map_colors = (3 if input_map.shape[2] >= 3 else 1) if with_map else 0
    ix = O.prepare_indices(colors, map_colors, alpha_t, alpha_c, with_map)
    tp = _pixmap(pixels, selection, ix.total_bpp, output_map if with_map else None, ix.map_start_bip)
    cp = _pixmap(corpus, corpus_selection, ix.total_bpp, input_map if with_map else None, ix.map_start_bip)
    return ix, tp, cp


def heal(pixels, selection, radius=50, direction=0, order=0, inner_edge=True, **kw):
    """Heal selection through the oracle engine (serial GRand): returns (Result, pixels out)."""
    if pixels.ndim == 2:
        pixels = pixels[..., None]
    corpus, csel, use_border = heal_corpus(pixels, selection, radius, direction, order, inner_edge)
    ix, tp, cp = plugin_maps(pixels, selection, corpus, csel)
    params = dict(tileH=0, tileV=0, matchContextType=use_border, mapWeight=0.0, sensitivity=0.117, patchSize=16,
                  maxProbeCount=500)
    r, out = O.engine(params, ix, np.ascontiguousarray(tp), np.ascontiguousarray(cp), **kw)
    return r, out[..., 1:1 + pixels.shape[2]]
