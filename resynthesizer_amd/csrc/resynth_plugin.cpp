// GIMP plug-in shim and heal-selection helper of libresynth_mi355x.so: the host side of the
// engine plug-in's run() (src/resynthesizer/resynthesizer.c:286-546 with its adapter,
// src/resynthesizer/adaptGimp.h) and of the heal-selection script
// (PluginScripts/plugin-heal-selection.py:36-152), on drawables passed as plain buffers.
// Both end in engine(), which runs on the device.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "resynth_mi355x.h"

namespace {

constexpr unsigned kMaxNeighbours = 64;   // IMAGE_SYNTH_MAX_NEIGHBORS, lib/imageSynthConstants.h
constexpr unsigned char kTotallySelected = 0xFF;   // MASK_TOTALLY_SELECTED

unsigned drawable_bpp(const ResynthDrawable &d) { return d.colorChannels + (d.hasAlpha ? 1u : 0u); }

bool valid_drawable(const ResynthDrawable *d)
{
  return d && d->pixels && d->width && d->height && (d->colorChannels == 1 || d->colorChannels == 3) &&
         d->rowBytes >= (size_t)d->width * drawable_bpp(*d) &&
         (!d->selection || d->selectionRowBytes >= d->width);
}

bool has_selection(const ResynthDrawable &d)
{
  if (!d.selection) return false;
  for (unsigned y = 0; y < d.height; y++) {
    const unsigned char *row = d.selection + (size_t)y * d.selectionRowBytes;
    for (unsigned x = 0; x < d.width; x++)
      if (row[x]) return true;
  }
  return false;
}

// An engine pixmap of the drawable (fetch_image_mask_map, src/resynthesizer/adaptGimp.h:
// 257-322): [mask, colour and alpha channels, map channels without the map's alpha],
// total_bpp bytes per pixel, unused bytes zero.  The mask is the selection, or totally
// selected when no selection intersects the drawable.
struct Pixmap {
  std::vector<unsigned char> bytes;
  GArray array;
  Map map;
  Pixmap(const ResynthDrawable &d, unsigned totalBpp, const ResynthDrawable *mapDrawable, unsigned mapStart)
      : bytes((size_t)d.width * d.height * totalBpp, 0)
  {
    const unsigned bpp = drawable_bpp(d);
    const bool sel = has_selection(d);
    for (unsigned y = 0; y < d.height; y++) {
      const unsigned char *src = d.pixels + (size_t)y * d.rowBytes;
      const unsigned char *mrow = sel ? d.selection + (size_t)y * d.selectionRowBytes : nullptr;
      unsigned char *dst = bytes.data() + (size_t)y * d.width * totalBpp;
      for (unsigned x = 0; x < d.width; x++) {
        unsigned char *p = dst + (size_t)x * totalBpp;
        p[0] = sel ? mrow[x] : kTotallySelected;
        std::memcpy(p + 1, src + (size_t)x * bpp, bpp);
      }
    }
    if (mapDrawable) {
      const unsigned mbpp = drawable_bpp(*mapDrawable), mcopy = mapDrawable->colorChannels;
      for (unsigned y = 0; y < d.height; y++) {
        const unsigned char *src = mapDrawable->pixels + (size_t)y * mapDrawable->rowBytes;
        unsigned char *dst = bytes.data() + (size_t)y * d.width * totalBpp;
        for (unsigned x = 0; x < d.width; x++)
          std::memcpy(dst + (size_t)x * totalBpp + mapStart, src + (size_t)x * mbpp, mcopy);
      }
    }
    array.data = (char *)bytes.data();
    array.len = (unsigned)((size_t)d.width * d.height);
    map.width = d.width;
    map.height = d.height;
    map.depth = totalBpp;
    map.data = &array;
  }
};

// pixmap_to_drawable (src/resynthesizer/adaptGimp.h:60-95): the drawable's channels back
// from the pixmap's bytes 1..bpp.
void post_results(const Pixmap &pm, ResynthDrawable &d)
{
  const unsigned bpp = drawable_bpp(d), depth = pm.map.depth;
  for (unsigned y = 0; y < d.height; y++) {
    unsigned char *dst = d.pixels + (size_t)y * d.rowBytes;
    const unsigned char *src = pm.bytes.data() + (size_t)y * d.width * depth;
    for (unsigned x = 0; x < d.width; x++) std::memcpy(dst + (size_t)x * bpp, src + (size_t)x * depth + 1, bpp);
  }
}

int fail(const char **message, const char *text)
{
  if (message) *message = text;
  return RESYNTH_ERROR_PLUGIN;
}

int plugin_run(const TGimpAdapterParameters &pp, ResynthDrawable &drawable, const ResynthDrawable &corpus,
               const ResynthDrawable *mapIn, const ResynthDrawable *mapOut, void (*progress)(int, void *),
               void *ctx, int *cancelFlag, const char **message)
{
  // src/resynthesizer/resynthesizer.c:357-421
  TGimpAdapterParameters p = pp;
  if (p.neighbours > (int)kMaxNeighbours) p.neighbours = (int)kMaxNeighbours;
  if (drawable.colorChannels != corpus.colorChannels)
    return fail(message, "The input texture and output image must have the same number of color channels.");
  const bool withMap = mapIn && mapOut;   // one map alone is ignored
  if (withMap) {
    if (mapIn->colorChannels != mapOut->colorChannels)
      return fail(message, "The input and output maps must have the same mode");
    if (mapIn->width != corpus.width || mapIn->height != corpus.height)
      return fail(message, "The input map should be the same size as the input texture image");
    if (mapOut->width != drawable.width || mapOut->height != drawable.height)
      return fail(message, "The output map should be the same size as the output image");
  }
  // :448-498
  TFormatIndices ix;
  prepareImageFormatIndices(&ix, drawable.colorChannels, withMap ? mapIn->colorChannels : 0, drawable.hasAlpha != 0,
                            corpus.hasAlpha != 0, withMap);
  Pixmap target(drawable, ix.total_bpp, withMap ? mapOut : nullptr, ix.map_start_bip);
  Pixmap source(corpus, ix.total_bpp, withMap ? mapIn : nullptr, ix.map_start_bip);
  // adaptPluginToLibraryParameters (src/resynthesizer/adaptParameters.c)
  TImageSynthParameters ep;
  ep.isMakeSeamlesslyTileableHorizontally = p.h_tile;
  ep.isMakeSeamlesslyTileableVertically = p.v_tile;
  ep.matchContextType = p.use_border;
  ep.mapWeight = p.map_weight;
  ep.sensitivityToOutliers = p.autism;
  ep.patchSize = (unsigned)p.neighbours;
  ep.maxProbeCount = (unsigned)p.trys;
  // :506-540
  const int result = engine(ep, &ix, &target.map, &source.map, progress, ctx, cancelFlag);
  if (result == IMAGE_SYNTH_ERROR_EMPTY_CORPUS) {
    if (message) *message = "The texture source is empty. Does any selection include non-transparent pixels?";
    return result;
  }
  if (result == IMAGE_SYNTH_ERROR_EMPTY_TARGET) {
    if (message) *message = "The output layer is empty. Does any selection have visible pixels in the active layer?";
    return result;
  }
  if (result) return result;
  post_results(target, drawable);
  return 0;
}

// Half-heights of the disc GIMP grows selections with: the disc's column at offset i - r
// (i in 0..2r) spans rows -circ[i]..circ[i], its height taken at the column edge nearer
// the centre.
std::vector<int> grow_disc(int r)
{
  std::vector<int> circ((size_t)(2 * r + 1));
  for (int i = 0; i <= 2 * r; i++) {
    const double t = i > r ? (i - r) - 0.5 : (i < r ? (r - i) - 0.5 : 0.0);
    circ[(size_t)i] = (int)std::floor(std::sqrt((double)r * r - t * t) + 0.5);
  }
  return circ;
}

struct Box {
  int x0, y0, x1, y1;   // x1, y1 exclusive; x0 >= x1 when empty
};

Box bounds(const std::vector<unsigned char> &m, int w, int h)
{
  Box b = {w, h, 0, 0};
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++)
      if (m[(size_t)y * w + x]) {
        b.x0 = std::min(b.x0, x);
        b.x1 = std::max(b.x1, x + 1);
        b.y0 = std::min(b.y0, y);
        b.y1 = std::max(b.y1, y + 1);
      }
  return b;
}

// Grey-level dilation of sel (w x h) by the disc of radius r: each pixel takes the largest
// selection value within the disc around it.  Per distinct half-height c, a sliding maximum
// over 2c+1 rows of each selected column (van Herk: block prefix and suffix maxima), then
// shifted by every column offset whose half-height is c.
std::vector<unsigned char> grow_selection(const std::vector<unsigned char> &sel, int w, int h, int r, const Box &sb)
{
  std::vector<unsigned char> out((size_t)w * h, 0);
  if (sb.x0 >= sb.x1) return out;
  if (r <= 0) return sel;
  const std::vector<int> circ = grow_disc(r);
  const int gy0 = std::max(0, sb.y0 - r), gy1 = std::min(h, sb.y1 + r), rows = gy1 - gy0;
  std::vector<int> cs(circ.begin(), circ.end());
  std::sort(cs.begin(), cs.end());
  cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
  std::vector<unsigned char> vm((size_t)rows), col, pre, suf;
  for (int c : cs) {
    const int k = 2 * c + 1, n = rows + 2 * c;
    const int nb = (n + k - 1) / k * k;
    col.assign((size_t)nb, 0);
    pre.resize((size_t)nb);
    suf.resize((size_t)nb);
    for (int x = sb.x0; x < sb.x1; x++) {
      // column x over rows gy0 - c .. gy1 + c (zero outside the image)
      for (int i = 0; i < n; i++) {
        const int y = gy0 - c + i;
        col[(size_t)i] = (y >= 0 && y < h) ? sel[(size_t)y * w + x] : 0;
      }
      for (int i = n; i < nb; i++) col[(size_t)i] = 0;
      for (int b = 0; b < nb; b += k) {
        pre[(size_t)b] = col[(size_t)b];
        for (int i = b + 1; i < b + k; i++) pre[(size_t)i] = std::max(pre[(size_t)i - 1], col[(size_t)i]);
        suf[(size_t)(b + k - 1)] = col[(size_t)(b + k - 1)];
        for (int i = b + k - 2; i >= b; i--) suf[(size_t)i] = std::max(suf[(size_t)i + 1], col[(size_t)i]);
      }
      // row y = gy0 + j takes the window col[j .. j + 2c]
      for (int j = 0; j < rows; j++) vm[(size_t)j] = std::max(suf[(size_t)j], pre[(size_t)(j + 2 * c)]);
      for (int i = 0; i <= 2 * r; i++) {
        if (circ[(size_t)i] != c) continue;
        const int xo = x - (i - r);   // output column whose offset i - r reaches column x
        if (xo < 0 || xo >= w) continue;
        for (int j = 0; j < rows; j++) {
          unsigned char &o = out[(size_t)(gy0 + j) * w + xo];
          o = std::max(o, vm[(size_t)j]);
        }
      }
    }
  }
  return out;
}

int heal_selection(ResynthDrawable &d, int radius, int direction, int order, void (*progress)(int, void *), void *ctx,
                   int *cancelFlag, const char **message)
{
  if (!has_selection(d)) return fail(message, "You must first select a region to heal.");
  const int w = (int)d.width, h = (int)d.height;
  std::vector<unsigned char> sel((size_t)w * h);
  for (int y = 0; y < h; y++) std::memcpy(&sel[(size_t)y * w], d.selection + (size_t)y * d.selectionRowBytes, w);
  const Box tb = bounds(sel, w, h);
  // the frisket: the grown selection less the original (a subtraction of channel values)
  std::vector<unsigned char> frisket = grow_selection(sel, w, h, radius, tb);
  for (size_t i = 0; i < frisket.size(); i++) frisket[i] = (unsigned char)std::max(0, (int)frisket[i] - (int)sel[i]);
  Box fb = bounds(frisket, w, h);
  if (fb.x0 >= fb.x1) fb = {0, 0, w, h};   // mask bounds of an empty selection: the drawable
  int nw, nh, nx, ny;
  if (direction == 0) {
    nw = fb.x1 - fb.x0; nh = fb.y1 - fb.y0; nx = fb.x0; ny = fb.y0;
  } else if (direction == 1) {
    nw = fb.x1 - fb.x0; nh = tb.y1 - tb.y0; nx = fb.x0; ny = tb.y0;
  } else {
    nw = tb.x1 - tb.x0; nh = fb.y1 - fb.y0; nx = tb.x0; ny = fb.y0;
  }
  nw = std::min(w - nx, nw);
  nh = std::min(h - ny, nh);
  // the corpus: a cropped copy of the drawable, its selection the cropped frisket
  const unsigned bpp = drawable_bpp(d);
  std::vector<unsigned char> cpix((size_t)nw * nh * bpp), csel((size_t)nw * nh);
  for (int y = 0; y < nh; y++) {
    std::memcpy(&cpix[(size_t)y * nw * bpp], d.pixels + (size_t)(ny + y) * d.rowBytes + (size_t)nx * bpp,
                (size_t)nw * bpp);
    std::memcpy(&csel[(size_t)y * nw], &frisket[(size_t)(ny + y) * w + nx], (size_t)nw);
  }
  ResynthDrawable corpus = d;
  corpus.pixels = cpix.data();
  corpus.width = (unsigned)nw;
  corpus.height = (unsigned)nh;
  corpus.rowBytes = (size_t)nw * bpp;
  corpus.selection = csel.data();
  corpus.selectionRowBytes = (size_t)nw;
  // order and direction in one match-context type: random, inwards (2 + direction),
  // outwards (5 + direction)
  const int useBorder = order == 0 ? 1 : (order == 1 ? direction + 2 : direction + 5);
  TGimpAdapterParameters p = {0, 0, useBorder, -1, -1, -1, 0.0, 0.117, 16, 500};
  return plugin_run(p, d, corpus, nullptr, nullptr, progress, ctx, cancelFlag, message);
}

}  // namespace

extern "C" {

int resynth_plugin_run(const TGimpAdapterParameters *parameters, ResynthDrawable *drawable,
                       const ResynthDrawable *corpus, const ResynthDrawable *inputMap,
                       const ResynthDrawable *outputMap, void (*progressCallback)(int, void *), void *contextInfo,
                       int *cancelFlag, const char **message)
{
  if (message) *message = nullptr;
  if (!parameters || !drawable) return RESYNTH_ERROR_INVALID_ARGUMENT;
  if (drawable->colorChannels != 1 && drawable->colorChannels != 3) return fail(message, "Incompatible image mode.");
  if (!corpus) corpus = drawable;
  for (const ResynthDrawable *d : {(const ResynthDrawable *)drawable, corpus})
    if (!valid_drawable(d)) return RESYNTH_ERROR_INVALID_ARGUMENT;
  for (const ResynthDrawable *m : {inputMap, outputMap})
    if (m && !valid_drawable(m)) return RESYNTH_ERROR_INVALID_ARGUMENT;
  try {
    return plugin_run(*parameters, *drawable, *corpus, inputMap, outputMap, progressCallback, contextInfo, cancelFlag,
                      message);
  } catch (const std::bad_alloc &) {
    return RESYNTH_ERROR_DEVICE;
  }
}

int resynth_heal_selection(ResynthDrawable *drawable, int samplingRadius, int direction, int order,
                           void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag,
                           const char **message)
{
  if (message) *message = nullptr;
  if (!drawable || !valid_drawable(drawable) || samplingRadius < 0 || samplingRadius > 4096 || direction < 0 ||
      direction > 2 || order < 0 || order > 2)
    return RESYNTH_ERROR_INVALID_ARGUMENT;
  try {
    return heal_selection(*drawable, samplingRadius, direction, order, progressCallback, contextInfo, cancelFlag,
                          message);
  } catch (const std::bad_alloc &) {
    return RESYNTH_ERROR_DEVICE;
  }
}

}  // extern "C"
