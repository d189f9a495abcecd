// Internal declarations shared by the host driver (resynth_host.cpp) and the gfx950
// kernels (resynth_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rs {

// A HIP call that failed inside an entry point.  It is thrown by hip_fail and caught at
// the C boundary, which returns RESYNTH_ERROR_DEVICE instead of ending the caller's
// process.
struct HipFailure {
  hipError_t err;
};
[[noreturn]] void hip_fail(hipError_t e, const char *expr, const char *file, int line);

// Device pixel word.  Byte 0 is a flag byte; bytes 1.. hold the matched channels,
// colours first, then map channels, each one byte (the engine's Pixelel).
//   target word: flag bit 0 = hasValue (lib/engine.c:152-168)
//   corpus word: flag byte 0xFF = selected in the corpus mask and opaque (a corpus point,
//                lib/engine.c:400-431), 0xFE = selected but fully transparent (still
//                matched, lib/engine.c:247-258), 0 = not selected
// A word is 32 bits when colours+maps <= 3 and 64 bits otherwise, so every pixel is
// read and written with one aligned access: a reader sees a pixel's value and its
// hasValue flag from the same store.

struct Geometry {
  int tw, th;             // target image
  int cw, ch;             // corpus image
  int nColors, nMaps;     // img_match_bpp, map_match_bpp
  int tileH, tileV;
  unsigned patchLimit;    // neighbours gathered: max(patchSize, 2), <= 64
  unsigned maxProbe;      // maxProbeCount
  unsigned nT, nC, nO;    // targets, corpus points, offsets
  uint32_t penalty;       // 65535*nColors + Tm[0]*nMaps: score of a clipped/masked neighbour
  int tilesW, tilesH;     // 8x8-pixel tiles of the target image
  unsigned rhoMax;        // largest radius indexed by SynthBuffers::firstIdx
  int skipEmpty;          // pass 0 without tiling: start scans past provably empty rings
  unsigned nVisited;      // batched pass 0: targets[0, nVisited) hold values (chunk start)
  int probeByPixel;       // batched probes by rejection over corpus pixels (dense corpora)
  int bands;              // batched: 0, or kBands (a workgroup on XCD x draws its random probes
                          // from corpus rows [x*ch/kBands, (x+1)*ch/kBands))
  unsigned drawScale;     // batched, probes by pixel: ceil(65536 * corpus pixels / corpus points), the
                          // draws a round makes per remaining valid probe (16.16 fixed point)
};

constexpr int kBands = 8;   // one band of corpus rows per XCD

// Corpus words are stored in tiles of 8 x 4 pixels (32 words: one 128-byte line of 32-bit
// words): the corpus is cut into rows of four pixel rows, and within one such row each
// pixel's four words (one per pixel row) lie side by side.  A candidate's patch of the ~30
// nearest neighbours (about 7 x 7 pixels) then touches about 4 lines, where rows of pixels
// would take 7 to 8.  The buffer covers the corpus rounded up to whole tiles
// (cword_words); the words past the corpus are never read.
constexpr int kCTileW = 8, kCTileH = 4;
__host__ __device__ inline size_t cword_words(int cw, int ch)
{
  return (size_t)((cw + kCTileW - 1) / kCTileW) * kCTileW * (size_t)((ch + kCTileH - 1) / kCTileH) * kCTileH;
}

constexpr int kTile = 8;

struct Tables {
  uint16_t T[512];
  uint32_t Tm[512];
};

// Counters written by the synthesis kernels; 64 slots each to spread atomics.
enum { CNT_SLOTS = 64 };
struct Counters {
  unsigned long long candidates[CNT_SLOTS];
  unsigned long long terms[CNT_SLOTS];
  unsigned long long termsWave[CNT_SLOTS];   // terms scored by the large-chunk (wavefront and
                                             // segmented) visit kernels only
  unsigned long long scanWave[CNT_SLOTS];    // offsets their neighbour scans read
  unsigned long long stageWave[CNT_SLOTS];   // neighbours they staged
  unsigned long long visitsWave[CNT_SLOTS];  // their visits
  unsigned int betters[CNT_SLOTS];
  unsigned long long keysScored;             // dot products of the learned fine search
};

struct SynthBuffers {
  void *tword;            // [th*tw] Word
  int2 *sourceOf;         // [th*tw]
  const void *cword;      // [ch*cw] Word
  uint32_t *prober;       // [ch*cw] recent prober (lib/engine.c:308-327)
  const int2 *targets;    // [nT] in synthesis order
  const int2 *corpusPts;  // [nC] row-major
  const int2 *offsets;    // [nO] sorted offsets
  uint32_t *lastBest;     // [th*tw]
  const Tables *tables;
  Counters *counters;
  uint8_t *tileHas;       // [tilesH*tilesW] 1 when the tile holds a pixel with a value
  const int32_t *ctxRadius; // [tilesH*tilesW] radius around each tile inside which no pixel of the
                            // occupancy before any visit (the context) lies; see k_ctx_radius
  const uint32_t *firstIdx; // [rhoMax+1] number of sorted offsets with x^2+y^2 < rho^2
  uint32_t *ctl;          // [WCTL_WORDS] sliding-window launch control
  const uint32_t *bandStart; // [kBands+1] index in corpusPts of each band's first point
  const int2 *props;      // learned proposals of targets propsK0.. ([16] each, (-1,-1) = none)
  unsigned propsK0;
};

// Control words of a sliding-window launch, each on its own 128-byte line.
enum { WCTL_TICKET = 0, WCTL_DONE = 32, WCTL_EXPIRED = 64, WCTL_WORDS = 96 };

// State of the exact (serial) PRNG stream consumed on the device.
struct StreamState {
  uint32_t pos;           // next raw output to consume
  uint32_t kDone;         // first visit not completed by the last launch
  uint32_t aborted;       // visit kDone needs more raw outputs than the buffer holds
  uint32_t pad;
};

// One request of a multi-request launch (device memory).
struct ReqDesc {
  Geometry g;             // skipEmpty marks a request whose pass-0 scans may skip empty rings
  SynthBuffers b;
  unsigned sizes[6];      // pass sizes (lib/passes.h:68-93)
  unsigned seed;          // probe-stream seed
  unsigned minW;          // pass-0 window floor (>= 1)
  unsigned ramp;          // pass-0 window growth, permille of the targets visited in the pass
  unsigned capW;          // window cap
  unsigned maxPasses;     // stop after this many passes (0: the refiner's rule alone)
};

// Scheduling state of one request, each hot word on its own 128-byte line.
struct ReqCtl {
  unsigned long long ctr;            // (pass << 32) | next ticket of that pass; pass 6 = finished
  unsigned long long passesRun;      // set when the request finishes
  unsigned long long pad0[14];
  struct {
    unsigned long long word;         // (betters << 32) | visits finished in the pass
    unsigned long long pad[15];
  } pass[6];
};

// Global words of a multi-request launch.
enum { MG_ACTIVE = 0, MG_EXPIRED = 32, MG_WORDS = 64 };

// Kernel shape of the multi-request launch, chosen by the autotuner.
struct ManyConfig {
  int waves;              // wavefronts per workgroup: 2, 4 or 8
  int firstGroup;         // loads issued together for a candidate's first neighbours: 2
  int laterGroup;         // loads issued together afterwards: 4
  int lanes;              // lanes per target visit (64 / targets per wave): 16, 32 or 64
};

// ---- kernel launchers (resynth_kernels.hip) ----
hipError_t launch_adapt_simple(hipStream_t s, bool wide, const unsigned char *image, size_t rowBytes, unsigned bpp,
                               const unsigned char *mask, size_t maskRowBytes, const unsigned char *mask2,
                               size_t mask2RowBytes, int w, int h, int nColors, int alphaIdx, int useContext,
                               void *tword, void *cword, uint8_t *tflag, uint8_t *cflag);
hipError_t launch_adapt_maps(hipStream_t s, bool wide, const unsigned char *pix, int w, int h, int depth,
                             int colorEnd, int mapStart, int mapEnd, int alphaIdx, int useContext, int isTarget,
                             void *word, uint8_t *flag);
size_t compact_temp_bytes(size_t n);
hipError_t launch_compact(hipStream_t s, const uint8_t *flags, int w, int h, int2 *out, unsigned *count,
                          void *temp);
// How much of the sorted offsets to build: the prefix with x^2 + y^2 <= lim2 (plan_offsets).
struct OffsetPlan {
  int w = 0, h = 0;        // offsets span [-w+1, w-1] x [-h+1, h-1]
  int qw = 0, qh = 0;      // quadrant corner sorted
  uint32_t lim2 = 0;       // largest x^2 + y^2 kept
  size_t nQ = 0;           // quadrant entries sorted (qw * qh)
  size_t nKeep = 0;        // of them inside the disc
  size_t nO = 0;           // offsets in the table
  bool whole = true;       // the whole table
};
OffsetPlan plan_offsets(int w, int h, uint64_t maxD2);
size_t offsets_temp_bytes(const OffsetPlan &p);
size_t sort_pairs_temp_bytes(size_t n);
hipError_t launch_sort_pairs(hipStream_t s, const uint32_t *kIn, uint32_t *kOut, const uint32_t *vIn, uint32_t *vOut,
                             size_t n, unsigned bits, void *temp, size_t tempBytes);
hipError_t launch_sorted_offsets(hipStream_t s, const OffsetPlan &p, int2 *out, void *temp, size_t tempBytes);
hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, size_t n);
hipError_t launch_band_starts(hipStream_t s, const int2 *corpusPts, const unsigned *nC, int ch, uint32_t *out);
hipError_t launch_ctx_radius(hipStream_t s, const uint8_t *mark, int W, int H, int32_t *scratch, int32_t *radius);
hipError_t launch_reach_radius(hipStream_t s, bool wide, const void *tword, const uint8_t *tflag, int tw, int th,
                               uint8_t *full, uint8_t *hasTarget, int32_t *scratch, int32_t *rhoFull, uint32_t *reach);
hipError_t launch_tile_init(hipStream_t s, bool wide, const void *tword, int tw, int th, uint8_t *tileHas);
hipError_t launch_first_idx(hipStream_t s, int w, int h, unsigned rhoMax, uint32_t *firstIdx);
hipError_t launch_gather_targets(hipStream_t s, const int2 *rm, const uint32_t *perm, int2 *out, unsigned n);
hipError_t launch_random_order(hipStream_t s, uint32_t seed, unsigned n, const int2 *rm, int2 *out);
// target ordering of match-context types 2..8 (resynth_order.hip)
size_t order_temp_bytes(unsigned n);
hipError_t launch_order_sort(hipStream_t s, const int2 *rm, unsigned n, int type, const uint32_t *axisRays,
                             void *temp, size_t tempBytes, const uint32_t **order);
hipError_t launch_order_compose(hipStream_t s, const int2 *rm, const uint32_t *order, const uint32_t *perm, int2 *out,
                                unsigned n);
// lpt: 0 = by chunk size and pass; 1 = one wavefront per target; 16/32/64 = segmented visits
hipError_t launch_synth_batched(hipStream_t s, bool wide, const Geometry &g, const SynthBuffers &b, unsigned k0,
                                unsigned k1, unsigned seed, unsigned pass, int lpt = 0);
unsigned synth_wave_min();   // chunks of at least this many targets use k_synth_wave
hipError_t launch_synth_window(hipStream_t s, bool wide, const Geometry &g, const SynthBuffers &b, unsigned k0,
                               unsigned k1, unsigned seed, unsigned pass, unsigned minW, unsigned rampPermille,
                               unsigned cap);
bool many_config_valid(const ManyConfig &cfg);
// The multi-request launch as an instantiated HIP graph (init kernel, persistent kernel),
// kept across calls: a call with the kernel shape of the last one updates the nodes'
// arguments in place; another shape rebuilds it (instantiated = true).
struct ManyGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t init = nullptr, many = nullptr;
  const void *fn = nullptr;   // the persistent kernel of the instantiated graph
};
hipError_t many_graph_prepare(ManyGraph &G, bool wide, const ManyConfig &cfg, bool allRgb, const ReqDesc *descs,
                              ReqCtl *ctls, unsigned nReq, uint32_t *global, bool &instantiated);
void many_graph_release(ManyGraph &G);
hipError_t launch_synth_serial(hipStream_t s, bool wide, const Geometry &g, const SynthBuffers &b, unsigned k0,
                               unsigned k1, int prngKind, const uint32_t *stream, uint32_t streamLen,
                               StreamState *state);
hipError_t launch_writeback_simple(hipStream_t s, bool wide, const void *tword, const int2 *targets, unsigned nT,
                                   unsigned char *image, size_t rowBytes, unsigned bpp, int nColors, int w);
hipError_t launch_writeback_pixmap(hipStream_t s, bool wide, const void *tword, const int2 *targets, unsigned nT,
                                   unsigned char *pix, int depth, int nColors, int w);
hipError_t launch_mean_error(hipStream_t s, const uint32_t *lastBest, const int2 *targets, unsigned nT, int w,
                             unsigned long long *sumCount);
hipError_t launch_score_frozen(hipStream_t s, bool wide, int cw, int ch, int nColors, int nMaps, uint32_t penalty,
                               const void *cword, const Tables *tables, unsigned nProblems, const unsigned *nbCount,
                               const int2 *nbOffset, const void *nbWord, const unsigned *candCount,
                               unsigned maxCands, const int2 *cands, int32_t *winner, uint32_t *bestSum,
                               int32_t *stopped);
hipError_t launch_pack_words(hipStream_t s, bool wide, const unsigned char *pix, int n, int w, int depth,
                             int colorEnd, int mapStart, int mapEnd, int corpusFlags, void *out);

}  // namespace rs
