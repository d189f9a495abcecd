/*
 * oracle.h -- CPU restatement of the resynthesizer synthesis engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py may use it, and only as the checker (or, for bench.py, as the
 * timed CPU baseline), never as the thing measured on the GPU.
 *
 * Every function restates the reference's algorithm from its source text; the
 * reference file:line each one follows is cited beside its definition in
 * resynth_oracle.c.  Parity pins (what checks each function against an
 * independent source of truth) are listed in that file's header.
 *
 * Structure layouts deliberately match the reference's public types so that a
 * test can hand the same bytes to the oracle and to the GPU library:
 *   OImageBuffer  == ImageBuffer           (lib/imageBuffer.h:12-18)
 *   OParams       == TImageSynthParameters (lib/engineParams.h:31-86)
 *   OFormatIndices== TFormatIndices        (lib/imageFormatIndicies.h:46-58)
 */
#ifndef RESYNTH_ORACLE_H
#define RESYNTH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  unsigned char *data;
  unsigned int width;
  unsigned int height;
  size_t rowBytes;
} OImageBuffer;

typedef struct {
  int tileH;
  int tileV;
  int matchContextType;
  double mapWeight;
  double sensitivity;
  unsigned int patchSize;
  unsigned int maxProbeCount;
} OParams;

typedef struct {
  uint8_t colorEndBip, alpha_bip, map_start_bip, map_end_bip;
  uint8_t img_match_bpp, map_match_bpp, total_bpp;
  int isAlphaTarget, isAlphaSource;
} OFormatIndices;

/* Unpadded interleaved pixmap: pixel (x,y) starts at data[(x + y*width)*depth]. */
typedef struct {
  unsigned int width, height, depth;
  uint8_t *data;
} OPixmap;

enum { ORC_PRNG_GRAND = 0, ORC_PRNG_LIBC = 1 };

typedef struct {
  int threads;   /* 0 or 1: serial refiner; >1: threaded refiner with that many threads */
  int prng;      /* ORC_PRNG_GRAND (GLib build) or ORC_PRNG_LIBC (stand-alone lib build) */
  int batch;     /* >0: model of batched synthesis (not a reference mode): each pass runs in
                    chunks of this many targets that all read the state at the chunk start;
                    pass 0 ramps its chunks up from 256 */
  int conflictRadius; /* batch model: >0 defers a target to a later chunk while a target of
                         the current chunk lies within this Chebyshev distance */
  int rampPermille;   /* batch model: >0 sizes pass-0 chunks as max(floor, k*rampPermille/1000)
                         where k is the number of targets already visited in the pass */
  unsigned seed;      /* 0: the engine's seed 1198472 (lib/engine.c:643); else a study seed */
  int minChunk;       /* batch model: floor of the pass-0 ramp (0: 64) */
} OOptions;

typedef struct {
  unsigned int nTargets, nCorpus, nOffsets;
  unsigned int passes;               /* passes actually run */
  unsigned int passSize[6];
  unsigned int betters[6];
  unsigned long long scorerCalls;    /* computeBestFit invocations */
  unsigned long long visits;         /* target visits */
  unsigned int progressCalls;
  int lastPercent;
  /* optional outputs, caller-owned, may be NULL */
  uint32_t *lastBest;   /* [target W*H] best sum of each target pixel's last visit, UINT32_MAX elsewhere */
  int32_t *sourceOf;    /* [target W*H*2] final source coordinates, (-1,-1) where none */
  double meanError;     /* mean of lastBest over target pixels that were scored */
} OStats;

/* ---- PRNG restatements ---- */
typedef struct { uint32_t mt[624]; int mti; } OGRand;
void orc_grand_seed(OGRand *r, uint32_t seed);
uint32_t orc_grand_int(OGRand *r);
int32_t orc_grand_int_range(OGRand *r, int32_t begin, int32_t end);

typedef struct { int32_t s[31]; int f, r; } OLibcRand;
void orc_libc_srand(OLibcRand *r, uint32_t seed);
int32_t orc_libc_rand(OLibcRand *r);
uint32_t orc_libc_int_range(OLibcRand *r, uint32_t lo, uint32_t hi);

/* ---- sort and tables ---- */
void orc_merge_sort(void *base, size_t n, size_t size, int (*cmp)(const void *, const void *));
/* offsets for (w,h) = min of corpus/target dims, sorted as the engine sorts them; out has (2w-1)(2h-1)*2 ints */
void orc_sorted_offsets(int w, int h, int32_t *out);
void orc_quantize(double sensitivity, double mapWeight, uint16_t T[512], uint32_t Tm[512]);
void orc_repetition(unsigned n, unsigned sizes[6]);

/* ---- frozen-state scorer ----
 * Evaluates candidates cands[0..nCand) in order against one neighbour set, with the
 * reference's sequential computeBestFit semantics (strict-less update, early out,
 * stop at the first perfect match).  Returns the index of the winning candidate or -1. */
int orc_best_of(const OPixmap *corpus, const OFormatIndices *ix,
                const uint16_t T[512], const uint32_t Tm[512],
                unsigned count, const int32_t *offs, const uint8_t *pix /* 8 bytes per neighbour */,
                unsigned nCand, const int32_t *cands,
                uint32_t *bestSum, int *stoppedAt);

/* ---- target ordering (exposed for tests) ---- */
int orc_order_targets(int matchContextType, int32_t *points, unsigned n, int prngKind,
                      uint64_t *drawsUsed);

/* ---- library entries ---- */
int oracle_engine(OParams params, const OFormatIndices *ix, OPixmap *target, const OPixmap *corpus,
                  void (*progress)(int, void *), void *ctx, int *cancelFlag,
                  const OOptions *opt, OStats *stats);

int oracle_imageSynth(OImageBuffer *image, OImageBuffer *mask, int format, OParams *params,
                      void (*progress)(int, void *), void *ctx, int *cancelFlag,
                      const OOptions *opt, OStats *stats);

int oracle_imageSynth2(OImageBuffer *image, OImageBuffer *mask, OImageBuffer *mask2, int format,
                       OParams *params, void (*progress)(int, void *), void *ctx, int *cancelFlag,
                       const OOptions *opt, OStats *stats);

unsigned oracle_count_pixelels(int format);
int oracle_format_indices(OFormatIndices *ix, int format);
void oracle_prepare_indices(OFormatIndices *ix, unsigned colors, unsigned maps,
                            int alphaTarget, int alphaSource, int isMap);

#ifdef __cplusplus
}
#endif
#endif
