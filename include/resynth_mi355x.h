/*
 * resynth_mi355x.h -- C ABI of libresynth_mi355x.so, an MI355X (gfx950) drop-in for
 * the resynthesizer synthesis library (reference: lib/ of bootchk/resynthesizer).
 *
 * Part 1 re-declares the reference's public types and entry points with identical
 * layouts, names, argument meaning and error codes, so that a caller compiled against
 * the reference's headers links against this library unchanged.
 * Part 2 declares extensions with no counterpart in the reference: mode selection,
 * device-resident inputs (for callers whose images already live in HBM), run
 * statistics, and the frozen-state scorer used by the parity tests.
 *
 * All entry points are synchronous and reentrant per call; calls on one process share
 * one HIP stream per device and are serialized by the library.
 */
#ifndef RESYNTH_MI355X_H
#define RESYNTH_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================= Part 1 */

/* lib/imageBuffer.h:12-18 -- rows padded to rowBytes, pixels unpadded */
typedef struct _ImageBuffer {
  unsigned char *data;
  unsigned int width;
  unsigned int height;
  size_t rowBytes;
} ImageBuffer;

/* lib/imageFormat.h:36-42 */
typedef enum ImageFormat { T_RGB, T_RGBA, T_Gray, T_GrayA } TImageFormat;

/* lib/engineParams.h:13-28 */
typedef enum ImageSynthError {
  IMAGE_SYNTH_SUCCESS,
  IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT,
  IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH,
  IMAGE_SYNTH_ERROR_PATCH_SIZE_EXCEEDED,
  IMAGE_SYNTH_ERROR_MATCH_CONTEXT_TYPE_RANGE,
  IMAGE_SYNTH_ERROR_EMPTY_TARGET,
  IMAGE_SYNTH_ERROR_EMPTY_CORPUS
} TImageSynthError;

/* lib/engineParams.h:31-86 */
typedef struct ImageSynthParametersStruct {
  int isMakeSeamlesslyTileableHorizontally;
  int isMakeSeamlesslyTileableVertically;
  int matchContextType;
  double mapWeight;
  double sensitivityToOutliers;
  unsigned int patchSize;
  unsigned int maxProbeCount;
} TImageSynthParameters;

/* lib/imageFormatIndicies.h:39-58 */
typedef unsigned char TPixelelIndex;
typedef struct indicesStruct {
  TPixelelIndex colorEndBip;
  TPixelelIndex alpha_bip;
  TPixelelIndex map_start_bip;
  TPixelelIndex map_end_bip;
  TPixelelIndex img_match_bpp;
  TPixelelIndex map_match_bpp;
  TPixelelIndex total_bpp;
  int isAlphaTarget;
  int isAlphaSource;
} TFormatIndices;

/* lib/map.h:28-46.  GArray is GLib's public prefix {data, len} (lib/glibProxy.h:86-92);
 * callers that include <glib.h> first get GLib's own declaration. */
#ifndef __G_ARRAY_H__
typedef struct _GArray {
  char *data;
  unsigned int len;
} GArray;
#endif
typedef struct {
  unsigned int width;
  unsigned int height;
  unsigned int depth;
  GArray *data;
} Map;
typedef struct {
  int x;
  int y;
} Coordinates;
typedef unsigned char Pixelel;

/* lib/imageSynth.h:32-40 (replaces lib/imageSynth.c:62-137) */
int imageSynth(ImageBuffer *imageBuffer, ImageBuffer *mask, TImageFormat imageFormat,
               TImageSynthParameters *parameters, void (*progressCallback)(int, void *),
               void *contextInfo, int *cancelFlag);

/* lib/imageSynth.h:43-52 (replaces lib/imageSynth.c:140-217) */
int imageSynth2(ImageBuffer *imageBuffer, ImageBuffer *mask, ImageBuffer *mask2, TImageFormat imageFormat,
                TImageSynthParameters *parameters, void (*progressCallback)(int, void *),
                void *contextInfo, int *cancelFlag);

/* lib/engine.h:4-12 (replaces lib/engine.c:539-690) */
int engine(TImageSynthParameters parameters, TFormatIndices *indices, Map *targetMap, Map *corpusMap,
           void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag);

/* lib/engineParams.h:93-95 (replaces lib/engineParams.c:9-18) */
void setDefaultParams(TImageSynthParameters *param);

/* lib/imageFormatIndicies.h:60-84 (replace lib/imageFormat.c:40-207) */
unsigned int countPixelelsPerPixelForFormat(TImageFormat format);
int prepareImageFormatIndicesFromFormatType(TFormatIndices *indices, TImageFormat format);
void prepareImageFormatIndices(TFormatIndices *indices, unsigned int count_color_channels_target,
                               unsigned int count_color_channels_map, int is_alpha_target,
                               int is_alpha_source, int isMap);
void prepareDefaultFormatIndices(TFormatIndices *formatIndices);

/* lib/map.h:49-99 (replace lib/mapOps.h:45-168) -- maps for engine() callers */
void free_map(Map *map);
void new_pixmap(Map *map, unsigned int width, unsigned int height, unsigned int depth);
void new_bytemap(Map *map, unsigned int width, unsigned int height);
void new_intmap(Map *map, unsigned int width, unsigned int height);
void new_coordmap(Map *map, unsigned int width, unsigned int height);
void set_bytemap(Map *map, unsigned char value);
void invert_bytemap(Map *map);
void interleave_mask(Map *pixmap, Map *mask);

/* ======================================================================= Part 2 */

/* Return codes beyond TImageSynthError.  No counterpart in the reference, whose engine
 * runs on the host and has no device to fail.  Every entry point returns one of these
 * instead of terminating the caller's process:
 *   RESYNTH_ERROR_INVALID_ARGUMENT: a Part 2 argument is invalid (null pointer, a row
 *     pitch shorter than the row, a kernel shape that does not exist, ...);
 *   RESYNTH_ERROR_DEVICE: no usable HIP device, or a HIP call failed (out of device
 *     memory, a failed launch); the caller's image is left as it was unless the failure
 *     came during the final write-back;
 *   RESYNTH_ERROR_TOO_LARGE: an image exceeds the library's size limits (fewer than 2^31
 *     pixels per image, sides below 2^24, fewer than 2^32 sorted offsets), or the
 *     allocation-free entry's maximum. */
#define RESYNTH_ERROR_INVALID_ARGUMENT (-1)
#define RESYNTH_ERROR_DEVICE 100
#define RESYNTH_ERROR_TOO_LARGE 101

/* Caps the device memory the library's grow-only engine buffers may hold (bytes; 0 = no
 * cap, the default).  An engine call whose buffers would exceed it fails with
 * RESYNTH_ERROR_DEVICE, as it would when hipMalloc runs out of memory.  For callers that
 * share a GPU with other work, and for tests of the out-of-memory path.  A cap below what
 * the library holds releases its idle engine buffers. */
void resynth_set_device_memory_limit(size_t bytes);

/* Synthesis modes.  No counterpart in the reference (which has one serial build and
 * one nondeterministic threaded build, lib/buildSwitches.h:79).
 *   RESYNTH_MODE_SERIAL:  the reference's serial visiting order and PRNG stream,
 *                         byte-identical output to the reference's serial build.
 *   RESYNTH_MODE_BATCHED: targets in conflict-tolerant batches, one wavefront per
 *                         target, per-target counter-based random probes; judged by
 *                         mean best-match error, like the reference's threaded build.
 * The default for Part 1 entry points is RESYNTH_MODE_BATCHED unless the environment
 * variable RESYNTH_MODE is "serial", or resynth_set_default_mode() was called.  Their PRNG
 * is GRand unless the environment variable RESYNTH_PRNG is "libc". */
enum { RESYNTH_MODE_SERIAL = 0, RESYNTH_MODE_BATCHED = 1 };

/* PRNG of the serial mode: GLib GRand (the GIMP build) or the C library's rand()
 * (the stand-alone build through lib/glibProxy.c). */
enum { RESYNTH_PRNG_GRAND = 0, RESYNTH_PRNG_LIBC = 1 };

/* Learned-proposal network (opaque; see resynth_model_* below).  No counterpart in the
 * reference: an extension of the batched mode that proposes candidates from corpus
 * embeddings so that most random probes can be cut; the exact scorer still judges. */
typedef struct ResynthModel ResynthModel;

typedef struct {
  int mode;            /* RESYNTH_MODE_* */
  int prng;            /* RESYNTH_PRNG_* (serial mode) */
  unsigned int seed;   /* 0: the engine's seed 1198472.  Batched mode: seed of the per-target
                          probe streams.  Serial mode: a nonzero seed replaces the engine's
                          seed (for statistics over seeds; not a reference behaviour) */
  unsigned int batch;  /* batched mode: largest number of concurrent target visits
                          (0 = automatic: nT/32 clamped to [64, 65536]) */
  int timing;          /* nonzero: time every synthesis kernel with HIP events */
  const ResynthModel *model;  /* batched mode: learned candidate proposals from this network
                                 (NULL: the plain method) */
  unsigned int learnedProbes; /* random probes per visit that has proposals (0 = 32); visits
                                 of the first, small pass-0 chunks keep maxProbeCount */
  unsigned int learnedPasses; /* passes whose visits get proposals (0 = all) */
  unsigned int learnedClusters; /* inverted-file clusters searched per query: 1 or 2 (0 = 2) */
  unsigned int learnedProbesLater; /* random probes per visit in passes without proposals when a
                                      model is set (0 = maxProbeCount) */
  unsigned int learnedRefresh;  /* 0: proposals for every chunk from its targets' current
                                   neighbourhoods; else a bit mask of passes at whose start the
                                   proposals of all targets are computed at once, then reused
                                   by every later visit (passes before the first set bit have
                                   none) */
  unsigned int maxPasses;       /* batched mode: stop after this many passes even if the
                                   refiner's rule would go on (0 = the rule alone) */
  unsigned int rampPermille;    /* batched mode: pass-0 chunks hold this many permille of the
                                   targets visited so far (0 = 60) */
  unsigned int lanesPerTarget;  /* batched mode: 0 = the library's choice (large chunks: wavefronts
                                   split into segments of 64, 32 or 16 lanes by the random-probe
                                   count, one target per segment); 16, 32 or 64 = every chunk
                                   visited with segments of that many lanes; 1 = every chunk by
                                   one wavefront per target.  A target's result does not depend on
                                   it: the choice trades issue slots against latency. */
  unsigned int firstChunk;      /* batched mode: targets in each of pass 0's first chunks (0 = automatic:
                                   nT/2048 clamped to [1, 64]).  Larger first chunks take fewer, fuller
                                   launches at the start of pass 0, where every visit scans far, at some
                                   cost in match error (their visits do not see each other). */
} ResynthOptions;

typedef struct {
  unsigned int nTargets, nCorpus;
  unsigned int nOffsets;              /* sorted offsets built: the prefix every neighbour scan of the run
                                         can reach (all (2w-1)(2h-1) when wrapping, or when the corpus is
                                         smaller than the target) */
  unsigned int passes;
  unsigned int passSize[6];
  unsigned int betters[6];
  unsigned long long visits;          /* target visits */
  unsigned long long candidates;      /* scorer invocations (computeBestFit calls) */
  unsigned long long patchTerms;      /* sum over scorer invocations of the patch size */
  double meanError;                   /* mean best-match sum of each target's last visit */
  unsigned int launches;              /* synthesis kernel launches */
  double synthKernelMs;               /* sum of synthesis kernel durations (timing != 0) */
  double prepMs, synthMs, totalMs;    /* host wall clock of the phases */
  unsigned int pixelWordBytes;        /* 4 or 8: bytes gathered per corpus pixel */
  unsigned int windowLaunches;        /* batched: persistent sliding-window launches */
  unsigned int windowWaitsExpired;    /* batched: window waits cut short by their spin limit
                                         (0 unless the device misbehaves) */
  unsigned int waveLaunches;          /* batched: launches of the wavefront-per-target kernel */
  double waveKernelMs;                /* their summed durations (timing != 0) */
  unsigned long long wavePatchTerms;  /* patch terms they scored */
  unsigned long long waveScanned;     /* patch offsets their neighbour scans read */
  unsigned long long waveStaged;      /* neighbours they staged for scoring */
  unsigned long long waveVisits;      /* their target visits */
  unsigned int learnedLaunches;       /* launches whose visits scored learned proposals */
  unsigned int learnedKeys;           /* corpus keys embedded (0: no proposals) */
  unsigned int learnedClusters;       /* inverted-file clusters */
  double learnedIndexMs;              /* host wall clock of embedding the corpus and indexing */
  double learnedProposeMs;            /* proposal kernels (features, embedding, search; timing != 0) */
  unsigned long long learnedKeysScored; /* query-key dot products of the fine search */
  unsigned long long scratchBytes;    /* resynth_image_synth_device_noalloc: device scratch the
                                         run used (0 for the other entries) */
} ResynthStats;

void resynth_set_default_mode(int mode);
void resynth_default_options(ResynthOptions *opt);

/* imageSynth / imageSynth2 with explicit options and statistics (stats may be NULL). */
int imageSynthEx(ImageBuffer *imageBuffer, ImageBuffer *mask, ImageBuffer *mask2, TImageFormat imageFormat,
                 TImageSynthParameters *parameters, void (*progressCallback)(int, void *), void *contextInfo,
                 int *cancelFlag, const ResynthOptions *options, ResynthStats *stats);

/* engine() with explicit options and statistics. */
int engineEx(TImageSynthParameters parameters, TFormatIndices *indices, Map *targetMap, Map *corpusMap,
             void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag,
             const ResynthOptions *options, ResynthStats *stats);

/* imageSynth on device-resident buffers: image (rows of rowBytes, synthesized in place),
 * mask (rows of maskRowBytes), optional mask2 (corpus mask as in imageSynth2, else NULL),
 * all device pointers on the library's device.  Same semantics and error codes as
 * imageSynth2/imageSynth.  For callers whose images already live in HBM. */
int resynth_image_synth_device(unsigned char *d_image, unsigned int width, unsigned int height, size_t rowBytes,
                               const unsigned char *d_mask, size_t maskRowBytes,
                               const unsigned char *d_mask2, size_t mask2RowBytes,
                               TImageFormat imageFormat, const TImageSynthParameters *parameters,
                               void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag,
                               const ResynthOptions *options, ResynthStats *stats);

/* resynth_image_synth_device without any memory allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and error codes as
 * resynth_image_synth_device, serial and batched modes, plain method only (options->model
 * must be NULL: RESYNTH_ERROR_INVALID_ARGUMENT).  Every device buffer of the run is carved
 * from a fixed RESYNTH_NOALLOC_SCRATCH_MIB MiB array in the library's device image, and
 * every host staging buffer from a fixed array in its host image; images above
 * RESYNTH_NOALLOC_MAX_PIXELS pixels return RESYNTH_ERROR_TOO_LARGE.  The first call on a
 * device looks up that device's copy of the array (and, if it is the library's first call,
 * sets up its stream); later calls allocate nothing.  All calls on one device share the
 * array, which keeps the last call's contents (each run initialises what it reads); calls
 * are serialized by the library's lock, and a caller that shares the array's device with
 * another process's copy of the library shares nothing.  options->timing is ignored. */
#define RESYNTH_NOALLOC_MAX_PIXELS (1u << 20)
#define RESYNTH_NOALLOC_SCRATCH_MIB 192
int resynth_image_synth_device_noalloc(unsigned char *d_image, unsigned int width, unsigned int height,
                                       size_t rowBytes, const unsigned char *d_mask, size_t maskRowBytes,
                                       const unsigned char *d_mask2, size_t mask2RowBytes,
                                       TImageFormat imageFormat, const TImageSynthParameters *parameters,
                                       void (*progressCallback)(int, void *), void *contextInfo,
                                       int *cancelFlag, const ResynthOptions *options, ResynthStats *stats);

/* ---- Many requests in one launch.  No counterpart in the reference (each imageSynth call
 * there is one request); for servers healing many images of uneven sizes at once.  Every
 * request is an imageSynth call (same contract, mask semantics and error codes), all run
 * in batched mode by one persistent multi-request launch, captured in a HIP graph. ---- */
typedef struct {
  ImageBuffer image;              /* pixels synthesized in place (host or device memory) */
  ImageBuffer mask;               /* image-sized selection; nonzero = synthesize */
  TImageFormat format;
  TImageSynthParameters params;
  int error;                      /* out: what imageSynth would return for this request */
  unsigned int nTargets;          /* out */
  unsigned int passes;            /* out */
  double meanError;               /* out: mean best-match sum of the targets' last visits */
  unsigned long long candidates;  /* out: scorer invocations */
} ResynthRequest;

typedef struct {
  int wavesPerBlock;              /* wavefronts per workgroup: 2, 4 or 8 */
  int firstGroup;                 /* corpus loads issued together for a candidate's first
                                     neighbours: 2 (0 = 2) */
  int laterGroup;                 /* loads issued together for the rest: 4 (0 = 4) */
  int lanesPerTarget;             /* lanes of a wavefront per target visit, i.e. candidates scored
                                     at once per target: 16, 32 or 64 (4, 2 or 1 targets per
                                     wavefront at a time; 0 = 64) */
} ResynthKernelConfig;

typedef struct {
  unsigned int requests;          /* requests synthesized (error 0) */
  unsigned long long targets;
  ResynthKernelConfig config;     /* kernel shape used */
  unsigned long long candidates, patchTerms;
  double prepMs;                  /* host wall clock of every request's preparation */
  double synthKernelMs;           /* the multi-request launch(es), HIP events */
  double totalMs;
  unsigned int graphLaunches;     /* one per pixel-word width present */
  unsigned int waitsExpired;      /* waits cut short by their spin limit (0 normally) */
  unsigned int graphInstantiations; /* graphs instantiated by this call: the instantiated graph
                                     of each pixel-word width is kept, and a call with its kernel
                                     shape only updates its nodes' arguments (0 then) */
} ResynthBatchStats;

/* Runs n requests.  deviceBuffers != 0: image.data and mask.data are device pointers.
 * options may be NULL (batched defaults; the mode field is ignored: always batched);
 * config may be NULL (the shape resynth_autotune_batch chose for the launch's class, else
 * 4 waves per block, 64 lanes per target).
 * Returns 0, RESYNTH_ERROR_INVALID_ARGUMENT, or RESYNTH_ERROR_DEVICE; per-request results
 * are in requests[i].error (RESYNTH_ERROR_INVALID_ARGUMENT for a request whose rowBytes is
 * shorter than width * bytes per pixel, or whose mask rowBytes is shorter than width) and
 * the other out fields.  Of the options, seed, batch, maxPasses and rampPermille apply per
 * request; mode, timing, model, lanesPerTarget, firstChunk and the learned* fields are
 * ignored (always the plain batched method, shaped by config). */
int resynth_image_synth_batch(ResynthRequest *requests, unsigned int n, int deviceBuffers,
                              const ResynthOptions *options, const ResynthKernelConfig *config,
                              ResynthBatchStats *stats);

/* Autotuner of the multi-request launch: times every kernel shape (wavesPerBlock {2,4,8},
 * i.e. targets per workgroup, x lanesPerTarget {16,32,64}, i.e. candidates per wavefront
 * step: 9 shapes), three times each on copies of the sample requests, and keeps the one
 * with the lowest median for later resynth_image_synth_batch calls of the same class, which
 * it returns in *best.  A class is a pixel-word width, a patch-size class (target-weighted
 * mean patchSize up to 16, up to 32, larger) and a size class (mean targets per request
 * below 16384, below 131072, larger); a launch of a class never tuned uses
 * {4, 2, 4, 64}.  msPerConfig (may be NULL) receives the 9 medians in the order waves
 * {2,4,8} x lanes {16,32,64}.  The sample's images are left unchanged.  Returns 0,
 * RESYNTH_ERROR_INVALID_ARGUMENT or RESYNTH_ERROR_DEVICE. */
int resynth_autotune_batch(const ResynthRequest *sample, unsigned int n, int deviceBuffers,
                           ResynthKernelConfig *best, double *msPerConfig);

/* Frozen-state scorer (the computeBestFit loop of lib/synthesize.h:266-400 over an
 * ordered candidate list, strict-less update, stop at the first perfect match) for
 * nProblems independent problems, evaluated on the GPU.  Host pointers:
 *   corpus   : pixmap [ch][cw][depth] in the engine's layout (mask byte first)
 *   T, Tm    : the 512-entry metric tables (lib/matchWeighting.h:195-204)
 *   nbCount  : [nProblems] neighbour counts (<= 64)
 *   nbOffset : [nProblems][64][2] neighbour offsets
 *   nbPixel  : [nProblems][64][8] neighbour pixels in the engine layout
 *   candCount: [nProblems]; cands: [nProblems][maxCands][2]
 * Outputs: winner[nProblems] (-1 if none accepted), bestSum[nProblems],
 *          stopped[nProblems] (candidates consumed before the stop).
 * Returns 0, or -1 on invalid arguments. */
int resynth_score_frozen(const unsigned char *corpus, unsigned int cw, unsigned int ch,
                         const TFormatIndices *indices, const uint16_t *T, const uint32_t *Tm,
                         unsigned int nProblems, const unsigned int *nbCount, const int32_t *nbOffset,
                         const unsigned char *nbPixel, const unsigned int *candCount, unsigned int maxCands,
                         const int32_t *cands, int32_t *winner, uint32_t *bestSum, int32_t *stopped);

/* Host-side pieces of the engine, exposed so that they can be checked without a GPU:
 * raw PRNG outputs and int_range draws of the engine's PRNG (seeded like lib/engine.c:643
 * when seed is 1198472), the metric tables (lib/matchWeighting.h:195-204), and the
 * target ordering of lib/orderTarget.h:269-343 applied in place to n (x,y) points. */
void resynth_prng_sample(int prng, unsigned int seed, unsigned int n, uint32_t *out);
void resynth_prng_ranges(int prng, unsigned int seed, unsigned int n, const uint32_t *dists, uint32_t *out);
void resynth_metric_tables(double sensitivity, double mapWeight, uint16_t *T, uint32_t *Tm);
int resynth_order_targets(int matchContextType, int prng, int32_t *points, unsigned int n);

/* ---- Learned candidate proposals (continued).  A small BF16 MLP
 * (FP32 accumulate) embeds a pixel's neighbourhood (the 30 nearest offsets: R, G, B / 255
 * and a presence flag each, 120 inputs zero-padded to 128) through 256 -> 256 -> 64 with
 * ReLU and L2 normalisation; it is trained with InfoNCE (shared weights for both towers)
 * by hand-written forward/backward GEMM kernels on the GPU. ---- */
/* Number of FP32 parameters: W1[128x256], b1[256], W2[256x256], b2[256], W3[256x64], b3[64],
 * weights row-major [in x out], in that order. */
unsigned int resynth_model_param_count(void);
/* He-uniform weights (seeded), zero biases; NULL without a device. */
ResynthModel *resynth_model_create(unsigned int seed);
void resynth_model_free(ResynthModel *model);
int resynth_model_get_params(const ResynthModel *model, float *out);
int resynth_model_set_params(ResynthModel *model, const float *in);   /* also resets Adam */
/* x: n rows of 128 bf16 bit patterns (host); out: n x 64 normalised embeddings (host). */
int resynth_model_embed(const ResynthModel *model, const uint16_t *x, unsigned int n, float *out);
/* InfoNCE loss over n pairs (n a multiple of 64; logits z_q z_k^T / tau, positives on the
 * diagonal) and its gradient with respect to the parameters (host arrays; grad may be NULL). */
int resynth_model_infonce(ResynthModel *model, const uint16_t *xq, const uint16_t *xk, unsigned int n, float tau,
                          float *loss, float *grad);
/* Adam (lr, betas 0.9/0.999) for `steps` steps on batches of `batch` pairs (a multiple of
 * 64) drawn by a seeded hash from the n host pairs; losses[steps] (may be NULL) gets each
 * step's loss. */
int resynth_model_train(ResynthModel *model, const uint16_t *xq, const uint16_t *xk, unsigned int n,
                        unsigned int steps, unsigned int batch, float lr, float tau, unsigned int seed,
                        float *losses);
/* The MFMA GEMM used by the network, on host arrays: C[M x N] (fp32) = op(A) op(B) with bf16
 * A ([M x K], or [K x M] when ta) and B ([K x N], or [N x K] when tb); K a multiple of 32,
 * M and N multiples of 8.  For tests. */
int resynth_gemm_bf16(int ta, int tb, int M, int N, int K, const uint16_t *A, const uint16_t *B, float *C);
/* The sorted offsets the engine builds on the device (lib/engine.c:465-497) for offset
 * ranges [-w+1, w-1] x [-h+1, h-1]: writes (2w-1)(2h-1) (x, y) pairs to out.  Returns 0, or
 * -1 on invalid arguments.  For tests. */
int resynth_sorted_offsets(int w, int h, int32_t *out);
/* The prefix of those sorted offsets with x^2 + y^2 <= maxD2, the part of the table an
 * engine run builds when no scan can reach farther: writes *count (x, y) pairs to out (room
 * for (2w-1)(2h-1) pairs suffices).  Returns 0, or -1 on invalid arguments.  For tests. */
int resynth_sorted_offsets_within(int w, int h, unsigned long long maxD2, int32_t *out, unsigned long long *count);
/* The learned index's query routing, on host arrays, through the same device sort and kernel
 * the engine uses.  queries: bf16 [nq x 64]; supers: [nq] super-cluster per query (a
 * negative one counts as 0); subCentroids: bf16 [S*Q x 64], super-cluster s owning rows
 * [s*Q, (s+1)*Q).  Writes clusters [nq x P]: the ids s*Q + c of the P (1 or 2) rows of the
 * query's super-cluster with the largest dot products, best first.  Returns 0, or -1 on
 * invalid arguments.  For tests. */
int resynth_learn_route(const uint16_t *queries, unsigned int nq, const int32_t *supers,
                        const uint16_t *subCentroids, unsigned int S, unsigned int Q, int P, int32_t *clusters);
/* The learned index's fine search, on host arrays, through the same device sort and kernel
 * the engine uses.  queries: bf16 [nq x 64]; clusters: [nq x P] cluster ids (-1: none),
 * P 1 or 2; keys: bf16 [nKeys x 64] in cluster order, cluster c holding keys
 * [start[c], start[c+1]) (start has nClusters + 1 entries, start[nClusters] = nKeys);
 * keyPts: [nKeys] (x, y).  Writes props [nq x 16] (x, y) pairs: 2 * keyPts of the query's
 * top-16 keys by dot product over its clusters, best first, (-1, -1) past the last.
 * Returns 0, or -1 on invalid arguments.  For tests. */
int resynth_learn_fine_top16(const uint16_t *queries, unsigned int nq, const int32_t *clusters, int P,
                             const uint16_t *keys, unsigned int nKeys, const uint32_t *start,
                             unsigned int nClusters, const int32_t *keyPts, int32_t *props);

/* Runs imageSynth in batched mode (the plain method, options' seed) on the image, which is
 * synthesized in place, and appends up to maxPairs training pairs to the model's store:
 * for targets drawn by a seeded hash, the target's neighbourhood in the result (each
 * neighbour kept with a per-pair probability in [0.2, 1), as a partly filled hole shows it)
 * and the neighbourhood of the corpus point the plain method chose for it.  Returns the
 * imageSynth error code (pairs are only logged on success). */
int resynth_model_log_pairs(ResynthModel *model, ImageBuffer *image, ImageBuffer *mask, TImageFormat format,
                            const TImageSynthParameters *params, unsigned int maxPairs, unsigned int seed);
unsigned int resynth_model_pair_count(const ResynthModel *model);
void resynth_model_clear_pairs(ResynthModel *model);
/* resynth_model_train on the logged pairs. */
int resynth_model_fit(ResynthModel *model, unsigned int steps, unsigned int batch, float lr, float tau,
                      unsigned int seed, float *losses);

/* sizeof of ResynthOptions, ResynthStats, ResynthRequest, ResynthKernelConfig and
 * ResynthBatchStats, then the offsets of ResynthRequest.params, ResynthRequest.error and
 * ResynthBatchStats.prepMs, into out[0..min(n, 8)); returns how many were written.  Lets
 * a binding check its mirror of these structs. */
unsigned int resynth_struct_sizes(size_t *out, unsigned int n);

/* ======================================================================= Part 3 */

/* The GIMP plug-in's engine call and the heal-selection helper, with GIMP's drawables
 * passed as plain buffers.  A GIMP build of the plug-in keeps its libgimp glue (PDB
 * registration, pixel regions, selection bounds, progress bar) and calls these in place of
 * its adapter and engine code.  Progress and cancel stay on the calling thread. */

/* Error of the plug-in's own checks (the plug-in returns a PDB execution error and a
 * message); *message then holds the reference's text. */
#define RESYNTH_ERROR_PLUGIN 102

/* A drawable as the plug-in reads it (src/resynthesizer/adaptGimp.h:200-322): colour
 * channels then an optional alpha channel, rows of rowBytes; and the image selection over
 * the drawable, one byte per pixel in rows of selectionRowBytes.  A NULL or all-zero
 * selection is no selection intersecting the drawable: the whole drawable is selected
 * (fetch_mask, src/resynthesizer/adaptGimp.h:178-256). */
typedef struct {
  unsigned char *pixels;
  unsigned int width;
  unsigned int height;
  size_t rowBytes;
  unsigned int colorChannels;   /* 3: RGB, 1: gray (count_color_channels) */
  int hasAlpha;                 /* gimp_drawable_has_alpha */
  const unsigned char *selection;
  size_t selectionRowBytes;
} ResynthDrawable;

/* src/resynth-parameters.h:28-42.  The drawable ids are not read: the drawables are passed
 * to resynth_plugin_run themselves. */
typedef struct GIMPAdapterParametersStruct {
  int h_tile;
  int v_tile;
  int use_border;
  int corpus_id, input_map_id, output_map_id;
  double map_weight;
  double autism;
  int neighbours;
  int trys;
} TGimpAdapterParameters;

/* The engine plug-in's run() (src/resynthesizer/resynthesizer.c:286-546) from its checks
 * on: image mode and base types, maps (used only when both are given; same mode, sizes of
 * the corpus and of the drawable), neighbours clamped to 64, format indices, the target
 * and corpus pixmaps with their masks and map channels (fetch_image_mask_map), engine()
 * with the caller's progress callback and cancel flag, and the result written back into the
 * drawable's pixels (colour and alpha channels; alpha is never changed).  corpus NULL:
 * the drawable itself.  engine() runs in the default mode (resynth_set_default_mode).
 * Returns 0, a TImageSynthError of engine(), RESYNTH_ERROR_PLUGIN, or a Part 2 code;
 * message (may be NULL) receives the plug-in's message for a failed check or an empty
 * target or corpus, else NULL. */
int resynth_plugin_run(const TGimpAdapterParameters *parameters, ResynthDrawable *drawable,
                       const ResynthDrawable *corpus, const ResynthDrawable *inputMap,
                       const ResynthDrawable *outputMap, void (*progressCallback)(int, void *),
                       void *contextInfo, int *cancelFlag, const char **message);

/* Heal selection (PluginScripts/plugin-heal-selection.py:36-152): the drawable's selection
 * is synthesized from a band around it.  The selection is grown by samplingRadius pixels
 * as GIMP grows selections (a disc), the original selection is subtracted, and the copy of
 * the drawable that serves as corpus is cropped to that band -- all of it (direction 0),
 * the rows of the selection (1: sides) or its columns (2: above and below).  The filling
 * order is random (order 0), inwards (1) or outwards (2), then the engine plug-in runs with
 * the script's parameters (no tiling, no maps, sensitivity 0.117, 16 neighbours, 500
 * probes).  Returns as resynth_plugin_run; an empty selection returns RESYNTH_ERROR_PLUGIN
 * ("You must first select a region to heal."), a direction or order out of range
 * RESYNTH_ERROR_INVALID_ARGUMENT. */
int resynth_heal_selection(ResynthDrawable *drawable, int samplingRadius, int direction, int order,
                           void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag,
                           const char **message);

/* Library/device information: returns 0 when a gfx950 device is usable. */
int resynth_device_info(char *name, size_t nameLen, int *computeUnits);
const char *resynth_version(void);

#ifdef __cplusplus
}
#endif
#endif
