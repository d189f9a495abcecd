/*
 * resynth_oracle.c -- CPU restatement of the resynthesizer engine (lib/ of the
 * reference).  TEST INFRASTRUCTURE ONLY: the checker for the GPU library and the
 * timed CPU baseline of bench.py.  Never linked into the shipped library.
 *
 * Written from the reference's source text; each function cites the file:line
 * it restates.  The reference itself is not buildable in this image (its serial
 * refiner is selected only by editing lib/buildSwitches.h:79, and the GLib build
 * needs the autotools-generated ../config.h, lib/engine.c:79), so it is pinned by:
 *   - GLib GRand / g_array_sort and glibc rand(): golden vectors written by
 *     tests/golden/make_golden.py from the system libraries via ctypes;
 *   - the metric tables: closed-form values (tests/test_oracle.py);
 *   - the whole engine: the known answers of src/testSynth.c:163-212, brute force
 *     on tiny images, and invariants (see tests/test_oracle.py).
 *
 * Third-party algorithms restated here (absent from the reference tree):
 *   GLib 2.72 GRand (MT19937 + g_rand_int_range, glib/grand.c), GLib g_array_sort
 *   (the glibc-derived top-down merge sort of glib/gqsort.c), glibc 2.35 random_r
 *   TYPE_3 (stdlib/random_r.c).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MAX_NEIGHBORS 64          /* lib/imageSynthConstants.h:26 */
#define MASK_TOTALLY_SELECTED 0xFF
#define MAX_WEIGHT 65535u          /* lib/imageSynthConstants.h:73 (G_MAXUSHORT) */
#define CALLBACK_MASK 4095u        /* lib/imageSynthConstants.h:99 */
#define ENGINE_SEED 1198472u       /* lib/engine.c:643 */

/* ------------------------------------------------------------------ PRNG */

/* GLib GRand: g_rand_set_seed (MT19937 init_genrand). */
void orc_grand_seed(OGRand *r, uint32_t seed)
{
  r->mt[0] = seed;
  for (int i = 1; i < 624; i++)
    r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
  r->mti = 624;
}

static void grand_twist(OGRand *r)
{
  uint32_t *mt = r->mt;
  int k;
  for (k = 0; k < 624 - 397; k++) {
    uint32_t y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
    mt[k] = mt[k + 397] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }
  for (; k < 623; k++) {
    uint32_t y = (mt[k] & 0x80000000u) | (mt[k + 1] & 0x7fffffffu);
    mt[k] = mt[k + 397 - 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }
  uint32_t y = (mt[623] & 0x80000000u) | (mt[0] & 0x7fffffffu);
  mt[623] = mt[396] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  r->mti = 0;
}

uint32_t orc_grand_int(OGRand *r)
{
  if (r->mti >= 624) grand_twist(r);
  uint32_t y = r->mt[r->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

/* GLib >= 2.2 g_rand_int_range: unbiased rejection of the top remainder. */
int32_t orc_grand_int_range(OGRand *r, int32_t begin, int32_t end)
{
  uint32_t dist = (uint32_t)end - (uint32_t)begin;
  uint32_t v = 0;
  if (dist != 0) {
    uint32_t maxvalue;
    if (dist <= 0x80000000u) {
      uint32_t leftover = (0x80000000u % dist) * 2u;
      if (leftover >= dist) leftover -= dist;
      maxvalue = 0xffffffffu - leftover;
    } else {
      maxvalue = dist - 1u;
    }
    do v = orc_grand_int(r); while (v > maxvalue);
    v %= dist;
  }
  return (int32_t)((uint32_t)begin + v);
}

/* glibc srandom_r for TYPE_3 (degree 31, separation 3) */
void orc_libc_srand(OLibcRand *r, uint32_t seed)
{
  if (seed == 0) seed = 1;
  r->s[0] = (int32_t)seed;
  long word = (int32_t)seed;
  for (int i = 1; i < 31; i++) {
    long hi = word / 127773, lo = word % 127773;
    word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r->s[i] = (int32_t)word;
  }
  r->f = 3;
  r->r = 0;
  for (int i = 0; i < 310; i++) (void)orc_libc_rand(r);
}

/* glibc random_r: additive lagged Fibonacci, low bit discarded */
int32_t orc_libc_rand(OLibcRand *r)
{
  uint32_t v = (uint32_t)r->s[r->f] + (uint32_t)r->s[r->r];
  r->s[r->f] = (int32_t)v;
  int32_t out = (int32_t)(v >> 1);
  if (++r->f >= 31) { r->f = 0; ++r->r; }
  else if (++r->r >= 31) r->r = 0;
  return out;
}

/* lib/glibProxy.c:37-49 s_rand_int_range, unsigned arithmetic as written there */
uint32_t orc_libc_int_range(OLibcRand *r, uint32_t lo, uint32_t hi)
{
  if (hi < 1) return 0;
  hi -= 1;
  return lo + (uint32_t)orc_libc_rand(r) / (2147483647u / (hi - lo + 1u) + 1u);
}

typedef struct { int kind; OGRand g; OLibcRand c; uint64_t draws; } Prng;

static void prng_init(Prng *p, int kind, uint32_t seed)
{
  p->kind = kind;
  p->draws = 0;
  if (kind == ORC_PRNG_LIBC) orc_libc_srand(&p->c, seed);
  else orc_grand_seed(&p->g, seed);
}

static uint32_t prng_range(Prng *p, uint32_t n)
{
  p->draws++;
  if (p->kind == ORC_PRNG_LIBC) return orc_libc_int_range(&p->c, 0, n);
  return (uint32_t)orc_grand_int_range(&p->g, 0, (int32_t)n);
}

/* ------------------------------------------------------------------ sort */

/* Top-down merge sort with the split and merge rule of GLib's g_qsort_with_data
 * (glibc msort_with_tmp): left half n/2, take from the left iff cmp <= 0. */
static void msort_rec(char *b, size_t n, size_t s, int (*cmp)(const void *, const void *), char *tmp)
{
  if (n <= 1) return;
  size_t n1 = n / 2, n2 = n - n1;
  char *b1 = b, *b2 = b + n1 * s, *t = tmp;
  msort_rec(b1, n1, s, cmp, tmp);
  msort_rec(b2, n2, s, cmp, tmp);
  while (n1 > 0 && n2 > 0) {
    if (cmp(b1, b2) <= 0) { memcpy(t, b1, s); b1 += s; n1--; }
    else { memcpy(t, b2, s); b2 += s; n2--; }
    t += s;
  }
  if (n1 > 0) memcpy(t, b1, n1 * s);
  memcpy(b, tmp, (n - n2) * s);
}

void orc_merge_sort(void *base, size_t n, size_t size, int (*cmp)(const void *, const void *))
{
  if (n <= 1) return;
  char *tmp = (char *)malloc(n * size);
  msort_rec((char *)base, n, size, cmp, tmp);
  free(tmp);
}

/* lib/engineTypes.h:51-56 to_sort_result and :73-80 lessCartesian: never 0 */
static int cmp_less_cartesian(const void *pa, const void *pb)
{
  const int32_t *a = (const int32_t *)pa, *b = (const int32_t *)pb;
  return (a[1] * a[1] + a[0] * a[0] < b[1] * b[1] + b[0] * b[0]) ? -1 : 1;
}
/* lib/engineTypes.h:113-141 horizontal / vertical, and their inversions (:58-62) */
static int cmp_less_horizontal(const void *pa, const void *pb)
{
  const int32_t *a = (const int32_t *)pa, *b = (const int32_t *)pb;
  return (a[0] * a[0] < b[0] * b[0]) ? -1 : 1;
}
static int cmp_more_horizontal(const void *a, const void *b) { return cmp_less_horizontal(a, b) == 1 ? -1 : 1; }
static int cmp_less_vertical(const void *pa, const void *pb)
{
  const int32_t *a = (const int32_t *)pa, *b = (const int32_t *)pb;
  return (a[1] * a[1] < b[1] * b[1]) ? -1 : 1;
}
static int cmp_more_vertical(const void *a, const void *b) { return cmp_less_vertical(a, b) == 1 ? -1 : 1; }

typedef struct { int32_t x, y; float prop; } SortElem;   /* lib/engineTypes.h:23-26 */
static int cmp_less_inward(const void *pa, const void *pb)
{
  const SortElem *a = (const SortElem *)pa, *b = (const SortElem *)pb;
  return (a->prop < b->prop) ? -1 : 1;
}
static int cmp_more_inward(const void *a, const void *b) { return cmp_less_inward(a, b) == 1 ? -1 : 1; }

/* lib/engine.c:465-497 prepareSortedOffsets */
void orc_sorted_offsets(int w, int h, int32_t *out)
{
  size_t k = 0;
  for (int y = -h + 1; y < h; y++)
    for (int x = -w + 1; x < w; x++) { out[k++] = x; out[k++] = y; }
  orc_merge_sort(out, k / 2, 2 * sizeof(int32_t), cmp_less_cartesian);
}

/* ------------------------------------------------------------------ tables */

/* lib/matchWeighting.h:50-60, 143-159, 167-178, 195-204.  The parameters reach
 * quantizeMetricFuncs as gfloat, so both are narrowed to float first. */
void orc_quantize(double sensitivity, double mapWeight, uint16_t T[512], uint32_t Tm[512])
{
  float c = (float)sensitivity;
  float mw = (float)mapWeight;
  float scale = c * 256.0f;               /* cauchyParam*LIMIT_DOMAIN in float */
  double denom = log(((double)256 / scale) * ((double)256 / scale) + 1.0);
  for (int i = -256; i < 256; i++) {
    double num = log(((double)i / scale) * ((double)i / scale) + 1.0);
    double v = num / denom * (double)(float)MAX_WEIGHT;
    T[256 + i] = (uint16_t)v;
    Tm[256 + i] = (uint32_t)((float)(i * i) * mw * 4.0);
  }
}

/* lib/passes.h:68-93 */
void orc_repetition(unsigned n, unsigned sizes[6])
{
  sizes[0] = n;
  for (int i = 1; i < 6; i++) {
    sizes[i] = n;
    n = (unsigned)n * 3 / 4;
  }
}

/* ------------------------------------------------------------------ format */

/* lib/imageFormat.c:40-56 */
unsigned oracle_count_pixelels(int format)
{
  switch (format) {
  case 0: return 3;
  case 1: return 4;
  case 2: return 1;
  case 3: return 2;
  default: return 0;
  }
}

/* lib/imageFormat.c:117-207 */
void oracle_prepare_indices(OFormatIndices *ix, unsigned colors, unsigned maps,
                            int alphaTarget, int alphaSource, int isMap)
{
  ix->img_match_bpp = (uint8_t)colors;
  ix->colorEndBip = (uint8_t)(1 + colors);
  if (alphaTarget || alphaSource) {
    ix->alpha_bip = ix->colorEndBip;
    ix->map_start_bip = (uint8_t)(ix->colorEndBip + 1);
  } else {
    ix->alpha_bip = 0;  /* undefined in the reference; never read */
    ix->map_start_bip = ix->colorEndBip;
  }
  ix->map_match_bpp = isMap ? (uint8_t)maps : 0;
  ix->map_end_bip = (uint8_t)(ix->map_start_bip + ix->map_match_bpp);
  ix->total_bpp = ix->map_end_bip;
  ix->isAlphaTarget = alphaTarget;
  ix->isAlphaSource = alphaSource;
}

/* lib/imageFormat.c:66-113 */
int oracle_format_indices(OFormatIndices *ix, int format)
{
  switch (format) {
  case 0: oracle_prepare_indices(ix, 3, 0, 0, 0, 0); return 0;
  case 1: oracle_prepare_indices(ix, 3, 0, 1, 1, 0); return 0;
  case 2: oracle_prepare_indices(ix, 1, 0, 0, 0, 0); return 0;
  case 3: oracle_prepare_indices(ix, 1, 0, 1, 1, 0); return 0;
  default: return 1;  /* IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT */
  }
}

/* ------------------------------------------------------------------ engine state */

typedef struct { uint8_t pix[8]; int32_t ox, oy, sx, sy; } Neighbor;   /* lib/synthesize.h:120-125 */

typedef struct {
  OParams P;
  const OFormatIndices *ix;
  OPixmap *tgt;
  const OPixmap *cor;
  uint8_t *hasValue;      /* target W*H */
  int32_t *sourceOf;      /* target W*H*2 */
  uint32_t *prober;       /* corpus W*H */
  int32_t *targets;       /* N_T*2, in synthesis order */
  int32_t *corpusPts;     /* N_C*2 */
  int32_t *offsets;       /* N_O*2 */
  unsigned nT, nC, nO;
  uint16_t T[512];
  uint32_t Tm[512];
  uint32_t *lastBest;     /* target W*H, may be NULL */
  /* what neighbour gathers read: the live maps, or chunk-start snapshots (batch model) */
  const uint8_t *rdHas;
  const uint8_t *rdTgt;
  const int32_t *rdSrc;
  const unsigned *kmap;   /* batch model: order index of each visited slot, else NULL */
  /* progress (lib/progress.c:53-65) */
  void (*progress)(int, void *);
  void *ctx;
  int *cancel;
  unsigned estimated, completed, priorPercent, progressCalls;
  pthread_mutex_t mutex;
  unsigned long long scorerCalls, visits;
} Engine;

/* lib/progress.c:53-65 deepProgressCallback */
static void deep_progress(Engine *E)
{
  E->completed += CALLBACK_MASK;
  unsigned pct = (unsigned)(((float)E->completed / (float)E->estimated) * 100);
  if (pct > E->priorPercent) {
    if (E->progress) E->progress((int)pct, E->ctx);
    E->priorPercent = pct;
    E->progressCalls++;
  }
}

/* lib/synthesize.h:81-113 */
static int clip_or_wrap(const Engine *E, int32_t *px, int32_t *py)
{
  int W = (int)E->tgt->width, H = (int)E->tgt->height;
  while (*px < 0) { if (E->P.tileH) *px += W; else return 0; }
  while (*px >= W) { if (E->P.tileH) *px -= W; else return 0; }
  while (*py < 0) { if (E->P.tileV) *py += H; else return 0; }
  while (*py >= H) { if (E->P.tileV) *py -= H; else return 0; }
  return 1;
}

/* lib/engine.c:505-517 clippedOrMaskedCorpus (mask byte is pixelel 0) */
static int clipped_or_masked(const Engine *E, int32_t x, int32_t y)
{
  const OPixmap *c = E->cor;
  if (x < 0 || y < 0 || x >= (int)c->width || y >= (int)c->height) return 1;
  return c->data[((size_t)x + (size_t)y * c->width) * c->depth] != MASK_TOTALLY_SELECTED;
}

/* lib/synthesize.h:158-178 new_neighbor */
static void add_neighbor(Engine *E, Neighbor *nb, int32_t ox, int32_t oy, int32_t px, int32_t py)
{
  size_t p = (size_t)px + (size_t)py * E->tgt->width;
  nb->ox = ox; nb->oy = oy;
  pthread_mutex_lock(&E->mutex);
  nb->sx = E->rdSrc[2 * p];
  nb->sy = E->rdSrc[2 * p + 1];
  memset(nb->pix, 0, 8);
  memcpy(nb->pix, E->rdTgt + p * E->tgt->depth, E->ix->total_bpp);
  pthread_mutex_unlock(&E->mutex);
}

/* lib/synthesize.h:189-241 prepare_neighbors */
static unsigned prepare_neighbors(Engine *E, int32_t tx, int32_t ty, Neighbor *nb)
{
  unsigned count = 0;
  add_neighbor(E, &nb[count++], E->offsets[0], E->offsets[1], tx, ty);
  for (unsigned j = 1; j < E->nO; j++) {
    int32_t ox = E->offsets[2 * j], oy = E->offsets[2 * j + 1];
    int32_t px = tx + ox, py = ty + oy;
    if (clip_or_wrap(E, &px, &py) && E->rdHas[(size_t)px + (size_t)py * E->tgt->width]) {
      add_neighbor(E, &nb[count++], ox, oy, px, py);
      if (count >= E->P.patchSize) break;
    }
  }
  return count;
}

/* lib/synthesize.h:266-400 computeBestFit.  Returns 1 on a perfect match. */
static int compute_best_fit(Engine *E, int32_t cx, int32_t cy, uint32_t *best, int32_t *bx, int32_t *by,
                            unsigned count, const Neighbor *nb, int *bettered)
{
  const OFormatIndices *ix = E->ix;
  const OPixmap *c = E->cor;
  uint32_t sum = 0;
  __atomic_fetch_add(&E->scorerCalls, 1ull, __ATOMIC_RELAXED);
  for (unsigned i = 0; i < count; i++) {
    int32_t qx = cx + nb[i].ox, qy = cy + nb[i].oy;
    if (clipped_or_masked(E, qx, qy)) {
      sum += MAX_WEIGHT * ix->img_match_bpp + E->Tm[0] * ix->map_match_bpp;
    } else {
      const uint8_t *cp = c->data + ((size_t)qx + (size_t)qy * c->width) * c->depth;
      const uint8_t *tp = nb[i].pix;
      if (i) {
        for (unsigned j = 1; j < ix->colorEndBip; j++) sum += E->T[256u + tp[j] - cp[j]];
      }
      if (ix->map_match_bpp > 0) {
        for (unsigned j = ix->map_start_bip; j < ix->map_end_bip; j++) sum += E->Tm[256u + tp[j] - cp[j]];
      }
    }
    if (sum >= *best) return 0;
  }
  *best = sum;
  *bettered = 1;
  *bx = cx;
  *by = cy;
  return sum == 0;
}

/* lib/synthesize.h:426-642 synthesize (one thread's stride over a pass prefix) */
static unsigned synthesize(Engine *E, unsigned start, unsigned threadIndex, unsigned stride, unsigned end,
                           Prng *prng)
{
  Neighbor nb[MAX_NEIGHBORS];
  unsigned betters = 0;
  const OFormatIndices *ix = E->ix;
  size_t W = E->tgt->width;
  for (unsigned k = start + threadIndex % stride; k < end; k += stride) {
    if ((k & CALLBACK_MASK) == 0) {
      deep_progress(E);
      if (*E->cancel) break;
    }
    __atomic_fetch_add(&E->visits, 1ull, __ATOMIC_RELAXED);
    int32_t tx = E->targets[2 * k], ty = E->targets[2 * k + 1];
    unsigned count = prepare_neighbors(E, tx, ty, nb);
    uint32_t best = UINT32_MAX;
    int32_t bx = 0, by = 0;
    int perfect = 0, bettered = 0;

    /* heuristic 1: continue the sources of the neighbours (:537-580) */
    for (unsigned j = 0; j < count && best != 0; j++) {
      if (nb[j].sx == -1) continue;
      int32_t cx = nb[j].sx - nb[j].ox, cy = nb[j].sy - nb[j].oy;
      if (clipped_or_masked(E, cx, cy)) continue;
      uint32_t *rp = &E->prober[(size_t)cx + (size_t)cy * E->cor->width];
      unsigned tidx = E->kmap ? E->kmap[k] : k;
      if (*rp == tidx) continue;
      perfect = compute_best_fit(E, cx, cy, &best, &bx, &by, count, nb, &bettered);
      if (perfect) break;
      *rp = tidx;
    }
    /* random corpus probes (:583-604) */
    if (!perfect) {
      for (unsigned j = 0; j < E->P.maxProbeCount; j++) {
        uint32_t idx = prng_range(prng, E->nC);
        perfect = compute_best_fit(E, E->corpusPts[2 * idx], E->corpusPts[2 * idx + 1], &best, &bx, &by,
                                   count, nb, &bettered);
        if (perfect) break;
      }
    }
    /* store (:620-639) */
    size_t p = (size_t)tx + (size_t)ty * W;
    if (bettered) {
      if (!(E->sourceOf[2 * p] == bx && E->sourceOf[2 * p + 1] == by)) {
        betters++;
        pthread_mutex_lock(&E->mutex);
        const uint8_t *src = E->cor->data + ((size_t)bx + (size_t)by * E->cor->width) * E->cor->depth;
        uint8_t *dst = E->tgt->data + p * E->tgt->depth;
        for (unsigned j = 1; j < ix->colorEndBip; j++) dst[j] = src[j];
        E->sourceOf[2 * p] = bx;
        E->sourceOf[2 * p + 1] = by;
        pthread_mutex_unlock(&E->mutex);
      }
    }
    if (E->lastBest) E->lastBest[p] = best;
    E->hasValue[p] = 1;
  }
  return betters;
}

typedef struct { Engine *E; unsigned idx, stride, end; Prng prng; unsigned betters; } ThreadArg;

static void *thread_main(void *a)
{
  ThreadArg *t = (ThreadArg *)a;
  t->betters = synthesize(t->E, 0, t->idx, t->stride, t->end, &t->prng);
  return NULL;
}

/* lib/refiner.h:43-122 (serial) and lib/refinerThreaded.h:260-381 (threaded).
 * In threaded mode every thread gets its own GRand stream: the reference shares one
 * GRand among its threads without a lock (lib/refinerThreaded.h:339), a data race
 * whose outcome cannot be restated; per-thread streams keep the same work per draw. */
static void refiner(Engine *E, Prng *prng, const OOptions *opt, OStats *st)
{
  unsigned sizes[6];
  orc_repetition(E->nT, sizes);
  E->estimated = 0;
  for (int i = 0; i < 6; i++) E->estimated += sizes[i];
  E->completed = 0;
  E->priorPercent = 0;
  int threads = (opt && opt->threads > 1) ? opt->threads : 1;
  for (unsigned pass = 0; pass < 6; pass++) {
    unsigned end = sizes[pass];
    unsigned long betters = 0;
    if (opt && opt->batch > 0) {
      size_t np = (size_t)E->tgt->width * E->tgt->height;
      unsigned len = pass == 0 ? (256u < (unsigned)opt->batch ? 256u : (unsigned)opt->batch) : (unsigned)opt->batch;
      const unsigned floor0 = opt->minChunk > 0 ? (unsigned)opt->minChunk : 64u;
      if (pass == 0 && opt->rampPermille > 0) len = floor0 < (unsigned)opt->batch ? floor0 : (unsigned)opt->batch;
      int R = opt->conflictRadius;
      int TW = (int)E->tgt->width, TH = (int)E->tgt->height;
      /* pending targets in synthesis order; each chunk takes up to len of them, skipping
         (deferring) those within R of a target already taken */
      unsigned *queue = (unsigned *)malloc((end ? end : 1) * sizeof(unsigned));
      unsigned *next = (unsigned *)malloc((end ? end : 1) * sizeof(unsigned));
      unsigned *chunk = (unsigned *)malloc((end ? end : 1) * sizeof(unsigned));
      unsigned *stamp = (unsigned *)calloc(np, sizeof(unsigned));
      unsigned nq = end, chunkId = 0;
      for (unsigned i = 0; i < end; i++) queue[i] = i;
      int32_t *savedTargets = E->targets;
      int32_t *tmpTargets = (int32_t *)malloc((end ? end : 1) * 2 * sizeof(int32_t));
      while (nq > 0) {
        unsigned nc = 0, nn = 0;
        chunkId++;
        for (unsigned q = 0; q < nq; q++) {
          unsigned k = queue[q];
          int tx = savedTargets[2 * k], ty = savedTargets[2 * k + 1];
          int ok = nc < len;
          if (ok && R > 0) {
            for (int dy = -R; dy <= R && ok; dy++)
              for (int dx = -R; dx <= R && ok; dx++) {
                int x = tx + dx, y = ty + dy;
                if (x >= 0 && y >= 0 && x < TW && y < TH && stamp[(size_t)y * TW + x] == chunkId) ok = 0;
              }
          }
          if (ok) {
            chunk[nc++] = k;
            stamp[(size_t)ty * TW + tx] = chunkId;
          } else {
            next[nn++] = k;
          }
        }
        memcpy((void *)E->rdHas, E->hasValue, np);
        memcpy((void *)E->rdTgt, E->tgt->data, np * E->tgt->depth);
        memcpy((void *)E->rdSrc, E->sourceOf, np * 2 * sizeof(int32_t));
        /* visit the chunk: lay its targets out contiguously; kmap keeps their order index
           for the prober heuristic */
        for (unsigned c = 0; c < nc; c++) {
          tmpTargets[2 * c] = savedTargets[2 * chunk[c]];
          tmpTargets[2 * c + 1] = savedTargets[2 * chunk[c] + 1];
        }
        E->targets = tmpTargets;
        E->kmap = chunk;
        betters += synthesize(E, 0, 0, 1, nc, prng);
        E->targets = savedTargets;
        E->kmap = NULL;
        unsigned *t = queue; queue = next; next = t;
        nq = nn;
        if (pass == 0 && opt->rampPermille > 0) {
          unsigned long long done = end - nq;
          unsigned long long l = done * (unsigned long long)opt->rampPermille / 1000ull;
          if (l < floor0) l = floor0;
          len = (unsigned)(l > (unsigned long long)opt->batch ? (unsigned long long)opt->batch : l);
        } else if (len < (unsigned)opt->batch) len *= 2;
        if (*E->cancel) break;
      }
      free(queue); free(next); free(chunk); free(stamp); free(tmpTargets);
    } else if (threads == 1) {
      betters = synthesize(E, 0, 0, 1, end, prng);
    } else {
      pthread_t th[64];
      ThreadArg args[64];
      for (int t = 0; t < threads && t < 64; t++) {
        args[t].E = E; args[t].idx = (unsigned)t; args[t].stride = (unsigned)threads; args[t].end = end;
        prng_init(&args[t].prng, prng->kind, ENGINE_SEED + 1u + (unsigned)t + 64u * pass);
        pthread_create(&th[t], NULL, thread_main, &args[t]);
      }
      for (int t = 0; t < threads && t < 64; t++) {
        pthread_join(th[t], NULL);
        betters += args[t].betters;
      }
    }
    if (st) {
      st->passSize[pass] = end;
      st->betters[pass] = (unsigned)betters;
      st->passes = pass + 1;
    }
    if ((double)((float)betters / (float)E->nT) < 0.1) break;
  }
}

/* ------------------------------------------------------------------ ordering */

/* lib/brushfire.h:63-67 grad */
static unsigned grad(int32_t x, int32_t y)
{
  return (unsigned)(atan2((float)y, (float)x) * 200 / 3.14159265358979323846 + 200);
}

/* lib/engineTypes.h:208-226 get_bounds / get_center */
static void get_center(const int32_t *pts, unsigned n, int32_t *cx, int32_t *cy)
{
  int32_t ulx = 2147483647, uly = 2147483647, lrx = 0, lry = 0;
  for (unsigned i = 0; i < n; i++) {
    if (pts[2 * i] < ulx) ulx = pts[2 * i];
    if (pts[2 * i + 1] < uly) uly = pts[2 * i + 1];
    if (pts[2 * i] > lrx) lrx = pts[2 * i];
    if (pts[2 * i + 1] > lry) lry = pts[2 * i + 1];
  }
  *cx = (lrx - ulx) / 2 + ulx;
  *cy = (lry - uly) / 2 + uly;
}

/* lib/orderTarget.h:58-79 randomizeBandsTargetPoints */
static void randomize_bands(int32_t *pts, unsigned n, Prng *prng)
{
  int last = (int)n - 1;
  int halfBand = (int)(n * 0.1);
  for (int i = 0; i <= last; i++) {
    int bandStart = (i - halfBand) > 0 ? i - halfBand : 0;
    int bandEnd = (i + halfBand) < last ? i + halfBand : last;
    int bandSize = bandEnd - bandStart;
    int j = bandStart + (int)prng_range(prng, (uint32_t)bandSize);
    int32_t tx = pts[2 * i], ty = pts[2 * i + 1];
    pts[2 * i] = pts[2 * j]; pts[2 * i + 1] = pts[2 * j + 1];
    pts[2 * j] = tx; pts[2 * j + 1] = ty;
  }
}

/* lib/brushfire.h:73-148: sort elements carrying the proportional distance along each ray */
static SortElem *to_sort_array(const int32_t *pts, unsigned n)
{
  unsigned maxc[401];
  memset(maxc, 0, sizeof maxc);
  for (unsigned i = 0; i < n; i++) {
    unsigned c = (unsigned)(pts[2 * i] * pts[2 * i] + pts[2 * i + 1] * pts[2 * i + 1]);
    unsigned ray = grad(pts[2 * i], pts[2 * i + 1]);
    if (c > maxc[ray]) maxc[ray] = c;
  }
  SortElem *s = (SortElem *)malloc((n ? n : 1) * sizeof(SortElem));
  for (unsigned i = 0; i < n; i++) {
    int32_t x = pts[2 * i], y = pts[2 * i + 1];
    s[i].x = x; s[i].y = y;
    s[i].prop = (float)(y * y + x * x) / (float)maxc[grad(x, y)];
  }
  return s;
}

/* lib/orderTarget.h:269-343 orderTargetPoints, with the helpers at :36-47, :154-262 */
static int order_targets(int type, int32_t *pts, unsigned n, Prng *prng)
{
  int32_t cx, cy;
  switch (type) {
  case 0:
  case 1:
    for (unsigned i = 0; i < n; i++) {
      unsigned j = prng_range(prng, n);
      int32_t tx = pts[2 * i], ty = pts[2 * i + 1];
      pts[2 * i] = pts[2 * j]; pts[2 * i + 1] = pts[2 * j + 1];
      pts[2 * j] = tx; pts[2 * j + 1] = ty;
    }
    return 0;
  case 2: case 5: case 8: {
    get_center(pts, n, &cx, &cy);
    for (unsigned i = 0; i < n; i++) { pts[2 * i] -= cx; pts[2 * i + 1] -= cy; }
    SortElem *s = to_sort_array(pts, n);
    /* type 8 (squeeze, :209-262) interleaves a copy of the offsets, then overwrites the
       result with the moreInward-sorted array: its net order is that of type 2. */
    orc_merge_sort(s, n, sizeof(SortElem), type == 5 ? cmp_less_inward : cmp_more_inward);
    for (unsigned i = 0; i < n; i++) { pts[2 * i] = s[i].x + cx; pts[2 * i + 1] = s[i].y + cy; }
    free(s);
    randomize_bands(pts, n, prng);
    return 0;
  }
  case 3: case 4: case 6: case 7: {
    int (*cmp)(const void *, const void *) =
        type == 3 ? cmp_more_horizontal : type == 4 ? cmp_more_vertical :
        type == 6 ? cmp_less_horizontal : cmp_less_vertical;
    get_center(pts, n, &cx, &cy);
    for (unsigned i = 0; i < n; i++) { pts[2 * i] -= cx; pts[2 * i + 1] -= cy; }
    orc_merge_sort(pts, n, 2 * sizeof(int32_t), cmp);
    for (unsigned i = 0; i < n; i++) { pts[2 * i] += cx; pts[2 * i + 1] += cy; }
    randomize_bands(pts, n, prng);
    return 0;
  }
  default:
    return 4;  /* IMAGE_SYNTH_ERROR_MATCH_CONTEXT_TYPE_RANGE */
  }
}

int orc_order_targets(int matchContextType, int32_t *points, unsigned n, int prngKind, uint64_t *drawsUsed)
{
  Prng p;
  prng_init(&p, prngKind, ENGINE_SEED);
  int e = order_targets(matchContextType, points, n, &p);
  if (drawsUsed) *drawsUsed = p.draws;
  return e;
}

/* ------------------------------------------------------------------ engine */

/* lib/engine.c:539-690 engine */
int oracle_engine(OParams P, const OFormatIndices *ix, OPixmap *tgt, const OPixmap *cor,
                  void (*progress)(int, void *), void *ctx, int *cancelFlag,
                  const OOptions *opt, OStats *st)
{
  if (P.patchSize > MAX_NEIGHBORS) return 3;
  Engine E;
  memset(&E, 0, sizeof E);
  E.P = P; E.ix = ix; E.tgt = tgt; E.cor = cor;
  E.progress = progress; E.ctx = ctx;
  int noCancel = 0;
  E.cancel = cancelFlag ? cancelFlag : &noCancel;
  pthread_mutex_init(&E.mutex, NULL);
  size_t TW = tgt->width, TH = tgt->height, CW = cor->width, CH = cor->height;

  /* prepareTargetPoints (:338-391) */
  E.hasValue = (uint8_t *)calloc((TW * TH) > 0 ? (TW * TH) : 1, 1);
  unsigned nT = 0;
  for (size_t i = 0; i < TW * TH; i++) if (tgt->data[i * tgt->depth] != 0) nT++;
  E.targets = (int32_t *)malloc((nT ? nT : 1) * 2 * sizeof(int32_t));
  nT = 0;
  for (size_t y = 0; y < TH; y++)
    for (size_t x = 0; x < TW; x++) {
      const uint8_t *px = tgt->data + (x + y * TW) * tgt->depth;
      int selected = px[0] != 0;
      int opaque = ix->isAlphaTarget ? px[ix->alpha_bip] != 0 : 1;
      E.hasValue[x + y * TW] = (uint8_t)(P.matchContextType && !selected && opaque);
      if (selected) { E.targets[2 * nT] = (int32_t)x; E.targets[2 * nT + 1] = (int32_t)y; nT++; }
    }
  E.nT = nT;
  if (st) st->nTargets = nT;
  if (!nT) { free(E.targets); free(E.hasValue); return 5; }

  /* prepare_target_sources (:207-224) */
  E.sourceOf = (int32_t *)malloc(TW * TH * 2 * sizeof(int32_t));
  for (size_t i = 0; i < TW * TH * 2; i++) E.sourceOf[i] = -1;

  /* prepareCorpusPoints (:400-431) */
  E.corpusPts = (int32_t *)malloc(((CW * CH) > 0 ? (CW * CH) : 1) * 2 * sizeof(int32_t));
  unsigned nC = 0;
  for (size_t y = 0; y < CH; y++)
    for (size_t x = 0; x < CW; x++) {
      const uint8_t *px = cor->data + (x + y * CW) * cor->depth;
      if (px[0] == MASK_TOTALLY_SELECTED && (ix->isAlphaSource ? px[ix->alpha_bip] != 0 : 1)) {
        E.corpusPts[2 * nC] = (int32_t)x; E.corpusPts[2 * nC + 1] = (int32_t)y; nC++;
      }
    }
  E.nC = nC;
  if (st) st->nCorpus = nC;
  if (!nC) { free(E.targets); free(E.hasValue); free(E.sourceOf); free(E.corpusPts); return 6; }

  /* prepareSortedOffsets (:465-497) */
  int w = (int)(CW < TW ? CW : TW), h = (int)(CH < TH ? CH : TH);
  E.nO = (unsigned)((2 * w - 1) * (2 * h - 1));
  E.offsets = (int32_t *)malloc((size_t)E.nO * 2 * sizeof(int32_t));
  orc_sorted_offsets(w, h, E.offsets);
  if (st) st->nOffsets = E.nO;

  orc_quantize(P.sensitivity, P.mapWeight, E.T, E.Tm);

  Prng prng;
  prng_init(&prng, opt ? opt->prng : ORC_PRNG_GRAND, (opt && opt->seed) ? opt->seed : ENGINE_SEED);
  int err = order_targets(P.matchContextType, E.targets, nT, &prng);
  if (err) {
    free(E.targets); free(E.hasValue); free(E.sourceOf); free(E.corpusPts); free(E.offsets);
    return err;
  }

  /* prepareRecentProber (:314-327) */
  E.prober = (uint32_t *)malloc(CW * CH * sizeof(uint32_t));
  for (size_t i = 0; i < CW * CH; i++) E.prober[i] = 0xFFFFFFFFu;

  E.rdHas = E.hasValue;
  E.rdTgt = E.tgt->data;
  E.rdSrc = E.sourceOf;
  uint8_t *snapHas = NULL, *snapTgt = NULL;
  int32_t *snapSrc = NULL;
  if (opt && opt->batch > 0) {
    snapHas = (uint8_t *)malloc(TW * TH);
    snapTgt = (uint8_t *)malloc(TW * TH * tgt->depth);
    snapSrc = (int32_t *)malloc(TW * TH * 2 * sizeof(int32_t));
    E.rdHas = snapHas;
    E.rdTgt = snapTgt;
    E.rdSrc = snapSrc;
  }
  if (st && st->lastBest) {
    E.lastBest = st->lastBest;
    for (size_t i = 0; i < TW * TH; i++) E.lastBest[i] = UINT32_MAX;
  }
  refiner(&E, &prng, opt, st);

  if (st) {
    st->scorerCalls = E.scorerCalls;
    st->visits = E.visits;
    st->progressCalls = E.progressCalls;
    st->lastPercent = (int)E.priorPercent;
    if (st->sourceOf) memcpy(st->sourceOf, E.sourceOf, TW * TH * 2 * sizeof(int32_t));
    if (E.lastBest) {
      double s = 0; unsigned long long c = 0;
      for (unsigned i = 0; i < nT; i++) {
        uint32_t v = E.lastBest[(size_t)E.targets[2 * i] + (size_t)E.targets[2 * i + 1] * TW];
        if (v != UINT32_MAX) { s += v; c++; }
      }
      st->meanError = c ? s / (double)c : 0.0;
    }
  }
  free(E.prober); free(E.targets); free(E.hasValue); free(E.sourceOf); free(E.corpusPts); free(E.offsets);
  free(snapHas); free(snapTgt); free(snapSrc);
  pthread_mutex_destroy(&E.mutex);
  return 0;
}

/* ------------------------------------------------------------------ SimpleAPI adapters */

/* lib/adaptSimple.h:48-96 adaptImage: copy `count` pixelels per pixel from a row-padded
 * buffer into an unpadded pixmap at pixelel offset `off`, walking the SOURCE buffer's
 * own width/height (so a mask of other dimensions fills the pixmap sequentially). */
static void adapt_image(const OImageBuffer *src, OPixmap *dst, unsigned off, unsigned count)
{
  size_t d = off;
  for (unsigned row = 0; row < src->height; row++) {
    size_t s = (size_t)row * src->rowBytes;
    for (unsigned col = 0; col < src->width; col++) {
      for (unsigned k = 0; k < count; k++) dst->data[d + k] = src->data[s + k];
      s += count;
      d += count + off;
    }
  }
}

/* lib/adaptSimple.h:111-168 antiAdaptImage */
static void anti_adapt(OImageBuffer *dst, const OPixmap *src, unsigned off, unsigned count)
{
  size_t s = off;
  for (unsigned row = 0; row < dst->height; row++) {
    size_t d = (size_t)row * dst->rowBytes;
    for (unsigned col = 0; col < dst->width; col++) {
      for (unsigned k = 0; k < count; k++) dst->data[d + k] = src->data[s + k];
      s += count + off;
      d += count;
    }
  }
}

/* lib/adaptSimple.h:226-249 adaptImageAndMask + lib/mapOps.h:153-168 interleave_mask
 * (+ lib/mapOps.h:129-144 invert_bytemap when invert is set) */
static int adapt_image_and_mask(const OImageBuffer *img, const OImageBuffer *mask, OPixmap *pm,
                                unsigned bpp, int invert)
{
  size_t n = (size_t)img->width * img->height;
  if ((size_t)mask->width * mask->height > n) return 2;  /* would overrun in the reference */
  pm->width = img->width; pm->height = img->height; pm->depth = bpp + 1;
  pm->data = (uint8_t *)calloc((n * pm->depth) > 0 ? (n * pm->depth) : 1, 1);
  adapt_image(img, pm, 1, bpp);
  OPixmap mk = { img->width, img->height, 1, (uint8_t *)calloc(n ? n : 1, 1) };
  adapt_image(mask, &mk, 0, 1);
  for (size_t i = 0; i < n; i++) pm->data[i * pm->depth] = invert ? (uint8_t)~mk.data[i] : mk.data[i];
  free(mk.data);
  return 0;
}

static int image_synth_common(OImageBuffer *image, OImageBuffer *mask, OImageBuffer *mask2, int format,
                              OParams *params, void (*progress)(int, void *), void *ctx, int *cancelFlag,
                              const OOptions *opt, OStats *st)
{
  /* lib/imageSynth.c:62-137 / 140-217 */
  if (image->width != mask->width || image->height != mask->height) return 2;
  OParams defaults = { 0, 0, 1, 0.5, 0.117, 30, 200 };   /* lib/engineParams.c:9-18 */
  if (!params) params = &defaults;
  OFormatIndices ix;
  if (oracle_format_indices(&ix, format)) return 1;
  unsigned bpp = oracle_count_pixelels(format);
  OPixmap tgt, cor;
  if (adapt_image_and_mask(image, mask, &tgt, bpp, 0)) return 2;
  if (adapt_image_and_mask(image, mask2 ? mask2 : mask, &cor, bpp, mask2 ? 0 : 1)) { free(tgt.data); return 2; }
  int noCancel = 0;
  if (!cancelFlag) cancelFlag = &noCancel;
  int err = oracle_engine(*params, &ix, &tgt, &cor, progress, ctx, cancelFlag, opt, st);
  if (!err && !*cancelFlag) anti_adapt(image, &tgt, 1, bpp);
  free(tgt.data);
  free(cor.data);
  return err;
}

int oracle_imageSynth(OImageBuffer *image, OImageBuffer *mask, int format, OParams *params,
                      void (*progress)(int, void *), void *ctx, int *cancelFlag,
                      const OOptions *opt, OStats *st)
{
  return image_synth_common(image, mask, NULL, format, params, progress, ctx, cancelFlag, opt, st);
}

int oracle_imageSynth2(OImageBuffer *image, OImageBuffer *mask, OImageBuffer *mask2, int format,
                       OParams *params, void (*progress)(int, void *), void *ctx, int *cancelFlag,
                       const OOptions *opt, OStats *st)
{
  return image_synth_common(image, mask, mask2, format, params, progress, ctx, cancelFlag, opt, st);
}

/* ------------------------------------------------------------------ frozen-state scorer */

int orc_best_of(const OPixmap *corpus, const OFormatIndices *ix,
                const uint16_t T[512], const uint32_t Tm[512],
                unsigned count, const int32_t *offs, const uint8_t *pix,
                unsigned nCand, const int32_t *cands, uint32_t *bestSum, int *stoppedAt)
{
  Engine E;
  memset(&E, 0, sizeof E);
  E.ix = ix;
  E.cor = corpus;
  memcpy(E.T, T, sizeof E.T);
  memcpy(E.Tm, Tm, sizeof E.Tm);
  Neighbor nb[MAX_NEIGHBORS];
  if (count > MAX_NEIGHBORS) count = MAX_NEIGHBORS;
  for (unsigned i = 0; i < count; i++) {
    nb[i].ox = offs[2 * i]; nb[i].oy = offs[2 * i + 1];
    memcpy(nb[i].pix, pix + 8 * i, 8);
    nb[i].sx = nb[i].sy = -1;
  }
  uint32_t best = UINT32_MAX;
  int32_t bx = 0, by = 0;
  int bettered = 0, win = -1;
  unsigned c;
  for (c = 0; c < nCand; c++) {
    int before = bettered;
    uint32_t prev = best;
    int perfect = compute_best_fit(&E, cands[2 * c], cands[2 * c + 1], &best, &bx, &by, count, nb, &bettered);
    if (best != prev || (!before && bettered)) win = (int)c;
    if (perfect) { c++; break; }
  }
  *bestSum = best;
  if (stoppedAt) *stoppedAt = (int)c;
  return win;
}
