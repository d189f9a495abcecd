// Host side of libresynth_mi355x.so: parameter checks, the reference's data contract,
// metric tables, the GLib / libc PRNG streams, target ordering, the pass loop with the
// reference's progress and cancel semantics, and the C ABI (include/resynth_mi355x.h).
// All image-sized work runs on the GPU (resynth_kernels.hip).
#include "../../include/resynth_mi355x.h"
#include "resynth_internal.h"
#include "resynth_learn.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

using namespace rs;

void rs::hip_fail(hipError_t e, const char *expr, const char *file, int line)
{
  std::fprintf(stderr, "resynth_mi355x: %s failed: %s (%s:%d)\n", expr, hipGetErrorString(e), file, line);
  throw HipFailure{e};
}

namespace {

constexpr unsigned kMaxNeighbors = 64;        // lib/imageSynthConstants.h:26
constexpr uint32_t kEngineSeed = 1198472u;     // lib/engine.c:643
constexpr unsigned kCallbackMask = 4095u;      // lib/imageSynthConstants.h:99

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A failed HIP call unwinds to the C entry point (see `guarded`), which returns
// RESYNTH_ERROR_DEVICE to the caller.
#define HIP_TRY(expr)                                                                                 \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) hip_fail(e_, #expr, __FILE__, __LINE__);                                    \
  } while (0)

// ------------------------------------------------------------------ PRNG streams

// GLib GRand: MT19937 seeded by g_rand_set_seed, tempered g_rand_int, and the rejection
// rule of g_rand_int_range (GLib >= 2.2).
struct MT19937 {
  uint32_t s[624];
  int i = 624;
  explicit MT19937(uint32_t seed)
  {
    s[0] = seed;
    for (int k = 1; k < 624; k++) s[k] = 1812433253u * (s[k - 1] ^ (s[k - 1] >> 30)) + (uint32_t)k;
  }
  void regenerate()
  {
    auto mixbits = [](uint32_t hi, uint32_t lo) {
      uint32_t y = (hi & 0x80000000u) | (lo & 0x7FFFFFFFu);
      return (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
    };
    // s[k] ^= ... in order k = 0..623 with indices taken mod 624, split so that no index
    // needs a modulo: s[k + 397] for k < 227, the already regenerated s[k - 227] after.
    int k = 0;
    for (; k < 227; k++) s[k] = s[k + 397] ^ mixbits(s[k], s[k + 1]);
    for (; k < 623; k++) s[k] = s[k - 227] ^ mixbits(s[k], s[k + 1]);
    s[623] = s[396] ^ mixbits(s[623], s[0]);
    i = 0;
  }
  uint32_t next()
  {
    if (i >= 624) regenerate();
    uint32_t y = s[i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9D2C5680u;
    y ^= (y << 15) & 0xEFC60000u;
    y ^= y >> 18;
    return y;
  }
};

// The C library's rand() of glibc (TYPE_3 additive feedback, r[i] = r[i-3] + r[i-31]),
// the PRNG of the stand-alone build (lib/glibProxy.c:30-49).
struct LibcRand {
  // ring of the last 31 raw words o[n-31..n-1]; the next word is o[n-31] + o[n-3]
  int32_t r[31];
  unsigned n = 0;
  explicit LibcRand(uint32_t seed)
  {
    if (seed == 0) seed = 1;
    int32_t o[344];
    o[0] = (int32_t)seed;
    for (int k = 1; k < 31; k++) {
      long long v = (16807LL * o[k - 1]) % 2147483647LL;
      if (v < 0) v += 2147483647LL;
      o[k] = (int32_t)v;
    }
    for (int k = 31; k < 34; k++) o[k] = o[k - 31];
    for (int k = 34; k < 344; k++) o[k] = (int32_t)((uint32_t)o[k - 31] + (uint32_t)o[k - 3]);
    for (int k = 0; k < 31; k++) r[k] = o[313 + k];
  }
  uint32_t next()
  {
    int32_t v = (int32_t)((uint32_t)r[n % 31] + (uint32_t)r[(n + 28) % 31]);
    r[n % 31] = v;
    n++;
    return (uint32_t)v >> 1;
  }
};

struct Prng {
  int kind;
  MT19937 mt;
  LibcRand lc;
  Prng(int k, uint32_t seed) : kind(k), mt(seed), lc(seed) {}
  uint32_t raw() { return kind == RESYNTH_PRNG_LIBC ? lc.next() : mt.next(); }
  // int_range(0, n) as the engine's PRNG computes it
  uint32_t range(uint32_t n)
  {
    if (kind == RESYNTH_PRNG_LIBC) {
      if (n < 1) return 0;
      return lc.next() / (2147483647u / n + 1u);
    }
    if (n == 0) return 0;
    if (n != rangeN) {   // the rejection bound of the last n (the band shuffle repeats n)
      rangeN = n;
      if (n <= 0x80000000u) {
        uint32_t leftover = (0x80000000u % n) * 2u;
        if (leftover >= n) leftover -= n;
        rangeMax = 0xFFFFFFFFu - leftover;
      } else {
        rangeMax = n - 1u;
      }
    }
    uint32_t v;
    do v = mt.next(); while (v > rangeMax);
    return v % n;
  }
  uint32_t rangeN = 0, rangeMax = 0;
  // The random target order of types 0 and 1 (lib/orderTarget.h:36-47): for i in 0..n-1,
  // swap(perm[i], perm[int_range(0, n)]), perm starting as the identity.  Every draw has the
  // same n, so the rejection bound and the reduction are set up once (v % n by Lemire's
  // exact 64-bit fastmod), and the draws are taken in blocks ahead of the swaps they feed.
  void shuffle(uint32_t *perm, uint32_t n)
  {
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    if (n == 0) return;
    constexpr uint32_t kBlock = 4096;
    uint32_t js[kBlock];
    if (kind == RESYNTH_PRNG_LIBC) {
      const uint32_t d = 2147483647u / n + 1u;
      for (uint32_t i0 = 0; i0 < n; i0 += kBlock) {
        const uint32_t m = std::min(kBlock, n - i0);
        for (uint32_t t = 0; t < m; t++) js[t] = lc.next() / d;
        for (uint32_t t = 0; t < m; t++) std::swap(perm[i0 + t], perm[js[t]]);
      }
      return;
    }
    uint32_t maxvalue;
    if (n <= 0x80000000u) {
      uint32_t leftover = (0x80000000u % n) * 2u;
      if (leftover >= n) leftover -= n;
      maxvalue = 0xFFFFFFFFu - leftover;
    } else {
      maxvalue = n - 1u;
    }
    const uint64_t M = ~(uint64_t)0 / n + 1;  // 0 for n = 1, which reduces every v to 0
    for (uint32_t i0 = 0; i0 < n; i0 += kBlock) {
      const uint32_t m = std::min(kBlock, n - i0);
      for (uint32_t t = 0; t < m; t++) {
        uint32_t v;
        do v = mt.next(); while (v > maxvalue);
        js[t] = (uint32_t)(((unsigned __int128)(M * v) * n) >> 64);
      }
      for (uint32_t t = 0; t < m; t++) std::swap(perm[i0 + t], perm[js[t]]);
    }
  }
};

// ------------------------------------------------------------------ tables

// lib/matchWeighting.h:50-60, 143-159, 167-178: the parameters arrive as float.
void quantize_tables(double sensitivity, double mapWeight, Tables &t)
{
  const float c = (float)sensitivity;
  const float mw = (float)mapWeight;
  const float scale = c * 256.0f;
  const double top = std::log(((double)256 / scale) * ((double)256 / scale) + 1.0);
  for (int i = -256; i < 256; i++) {
    const double num = std::log(((double)i / scale) * ((double)i / scale) + 1.0);
    t.T[256 + i] = (uint16_t)(num / top * (double)(float)65535);
    t.Tm[256 + i] = (uint32_t)((float)(i * i) * mw * 4.0);
  }
}

// ------------------------------------------------------------------ ordering

// The banded shuffle of lib/orderTarget.h:58-79 (halfBand = (gint)(n * 0.1)) applied to
// the identity: out[i] is the sorted position that ends at position i.  It consumes the
// engine's PRNG exactly as the reference's loop does; the device sorts the targets
// meanwhile (types 2..8, resynth_order.hip) and composes the two.
void band_permutation(uint32_t *perm, uint32_t n, Prng &prng)
{
  for (uint32_t i = 0; i < n; i++) perm[i] = i;
  if (n == 0) return;
  const int last = (int)n - 1;
  const int half = (int)(n * 0.1);
  for (int i = 0; i <= last; i++) {
    const int start = std::max(i - half, 0);
    const int end = std::min(i + half, last);
    const int j = start + (int)prng.range((uint32_t)(end - start));
    std::swap(perm[i], perm[j]);
  }
}

// grad (lib/brushfire.h:63-67) of the nine directions where the ray index falls exactly on
// a boundary, in the order the device expects: centre, +x, -x, +y, -y, (+,+), (-,+),
// (-,-), (+,-).  Computed with this host's C library, as the reference computes them.
void axis_rays(uint32_t out[9])
{
  static const int d[9][2] = {{0, 0}, {1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
  for (int k = 0; k < 9; k++)
    out[k] = (uint32_t)(std::atan2((double)(float)d[k][1], (double)(float)d[k][0]) * 200 / 3.14159265358979323846 +
                        200);
}

// ------------------------------------------------------------------ device state

// Bump allocator over a fixed region (the allocation-free entry's scratch): carve() hands
// out 256-byte aligned pieces until the region is spent, then fails like hipMalloc.
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
  void *carve(size_t bytes)
  {
    const size_t at = (used + 255) & ~(size_t)255;
    if (!base || at > size || bytes > size - at) hip_fail(hipErrorOutOfMemory, "scratch carve", __FILE__, __LINE__);
    used = at + bytes;
    return base + at;
  }
};

// While set (by the allocation-free entry, under the library lock), Buf and PinnedBuf carve
// from these regions instead of allocating.
Arena *g_devArena = nullptr, *g_hostArena = nullptr;

// Device bytes held by grow-only buffers, and the cap set by resynth_set_device_memory_limit.
size_t g_devBytes = 0, g_devLimit = 0;

struct Buf {
  void *p = nullptr;
  size_t cap = 0;
  bool carved = false;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  ~Buf() { release(); }
  void *get(size_t bytes)
  {
    if (bytes > cap) {
      if (g_devArena) {
        release();
        p = g_devArena->carve(bytes);
        cap = bytes;
        carved = true;
        return p;
      }
      const size_t c = std::max(bytes, cap + cap / 2);
      release();
      if (g_devLimit && g_devBytes + c > g_devLimit)
        hip_fail(hipErrorOutOfMemory, "device memory limit", __FILE__, __LINE__);
      HIP_TRY(hipMalloc(&p, c));
      cap = c;
      g_devBytes += c;
    }
    return p;
  }
  void release()
  {
    if (p && !carved) {
      (void)hipFree(p);
      g_devBytes -= cap;
    }
    p = nullptr;
    cap = 0;
    carved = false;
  }
};

struct PinnedBuf {
  void *p = nullptr;
  size_t cap = 0;
  bool carved = false;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf()
  {
    if (p && !carved) (void)hipHostFree(p);
  }
  void *get(size_t bytes)
  {
    if (bytes > cap) {
      if (p && !carved) (void)hipHostFree(p);
      p = nullptr;
      cap = 0;
      carved = false;
      if (g_hostArena) {
        p = g_hostArena->carve(bytes);
        cap = bytes;
        carved = true;
        return p;
      }
      const size_t c = std::max(bytes, cap + cap / 2);
      HIP_TRY(hipHostMalloc(&p, c, hipHostMallocDefault));
      cap = c;
    }
    return p;
  }
  // Grows to `bytes` keeping the first `keep` bytes of the contents.
  void *grow_keep(size_t bytes, size_t keep)
  {
    if (bytes <= cap) return p;
    void *old = p;
    const bool oldCarved = carved;
    void *q = nullptr;
    if (g_hostArena) {
      q = g_hostArena->carve(bytes);
    } else {
      HIP_TRY(hipHostMalloc(&q, bytes, hipHostMallocDefault));
    }
    if (old && keep) std::memcpy(q, old, std::min(keep, cap));
    if (old && !oldCarved) (void)hipHostFree(old);
    p = q;
    cap = bytes;
    carved = g_hostArena != nullptr;
    return p;
  }
};

// Device memory owned by one call (test hooks, the autotuner's sample copies).
struct TmpBuf {
  void *p = nullptr;
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes < 256 ? 256 : bytes); }
  TmpBuf() = default;
  TmpBuf(const TmpBuf &) = delete;
  TmpBuf &operator=(const TmpBuf &) = delete;
  ~TmpBuf()
  {
    if (p) (void)hipFree(p);
  }
};

// Device and pinned buffers of one engine run (grow-only, reused across calls).
struct Workspace {
  Buf tword, cword, tflag, cflag, sourceOf, prober, targetsRM, targets, corpusPts, offsets, offTemp, lastBest,
      tables, counters, compactTemp, perm, meanSum, image, mask, mask2, tpix, cpix, tileHas, tileFull, ctxRadius, ringTemp,
      firstIdx, wctl,
      bandStart;
  // learned proposals: corpus keys, their embeddings and inverted file, query scratch
  Buf keyFlags, keyPts, keyTemp, keyX, keyH1, keyH2, keyE, keyZ, cent, keyCl, bCount, bStart, bCursor,
      bMembers, qX, qH1, qH2, qE, qZ, qCl, props, propsAll, pairPts, pairSrc, superCent, superId, sCount, sStart,
      sCursor, sMembers, subCent, tileStart, qSuper, sortK0, sortV0, sortK1, sortV1, sortTemp, keyZs, keyPtsS, qK0,
      qV0, qK1, qV1, qSortTemp;
  // serial mode: raw PRNG outputs and the stream state
  Buf streamBuf, streamState;
  // target ordering of types 2..8
  Buf orderRays, orderTemp;
  PinnedBuf pinCounts, pinTables, pinPerm, pinStream, pinPass;
};

struct DeviceState {
  bool ready = false;
  bool ok = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;                     // learned index builds, beside pass 0
  hipEvent_t indexReady = nullptr;                // recorded on `side` when an index is built
  hipEvent_t weightsReady = nullptr;              // recorded on `side` after the network's weights are laid out
  hipEvent_t passCounts = nullptr;                // a pass's counters on the host
  Workspace ws;                                   // single-request entries
  std::vector<std::unique_ptr<Workspace>> pool;   // one per request of a multi-request call
  Buf manyDesc, manyCtl, manyGlobal;
  ManyGraph manyGraph[2];                         // per pixel-word width, kept across calls
  PinnedBuf pinDesc, pinCtl;
  hipEvent_t countsReady = nullptr;               // prepare_engine: point counts on the host
};

std::mutex g_mutex;
// Never destroyed: its buffers must not be freed after the HIP runtime has shut down.
DeviceState &g_dev = *new DeviceState;
int g_defaultMode = -1;

bool ensure_device()
{
  if (g_dev.ready) return g_dev.ok;
  g_dev.ready = true;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    std::fprintf(stderr, "resynth_mi355x: no HIP device is available\n");
    return false;
  }
  if (hipGetDevice(&g_dev.device) != hipSuccess ||
      hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&g_dev.side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&g_dev.indexReady, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g_dev.weightsReady, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g_dev.passCounts, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&g_dev.countsReady, hipEventDisableTiming) != hipSuccess) {
    std::fprintf(stderr, "resynth_mi355x: the HIP device could not be set up\n");
    return false;
  }
  g_dev.ok = true;
  return true;
}

// Clears what a failed HIP call may have left: work still queued on the library's stream
// and a non-sticky error code.
void recover_after_failure()
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_dev.ok) {
    (void)hipStreamSynchronize(g_dev.stream);
    (void)hipStreamSynchronize(g_dev.side);
  }
  (void)hipGetLastError();
}

// Runs the body of a C entry point; a HIP failure (hip_fail) or host memory exhaustion
// inside becomes RESYNTH_ERROR_DEVICE.
template <typename F>
int guarded(F &&body)
{
  try {
    return body();
  } catch (const HipFailure &) {
    recover_after_failure();
    return RESYNTH_ERROR_DEVICE;
  } catch (const std::bad_alloc &) {
    recover_after_failure();
    return RESYNTH_ERROR_DEVICE;
  }
}

int default_mode()
{
  if (g_defaultMode >= 0) return g_defaultMode;
  const char *m = std::getenv("RESYNTH_MODE");
  if (m && (std::strcmp(m, "serial") == 0 || std::strcmp(m, "0") == 0)) return RESYNTH_MODE_SERIAL;
  return RESYNTH_MODE_BATCHED;
}

// ------------------------------------------------------------------ allocation-free scratch

// Device scratch of resynth_image_synth_device_noalloc: every buffer of its engine run is
// carved from this array, so the entry allocates nothing.  Each device the library runs on
// holds its own copy until the process ends.  Host code must use the address that
// hipGetSymbolAddress returns for the current device (scratch_address), never the name.
constexpr size_t kScratchBytes = (size_t)RESYNTH_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Host staging of the same runs (point counts, tables, target permutation, serial PRNG
// stream): pageable memory in the library image, used only under the library lock.
constexpr size_t kNoAllocHostBytes = (size_t)12 << 20;
alignas(256) unsigned char g_noallocHost[kNoAllocHostBytes];

constexpr int kMaxDevices = 64;
char *g_scratchPtr[kMaxDevices] = {};

// The scratch array's address on the library's device, looked up once per device.
char *scratch_address()
{
  const int dev = g_dev.device;
  if (dev < 0 || dev >= kMaxDevices) hip_fail(hipErrorInvalidDevice, "scratch device index", __FILE__, __LINE__);
  if (!g_scratchPtr[dev]) {
    int cur = -1;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != dev) HIP_TRY(hipSetDevice(dev));
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != dev) HIP_TRY(hipSetDevice(cur));
    HIP_TRY(e);
    g_scratchPtr[dev] = (char *)ptr;
  }
  return g_scratchPtr[dev];
}

// Makes Buf and PinnedBuf carve from the two arenas for the lifetime of the scope.
struct ArenaScope {
  ArenaScope(Arena *d, Arena *h)
  {
    g_devArena = d;
    g_hostArena = h;
  }
  ~ArenaScope()
  {
    g_devArena = nullptr;
    g_hostArena = nullptr;
  }
};

// ------------------------------------------------------------------ engine job

// Where the two images come from.
struct Source {
  // SimpleAPI form (device pointers)
  bool simple = false;
  const unsigned char *image = nullptr;   // device
  size_t rowBytes = 0;
  unsigned bpp = 0;
  const unsigned char *mask = nullptr;    // device, image dims
  size_t maskRowBytes = 0;
  const unsigned char *mask2 = nullptr;   // device or null (inverse of mask)
  size_t mask2RowBytes = 0;
  int alphaIdx = -1;                      // index of alpha within a pixel, -1 if none
  // engine-layout form (device pixmaps)
  const unsigned char *tpix = nullptr;
  const unsigned char *cpix = nullptr;
  int tw = 0, th = 0, cw = 0, ch = 0;
};

struct Progress {
  void (*cb)(int, void *) = nullptr;
  void *ctx = nullptr;
  int *cancel = nullptr;
  unsigned estimated = 0, completed = 0, prior = 0;
  // lib/progress.c:53-65
  void tick()
  {
    completed += kCallbackMask;
    unsigned pct = (unsigned)(((float)completed / (float)estimated) * 100);
    if (pct > prior) {
      if (cb) cb((int)pct, ctx);
      prior = pct;
    }
  }
  bool cancelled() const { return cancel && *cancel; }
};

struct Job {
  TImageSynthParameters P;
  TFormatIndices ix;
  Source src;
  ResynthOptions opt;
  ResynthStats st;
  Progress prog;
  bool wide = false;
  int nColors = 0, nMaps = 0;
  Workspace *W = nullptr;
};

// A synthesis launch bracketed by HIP events (ResynthOptions.timing); `wave` marks the
// one-wavefront-per-target kernel, the dominant one.
struct TimedLaunch {
  hipEvent_t a = nullptr, b = nullptr;
  bool wave = false;
};

void record_event(Job &job, std::vector<TimedLaunch> &ev, bool start, bool wave = false)
{
  if (!job.opt.timing) return;
  hipEvent_t e;
  HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipEventRecord(e, g_dev.stream));
  if (start) ev.push_back({e, nullptr, wave});
  else ev.back().b = e;
}

// An engine run prepared on the device: point lists, sorted offsets, tables, target
// order and the state maps, described by g and b.
struct Prepared {
  Geometry g = {};
  SynthBuffers b = {};
  unsigned sizes[6] = {};
  unsigned nT = 0;
  size_t wb = 4;
  Prng prng{RESYNTH_PRNG_GRAND, kEngineSeed};   // the engine's PRNG after the ordering draws
  // Types 0 and 1 with a deferred order: the shuffle runs on a host thread (it needs only
  // the target count) while the device works; finish_order joins it and gathers the
  // row-major targets in that order.
  std::thread orderThread;
  bool orderPending = false;
  uint32_t *hPerm = nullptr, *dPerm = nullptr;
  const int2 *targetsRM = nullptr;
  const uint32_t *sortedIdx = nullptr;   // types 2..8: row-major indices in sorted order (device)
  ~Prepared()
  {
    if (orderThread.joinable()) orderThread.join();
  }
};

void finish_order(Prepared &pr)
{
  if (!pr.orderPending) return;
  if (pr.orderThread.joinable()) pr.orderThread.join();
  pr.orderPending = false;
  hipStream_t s = g_dev.stream;
  HIP_TRY(hipMemcpyAsync(pr.dPerm, pr.hPerm, (size_t)pr.nT * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  if (pr.sortedIdx)
    HIP_TRY(launch_order_compose(s, pr.targetsRM, pr.sortedIdx, pr.dPerm, const_cast<int2 *>(pr.b.targets), pr.nT));
  else
    HIP_TRY(launch_gather_targets(s, pr.targetsRM, pr.dPerm, const_cast<int2 *>(pr.b.targets), pr.nT));
}

// Preparation of an engine run (lib/engine.c:539-660) on the job's workspace.  Returns a
// TImageSynthError.  deferOrder: a random target order (types 0 and 1) is left to
// finish_order, so that its host shuffle overlaps the device work queued before the
// targets are needed.
int prepare_engine(Job &job, Prepared &pr, bool deferOrder = false)
{
  hipStream_t s = g_dev.stream;
  Workspace &W = *job.W;
  const TImageSynthParameters &P = job.P;
  const TFormatIndices &ix = job.ix;
  const double t0 = now_ms();
  if (P.patchSize > kMaxNeighbors) return IMAGE_SYNTH_ERROR_PATCH_SIZE_EXCEEDED;

  Source &S = job.src;
  const int tw = S.tw, th = S.th, cw = S.cw, ch = S.ch;
  const size_t nTpix = (size_t)tw * th, nCpix = (size_t)cw * ch;
  // The kernels index pixels with 32-bit words and multiply rows by __umul24 (sides below
  // 2^24), and count the sorted offsets in 32 bits.
  {
    const uint64_t ow = (uint64_t)std::min(cw, tw), oh = (uint64_t)std::min(ch, th);
    if (nTpix >= (1u << 31) || nCpix >= (1u << 31) || tw >= (1 << 24) || th >= (1 << 24) || cw >= (1 << 24) ||
        ch >= (1 << 24) || (2 * ow - 1) * (2 * oh - 1) >= (1ull << 32))
      return RESYNTH_ERROR_TOO_LARGE;
    // the inward/outward orders admit one NaN ratio (the centre pixel) while offsets from
    // the target's centre stay below 65536
    const int t = job.P.matchContextType;
    if ((t == 2 || t == 5 || t == 8) && (tw > 65536 || th > 65536)) return RESYNTH_ERROR_TOO_LARGE;
  }
  job.nColors = ix.img_match_bpp;
  job.nMaps = ix.map_match_bpp;
  job.wide = job.nColors + job.nMaps > 3;
  const size_t wb = job.wide ? 8 : 4;
  const int useContext = P.matchContextType != 0;
  const bool serialMode = job.opt.mode == RESYNTH_MODE_SERIAL;

  void *tword = W.tword.get(nTpix * wb);
  void *cword = W.cword.get(std::max(cword_words(cw, ch), nTpix) * wb);
  uint8_t *tflag = (uint8_t *)W.tflag.get(nTpix + 64);
  uint8_t *cflag = (uint8_t *)W.cflag.get(std::max(nCpix, nTpix) + 64);
  if (S.simple) {
    HIP_TRY(launch_adapt_simple(s, job.wide, S.image, S.rowBytes, S.bpp, S.mask, S.maskRowBytes, S.mask2,
                                S.mask2RowBytes, tw, th, job.nColors, S.alphaIdx, useContext, tword, cword, tflag,
                                cflag));
  } else {
    HIP_TRY(launch_adapt_maps(s, job.wide, S.tpix, tw, th, ix.total_bpp, ix.colorEndBip, ix.map_start_bip,
                              ix.map_end_bip, ix.isAlphaTarget ? ix.alpha_bip : -1, useContext, 1, tword, tflag));
    HIP_TRY(launch_adapt_maps(s, job.wide, S.cpix, cw, ch, ix.total_bpp, ix.colorEndBip, ix.map_start_bip,
                              ix.map_end_bip, ix.isAlphaSource ? ix.alpha_bip : -1, useContext, 0, cword, cflag));
  }
  // row-major target and corpus point lists (lib/engine.c:338-431)
  int2 *targetsRM = (int2 *)W.targetsRM.get(nTpix * sizeof(int2) + 64);
  int2 *corpusPts = (int2 *)W.corpusPts.get(nCpix * sizeof(int2) + 64);
  size_t ctb = std::max(compact_temp_bytes(nTpix), compact_temp_bytes(nCpix));
  char *ctemp = (char *)W.compactTemp.get(2 * ctb + 256);
  unsigned *dCounts = (unsigned *)(ctemp + 2 * ctb);
  HIP_TRY(launch_compact(s, tflag, tw, th, targetsRM, dCounts + 0, ctemp));
  HIP_TRY(launch_compact(s, cflag, cw, ch, corpusPts, dCounts + 1, ctemp + ctb));
  uint32_t *bandStart = (uint32_t *)W.bandStart.get((kBands + 1) * sizeof(uint32_t));
  HIP_TRY(launch_band_starts(s, corpusPts, dCounts + 1, ch, bandStart));
  // how far a neighbour scan can reach (k_tile_context), read back with the counts
  const int tilesW = (tw + kTile - 1) / kTile, tilesH = (th + kTile - 1) / kTile;
  const size_t nTiles = (size_t)tilesW * tilesH;
  uint8_t *tileFull = (uint8_t *)W.tileFull.get(2 * nTiles + 64);
  int32_t *ringTemp = (int32_t *)W.ringTemp.get(nTiles * sizeof(int32_t) + 64);
  int32_t *ctxRadius = (int32_t *)W.ctxRadius.get(nTiles * sizeof(int32_t) + 64);
  HIP_TRY(launch_reach_radius(s, job.wide, tword, tflag, tw, th, tileFull, tileFull + nTiles, ringTemp, ctxRadius,
                              dCounts + 2));
  unsigned *hCounts = (unsigned *)W.pinCounts.get(64);
  HIP_TRY(hipMemcpyAsync(hCounts, dCounts, 3 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
  hipEvent_t countsReady = g_dev.countsReady;
  HIP_TRY(hipEventRecord(countsReady, s));

  // Work that does not depend on the counts goes to the device now and overlaps the
  // host's wait: the metric tables and the state maps.
  Tables *hT = (Tables *)W.pinTables.get(sizeof(Tables));
  quantize_tables(P.sensitivityToOutliers, P.mapWeight, *hT);
  Tables *dT = (Tables *)W.tables.get(sizeof(Tables));
  HIP_TRY(hipMemcpyAsync(dT, hT, sizeof(Tables), hipMemcpyHostToDevice, s));
  int2 *sourceOf = (int2 *)W.sourceOf.get(nTpix * sizeof(int2));
  // the recent-prober map is indexed like the corpus words (tiled)
  uint32_t *prober = (uint32_t *)W.prober.get(cword_words(cw, ch) * sizeof(uint32_t));
  uint32_t *lastBest = (uint32_t *)W.lastBest.get(nTpix * sizeof(uint32_t));
  Counters *cnt = (Counters *)W.counters.get(sizeof(Counters));
  HIP_TRY(launch_fill_u32(s, (uint32_t *)sourceOf, 0xFFFFFFFFu, nTpix * 2));
  HIP_TRY(launch_fill_u32(s, prober, 0xFFFFFFFFu, cword_words(cw, ch)));
  HIP_TRY(launch_fill_u32(s, lastBest, 0xFFFFFFFFu, nTpix));
  HIP_TRY(launch_fill_u32(s, (uint32_t *)cnt, 0u, sizeof(Counters) / 4));
  // occupancy of 8x8 tiles, used to skip provably empty offsets in pass-0 scans, and the
  // context radius of the context-only occupancy, before any visit updates it
  uint8_t *tileHas = (uint8_t *)W.tileHas.get((nTiles + 3) / 4 * 4);
  HIP_TRY(launch_fill_u32(s, (uint32_t *)tileHas, 0u, (nTiles + 3) / 4));
  HIP_TRY(launch_tile_init(s, job.wide, tword, tw, th, tileHas));
  HIP_TRY(launch_ctx_radius(s, tileHas, tilesW, tilesH, ringTemp, ctxRadius));

  HIP_TRY(hipEventSynchronize(countsReady));
  // The sorted offsets (lib/engine.c:465-497), as far as a scan can reach: every target's
  // scan gathers its patch before passing an offset farther than reach + 21 from it
  // (k_tile_context), so the table's prefix up to that distance is all any visit reads.
  // Wrapped neighbourhoods (tiling) and a corpus narrower or shorter than the target (a
  // table that does not span the image) take the whole table.
  const int ow = std::min(cw, tw), oh = std::min(ch, th);
  const unsigned reach = hCounts[2];
  uint64_t maxD2 = ~0ull;
  if (!P.isMakeSeamlesslyTileableHorizontally && !P.isMakeSeamlesslyTileableVertically && ow == tw && oh == th &&
      reach < (1u << 30))
    maxD2 = (uint64_t)(reach + 21) * (reach + 21);
  const OffsetPlan plan = plan_offsets(ow, oh, maxD2);
  const size_t nO = plan.nO;
  job.st.nOffsets = (unsigned)nO;
  int2 *offsets = (int2 *)W.offsets.get(nO * sizeof(int2) + 64);
  const size_t otb = offsets_temp_bytes(plan);
  HIP_TRY(launch_sorted_offsets(s, plan, offsets, W.offTemp.get(otb), otb));
  // the start of each distance in the sorted offsets, up to the largest one the table holds
  const unsigned rhoFar = (unsigned)std::ceil(std::sqrt((double)(ow - 1) * (ow - 1) + (double)(oh - 1) * (oh - 1))) + 1;
  const unsigned rhoMax = plan.whole ? rhoFar : std::min(rhoFar, (unsigned)std::sqrt((double)plan.lim2));
  uint32_t *firstIdx = (uint32_t *)W.firstIdx.get(((size_t)rhoMax + 1) * sizeof(uint32_t));
  HIP_TRY(launch_first_idx(s, ow, oh, rhoMax, firstIdx));
  const unsigned nT = hCounts[0], nC = hCounts[1];
  job.st.nTargets = nT;
  job.st.nCorpus = nC;
  pr.nT = nT;
  if (!nT || !nC) {
    HIP_TRY(hipStreamSynchronize(s));
    return !nT ? IMAGE_SYNTH_ERROR_EMPTY_TARGET : IMAGE_SYNTH_ERROR_EMPTY_CORPUS;
  }

  // target order (lib/orderTarget.h:269-343) with the engine's PRNG
  // the engine's seed (lib/engine.c:643); a nonzero option seed gives the serial mode
  // another stream, for statistics over seeds
  pr.prng = Prng(job.opt.prng, (serialMode && job.opt.seed) ? job.opt.seed : kEngineSeed);
  Prng &prng = pr.prng;
  int2 *targets = (int2 *)W.targets.get((size_t)nT * sizeof(int2) + 64);
  const int type = P.matchContextType;
  if (type < 0 || type > 8) {
    HIP_TRY(hipStreamSynchronize(s));
    return IMAGE_SYNTH_ERROR_MATCH_CONTEXT_TYPE_RANGE;
  }
  // Types 0/1 are a shuffle of the row-major list; types 2..8 a device sort by direction
  // followed by a banded shuffle.  The shuffles need only nT and the PRNG: a host thread
  // draws them (as an index permutation) while the device works, and finish_order applies
  // them.  The allocation-free entry draws them inline.
  // The batched mode is judged by its mean match error, not by the reference's visiting
  // order: its random order (types 0/1) is a permutation drawn on the device from the probe
  // seed (k_random_order), which keeps the host shuffle (about 3 ns per target) off the
  // critical path.
  pr.targetsRM = targetsRM;
  if (!serialMode && type <= 1) {
    HIP_TRY(launch_random_order(s, (job.opt.seed ? job.opt.seed : kEngineSeed) ^ 0x5EED0DE5u, nT, targetsRM, targets));
    pr.orderPending = false;
  } else {
    pr.dPerm = (uint32_t *)W.perm.get((size_t)nT * sizeof(uint32_t) + 64);
    pr.hPerm = (uint32_t *)W.pinPerm.get((size_t)nT * sizeof(uint32_t) + 64);
    if (type >= 2) {
      static uint32_t hRays[9];
      static std::once_flag raysOnce;
      std::call_once(raysOnce, [] { axis_rays(hRays); });
      uint32_t *dRays = (uint32_t *)W.orderRays.get(64);
      HIP_TRY(hipMemcpyAsync(dRays, hRays, sizeof hRays, hipMemcpyHostToDevice, s));
      const size_t otb = order_temp_bytes(nT);
      HIP_TRY(launch_order_sort(s, targetsRM, nT, type, dRays, W.orderTemp.get(otb), otb, &pr.sortedIdx));
    }
    Prng *pp = &prng;
    uint32_t *hPerm = pr.hPerm;
    auto draw = [pp, hPerm, nT, type] {
      if (type <= 1) pp->shuffle(hPerm, nT);
      else band_permutation(hPerm, nT, *pp);
    };
    if (g_devArena) draw();
    else pr.orderThread = std::thread(draw);
    pr.orderPending = true;
  }

  Geometry g = {};
  g.tw = tw; g.th = th; g.cw = cw; g.ch = ch;
  g.nColors = job.nColors;
  g.nMaps = job.nMaps;
  g.tileH = P.isMakeSeamlesslyTileableHorizontally != 0;
  g.tileV = P.isMakeSeamlesslyTileableVertically != 0;
  g.patchLimit = std::max(P.patchSize, 2u);
  g.maxProbe = P.maxProbeCount;
  g.nT = nT; g.nC = nC; g.nO = (unsigned)nO;
  g.penalty = 65535u * (uint32_t)job.nColors + hT->Tm[0] * (uint32_t)job.nMaps;
  g.tilesW = tilesW;
  g.tilesH = tilesH;
  g.rhoMax = rhoMax;
  g.skipEmpty = 0;
  static const int probeKnob = [] {
    const char *e = std::getenv("RESYNTH_PROBE_BY_PIXEL");  // 0: never, 1: always, unset: by density
    return e ? std::atoi(e) : -1;
  }();
  g.probeByPixel = probeKnob >= 0 ? (probeKnob ? 1 : 0) : ((uint64_t)nC * 2 >= (uint64_t)nCpix ? 1 : 0);
  g.drawScale = (unsigned)std::min<uint64_t>(((uint64_t)nCpix << 16) / nC + 1, 64ull << 16);
  // RESYNTH_XCD_BANDS=1: batched random probes come from the XCD's band of corpus rows
  // when the corpus words outgrow an XCD's L2.  Off by default: on a 2048x2048 heal it cuts
  // the visit kernel's HBM fetches about 9x but not its time (the scorer is bound by
  // issue and latency, not by fetches), and it narrows each visit's probe distribution.
  const char *bandEnv = std::getenv("RESYNTH_XCD_BANDS");
  const bool bandsOn = bandEnv ? std::atoi(bandEnv) != 0 : false;
  g.bands = (!serialMode && bandsOn && nCpix * wb > (4u << 20) && ch >= 8 * kBands) ? kBands : 0;

  SynthBuffers b = {};
  b.tword = tword;
  b.sourceOf = sourceOf;
  b.cword = cword;
  b.prober = prober;
  b.targets = targets;
  b.corpusPts = corpusPts;
  b.offsets = offsets;
  b.lastBest = lastBest;
  b.tables = dT;
  b.counters = cnt;
  b.tileHas = tileHas;
  b.ctxRadius = ctxRadius;
  b.firstIdx = firstIdx;
  b.ctl = (uint32_t *)W.wctl.get(WCTL_WORDS * sizeof(uint32_t));
  b.bandStart = bandStart;

  // pass schedule (lib/passes.h:68-93) and progress estimate (lib/progress.c:31-37)
  unsigned sizes[6];
  {
    unsigned n = nT;
    sizes[0] = n;
    for (int i = 1; i < 6; i++) { sizes[i] = n; n = (unsigned)n * 3 / 4; }
  }
  job.prog.estimated = 0;
  for (int i = 0; i < 6; i++) job.prog.estimated += sizes[i];
  HIP_TRY(hipStreamSynchronize(s));
  job.st.prepMs = now_ms() - t0;
  job.st.pixelWordBytes = (unsigned)wb;
  pr.g = g;
  pr.b = b;
  std::copy(sizes, sizes + 6, pr.sizes);
  pr.wb = wb;
  if (!deferOrder) finish_order(pr);
  return 0;
}

// ------------------------------------------------------------------ learned proposals

constexpr int kKeyBlock = 1 << 20;   // corpus keys embedded per block (256 B of features each)

// Inverted-file index of the corpus for learned proposals: keys at every corpus point of
// even coordinates, their embeddings, centroids (every (nK/C)-th key) and the keys bucketed
// by their nearest centroid.
struct LearnedIndex {
  bool ok = false;
  bool pending = false;   // built on the side stream; the library stream has not waited yet
  std::thread builder;    // enqueues the build on the side stream (learned_build)
  std::exception_ptr builderError;
  ~LearnedIndex()
  {
    if (builder.joinable()) builder.join();
  }
  unsigned nK = 0, C = 0, S = 0, Q = 0;
  int2 *keyPts = nullptr, *keyPtsS = nullptr;      // keyPtsS: in cluster order
  uint16_t *keyZ = nullptr, *keyZs = nullptr;      // keyZs: embeddings in cluster order
  uint16_t *superCent = nullptr, *subCent = nullptr;
  uint32_t *start = nullptr;                       // [C+1] cluster c = keyZs rows start[c]..start[c+1]
};

unsigned bits_for(unsigned v)
{
  unsigned b = 1;
  while (b < 32 && (v >> b)) b++;
  return b;
}

// Share (permille) of the compute units the key embeddings of the index build take: each of
// their workgroups holds a whole CU's registers, and the rest of the CUs stay free for pass
// 0's first chunks on the library stream, which have no other way to start
// (RESYNTH_INDEX_CU_PERMILLE overrides it).
int index_cu_share()
{
  static const int v = [] {
    const char *e = std::getenv("RESYNTH_INDEX_CU_PERMILLE");
    const long x = e ? std::atol(e) : 600;
    return (int)std::min(1000L, std::max(1L, x));
  }();
  return v;
}

// Two-level inverted file over the corpus keys: S super-centroids (every (nK/S)-th key),
// Q = 128 sub-centroids inside each, about 240 keys per cluster.
void learned_build(Job &job, const Prepared &pr, LearnedIndex &L, hipEvent_t builtEvent)
{
  hipStream_t s = g_dev.stream;
  Workspace &W = *job.W;
  const ResynthModel *M = job.opt.model;
  const Geometry &g = pr.g;
  const int w2 = (g.cw + 1) / 2, h2 = (g.ch + 1) / 2;
  const size_t nGrid = (size_t)w2 * h2;
  uint8_t *flags = (uint8_t *)W.keyFlags.get(nGrid + 64);
  HIP_TRY(learn_key_flags(s, job.wide, pr.b.cword, g.cw, g.ch, flags));
  L.keyPts = (int2 *)W.keyPts.get(nGrid * sizeof(int2) + 64);
  unsigned *dCount = (unsigned *)W.sCount.get(4096 + 64);
  const size_t tb = compact_temp_bytes(nGrid);
  HIP_TRY(launch_compact(s, flags, w2, h2, L.keyPts, dCount, W.keyTemp.get(tb + 64)));
  unsigned nK = 0;
  HIP_TRY(hipMemcpyAsync(&nK, dCount, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  L.nK = nK;
  if (nK < 64) return;
  // From here the build reads only the corpus words, which synthesis never writes, and
  // buffers of its own: it runs on the side stream, overlapping pass 0 (whose early chunks
  // leave most compute units idle), and the first proposal's search waits for indexReady
  // (its queries' embedding only for the network's weights, weightsReady).  The
  // synchronization above has completed everything it depends on.  Its buffers are sized
  // here; a helper thread enqueues its kernels, so that the host goes on to enqueue pass 0
  // at once rather than after the build's some fifty launches.
  s = g_dev.side;
  const float *w = model_weights(M);
  const uint16_t *wt = model_weights_t(M, s);
  HIP_TRY(hipEventRecord(g_dev.weightsReady, s));
  L.keyZ = (uint16_t *)W.keyZ.get((size_t)nK * kEmb * 2 + 64);
  uint16_t *x = (uint16_t *)W.keyX.get((size_t)kKeyBlock * kIn * 2);
  const unsigned Q = 128;
  const unsigned S = std::max(1u, std::min(128u, nK / (Q * 240u)));
  L.S = S;
  L.Q = Q;
  L.C = S * Q;
  // super-centroids: keys c * nK / S, c < S (gathered on the device: a copy from pageable
  // host memory would hold the host until the side stream reached it)
  L.superCent = (uint16_t *)W.superCent.get((size_t)S * kEmb * 2 + 64);
  int32_t *sid = (int32_t *)W.superId.get((size_t)nK * 4 + 64);
  // keys sorted by super-cluster, then by cluster (radix sorts: no atomics on hot counters)
  uint32_t *k0 = (uint32_t *)W.sortK0.get((size_t)nK * 4 + 64), *v0 = (uint32_t *)W.sortV0.get((size_t)nK * 4 + 64);
  uint32_t *k1 = (uint32_t *)W.sortK1.get((size_t)nK * 4 + 64), *v1 = (uint32_t *)W.sortV1.get((size_t)nK * 4 + 64);
  const size_t stb = sort_pairs_temp_bytes(nK);
  void *stemp = W.sortTemp.get(stb);
  uint32_t *startS = (uint32_t *)W.sStart.get(((size_t)S + 65) * 4);
  // sub-centroids and final clusters
  L.subCent = (uint16_t *)W.subCent.get((size_t)L.C * kEmb * 2 + 64);
  int32_t *cl = (int32_t *)W.keyCl.get((size_t)nK * 4 + 64);
  uint32_t *tileStart = (uint32_t *)W.tileStart.get(((size_t)S + 65) * 4);
  L.start = (uint32_t *)W.bStart.get(((size_t)L.C + 65) * 4);
  L.keyZs = (uint16_t *)W.keyZs.get((size_t)nK * kEmb * 2 + 64);
  L.keyPtsS = (int2 *)W.keyPtsS.get((size_t)nK * sizeof(int2) + 64);
  const bool wide = job.wide;
  const void *cword = pr.b.cword;
  const int2 *offsets = pr.b.offsets;
  const int share = index_cu_share();
  L.ok = true;
  L.pending = true;
  LearnedIndex *Lp = &L;
  L.builder = std::thread([=]() {
    try {
      for (unsigned b0 = 0; b0 < nK; b0 += kKeyBlock) {
        const int nb = (int)std::min<unsigned>(kKeyBlock, nK - b0);
        HIP_TRY(learn_features(s, wide, cword, g.cw, g.ch, 0, 0, g.nColors, offsets, g.nO, Lp->keyPts + b0, nb, true,
                               true, 0, x));
        HIP_TRY(mlp_embed_fused(s, w, wt, x, nb, Lp->keyZ + (size_t)b0 * kEmb, nullptr, share));
      }
      HIP_TRY(gather_spread_rows_bf16(s, Lp->keyZ, kEmb, nK, (int)S, Lp->superCent));
      HIP_TRY(learn_topk_cols(s, Lp->keyZ, (int)nK, Lp->superCent, (int)S, 1, sid));
      HIP_TRY(learn_sort_keys(s, sid, 1, (int)nK, k0, v0));
      HIP_TRY(launch_sort_pairs(s, k0, k1, v0, v1, nK, bits_for(S), stemp, stb));
      HIP_TRY(learn_bucket_starts(s, k1, (int)nK, (int)S, startS));
      HIP_TRY(learn_sub_index(s, startS, v1, (int)S, (int)Q, (int)nK, Lp->keyZ, Lp->subCent, tileStart, cl));
      HIP_TRY(learn_sort_keys(s, cl, 1, (int)nK, k0, v0));
      HIP_TRY(launch_sort_pairs(s, k0, k1, v0, v1, nK, bits_for(Lp->C), stemp, stb));
      HIP_TRY(learn_bucket_starts(s, k1, (int)nK, (int)Lp->C, Lp->start));
      HIP_TRY(learn_permute_keys(s, v1, (int)nK, Lp->keyZ, Lp->keyPts, Lp->keyZs, Lp->keyPtsS));
      HIP_TRY(hipEventRecord(g_dev.indexReady, s));
      if (builtEvent) HIP_TRY(hipEventRecord(builtEvent, s));
    } catch (...) {
      Lp->builderError = std::current_exception();
    }
  });
}

// Makes the library stream wait for an index still building on the side stream (after its
// helper thread has enqueued the build, or rethrows what stopped it).
void learned_wait(LearnedIndex &L)
{
  if (L.builder.joinable()) L.builder.join();
  if (L.builderError) {
    std::exception_ptr e = L.builderError;
    L.builderError = nullptr;
    L.pending = false;
    std::rethrow_exception(e);
  }
  if (!L.pending) return;
  HIP_TRY(hipStreamWaitEvent(g_dev.stream, g_dev.indexReady, 0));
  L.pending = false;
}

// Proposals (16 corpus points each) for targets k0..k1-1 from their current neighbourhoods.
// Queries embedded and searched per block.  Large blocks keep the searches efficient: the
// queries are sorted by cluster and taken 16 per wavefront, so with many queries per
// cluster a wavefront's 16 mostly share one cluster and walk its keys once (a block of 2^20
// queries holds about 64 per cluster of a 16 384-cluster index; one of 2^16 about 4).
// Scratch per block: about 400 bytes per query.
constexpr unsigned kQueryBlock = 1u << 20;
constexpr unsigned kLearnMinChunk = 256;  // per-chunk proposals for chunks of at least this size

void learned_propose_block(Job &job, const Prepared &pr, LearnedIndex &L, unsigned k0, unsigned k1, int P,
                           int2 *props);

// Proposals for targets k0..k1-1, 16 each, written to out (or to the workspace when out is
// null); returns where they are.
const int2 *learned_propose(Job &job, const Prepared &pr, LearnedIndex &L, unsigned k0, unsigned k1, int P,
                            int2 *out = nullptr)
{
  Workspace &W = *job.W;
  if (!out) out = (int2 *)W.props.get((size_t)(k1 - k0) * 16 * sizeof(int2) + 64);
  for (unsigned b = k0; b < k1; b += kQueryBlock)
    learned_propose_block(job, pr, L, b, std::min(k1, b + kQueryBlock), P, out + (size_t)(b - k0) * 16);
  return out;
}

void learned_propose_block(Job &job, const Prepared &pr, LearnedIndex &L, unsigned k0, unsigned k1, int P,
                           int2 *props)
{
  hipStream_t s = g_dev.stream;
  Workspace &W = *job.W;
  const Geometry &g = pr.g;
  const int n = (int)(k1 - k0);
  uint16_t *x = (uint16_t *)W.qX.get((size_t)n * kIn * 2 + 64);
  uint16_t *z = (uint16_t *)W.qZ.get((size_t)n * kEmb * 2 + 64);
  int32_t *qcl = (int32_t *)W.qCl.get((size_t)n * 2 * 4 + 64);
  HIP_TRY(learn_features(s, job.wide, pr.b.tword, g.tw, g.th, g.tileH, g.tileV, g.nColors, pr.b.offsets,
                         g.nO, pr.b.targets + k0, n, false, false, 0, x));
  // The queries' features and embedding need only the targets' state and the network, so
  // they run while the index build finishes on the side stream; the search waits for it.
  HIP_TRY(hipStreamWaitEvent(s, g_dev.weightsReady, 0));
  HIP_TRY(mlp_embed_fused(s, model_weights(job.opt.model), model_weights_t(job.opt.model, s), x, n, z, nullptr));
  learned_wait(L);
  int32_t *qs = (int32_t *)W.qSuper.get((size_t)n * 4 + 64);
  HIP_TRY(learn_topk_cols(s, z, n, L.superCent, (int)L.S, 1, qs));
  // Both searches visit the queries sorted by the cluster they search, so that a tile of 16
  // queries mostly shares its rows and consecutive tiles read them from L2.  The sort keys
  // only need the bits that tell clusters apart (ids below 2^bits; a -1 may land anywhere,
  // which costs locality only).
  uint32_t *qk0 = (uint32_t *)W.qK0.get((size_t)n * 4 + 64), *qv0 = (uint32_t *)W.qV0.get((size_t)n * 4 + 64);
  uint32_t *qk1 = (uint32_t *)W.qK1.get((size_t)n * 4 + 64), *qv1 = (uint32_t *)W.qV1.get((size_t)n * 4 + 64);
  const size_t qtb = sort_pairs_temp_bytes(n);
  void *qtemp = W.qSortTemp.get(qtb);
  HIP_TRY(learn_sort_keys(s, qs, 1, n, qk0, qv0));
  HIP_TRY(launch_sort_pairs(s, qk0, qk1, qv0, qv1, (size_t)n, bits_for(L.S), qtemp, qtb));
  HIP_TRY(learn_query_sub(s, z, n, qv1, qs, (int)L.Q, L.subCent, P, qcl));
  HIP_TRY(learn_count_scored(s, qcl, n, P, L.start, &pr.b.counters->keysScored));
  HIP_TRY(learn_sort_keys(s, qcl, P, n, qk0, qv0));
  HIP_TRY(launch_sort_pairs(s, qk0, qk1, qv0, qv1, (size_t)n, bits_for(L.C), qtemp, qtb));
  HIP_TRY(learn_fine_top16_sorted(s, z, n, qv1, qcl, P, L.keyZs, L.start, L.keyPtsS, props));
}

// The refiner (lib/refiner.h:43-122) over a prepared run: serial or batched passes with
// the reference's progress, cancel and stop rule, then the run's statistics.
int refine_engine(Job &job, Prepared &pr)
{
  hipStream_t s = g_dev.stream;
  Workspace &W = *job.W;
  const double t1 = now_ms();
  const Geometry &g0 = pr.g;
  Geometry g = g0;
  SynthBuffers &b = pr.b;
  const unsigned nT = pr.nT;
  const unsigned *sizes = pr.sizes;
  const bool serial = job.opt.mode == RESYNTH_MODE_SERIAL;
  // without a learned index to build first, nothing else can overlap the host shuffle
  if (serial || !job.opt.model) finish_order(pr);
  Prng &prng = pr.prng;
  Counters *cnt = b.counters;
  const uint32_t *lastBest = b.lastBest;
  const int2 *targets = b.targets;
  const int tw = g.tw;

  std::vector<TimedLaunch> ev;
  unsigned launches = 0;

  // serial mode: raw PRNG outputs after the ordering draws, streamed to the device
  // raw outputs buffered on the device; RESYNTH_SERIAL_STREAM_CAP shrinks the buffer so
  // that tests exercise the refill path
  uint32_t streamCap = g_devArena ? (1u << 20) : (1u << 22);   // the arena's share is 4 MB
  if (const char *capEnv = std::getenv("RESYNTH_SERIAL_STREAM_CAP")) {
    long v = std::atol(capEnv);
    if (v >= 1024) streamCap = (uint32_t)std::min<long>(v, 1L << 30);
  }
  uint32_t *dStream = nullptr, *hStream = nullptr;
  StreamState *dState = nullptr, *hState = nullptr;
  uint32_t streamLen = 0, streamPos = 0;
  if (serial) {
    dStream = (uint32_t *)W.streamBuf.get((size_t)streamCap * 4);
    hStream = (uint32_t *)W.pinStream.get((size_t)streamCap * 4 + sizeof(StreamState) * 2);
    hState = (StreamState *)(hStream + streamCap);
    dState = (StreamState *)W.streamState.get(sizeof(StreamState));
  }
  auto refill = [&]() {
    // keep the unconsumed tail, append fresh outputs, upload
    uint32_t tail = streamLen - streamPos;
    HIP_TRY(hipStreamSynchronize(s));
    std::memmove(hStream, hStream + streamPos, (size_t)tail * 4);
    for (uint32_t i = tail; i < streamCap; i++) hStream[i] = prng.raw();
    streamLen = streamCap;
    streamPos = 0;
    HIP_TRY(hipMemcpyAsync(dStream, hStream, (size_t)streamCap * 4, hipMemcpyHostToDevice, s));
    hState->pos = 0;
    hState->kDone = 0;
    hState->aborted = 0;
    HIP_TRY(hipMemcpyAsync(dState, hState, sizeof(StreamState), hipMemcpyHostToDevice, s));
  };
  if (serial) refill();

  // Batched schedule.  Concurrent visits do not see each other's results, so a chunk is
  // kept small relative to what the pass has already written: in pass 0 a chunk holds 6%
  // of the targets visited so far (at least nT/2048, within [1, 64]); later passes use
  // the cap, nT/32 clamped to [64, 65536] unless the caller sets it (the mean match error
  // grows with the cap's share of the targets: about +1.5% at nT/64, +2% at nT/32 and
  // +6% at nT/8 over the serial engine on 512x512 heals).
  const unsigned cap = job.opt.batch ? job.opt.batch : std::min(65536u, std::max(64u, nT / 32u));
  // RESYNTH_MIN_CHUNK overrides the smallest pass-0 chunk (for tuning studies)
  static const unsigned minChunkEnv = [] {
    const char *e = std::getenv("RESYNTH_MIN_CHUNK");
    return e ? (unsigned)std::max(1L, std::atol(e)) : 0u;
  }();
  const unsigned minChunk =
      std::min(cap, job.opt.firstChunk   ? job.opt.firstChunk
                    : minChunkEnv        ? minChunkEnv
                                         : std::max(1u, std::min(64u, nT / 2048u)));
  static const unsigned rampDefault = [] {
    const char *e = std::getenv("RESYNTH_RAMP_PERMILLE");
    long v = e ? std::atol(e) : 60;
    return (unsigned)std::min(1000L, std::max(1L, v));
  }();
  const unsigned rampPermille = job.opt.rampPermille ? std::min(1000u, job.opt.rampPermille) : rampDefault;
  // Sliding-window launches (RESYNTH_WINDOW=1 turns them on): once a chunk would reach
  // RESYNTH_WINDOW_MIN targets, the rest of the pass runs as one persistent launch whose
  // window is half the chunk the schedule above would use (a visit in a chunk misses half
  // the chunk on average, a visit in a window up to the whole window).  Off by default:
  // the two device-wide counters every visit updates cost more than the chunk tails.
  const int windowOn = [] {
    const char *e = std::getenv("RESYNTH_WINDOW");
    return e ? std::atoi(e) : 0;
  }();
  const unsigned windowMin = [] {
    const char *e = std::getenv("RESYNTH_WINDOW_MIN");
    long v = e ? std::atol(e) : 2048;
    return (unsigned)std::max(64L, v);
  }();
  unsigned windowLaunches = 0, windowExpired = 0;
  LearnedIndex learned;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> proposeEvents;
  int2 *propsAll = nullptr;
  std::pair<hipEvent_t, hipEvent_t> indexEvents{nullptr, nullptr};
  if (job.opt.model && !serial) {
    const double tl = now_ms();
    if (job.opt.timing) {
      HIP_TRY(hipEventCreate(&indexEvents.first));
      HIP_TRY(hipEventCreate(&indexEvents.second));
      HIP_TRY(hipEventRecord(indexEvents.first, s));
    }
    learned_build(job, pr, learned, indexEvents.second);
    if (job.opt.timing && !learned.pending) HIP_TRY(hipEventRecord(indexEvents.second, s));
    job.st.learnedKeys = learned.nK;
    job.st.learnedClusters = learned.ok ? learned.C : 0;
    job.st.learnedIndexMs = now_ms() - tl;   // host time; the device time replaces it when timed
  }
  finish_order(pr);
  unsigned pass;
  unsigned passesRun = 0;
  unsigned bettersBefore = 0;
  bool lastBettersPending = false;
  // pinned, so that the counter copies below do not hold the host
  Counters *hcPass = (Counters *)W.pinPass.get(2 * sizeof(Counters) + 64);
  unsigned long long *hSum = (unsigned long long *)(hcPass + 1);
  // proposals of every target from the state at the start of pass p
  unsigned proposedFor = ~0u;
  auto refreshes = [&](unsigned p) { return learned.ok && p < 6 && ((job.opt.learnedRefresh >> p) & 1u); };
  auto propose_pass = [&](unsigned p) {
    hipEvent_t pa = nullptr, pb = nullptr;
    if (job.opt.timing) {
      HIP_TRY(hipEventCreate(&pa));
      HIP_TRY(hipEventCreate(&pb));
      HIP_TRY(hipEventRecord(pa, s));
    }
    propsAll = (int2 *)W.propsAll.get((size_t)nT * 16 * sizeof(int2) + 64);
    learned_propose(job, pr, learned, 0, nT, job.opt.learnedClusters == 1 ? 1 : 2, propsAll);
    if (job.opt.timing) {
      HIP_TRY(hipEventRecord(pb, s));
      proposeEvents.push_back({pa, pb});
    }
    proposedFor = p;
  };
  for (pass = 0; pass < 6; pass++) {
    const unsigned end = sizes[pass];
    // every target has a value after pass 0; scans can skip empty rings only in pass 0,
    // and wrapped neighbourhoods (tiling) are not covered by the tile walk
    g.skipEmpty = (pass == 0 && !g.tileH && !g.tileV) ? 1 : 0;
    if (refreshes(pass) && proposedFor != pass) propose_pass(pass);
    // lib/synthesize.h:493-497: progress + cancel poll before target k when k & 4095 == 0
    unsigned k = 0;
    bool stop = false;
    while (k < end && !stop) {
      unsigned k1;
      unsigned windowLen = 0;
      if (serial) {
        k1 = std::min(end, (k & ~kCallbackMask) + 4096u);
      } else {
        unsigned len = (pass == 0) ? std::max(minChunk, (unsigned)((unsigned long long)k * rampPermille / 1000u)) : cap;
        len = std::min(len, cap);
        k1 = std::min(end, k + len);
        if (windowOn && len >= windowMin) {
          k1 = end;
          windowLen = len;
        }
      }
      // progress points inside [k, k1)
      unsigned firstPoint = (k + kCallbackMask) & ~kCallbackMask;
      for (unsigned pt = firstPoint; pt < k1; pt += 4096u) {
        job.prog.tick();
        if (job.prog.cancelled()) {
          k1 = pt;
          stop = true;
          break;
        }
      }
      if (k1 > k) {
        if (serial) {
          unsigned kk = k;
          while (kk < k1) {
            if (streamLen - streamPos < streamCap / 2) refill();
            record_event(job, ev, true);
            HIP_TRY(launch_synth_serial(s, job.wide, g, b, kk, k1, job.opt.prng, dStream, streamLen, dState));
            record_event(job, ev, false);
            launches++;
            HIP_TRY(hipMemcpyAsync(hState, dState, sizeof(StreamState), hipMemcpyDeviceToHost, s));
            HIP_TRY(hipStreamSynchronize(s));
            streamPos = hState->pos;
            // read before a grow below moves hState to the new (unwritten) pinned buffer
            const uint32_t kDone = hState->kDone;
            if (kDone == kk && streamPos == 0) {
              // one visit needs more draws than the stream holds: grow it
              if (streamCap >= (1u << 30)) {
                std::fprintf(stderr, "resynth_mi355x: maxProbeCount too large for the serial stream\n");
                return RESYNTH_ERROR_TOO_LARGE;
              }
              streamCap *= 2;
              dStream = (uint32_t *)W.streamBuf.get((size_t)streamCap * 4);
              hStream = (uint32_t *)W.pinStream.grow_keep((size_t)streamCap * 4 + sizeof(StreamState) * 2,
                                                          (size_t)streamLen * 4);
              hState = (StreamState *)(hStream + streamCap);
            }
            kk = kDone;
            if (kk < k1) refill();   // stopped for want of PRNG outputs
          }
        } else if (windowLen) {
          record_event(job, ev, true);
          HIP_TRY(launch_synth_window(s, job.wide, g, b, k, k1, job.opt.seed ? job.opt.seed : kEngineSeed, pass,
                                      windowLen / 2, pass == 0 ? rampPermille / 2 : 0, std::max(1u, cap / 2)));
          record_event(job, ev, false);
          launches++;
          windowLaunches++;
        } else {
          Geometry gl = g;
          gl.nVisited = g.skipEmpty ? k : 0u;   // pass 0: targets before the chunk hold values
          SynthBuffers bl = b;
          const bool learnPass = !job.opt.learnedPasses || pass < job.opt.learnedPasses;
          if (learned.ok && job.opt.learnedRefresh && propsAll) {
            bl.props = propsAll;
            bl.propsK0 = 0;
            gl.maxProbe = job.opt.learnedProbes ? job.opt.learnedProbes : 32u;
            job.st.learnedLaunches++;
          } else if (learned.ok && job.opt.learnedRefresh) {
            if (job.opt.learnedProbesLater) gl.maxProbe = job.opt.learnedProbesLater;
          } else if (learned.ok && learnPass && k1 - k >= kLearnMinChunk) {
            hipEvent_t pa = nullptr, pb = nullptr;
            if (job.opt.timing) {
              HIP_TRY(hipEventCreate(&pa));
              HIP_TRY(hipEventCreate(&pb));
              HIP_TRY(hipEventRecord(pa, s));
            }
            bl.props = learned_propose(job, pr, learned, k, k1, job.opt.learnedClusters == 1 ? 1 : 2);
            if (job.opt.timing) {
              HIP_TRY(hipEventRecord(pb, s));
              proposeEvents.push_back({pa, pb});
            }
            bl.propsK0 = k;
            gl.maxProbe = job.opt.learnedProbes ? job.opt.learnedProbes : 32u;
            job.st.learnedLaunches++;
          } else if (learned.ok && !learnPass && job.opt.learnedProbesLater) {
            gl.maxProbe = job.opt.learnedProbesLater;
          }
          record_event(job, ev, true, job.opt.lanesPerTarget || k1 - k >= synth_wave_min() || bl.props);
          HIP_TRY(launch_synth_batched(s, job.wide, gl, bl, k, k1, job.opt.seed ? job.opt.seed : kEngineSeed, pass,
                                       (int)job.opt.lanesPerTarget));
          record_event(job, ev, false);
          launches++;
        }
      }
      k = k1;
    }
    // A pass after which the loop ends whatever the stop rule says (the pass cap, or the
    // sixth pass) leaves its counters to the statistics' read below: one host round trip less.
    const bool capped = pass == 5 || (!serial && job.opt.maxPasses && pass + 1 >= job.opt.maxPasses);
    if (capped && !windowLaunches) {
      job.st.passSize[pass] = end;
      passesRun = pass + 1;
      job.st.visits += k;
      lastBettersPending = true;
      break;
    }
    uint32_t expired = 0;
    HIP_TRY(hipMemcpyAsync(hcPass, cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
    if (windowLaunches) HIP_TRY(hipMemcpyAsync(&expired, b.ctl + WCTL_EXPIRED, 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipEventRecord(g_dev.passCounts, s));
    // The next pass's proposals read only this pass's result: they go in before the host
    // waits for the stop rule's counters (and are dropped if the rule ends the run).
    if (refreshes(pass + 1)) propose_pass(pass + 1);
    HIP_TRY(hipEventSynchronize(g_dev.passCounts));
    const Counters &hc = *hcPass;
    if (expired) windowExpired++;
    unsigned bettersNow = 0;
    for (int i = 0; i < CNT_SLOTS; i++) bettersNow += hc.betters[i];
    const unsigned betters = bettersNow - bettersBefore;
    bettersBefore = bettersNow;
    job.st.passSize[pass] = end;
    job.st.betters[pass] = betters;
    passesRun = pass + 1;
    job.st.visits += k;
    // lib/refiner.h:111-115
    if ((double)((float)betters / (float)nT) < 0.1) break;
    if (!serial && job.opt.maxPasses && passesRun >= job.opt.maxPasses) break;
  }
  learned_wait(learned);   // an index no pass used still finishes before the call returns
  job.st.passes = passesRun;
  job.st.launches = launches;
  job.st.windowLaunches = windowLaunches;
  job.st.windowWaitsExpired = windowExpired;
  if (windowExpired)
    std::fprintf(stderr, "resynth_mi355x: %u sliding-window pass(es) cut a wait short at its spin limit\n",
                 windowExpired);

  // statistics
  {
    unsigned long long *dSum = (unsigned long long *)W.meanSum.get(16);
    HIP_TRY(launch_fill_u32(s, (uint32_t *)dSum, 0u, 4));
    HIP_TRY(launch_mean_error(s, lastBest, pr.targetsRM ? pr.targetsRM : targets, nT, tw, dSum));
    unsigned long long *hs = hSum;
    HIP_TRY(hipMemcpyAsync(hs, dSum, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hcPass, cnt, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const Counters &hc = *hcPass;
    if (lastBettersPending) {
      unsigned bettersNow = 0;
      for (int i = 0; i < CNT_SLOTS; i++) bettersNow += hc.betters[i];
      job.st.betters[passesRun - 1] = bettersNow - bettersBefore;
    }
    job.st.meanError = hs[1] ? (double)hs[0] / (double)hs[1] : 0.0;
    for (int i = 0; i < CNT_SLOTS; i++) {
      job.st.candidates += hc.candidates[i];
      job.st.patchTerms += hc.terms[i];
      job.st.wavePatchTerms += hc.termsWave[i];
      job.st.waveScanned += hc.scanWave[i];
      job.st.waveStaged += hc.stageWave[i];
      job.st.waveVisits += hc.visitsWave[i];
    }
    job.st.learnedKeysScored = hc.keysScored;
    double ms = 0;
    double wms = 0;
    for (auto &e : ev) {
      float f = 0;
      HIP_TRY(hipEventElapsedTime(&f, e.a, e.b));
      ms += f;
      if (e.wave) {
        wms += f;
        job.st.waveLaunches++;
      }
      HIP_TRY(hipEventDestroy(e.a));
      HIP_TRY(hipEventDestroy(e.b));
    }
    job.st.synthKernelMs = ms;
    job.st.waveKernelMs = wms;
    for (auto &e : proposeEvents) {
      float f = 0;
      HIP_TRY(hipEventElapsedTime(&f, e.first, e.second));
      job.st.learnedProposeMs += f;
      HIP_TRY(hipEventDestroy(e.first));
      HIP_TRY(hipEventDestroy(e.second));
    }
    if (indexEvents.first) {
      float f = 0;
      HIP_TRY(hipEventElapsedTime(&f, indexEvents.first, indexEvents.second));
      job.st.learnedIndexMs = f;
      HIP_TRY(hipEventDestroy(indexEvents.first));
      HIP_TRY(hipEventDestroy(indexEvents.second));
    }
  }
  job.st.synthMs = now_ms() - t1;
  return 0;
}

// The engine (lib/engine.c:539-690) on device buffers.  Returns a TImageSynthError;
// the synthesized colours are in the target words of the job's workspace on success.
int run_engine(Job &job, unsigned &nTargetsOut)
{
  Prepared pr;
  const int err = prepare_engine(job, pr, true);
  nTargetsOut = pr.nT;
  if (err) return err;
  return refine_engine(job, pr);
}

void init_job(Job &job, const TImageSynthParameters *params, const ResynthOptions *opt, void (*cb)(int, void *),
              void *ctx, int *cancel)
{
  if (params) job.P = *params;
  else setDefaultParams(&job.P);
  if (opt) job.opt = *opt;
  else resynth_default_options(&job.opt);
  std::memset(&job.st, 0, sizeof job.st);
  job.prog.cb = cb;
  job.prog.ctx = ctx;
  job.prog.cancel = cancel;
  job.W = &g_dev.ws;
}

// Builds the corpus mask of imageSynth2 as lib/adaptSimple.h:48-96 fills it: mask2's own
// width x height bytes, in its own row order, written consecutively into an image-sized
// bytemap that starts out all zero.
bool remap_mask2(const ImageBuffer *image, const ImageBuffer *mask2, std::vector<unsigned char> &out)
{
  const size_t n = (size_t)image->width * image->height;
  if ((size_t)mask2->width * mask2->height > n) return false;
  out.assign(n, 0);
  size_t d = 0;
  for (unsigned r = 0; r < mask2->height; r++)
    for (unsigned c = 0; c < mask2->width; c++) out[d++] = mask2->data[(size_t)r * mask2->rowBytes + c];
  return true;
}

// `after` (may be null) runs under the library lock once the engine has succeeded, with the
// run's device state still in the workspace.
int image_synth_impl(ImageBuffer *image, ImageBuffer *mask, ImageBuffer *mask2, TImageFormat format,
                     TImageSynthParameters *params, void (*cb)(int, void *), void *ctx, int *cancelFlag,
                     const ResynthOptions *opt, ResynthStats *stats,
                     const std::function<void(Job &, unsigned)> *after)
{
  // lib/imageSynth.c:62-137 / 140-217
  if (image->width != mask->width || image->height != mask->height) return IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH;
  TFormatIndices ix;
  if (prepareImageFormatIndicesFromFormatType(&ix, format)) return IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT;
  // rows shorter than their pixels cannot be what the sizes say
  if ((size_t)image->width * countPixelelsPerPixelForFormat(format) > image->rowBytes ||
      (size_t)mask->width > mask->rowBytes || (image->width && image->height && (!image->data || !mask->data)))
    return RESYNTH_ERROR_INVALID_ARGUMENT;
  if (mask2 && (!mask2->data || (size_t)mask2->width > mask2->rowBytes)) return RESYNTH_ERROR_INVALID_ARGUMENT;
  std::vector<unsigned char> m2;
  if (mask2 && !remap_mask2(image, mask2, m2)) return IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
  const double t0 = now_ms();
  hipStream_t s = g_dev.stream;
  const unsigned w = image->width, h = image->height;
  const unsigned bpp = countPixelelsPerPixelForFormat(format);
  Job job;
  init_job(job, params, opt, cb, ctx, cancelFlag);
  job.ix = ix;
  Source &S = job.src;
  S.simple = true;
  S.tw = S.cw = (int)w;
  S.th = S.ch = (int)h;
  S.bpp = bpp;
  S.rowBytes = (size_t)w * bpp;
  S.maskRowBytes = w;
  S.mask2RowBytes = w;
  S.alphaIdx = (format == T_RGBA) ? 3 : (format == T_GrayA) ? 1 : -1;
  const size_t npx = (size_t)w * h;
  unsigned char *dImg = (unsigned char *)g_dev.ws.image.get(npx * bpp + 64);
  unsigned char *dMask = (unsigned char *)g_dev.ws.mask.get(npx + 64);
  if (npx) {
    HIP_TRY(hipMemcpy2DAsync(dImg, S.rowBytes, image->data, image->rowBytes, S.rowBytes, h, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpy2DAsync(dMask, w, mask->data, mask->rowBytes, w, h, hipMemcpyHostToDevice, s));
  }
  S.image = dImg;
  S.mask = dMask;
  if (mask2) {
    unsigned char *dM2 = (unsigned char *)g_dev.ws.mask2.get(npx + 64);
    HIP_TRY(hipMemcpyAsync(dM2, m2.data(), npx, hipMemcpyHostToDevice, s));
    S.mask2 = dM2;
  }
  unsigned nT = 0;
  int err = npx ? run_engine(job, nT) : IMAGE_SYNTH_ERROR_EMPTY_TARGET;
  if (!err && after) (*after)(job, nT);
  if (!err && !(cancelFlag && *cancelFlag)) {
    HIP_TRY(launch_writeback_simple(s, job.wide, g_dev.ws.tword.p, (const int2 *)g_dev.ws.targets.p, nT, dImg,
                                    S.rowBytes, bpp, job.nColors, (int)w));
    HIP_TRY(hipMemcpy2DAsync(image->data, image->rowBytes, dImg, S.rowBytes, S.rowBytes, h, hipMemcpyDeviceToHost,
                             s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  job.st.totalMs = now_ms() - t0;
  if (stats) *stats = job.st;
  return err;
}

int engine_impl(TImageSynthParameters parameters, TFormatIndices *indices, Map *targetMap, Map *corpusMap,
                void (*cb)(int, void *), void *ctx, int *cancelFlag, const ResynthOptions *opt, ResynthStats *stats)
{
  if (parameters.patchSize > kMaxNeighbors) return IMAGE_SYNTH_ERROR_PATCH_SIZE_EXCEEDED;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
  const double t0 = now_ms();
  hipStream_t s = g_dev.stream;
  Job job;
  init_job(job, &parameters, opt, cb, ctx, cancelFlag);
  job.ix = *indices;
  Source &S = job.src;
  S.tw = (int)targetMap->width;
  S.th = (int)targetMap->height;
  S.cw = (int)corpusMap->width;
  S.ch = (int)corpusMap->height;
  const size_t depth = indices->total_bpp;
  const size_t tbytes = (size_t)S.tw * S.th * depth, cbytes = (size_t)S.cw * S.ch * depth;
  unsigned char *dT = (unsigned char *)g_dev.ws.tpix.get(tbytes + 64);
  unsigned char *dC = (unsigned char *)g_dev.ws.cpix.get(cbytes + 64);
  if (tbytes) HIP_TRY(hipMemcpyAsync(dT, targetMap->data->data, tbytes, hipMemcpyHostToDevice, s));
  if (cbytes) HIP_TRY(hipMemcpyAsync(dC, corpusMap->data->data, cbytes, hipMemcpyHostToDevice, s));
  S.tpix = dT;
  S.cpix = dC;
  unsigned nT = 0;
  int err = (tbytes && cbytes) ? run_engine(job, nT)
                               : (tbytes ? IMAGE_SYNTH_ERROR_EMPTY_CORPUS : IMAGE_SYNTH_ERROR_EMPTY_TARGET);
  if (!err) {
    // lib/engine.c:675-676: results are left in targetMap
    HIP_TRY(launch_writeback_pixmap(s, job.wide, g_dev.ws.tword.p, (const int2 *)g_dev.ws.targets.p, nT, dT, (int)depth,
                                    job.nColors, S.tw));
    HIP_TRY(hipMemcpyAsync(targetMap->data->data, dT, tbytes, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  job.st.totalMs = now_ms() - t0;
  if (stats) *stats = job.st;
  return err;
}

int image_synth_host(ImageBuffer *image, ImageBuffer *mask, ImageBuffer *mask2, TImageFormat format,
                     TImageSynthParameters *params, void (*cb)(int, void *), void *ctx, int *cancelFlag,
                     const ResynthOptions *opt, ResynthStats *stats,
                     const std::function<void(Job &, unsigned)> *after = nullptr)
{
  if (!image || !mask) return RESYNTH_ERROR_INVALID_ARGUMENT;
  return guarded([&] { return image_synth_impl(image, mask, mask2, format, params, cb, ctx, cancelFlag, opt, stats, after); });
}

int engine_host(TImageSynthParameters parameters, TFormatIndices *indices, Map *targetMap, Map *corpusMap,
                void (*cb)(int, void *), void *ctx, int *cancelFlag, const ResynthOptions *opt, ResynthStats *stats)
{
  if (!indices || !targetMap || !corpusMap) return RESYNTH_ERROR_INVALID_ARGUMENT;
  return guarded(
      [&] { return engine_impl(parameters, indices, targetMap, corpusMap, cb, ctx, cancelFlag, opt, stats); });
}

// imageSynth on device buffers with workspace W (the library lock held, device ready).
int image_synth_device_impl(Workspace &W, unsigned char *d_image, unsigned int width, unsigned int height,
                            size_t rowBytes, const unsigned char *d_mask, size_t maskRowBytes,
                            const unsigned char *d_mask2, size_t mask2RowBytes, TImageFormat imageFormat,
                            const TImageSynthParameters *parameters, void (*progressCallback)(int, void *),
                            void *contextInfo, int *cancelFlag, const ResynthOptions *options, ResynthStats *stats)
{
  TFormatIndices ix;
  if (prepareImageFormatIndicesFromFormatType(&ix, imageFormat)) return IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT;
  const unsigned bpp = countPixelelsPerPixelForFormat(imageFormat);
  if ((size_t)width * height && (!d_image || !d_mask || rowBytes < (size_t)width * bpp || maskRowBytes < width ||
                                 (d_mask2 && mask2RowBytes < width)))
    return RESYNTH_ERROR_INVALID_ARGUMENT;
  const double t0 = now_ms();
  Job job;
  init_job(job, parameters, options, progressCallback, contextInfo, cancelFlag);
  job.W = &W;
  job.ix = ix;
  Source &S = job.src;
  S.simple = true;
  S.tw = S.cw = (int)width;
  S.th = S.ch = (int)height;
  S.bpp = bpp;
  S.image = d_image;
  S.rowBytes = rowBytes;
  S.mask = d_mask;
  S.maskRowBytes = maskRowBytes;
  S.mask2 = d_mask2;
  S.mask2RowBytes = mask2RowBytes;
  S.alphaIdx = (imageFormat == T_RGBA) ? 3 : (imageFormat == T_GrayA) ? 1 : -1;
  unsigned nT = 0;
  int err = ((size_t)width * height) ? run_engine(job, nT) : IMAGE_SYNTH_ERROR_EMPTY_TARGET;
  if (!err && !(cancelFlag && *cancelFlag))
    HIP_TRY(launch_writeback_simple(g_dev.stream, job.wide, W.tword.p, (const int2 *)W.targets.p, nT, d_image,
                                    rowBytes, S.bpp, job.nColors, (int)width));
  HIP_TRY(hipStreamSynchronize(g_dev.stream));
  job.st.totalMs = now_ms() - t0;
  if (stats) *stats = job.st;
  return err;
}

// ------------------------------------------------------------------ many requests

const ManyConfig kDefaultMany = {4, 2, 4, 64};

// Tuned kernel shapes per launch class: pixel-word width, patch-size class and size class
// (zero waves: not tuned).
struct ManyClass {
  int wide, patch, size;
};
ManyConfig g_tuned[2][3][3] = {};

ManyClass many_class(bool wide, const std::vector<unsigned> &idx, const std::vector<Job> &jobs,
                     const std::vector<Prepared> &prs)
{
  double targets = 0, weighted = 0;
  for (unsigned i : idx) {
    targets += prs[i].nT;
    weighted += (double)prs[i].nT * jobs[i].P.patchSize;
  }
  const double patch = targets > 0 ? weighted / targets : 30.0, mean = idx.empty() ? 0 : targets / idx.size();
  return {wide ? 1 : 0, patch <= 16 ? 0 : (patch <= 32 ? 1 : 2), mean < 16384 ? 0 : (mean < 131072 ? 1 : 2)};
}

unsigned env_permille(const char *name, long dflt)
{
  const char *e = std::getenv(name);
  long v = e ? std::atol(e) : dflt;
  return (unsigned)std::min(1000L, std::max(1L, v));
}

// Sliding-window parameters of a request: the single-request schedule's chunk sizes divided
// by RESYNTH_MANY_WINDOW_DIV0 in pass 0 (default 3) and RESYNTH_MANY_WINDOW_DIV later
// (default 8), with the options' pass-0 ramp.  On the 96 requests of the uneven-batch
// workload (every fourth request, three probe seeds) the mean match error against the
// serial engine's, and the launch time, were: 2/2 1.025, 526 ms; 3/3 1.012, 584 ms;
// 3/8 1.004, 597 ms; 8/2 1.003, 735 ms; 6/6 0.998, 653 ms.  Windows in the later passes
// are cheap to narrow: every visit there is short.
void request_window(const ResynthOptions &o, unsigned nT, unsigned &minW, unsigned &ramp, unsigned &capW)
{
  static const unsigned div = [] {
    const char *e = std::getenv("RESYNTH_MANY_WINDOW_DIV");
    const long v = e ? std::atol(e) : 8;
    return (unsigned)std::min(64L, std::max(1L, v));
  }();
  static const unsigned div0 = [] {
    const char *e = std::getenv("RESYNTH_MANY_WINDOW_DIV0");
    const long v = e ? std::atol(e) : 3;
    return (unsigned)std::min(64L, std::max(1L, v));
  }();
  const unsigned cap = o.batch ? o.batch : std::min(65536u, std::max(64u, nT / 32u));
  const unsigned minChunk = std::min(cap, std::max(1u, std::min(64u, nT / 2048u)));
  minW = std::max(1u, minChunk / div0);
  const unsigned rampPermille =
      o.rampPermille ? std::min(1000u, o.rampPermille) : env_permille("RESYNTH_RAMP_PERMILLE", 60);
  ramp = std::max(1u, rampPermille / div0);
  capW = std::max(1u, cap / div);
}

// A request whose buffers cannot be what its sizes say: null data, or a row pitch shorter
// than a row of pixels (image) or of mask bytes.
bool request_invalid(const ResynthRequest &q)
{
  const unsigned bpp = countPixelelsPerPixelForFormat(q.format);
  if (!bpp || !q.image.width || !q.image.height) return false;   // reported by the usual checks
  return !q.image.data || !q.mask.data || q.image.rowBytes < (size_t)q.image.width * bpp ||
         (q.mask.width == q.image.width && q.mask.height == q.image.height && q.mask.rowBytes < q.mask.width);
}

int run_batch(ResynthRequest *reqs, unsigned n, bool device, const ResynthOptions *opt, const ManyConfig *force,
              ResynthBatchStats *st, std::vector<ManyClass> *classesOut = nullptr)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
  const double t0 = now_ms();
  hipStream_t s = g_dev.stream;
  while (g_dev.pool.size() < n) g_dev.pool.emplace_back(new Workspace());
  ResynthOptions o;
  if (opt) o = *opt;
  else resynth_default_options(&o);
  o.mode = RESYNTH_MODE_BATCHED;
  o.timing = 0;
  std::vector<Job> jobs(n);
  std::vector<Prepared> prs(n);
  ResynthBatchStats bs;
  std::memset(&bs, 0, sizeof bs);

  // every request's preparation (lib/imageSynth.c:62-137, lib/engine.c:539-660)
  for (unsigned i = 0; i < n; i++) {
    ResynthRequest &q = reqs[i];
    q.error = 0;
    q.nTargets = 0;
    q.passes = 0;
    q.meanError = 0;
    q.candidates = 0;
    Job &job = jobs[i];
    init_job(job, &q.params, &o, nullptr, nullptr, nullptr);
    job.W = g_dev.pool[i].get();
    const ImageBuffer *image = &q.image, *mask = &q.mask;
    if (image->width != mask->width || image->height != mask->height) {
      q.error = IMAGE_SYNTH_ERROR_IMAGE_MASK_MISMATCH;
      continue;
    }
    if (request_invalid(q)) {
      q.error = RESYNTH_ERROR_INVALID_ARGUMENT;
      continue;
    }
    if (prepareImageFormatIndicesFromFormatType(&job.ix, q.format)) {
      q.error = IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT;
      continue;
    }
    const unsigned w = image->width, h = image->height;
    const unsigned bpp = countPixelelsPerPixelForFormat(q.format);
    const size_t npx = (size_t)w * h;
    if (!npx) {
      q.error = IMAGE_SYNTH_ERROR_EMPTY_TARGET;
      continue;
    }
    Source &S = job.src;
    S.simple = true;
    S.tw = S.cw = (int)w;
    S.th = S.ch = (int)h;
    S.bpp = bpp;
    S.alphaIdx = (q.format == T_RGBA) ? 3 : (q.format == T_GrayA) ? 1 : -1;
    if (device) {
      S.image = image->data;
      S.rowBytes = image->rowBytes;
      S.mask = mask->data;
      S.maskRowBytes = mask->rowBytes;
    } else {
      Workspace &W = *job.W;
      S.rowBytes = (size_t)w * bpp;
      S.maskRowBytes = w;
      unsigned char *dImg = (unsigned char *)W.image.get(npx * bpp + 64);
      unsigned char *dMask = (unsigned char *)W.mask.get(npx + 64);
      HIP_TRY(hipMemcpy2DAsync(dImg, S.rowBytes, image->data, image->rowBytes, S.rowBytes, h, hipMemcpyHostToDevice,
                               s));
      HIP_TRY(hipMemcpy2DAsync(dMask, w, mask->data, mask->rowBytes, w, h, hipMemcpyHostToDevice, s));
      S.image = dImg;
      S.mask = dMask;
    }
    q.error = prepare_engine(job, prs[i]);
    q.nTargets = prs[i].nT;
  }
  bs.prepMs = now_ms() - t0;

  // one multi-request launch per pixel-word width, captured in a graph
  std::vector<unsigned> passesRun(n, 0);
  for (int wide = 0; wide < 2; wide++) {
    std::vector<unsigned> idx;
    for (unsigned i = 0; i < n; i++)
      if (!reqs[i].error && jobs[i].wide == (wide != 0)) idx.push_back(i);
    if (idx.empty()) continue;
    const unsigned m = (unsigned)idx.size();
    ReqDesc *hD = (ReqDesc *)g_dev.pinDesc.get((size_t)m * sizeof(ReqDesc));
    for (unsigned j = 0; j < m; j++) {
      const Prepared &pr = prs[idx[j]];
      const Job &job = jobs[idx[j]];
      ReqDesc &d = hD[j];
      std::memset(&d, 0, sizeof d);
      d.g = pr.g;
      d.g.skipEmpty = (!pr.g.tileH && !pr.g.tileV) ? 1 : 0;
      d.b = pr.b;
      std::copy(pr.sizes, pr.sizes + 6, d.sizes);
      d.seed = job.opt.seed ? job.opt.seed : kEngineSeed;
      request_window(job.opt, pr.nT, d.minW, d.ramp, d.capW);
      d.maxPasses = job.opt.maxPasses;
    }
    ReqDesc *dD = (ReqDesc *)g_dev.manyDesc.get((size_t)m * sizeof(ReqDesc));
    ReqCtl *dC = (ReqCtl *)g_dev.manyCtl.get((size_t)m * sizeof(ReqCtl));
    uint32_t *dG = (uint32_t *)g_dev.manyGlobal.get(MG_WORDS * sizeof(uint32_t));
    HIP_TRY(hipMemcpyAsync(dD, hD, (size_t)m * sizeof(ReqDesc), hipMemcpyHostToDevice, s));
    const ManyClass mc = many_class(wide != 0, idx, jobs, prs);
    if (classesOut) classesOut->push_back(mc);
    const ManyConfig &tuned = g_tuned[mc.wide][mc.patch][mc.size];
    const ManyConfig cfg = force ? *force : (tuned.waves ? tuned : kDefaultMany);
    bs.config.wavesPerBlock = cfg.waves;
    bs.config.firstGroup = cfg.firstGroup;
    bs.config.laterGroup = cfg.laterGroup;
    bs.config.lanesPerTarget = cfg.lanes;
    bool allRgb = true;
    for (unsigned j = 0; j < m; j++) allRgb = allRgb && hD[j].g.nColors == 3 && hD[j].g.nMaps == 0;
    bool instantiated = false;
    HIP_TRY(many_graph_prepare(g_dev.manyGraph[wide], wide != 0, cfg, allRgb, dD, dC, m, dG, instantiated));
    if (instantiated) bs.graphInstantiations++;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s));
    HIP_TRY(hipGraphLaunch(g_dev.manyGraph[wide].exec, s));
    HIP_TRY(hipEventRecord(e1, s));
    ReqCtl *hC = (ReqCtl *)g_dev.pinCtl.get((size_t)m * sizeof(ReqCtl) + MG_WORDS * sizeof(uint32_t));
    uint32_t *hG = (uint32_t *)(hC + m);
    HIP_TRY(hipMemcpyAsync(hC, dC, (size_t)m * sizeof(ReqCtl), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(hG, dG, MG_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    bs.synthKernelMs += ms;
    bs.graphLaunches++;
    if (hG[MG_EXPIRED]) bs.waitsExpired++;
    for (unsigned j = 0; j < m; j++) passesRun[idx[j]] = (unsigned)hC[j].passesRun;
    HIP_TRY(hipEventDestroy(e0));
    HIP_TRY(hipEventDestroy(e1));
  }
  if (bs.waitsExpired)
    std::fprintf(stderr, "resynth_mi355x: a multi-request launch cut a wait short at its spin limit\n");

  // results: mean error, counters and the write-back of every request
  std::vector<Counters> hc(n);
  std::vector<unsigned long long> sums(2 * (size_t)n, 0);
  for (unsigned i = 0; i < n; i++) {
    if (reqs[i].error) continue;
    Job &job = jobs[i];
    Prepared &pr = prs[i];
    Workspace &W = *job.W;
    unsigned long long *dSum = (unsigned long long *)W.meanSum.get(16);
    HIP_TRY(launch_fill_u32(s, (uint32_t *)dSum, 0u, 4));
    HIP_TRY(launch_mean_error(s, pr.b.lastBest, pr.targetsRM ? pr.targetsRM : pr.b.targets, pr.nT, pr.g.tw, dSum));
    HIP_TRY(hipMemcpyAsync(&sums[2 * i], dSum, 16, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&hc[i], pr.b.counters, sizeof(Counters), hipMemcpyDeviceToHost, s));
    const Source &S = job.src;
    unsigned char *dImg = const_cast<unsigned char *>(S.image);
    HIP_TRY(launch_writeback_simple(s, job.wide, pr.b.tword, pr.b.targets, pr.nT, dImg, S.rowBytes, S.bpp,
                                    job.nColors, S.tw));
    if (!device)
      HIP_TRY(hipMemcpy2DAsync(reqs[i].image.data, reqs[i].image.rowBytes, dImg, S.rowBytes, S.rowBytes,
                               (size_t)S.th, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipStreamSynchronize(s));
  for (unsigned i = 0; i < n; i++) {
    ResynthRequest &q = reqs[i];
    if (q.error) continue;
    q.passes = passesRun[i];
    q.meanError = sums[2 * i + 1] ? (double)sums[2 * i] / (double)sums[2 * i + 1] : 0.0;
    for (int k = 0; k < CNT_SLOTS; k++) {
      q.candidates += hc[i].candidates[k];
      bs.patchTerms += hc[i].terms[k];
    }
    bs.candidates += q.candidates;
    bs.requests++;
    bs.targets += q.nTargets;
  }
  bs.totalMs = now_ms() - t0;
  if (st) *st = bs;
  return 0;
}

// Times every kernel shape on private device copies of the sample (the median of three runs
// each) and keeps the fastest for the sample's launch classes.
int autotune_batch(const ResynthRequest *sample, unsigned n, bool device, ResynthKernelConfig *best, double *msOut)
{
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
  }
  hipStream_t s = g_dev.stream;
  std::vector<ResynthRequest> reqs(sample, sample + n);
  std::vector<TmpBuf> pristine(n), work(n), masks(n);
  std::vector<size_t> bytes(n, 0), mbytes(n, 0);
  for (unsigned i = 0; i < n; i++) {
    const ResynthRequest &q = sample[i];
    if (request_invalid(q)) {
      // run_batch reports it; no pointer of the caller's is handed to the device
      reqs[i].image.data = nullptr;
      reqs[i].mask.data = nullptr;
      continue;
    }
    bytes[i] = q.image.rowBytes * q.image.height;
    mbytes[i] = q.mask.rowBytes * q.mask.height;
    HIP_TRY(pristine[i].alloc(bytes[i]));
    HIP_TRY(work[i].alloc(bytes[i]));
    HIP_TRY(masks[i].alloc(mbytes[i]));
    const hipMemcpyKind k = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    HIP_TRY(hipMemcpyAsync(pristine[i].p, q.image.data, bytes[i], k, s));
    HIP_TRY(hipMemcpyAsync(masks[i].p, q.mask.data, mbytes[i], k, s));
    reqs[i].image.data = (unsigned char *)work[i].p;
    reqs[i].mask.data = (unsigned char *)masks[i].p;
  }
  auto restore = [&]() {
    for (unsigned i = 0; i < n; i++)
      if (work[i].p) HIP_TRY(hipMemcpyAsync(work[i].p, pristine[i].p, bytes[i], hipMemcpyDeviceToDevice, s));
  };
  double bestMs = 1e30;
  ManyConfig bestCfg = kDefaultMany;
  int c = 0, rc = 0;
  std::vector<ManyClass> classes;
  restore();
  rc = run_batch(reqs.data(), n, true, nullptr, &kDefaultMany, nullptr, &classes);   // warm-up
  const int wavesList[3] = {2, 4, 8}, lanesList[3] = {16, 32, 64};
  for (int wv : wavesList)
    for (int ln : lanesList) {
      if (rc) break;
      const ManyConfig cfg = {wv, 2, 4, ln};
      double t[3];
      for (int rep = 0; rep < 3 && !rc; rep++) {
        restore();
        ResynthBatchStats st;
        rc = run_batch(reqs.data(), n, true, nullptr, &cfg, &st);
        t[rep] = st.synthKernelMs;
      }
      if (rc) break;
      std::sort(t, t + 3);
      if (msOut) msOut[c] = t[1];
      if (t[1] < bestMs) {
        bestMs = t[1];
        bestCfg = cfg;
      }
      c++;
    }
  HIP_TRY(hipStreamSynchronize(s));
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    for (const ManyClass &mc : classes) g_tuned[mc.wide][mc.patch][mc.size] = bestCfg;
  }
  if (best) {
    best->wavesPerBlock = bestCfg.waves;
    best->firstGroup = bestCfg.firstGroup;
    best->laterGroup = bestCfg.laterGroup;
    best->lanesPerTarget = bestCfg.lanes;
  }
  return 0;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

void setDefaultParams(TImageSynthParameters *param)
{
  // lib/engineParams.c:9-18
  param->isMakeSeamlesslyTileableHorizontally = 0;
  param->isMakeSeamlesslyTileableVertically = 0;
  param->matchContextType = 1;
  param->mapWeight = 0.5;
  param->sensitivityToOutliers = 0.117;
  param->patchSize = 30;
  param->maxProbeCount = 200;
}

unsigned int countPixelelsPerPixelForFormat(TImageFormat format)
{
  switch (format) {
  case T_RGB: return 3;
  case T_RGBA: return 4;
  case T_Gray: return 1;
  case T_GrayA: return 2;
  default: return 0;
  }
}

void prepareImageFormatIndices(TFormatIndices *ix, unsigned int colors, unsigned int maps, int alphaTarget,
                               int alphaSource, int isMap)
{
  // lib/imageFormat.c:117-207: [mask][colours][alpha if either image has one][maps]
  ix->img_match_bpp = (TPixelelIndex)colors;
  ix->colorEndBip = (TPixelelIndex)(1 + colors);
  if (alphaTarget || alphaSource) {
    ix->alpha_bip = ix->colorEndBip;
    ix->map_start_bip = (TPixelelIndex)(ix->colorEndBip + 1);
  } else {
    ix->alpha_bip = 0;
    ix->map_start_bip = ix->colorEndBip;
  }
  ix->map_match_bpp = isMap ? (TPixelelIndex)maps : 0;
  ix->map_end_bip = (TPixelelIndex)(ix->map_start_bip + ix->map_match_bpp);
  ix->total_bpp = ix->map_end_bip;
  ix->isAlphaTarget = alphaTarget;
  ix->isAlphaSource = alphaSource;
}

int prepareImageFormatIndicesFromFormatType(TFormatIndices *ix, TImageFormat format)
{
  switch (format) {
  case T_RGB: prepareImageFormatIndices(ix, 3, 0, 0, 0, 0); return 0;
  case T_RGBA: prepareImageFormatIndices(ix, 3, 0, 1, 1, 0); return 0;
  case T_Gray: prepareImageFormatIndices(ix, 1, 0, 0, 0, 0); return 0;
  case T_GrayA: prepareImageFormatIndices(ix, 1, 0, 1, 1, 0); return 0;
  default: return IMAGE_SYNTH_ERROR_INVALID_IMAGE_FORMAT;
  }
}

void prepareDefaultFormatIndices(TFormatIndices *ix)
{
  // lib/imageFormat.c:209-240: RGBA, no maps
  ix->img_match_bpp = 3;
  ix->colorEndBip = 4;
  ix->alpha_bip = 4;
  ix->map_match_bpp = 0;
  ix->map_start_bip = 5;
  ix->map_end_bip = 5;
  ix->total_bpp = 5;
  ix->isAlphaTarget = 1;
  ix->isAlphaSource = 1;
}

// ---- maps (lib/mapOps.h:45-168) ----
void free_map(Map *map)
{
  if (map->data) {
    std::free(map->data->data);
    std::free(map->data);
  }
  map->data = nullptr;
}

static void new_map(Map *map, unsigned w, unsigned h, unsigned depth, unsigned eltSize)
{
  map->width = w;
  map->height = h;
  map->depth = depth;
  map->data = (GArray *)std::calloc(1, sizeof(GArray));
  size_t bytes = (size_t)w * h * eltSize;
  map->data->data = (char *)std::calloc(bytes ? bytes : 1, 1);
  map->data->len = 0;
}

void new_pixmap(Map *map, unsigned int w, unsigned int h, unsigned int depth) { new_map(map, w, h, depth, depth); }
void new_bytemap(Map *map, unsigned int w, unsigned int h) { new_map(map, w, h, 1, 1); }
void new_intmap(Map *map, unsigned int w, unsigned int h) { new_map(map, w, h, sizeof(unsigned), sizeof(unsigned)); }
void new_coordmap(Map *map, unsigned int w, unsigned int h)
{
  new_map(map, w, h, sizeof(Coordinates), sizeof(Coordinates));
}

void set_bytemap(Map *map, unsigned char value)
{
  std::memset(map->data->data, value, (size_t)map->width * map->height);
}

void invert_bytemap(Map *map)
{
  unsigned char *d = (unsigned char *)map->data->data;
  for (size_t i = 0; i < (size_t)map->width * map->height; i++) d[i] = (unsigned char)~d[i];
}

void interleave_mask(Map *pixmap, Map *mask)
{
  size_t n = (size_t)pixmap->width * pixmap->height;
  unsigned char *p = (unsigned char *)pixmap->data->data;
  const unsigned char *m = (const unsigned char *)mask->data->data;
  for (size_t i = 0; i < n; i++) p[i * pixmap->depth] = m[i * mask->depth];
}

// ---- engine entries ----
void resynth_set_default_mode(int mode) { g_defaultMode = mode; }

void resynth_set_device_memory_limit(size_t bytes)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  g_devLimit = bytes;
  if (bytes && g_devBytes > bytes) {
    // release the idle engine buffers, so that later calls grow under the new cap
    if (g_dev.ok) (void)hipStreamSynchronize(g_dev.stream);
    g_dev.ws.~Workspace();
    new (&g_dev.ws) Workspace();
    g_dev.pool.clear();
    g_dev.manyDesc.release();
    g_dev.manyCtl.release();
    g_dev.manyGlobal.release();
  }
}

void resynth_default_options(ResynthOptions *opt)
{
  std::memset(opt, 0, sizeof *opt);
  opt->mode = default_mode();
  // the GIMP build's GRand, or with RESYNTH_PRNG=libc the stand-alone build's rand()
  // (lib/glibProxy.c:30-49)
  const char *p = std::getenv("RESYNTH_PRNG");
  opt->prng = (p && std::strcmp(p, "libc") == 0) ? RESYNTH_PRNG_LIBC : RESYNTH_PRNG_GRAND;
}

int imageSynth(ImageBuffer *imageBuffer, ImageBuffer *mask, TImageFormat imageFormat,
               TImageSynthParameters *parameters, void (*progressCallback)(int, void *), void *contextInfo,
               int *cancelFlag)
{
  return image_synth_host(imageBuffer, mask, nullptr, imageFormat, parameters, progressCallback, contextInfo,
                          cancelFlag, nullptr, nullptr);
}

int imageSynth2(ImageBuffer *imageBuffer, ImageBuffer *mask, ImageBuffer *mask2, TImageFormat imageFormat,
                TImageSynthParameters *parameters, void (*progressCallback)(int, void *), void *contextInfo,
                int *cancelFlag)
{
  return image_synth_host(imageBuffer, mask, mask2, imageFormat, parameters, progressCallback, contextInfo,
                          cancelFlag, nullptr, nullptr);
}

int imageSynthEx(ImageBuffer *imageBuffer, ImageBuffer *mask, ImageBuffer *mask2, TImageFormat imageFormat,
                 TImageSynthParameters *parameters, void (*progressCallback)(int, void *), void *contextInfo,
                 int *cancelFlag, const ResynthOptions *options, ResynthStats *stats)
{
  return image_synth_host(imageBuffer, mask, mask2, imageFormat, parameters, progressCallback, contextInfo,
                          cancelFlag, options, stats);
}

int engine(TImageSynthParameters parameters, TFormatIndices *indices, Map *targetMap, Map *corpusMap,
           void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag)
{
  return engine_host(parameters, indices, targetMap, corpusMap, progressCallback, contextInfo, cancelFlag, nullptr,
                     nullptr);
}

int engineEx(TImageSynthParameters parameters, TFormatIndices *indices, Map *targetMap, Map *corpusMap,
             void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag,
             const ResynthOptions *options, ResynthStats *stats)
{
  return engine_host(parameters, indices, targetMap, corpusMap, progressCallback, contextInfo, cancelFlag, options,
                     stats);
}

int resynth_image_synth_device(unsigned char *d_image, unsigned int width, unsigned int height, size_t rowBytes,
                               const unsigned char *d_mask, size_t maskRowBytes, const unsigned char *d_mask2,
                               size_t mask2RowBytes, TImageFormat imageFormat,
                               const TImageSynthParameters *parameters, void (*progressCallback)(int, void *),
                               void *contextInfo, int *cancelFlag, const ResynthOptions *options,
                               ResynthStats *stats)
{
  return guarded([&]() -> int {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
    return image_synth_device_impl(g_dev.ws, d_image, width, height, rowBytes, d_mask, maskRowBytes, d_mask2,
                                   mask2RowBytes, imageFormat, parameters, progressCallback, contextInfo, cancelFlag,
                                   options, stats);
  });
}

int resynth_image_synth_device_noalloc(unsigned char *d_image, unsigned int width, unsigned int height,
                                       size_t rowBytes, const unsigned char *d_mask, size_t maskRowBytes,
                                       const unsigned char *d_mask2, size_t mask2RowBytes, TImageFormat imageFormat,
                                       const TImageSynthParameters *parameters,
                                       void (*progressCallback)(int, void *), void *contextInfo, int *cancelFlag,
                                       const ResynthOptions *options, ResynthStats *stats)
{
  if ((size_t)width * height > (size_t)RESYNTH_NOALLOC_MAX_PIXELS) return RESYNTH_ERROR_TOO_LARGE;
  if (options && options->model) return RESYNTH_ERROR_INVALID_ARGUMENT;   // the learned index is not bounded
  return guarded([&]() -> int {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
    Arena dev{scratch_address(), kScratchBytes, 0};
    Arena host{(char *)g_noallocHost, kNoAllocHostBytes, 0};
    ResynthOptions o;
    if (options) o = *options;
    else resynth_default_options(&o);
    o.timing = 0;   // timing events are created per launch
    ArenaScope scope(&dev, &host);
    Workspace W;    // every buffer of the run is carved from the arenas
    const int err = image_synth_device_impl(W, d_image, width, height, rowBytes, d_mask, maskRowBytes, d_mask2,
                                            mask2RowBytes, imageFormat, parameters, progressCallback, contextInfo,
                                            cancelFlag, &o, stats);
    if (stats) stats->scratchBytes = dev.used;
    return err;
  });
}

int resynth_image_synth_batch(ResynthRequest *requests, unsigned int n, int deviceBuffers,
                              const ResynthOptions *options, const ResynthKernelConfig *config,
                              ResynthBatchStats *stats)
{
  if (!requests && n) return -1;
  ManyConfig cfg;
  const ManyConfig *force = nullptr;
  if (config && config->wavesPerBlock) {
    cfg = {config->wavesPerBlock, config->firstGroup ? config->firstGroup : 2, config->laterGroup ? config->laterGroup : 4,
           config->lanesPerTarget ? config->lanesPerTarget : 64};
    if (!many_config_valid(cfg)) return -1;
    force = &cfg;
  }
  return guarded([&]() -> int {
    return run_batch(requests, n, deviceBuffers != 0, options, force, stats);
  });
}

int resynth_model_log_pairs(ResynthModel *model, ImageBuffer *image, ImageBuffer *mask, TImageFormat format,
                            const TImageSynthParameters *params, unsigned int maxPairs, unsigned int seed)
{
  if (!model || !image || !mask) return -1;
  ResynthOptions o;
  resynth_default_options(&o);
  o.mode = RESYNTH_MODE_BATCHED;
  o.seed = seed;
  TImageSynthParameters p;
  if (params) p = *params;
  else setDefaultParams(&p);
  const std::function<void(Job &, unsigned)> log = [&](Job &job, unsigned nT) {
    if (!maxPairs || !nT) return;
    hipStream_t s = g_dev.stream;
    Workspace &W = *job.W;
    const int tw = job.src.tw, th = job.src.th;
    const unsigned n = std::min(maxPairs, nT);
    int2 *pts = (int2 *)W.pairPts.get((size_t)n * sizeof(int2) + 64);
    int2 *src = (int2 *)W.pairSrc.get((size_t)n * sizeof(int2) + 64);
    const uint64_t key = ((uint64_t)seed << 20) ^ 0x51ED27ull;
    HIP_TRY(learn_sample_pairs(s, (const int2 *)W.targets.p, nT, (const int2 *)W.sourceOf.p, tw, (int)n, key, pts,
                               src));
    uint16_t *pq, *pk;
    model_reserve_pairs(model, n, &pq, &pk);
    HIP_TRY(learn_features(s, job.wide, W.tword.p, tw, th, p.isMakeSeamlesslyTileableHorizontally,
                           p.isMakeSeamlesslyTileableVertically, job.nColors, (const int2 *)W.offsets.p, job.st.nOffsets,
                           pts, (int)n, false, false, key + 1, pq));
    HIP_TRY(learn_features(s, job.wide, W.cword.p, job.src.cw, job.src.ch, 0, 0, job.nColors,
                           (const int2 *)W.offsets.p, job.st.nOffsets, src, (int)n, true, false, 0, pk));
    HIP_TRY(hipStreamSynchronize(s));
    model_commit_pairs(model, n);
  };
  return image_synth_host(image, mask, nullptr, format, &p, nullptr, nullptr, nullptr, &o, nullptr, &log);
}

unsigned int resynth_struct_sizes(size_t *out, unsigned int n)
{
  const size_t v[8] = {sizeof(ResynthOptions),          sizeof(ResynthStats),
                       sizeof(ResynthRequest),          sizeof(ResynthKernelConfig),
                       sizeof(ResynthBatchStats),       offsetof(ResynthRequest, params),
                       offsetof(ResynthRequest, error), offsetof(ResynthBatchStats, prepMs)};
  unsigned k = 0;
  for (; k < n && k < 8; k++) out[k] = v[k];
  return k;
}

int resynth_autotune_batch(const ResynthRequest *sample, unsigned int n, int deviceBuffers,
                           ResynthKernelConfig *best, double *msPerConfig)
{
  if (!sample || !n) return -1;
  return guarded([&]() -> int {
    return autotune_batch(sample, n, deviceBuffers != 0, best, msPerConfig);
  });
}

int resynth_score_frozen(const unsigned char *corpus, unsigned int cw, unsigned int ch,
                         const TFormatIndices *indices, const uint16_t *T, const uint32_t *Tm,
                         unsigned int nProblems, const unsigned int *nbCount, const int32_t *nbOffset,
                         const unsigned char *nbPixel, const unsigned int *candCount, unsigned int maxCands,
                         const int32_t *cands, int32_t *winner, uint32_t *bestSum, int32_t *stopped)
{
  if (!corpus || !indices || !T || !Tm || !nbCount || !nbOffset || !nbPixel || !candCount || !cands || !winner ||
      !bestSum || !stopped)
    return -1;
  for (unsigned i = 0; i < nProblems; i++)
    if (nbCount[i] > 64 || candCount[i] > maxCands) return -1;
  return guarded([&]() -> int {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
    hipStream_t s = g_dev.stream;
    const int nColors = indices->img_match_bpp, nMaps = indices->map_match_bpp;
    const bool wide = nColors + nMaps > 3;
    const size_t wb = wide ? 8 : 4;
    const int depth = indices->total_bpp;
    const size_t ncp = (size_t)cw * ch, nnb = (size_t)nProblems * 64;
    // corpus pixmap and neighbour pixels -> words on the device
    unsigned char *dPix = (unsigned char *)g_dev.ws.cpix.get(ncp * depth + nnb * 8 + 64);
    unsigned char *dNbPix = dPix + ncp * depth;
    HIP_TRY(hipMemcpyAsync(dPix, corpus, ncp * depth, hipMemcpyHostToDevice, s));
    // neighbour pixels arrive as 8-byte records; repack them to `depth` bytes each
    std::vector<unsigned char> nb(nnb * depth);
    for (size_t i = 0; i < nnb; i++) std::memcpy(&nb[i * depth], nbPixel + i * 8, depth);
    HIP_TRY(hipMemcpyAsync(dNbPix, nb.data(), nb.size(), hipMemcpyHostToDevice, s));
    void *dC = g_dev.ws.cword.get(cword_words((int)cw, (int)ch) * wb + 64);
    void *dNbW = g_dev.ws.tword.get(nnb * wb + 64);
    HIP_TRY(launch_pack_words(s, wide, dPix, (int)ncp, (int)cw, depth, indices->colorEndBip, indices->map_start_bip,
                              indices->map_end_bip, 1, dC));
    HIP_TRY(launch_pack_words(s, wide, dNbPix, (int)nnb, 1, depth, indices->colorEndBip, indices->map_start_bip,
                              indices->map_end_bip, 0, dNbW));
    Tables ht;
    std::memcpy(ht.T, T, sizeof ht.T);
    std::memcpy(ht.Tm, Tm, sizeof ht.Tm);
    Tables *dT = (Tables *)g_dev.ws.tables.get(sizeof(Tables));
    HIP_TRY(hipMemcpyAsync(dT, &ht, sizeof ht, hipMemcpyHostToDevice, s));
    size_t szCnt = nProblems * sizeof(unsigned), szOff = nnb * sizeof(int2);
    size_t szCand = (size_t)nProblems * maxCands * sizeof(int2), szOut = nProblems * 4;
    char *dArgs = (char *)g_dev.ws.compactTemp.get(2 * szCnt + szOff + szCand + 3 * szOut + 512);
    unsigned *dNbCount = (unsigned *)dArgs;
    unsigned *dCandCount = (unsigned *)(dArgs + szCnt);
    int2 *dOff = (int2 *)(dArgs + 2 * szCnt);
    int2 *dCands = (int2 *)((char *)dOff + szOff);
    int32_t *dWin = (int32_t *)((char *)dCands + szCand);
    uint32_t *dBest = (uint32_t *)((char *)dWin + szOut);
    int32_t *dStop = (int32_t *)((char *)dBest + szOut);
    HIP_TRY(hipMemcpyAsync(dNbCount, nbCount, szCnt, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dCandCount, candCount, szCnt, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dOff, nbOffset, szOff, hipMemcpyHostToDevice, s));
    if (szCand) HIP_TRY(hipMemcpyAsync(dCands, cands, szCand, hipMemcpyHostToDevice, s));
    const uint32_t penalty = 65535u * (uint32_t)nColors + Tm[0] * (uint32_t)nMaps;
    HIP_TRY(launch_score_frozen(s, wide, (int)cw, (int)ch, nColors, nMaps, penalty, dC, dT, nProblems, dNbCount, dOff,
                                dNbW, dCandCount, maxCands, dCands, dWin, dBest, dStop));
    HIP_TRY(hipMemcpyAsync(winner, dWin, szOut, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(bestSum, dBest, szOut, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(stopped, dStop, szOut, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
  });
}

}  // extern "C"

extern "C" {

int resynth_learn_fine_top16(const uint16_t *queries, unsigned int nq, const int32_t *clusters, int P,
                             const uint16_t *keys, unsigned int nKeys, const uint32_t *start,
                             unsigned int nClusters, const int32_t *keyPts, int32_t *props)
{
  if (!queries || !clusters || !keys || !start || !keyPts || !props || P < 1 || P > 2 || !nq || !nClusters)
    return -1;
  if (start[0] != 0 || start[nClusters] != nKeys) return -1;
  for (unsigned c = 0; c < nClusters; c++)
    if (start[c] > start[c + 1]) return -1;
  for (size_t i = 0; i < (size_t)nq * P; i++)
    if (clusters[i] >= (int32_t)nClusters) return -1;
  return guarded([&]() -> int {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
    hipStream_t s = g_dev.stream;
    const size_t n = nq;
    const size_t tb = sort_pairs_temp_bytes(n);
    TmpBuf dq, dcl, dk, dst, dpts, dout, dk0, dv0, dk1, dv1, dtemp;
    HIP_TRY(dq.alloc(n * kEmb * 2));
    HIP_TRY(dcl.alloc(n * P * 4));
    HIP_TRY(dk.alloc((size_t)std::max(1u, nKeys) * kEmb * 2));
    HIP_TRY(dst.alloc(((size_t)nClusters + 1) * 4));
    HIP_TRY(dpts.alloc((size_t)std::max(1u, nKeys) * 8));
    HIP_TRY(dout.alloc(n * 16 * 8));
    HIP_TRY(dk0.alloc(n * 4));
    HIP_TRY(dv0.alloc(n * 4));
    HIP_TRY(dk1.alloc(n * 4));
    HIP_TRY(dv1.alloc(n * 4));
    HIP_TRY(dtemp.alloc(tb));
    HIP_TRY(hipMemcpyAsync(dq.p, queries, n * kEmb * 2, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dcl.p, clusters, n * P * 4, hipMemcpyHostToDevice, s));
    if (nKeys) {
      HIP_TRY(hipMemcpyAsync(dk.p, keys, (size_t)nKeys * kEmb * 2, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(dpts.p, keyPts, (size_t)nKeys * 8, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipMemcpyAsync(dst.p, start, ((size_t)nClusters + 1) * 4, hipMemcpyHostToDevice, s));
    const int32_t *qcl = (const int32_t *)dcl.p;
    uint32_t *k0 = (uint32_t *)dk0.p, *v0 = (uint32_t *)dv0.p, *k1 = (uint32_t *)dk1.p, *v1 = (uint32_t *)dv1.p;
    HIP_TRY(learn_sort_keys(s, qcl, P, (int)n, k0, v0));
    HIP_TRY(launch_sort_pairs(s, k0, k1, v0, v1, n, 32, dtemp.p, tb));
    HIP_TRY(learn_fine_top16_sorted(s, (const uint16_t *)dq.p, (int)n, v1, qcl, P, (const uint16_t *)dk.p,
                                    (const uint32_t *)dst.p, (const int2 *)dpts.p, (int2 *)dout.p));
    HIP_TRY(hipMemcpyAsync(props, dout.p, n * 16 * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
  });
}

int resynth_sorted_offsets(int w, int h, int32_t *out)
{
  unsigned long long n = 0;
  return resynth_sorted_offsets_within(w, h, ~0ull, out, &n);
}

int resynth_sorted_offsets_within(int w, int h, unsigned long long maxD2, int32_t *out, unsigned long long *count)
{
  if (w < 1 || h < 1 || !out || !count || (size_t)w * (size_t)h > (1u << 28)) return -1;
  return guarded([&]() -> int {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
    hipStream_t s = g_dev.stream;
    const OffsetPlan plan = plan_offsets(w, h, maxD2);
    const size_t tb = offsets_temp_bytes(plan);
    TmpBuf dout, dtemp;
    HIP_TRY(dout.alloc(plan.nO * sizeof(int2)));
    HIP_TRY(dtemp.alloc(tb));
    HIP_TRY(launch_sorted_offsets(s, plan, (int2 *)dout.p, dtemp.p, tb));
    HIP_TRY(hipMemcpyAsync(out, dout.p, plan.nO * sizeof(int2), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *count = plan.nO;
    return 0;
  });
}

int resynth_learn_route(const uint16_t *queries, unsigned int nq, const int32_t *supers,
                        const uint16_t *subCentroids, unsigned int S, unsigned int Q, int P, int32_t *clusters)
{
  if (!queries || !supers || !subCentroids || !clusters || !nq || !S || !Q || P < 1 || P > 2 || (unsigned)P > Q)
    return -1;
  for (unsigned i = 0; i < nq; i++)
    if (supers[i] >= (int32_t)S) return -1;
  return guarded([&]() -> int {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
    hipStream_t s = g_dev.stream;
    const size_t n = nq;
    const size_t tb = sort_pairs_temp_bytes(n);
    TmpBuf dq, dsup, dsub, dout, dk0, dv0, dk1, dv1, dtemp;
    HIP_TRY(dq.alloc(n * kEmb * 2));
    HIP_TRY(dsup.alloc(n * 4));
    HIP_TRY(dsub.alloc((size_t)S * Q * kEmb * 2));
    HIP_TRY(dout.alloc(n * P * 4));
    HIP_TRY(dk0.alloc(n * 4));
    HIP_TRY(dv0.alloc(n * 4));
    HIP_TRY(dk1.alloc(n * 4));
    HIP_TRY(dv1.alloc(n * 4));
    HIP_TRY(dtemp.alloc(tb));
    HIP_TRY(hipMemcpyAsync(dq.p, queries, n * kEmb * 2, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dsup.p, supers, n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dsub.p, subCentroids, (size_t)S * Q * kEmb * 2, hipMemcpyHostToDevice, s));
    const int32_t *qs = (const int32_t *)dsup.p;
    uint32_t *k0 = (uint32_t *)dk0.p, *v0 = (uint32_t *)dv0.p, *k1 = (uint32_t *)dk1.p, *v1 = (uint32_t *)dv1.p;
    HIP_TRY(learn_sort_keys(s, qs, 1, (int)n, k0, v0));
    HIP_TRY(launch_sort_pairs(s, k0, k1, v0, v1, n, bits_for(S), dtemp.p, tb));
    HIP_TRY(learn_query_sub(s, (const uint16_t *)dq.p, (int)n, v1, qs, (int)Q, (const uint16_t *)dsub.p, P,
                            (int32_t *)dout.p));
    HIP_TRY(hipMemcpyAsync(clusters, dout.p, n * P * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
  });
}

int resynth_device_info(char *name, size_t nameLen, int *computeUnits)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, g_dev.device) != hipSuccess) return -1;
  if (name && nameLen) {
    std::snprintf(name, nameLen, "%s", p.gcnArchName);
  }
  if (computeUnits) *computeUnits = p.multiProcessorCount;
  return std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 0 : 1;
}

const char *resynth_version(void) { return "resynth_mi355x 0.1 (gfx950)"; }

// ---- host-side helpers (no device work) ----
void resynth_prng_sample(int prng, unsigned int seed, unsigned int n, uint32_t *out)
{
  Prng p(prng, seed);
  for (unsigned i = 0; i < n; i++) out[i] = p.raw();
}

void resynth_prng_ranges(int prng, unsigned int seed, unsigned int n, const uint32_t *dists, uint32_t *out)
{
  Prng p(prng, seed);
  for (unsigned i = 0; i < n; i++) out[i] = p.range(dists[i]);
}

void resynth_metric_tables(double sensitivity, double mapWeight, uint16_t *T, uint32_t *Tm)
{
  Tables t;
  quantize_tables(sensitivity, mapWeight, t);
  std::memcpy(T, t.T, sizeof t.T);
  std::memcpy(Tm, t.Tm, sizeof t.Tm);
}

int resynth_order_targets(int matchContextType, int prng, int32_t *points, unsigned int n)
{
  if (matchContextType < 0 || matchContextType > 8) return IMAGE_SYNTH_ERROR_MATCH_CONTEXT_TYPE_RANGE;
  if (!points && n) return RESYNTH_ERROR_INVALID_ARGUMENT;
  return guarded([&]() -> int {
    Prng p(prng, kEngineSeed);
    std::vector<uint32_t> perm(n);
    std::vector<int32_t> rm(points, points + 2 * (size_t)n);
    if (matchContextType <= 1) {
      // the host shuffle the engine runs for types 0 and 1
      p.shuffle(perm.data(), n);
      for (unsigned i = 0; i < n; i++) {
        points[2 * i] = rm[2 * (size_t)perm[i]];
        points[2 * i + 1] = rm[2 * (size_t)perm[i] + 1];
      }
      return 0;
    }
    // types 2..8: the engine's device sort, the host's banded shuffle, the device composition
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!ensure_device()) return RESYNTH_ERROR_DEVICE;
    hipStream_t s = g_dev.stream;
    uint32_t hRays[9];
    axis_rays(hRays);
    const size_t otb = order_temp_bytes(n);
    TmpBuf drm, dtemp, dperm, dout, drays;
    HIP_TRY(drm.alloc((size_t)n * 8));
    HIP_TRY(dtemp.alloc(otb));
    HIP_TRY(dperm.alloc((size_t)n * 4));
    HIP_TRY(dout.alloc((size_t)n * 8));
    HIP_TRY(drays.alloc(sizeof hRays));
    HIP_TRY(hipMemcpyAsync(drm.p, rm.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(drays.p, hRays, sizeof hRays, hipMemcpyHostToDevice, s));
    const uint32_t *sorted = nullptr;
    HIP_TRY(launch_order_sort(s, (const int2 *)drm.p, n, matchContextType, (const uint32_t *)drays.p, dtemp.p, otb,
                              &sorted));
    band_permutation(perm.data(), n, p);
    HIP_TRY(hipMemcpyAsync(dperm.p, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_order_compose(s, (const int2 *)drm.p, sorted, (const uint32_t *)dperm.p, (int2 *)dout.p, n));
    HIP_TRY(hipMemcpyAsync(points, dout.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
  });
}

}  // extern "C"
