// gfx950 kernels of the synthesis engine: image adaptation, point compaction, sorted
// offsets, the per-target visit (neighbour gather + candidate scoring) in its serial
// and batched forms, write-back and statistics.
//
// Every wavefront (64 lanes) owns one target visit at a time.  The neighbour patch
// (<= 64 entries) is gathered by a ballot/prefix-count scan over the sorted offsets and
// staged in LDS; candidates are scored one per lane against it, each lane gathering
// one corpus pixel word per neighbour, with the metric tables in LDS.  The sequential
// strict-less update of lib/synthesize.h:266-400 is recovered exactly from the per-lane
// (final sum, max prefix sum) pairs by a ballot loop over the lanes in candidate order.
#include "resynth_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace rs {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename Word>
__device__ __forceinline__ uint32_t chan(Word w, int c)
{
  return (uint32_t)(w >> (8 * (1 + c))) & 0xFFu;
}

// Word index of corpus pixel (x, y) in the tiled corpus layout (cword_words): rows of four
// pixel rows, each pixel's four words (one per row) side by side, so that the 32 words of
// a 128-byte line hold an 8 x 4 tile.  One 24-bit multiply-add and two bit operations.
__device__ __forceinline__ unsigned cidx(int cw, unsigned x, unsigned y)
{
  const unsigned rowOf4 = (((unsigned)cw + kCTileW - 1) & ~(unsigned)(kCTileW - 1)) * kCTileH;
  return __umul24(y >> 2, rowOf4) + ((x << 2) | (y & 3u));
}

// Score contribution of one in-corpus neighbour (lib/synthesize.h:328-355): colours
// except for neighbour 0 (the target itself), maps for every neighbour.
template <typename Word>
__device__ __forceinline__ uint32_t neighbour_term(Word t, Word q, bool first, int nColors, int nMaps,
                                                   const uint16_t *T, const uint32_t *Tm)
{
  uint32_t s = 0;
  if (!first) {
    for (int c = 0; c < nColors; c++) s += T[256u + chan(t, c) - chan(q, c)];
  }
  for (int m = 0; m < nMaps; m++) s += Tm[256u + chan(t, nColors + m) - chan(q, nColors + m)];
  return s;
}

// Accumulates neighbours [base, base+G) of candidate (cx,cy): the G corpus words are
// loaded first, then summed in neighbour order with the early out below.  FIRST is false
// for groups that start past neighbour 0, which then need no test for it.
template <typename Word, int G, bool FIRST>
__device__ __forceinline__ void score_group(const Geometry &g, const Word *__restrict__ cword, int cx, int cy,
                                            unsigned base, unsigned count, const int2 *nbOff, const Word *nbPix,
                                            const uint16_t *T, const uint32_t *Tm, uint32_t B0, uint32_t &acc,
                                            uint32_t &mx, bool &dead)
{
  Word q[G];
  bool in[G];
#pragma unroll
  for (int u = 0; u < G; u++) {
    const unsigned i = base + u;
    const bool live = i < count;
    const int2 o = nbOff[live ? i : 0u];
    const int qx = cx + o.x, qy = cy + o.y;
    in[u] = live && (unsigned)qx < (unsigned)g.cw && (unsigned)qy < (unsigned)g.ch;
    // unconditional load, word 0 standing in for an outside neighbour (see score_group_rgb)
    const unsigned idx = in[u] ? cidx(g.cw, (unsigned)qx, (unsigned)qy) : 0u;
    q[u] = cword[idx];
  }
#pragma unroll
  for (int u = 0; u < G; u++) {
    unsigned i = base + u;
    if (i < count && !dead) {
      uint32_t term = (in[u] && (q[u] & 0xFEu) == 0xFEu)
                          ? neighbour_term<Word>(nbPix[i], q[u], FIRST && i == 0, g.nColors, g.nMaps, T, Tm)
                          : g.penalty;
      acc += term;
      mx = acc > mx ? acc : mx;
      if (acc >= B0) dead = true;
    }
  }
}

// The same accumulation without branches, for NC colour channels and no map channels:
// every term is computed (table indices stay in [1, 511] whatever the loaded word holds)
// and selected, so the wave does not diverge per neighbour.  The sums are identical.
template <typename Word, int G, int NC, bool FIRST>
__device__ __forceinline__ void score_group_rgb(const Geometry &g, const Word *__restrict__ cword, int cx, int cy,
                                                unsigned base, unsigned count, const int2 *nbOff, const Word *nbPix,
                                                const uint16_t *T, uint32_t B0, uint32_t &acc, uint32_t &mx,
                                                bool &dead)
{
  Word q[G];
  bool in[G];
#pragma unroll
  for (int u = 0; u < G; u++) {
    const unsigned i = base + u;
    const bool live = i < count;
    const int2 o = nbOff[live ? i : 0u];
    const int qx = cx + o.x, qy = cy + o.y;
    in[u] = live && (unsigned)qx < (unsigned)g.cw && (unsigned)qy < (unsigned)g.ch;
    // Every lane loads (word 0 stands in for an outside neighbour, whose term is the
    // penalty), so the loads need no branch around them; the index is 32-bit (corpus
    // sides below 2^24 pixels, fewer than 2^31 words).
    const unsigned idx = in[u] ? cidx(g.cw, (unsigned)qx, (unsigned)qy) : 0u;
    q[u] = cword[idx];
  }
#pragma unroll
  for (int u = 0; u < G; u++) {
    const unsigned i = base + u;
    const Word t = nbPix[i < count ? i : 0u];
    uint32_t cs = 0;
#pragma unroll
    for (int c = 0; c < NC; c++) cs += T[256u + chan(t, c) - chan(q[u], c)];
    if constexpr (FIRST) cs = (i == 0) ? 0u : cs;
    const bool valid = in[u] && (q[u] & 0xFEu) == 0xFEu;
    const uint32_t term = valid ? cs : g.penalty;
    const bool act = i < count && !dead;
    acc = act ? acc + term : acc;
    mx = acc > mx ? acc : mx;
    dead = dead || (act && acc >= B0);
  }
}

// Scores candidate (cx,cy) against the staged patch.  The lane stops (dead) at the first
// prefix sum >= B0, where B0 is the best sum before this batch of candidates: such a
// candidate cannot be accepted by the sequential update, whose best only decreases.
// For survivors F is the full sum and M the largest prefix sum (equal to F unless the
// u32 sum wrapped); the sequential rule accepts a survivor iff M < best-at-its-turn.
// skip0: the candidate is known to be a selected in-bounds corpus pixel and there are no
// map channels, so neighbour 0 (the candidate itself, colours not compared) adds 0; with
// B0 > 0 its early-out test cannot fire either.  G0 is the size of the first group of
// loads, which is where most random candidates die.
// NC > 0: the branch-free path for NC colours and no maps (the caller guarantees it).
template <typename Word, int G0 = 8, int GL = 8, int NC = 0>
__device__ __forceinline__ void score_candidate(const Geometry &g, const Word *__restrict__ cword, int cx, int cy,
                                                unsigned count, const int2 *nbOff, const Word *nbPix,
                                                const uint16_t *T, const uint32_t *Tm, uint32_t B0, uint32_t &F,
                                                uint32_t &M, bool &alive, bool skip0 = false)
{
  uint32_t acc = 0, mx = 0;
  bool dead = false;
  unsigned base = (skip0 && B0 > 0) ? 1u : 0u;
  if constexpr (NC > 0) {
    score_group_rgb<Word, G0, NC, true>(g, cword, cx, cy, base, count, nbOff, nbPix, T, B0, acc, mx, dead);
    for (base += G0; base < count && !dead; base += GL)
      score_group_rgb<Word, GL, NC, false>(g, cword, cx, cy, base, count, nbOff, nbPix, T, B0, acc, mx, dead);
  } else {
    score_group<Word, G0, true>(g, cword, cx, cy, base, count, nbOff, nbPix, T, Tm, B0, acc, mx, dead);
    for (base += G0; base < count && !dead; base += GL)
      score_group<Word, GL, false>(g, cword, cx, cy, base, count, nbOff, nbPix, T, Tm, B0, acc, mx, dead);
  }
  F = acc;
  M = mx;
  alive = !dead;
}

// Sequential strict-less resolution over the lanes of one wave, in lane order.
// On return best/win are updated; stop is the lane of the first perfect match or 64.
__device__ __forceinline__ void resolve(bool cand, uint32_t M, uint32_t F, uint32_t &best, int &win, int &stop)
{
  uint64_t rem = __ballot(cand);
  stop = 64;
  while (rem) {
    uint64_t m = __ballot(cand && M < best) & rem;
    if (!m) break;
    int i = __ffsll((unsigned long long)m) - 1;
    best = __shfl(F, i);
    win = i;
    rem &= (i == 63) ? 0ull : ~((2ull << i) - 1ull);
    if (best == 0) {
      stop = i;
      break;
    }
  }
}

// First sorted-offset index worth scanning for target (tx,ty): with R empty Chebyshev
// rings of tiles around the target's tile, every pixel holding a value is at least 8R+1
// pixels away, so the offsets with x^2+y^2 < (8R+1)^2 cannot hold one.  `lanes` threads
// (a wave or a workgroup) share the ring walk; `anyTile` reduces over them.
template <typename AnyFn>
__device__ __forceinline__ int empty_rings(const Geometry &g, const uint8_t *tileHas, int tx, int ty, unsigned tid,
                                           unsigned lanes, AnyFn anyTile)
{
  const int TX = tx / kTile, TY = ty / kTile;
  const int maxR = max(max(TX, g.tilesW - 1 - TX), max(TY, g.tilesH - 1 - TY));
  int R = -1;   // number of empty rings found
  for (int r = 0; r <= maxR; r++) {
    const unsigned n = r == 0 ? 1u : 8u * (unsigned)r;
    bool hit = false;
    for (unsigned i0 = 0; i0 < n; i0 += lanes) {
      bool mine = false;
      unsigned i = i0 + tid;
      if (i < n) {
        int X, Y;
        if (r == 0) { X = TX; Y = TY; }
        else {
          // walk the ring's perimeter: top row, right column, bottom row, left column
          unsigned side = i / (2u * r), o = i % (2u * r);
          if (side == 0) { X = TX - r + (int)o; Y = TY - r; }
          else if (side == 1) { X = TX + r; Y = TY - r + (int)o; }
          else if (side == 2) { X = TX + r - (int)o; Y = TY + r; }
          else { X = TX - r; Y = TY + r - (int)o; }
        }
        if (X >= 0 && Y >= 0 && X < g.tilesW && Y < g.tilesH) mine = tileHas[(size_t)Y * g.tilesW + X] != 0;
      }
      if (anyTile(mine)) { hit = true; break; }
    }
    if (hit) break;
    R = r;
  }
  return R;
}

// Index of the first sorted offset with x^2 + y^2 >= rho^2 (at least 1: offset 0 is the
// target itself), g.nO past the table.
__device__ __forceinline__ unsigned offset_index(const Geometry &g, const uint32_t *firstIdx, unsigned rho)
{
  if (rho > g.rhoMax) return g.nO;
  const unsigned j = firstIdx[rho];
  return j < 1u ? 1u : j;
}

// Index of the first sorted offset worth scanning (see empty_rings).
template <typename AnyFn>
__device__ __forceinline__ unsigned scan_start(const Geometry &g, const uint8_t *tileHas, const uint32_t *firstIdx,
                                               int tx, int ty, unsigned tid, unsigned lanes, AnyFn anyTile)
{
  const int R = empty_rings(g, tileHas, tx, ty, tid, lanes, anyTile);
  return R < 0 ? 1u : offset_index(g, firstIdx, (unsigned)(kTile * R + 1));
}

// Context radius of every tile of a static occupancy map (the context-only snapshot): the
// largest rho such that no pixel of a marked tile lies nearer than rho to any pixel of the
// tile -- floor of the least Euclidean distance between the tile's box and a marked tile's
// box (boxes of tiles t apart in a row or column are (t - 1) * kTile + 1 pixels apart); 0
// for a marked tile, and past every offset when nothing is marked.  Rows first:
// h = horizontal distance in tiles to the nearest marked tile of the row.
__global__ void k_ring_rows(const uint8_t *mark, int W, int H, int32_t *h)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  constexpr int kFar = 1 << 30;
  const int x = i % W;
  const uint8_t *row = mark + (size_t)(i / W) * W;
  int d = kFar;
  for (int dist = 0; x - dist >= 0 || x + dist < W; dist++)
    if ((x - dist >= 0 && row[x - dist]) || (x + dist < W && row[x + dist])) {
      d = dist;
      break;
    }
  h[i] = d;
}

__device__ __forceinline__ uint64_t box_gap(int tiles)
{
  return tiles <= 0 ? 0u : (uint64_t)(tiles - 1) * kTile + 1u;
}

// Then columns: d2(x, y) = min over rows y' of gap(h(x, y'))^2 + gap(|y - y'|)^2, searched
// outward from y and stopped once the row gap alone reaches the best found.
__global__ void k_ctx_radius(const int32_t *h, int W, int H, int32_t *radius)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  constexpr int kFar = 1 << 30;
  const int x = i % W, y = i / W;
  uint64_t best = ~0ull;
  for (int dist = 0; y - dist >= 0 || y + dist < H; dist++) {
    const uint64_t gy = box_gap(dist);
    if (gy * gy >= best) break;
    for (int side = 0; side < (dist ? 2 : 1); side++) {
      const int yy = side ? y + dist : y - dist;
      if (yy < 0 || yy >= H) continue;
      const int hd = h[(size_t)yy * W + x];
      if (hd >= kFar) continue;
      const uint64_t gx = box_gap(hd), d2 = gx * gx + gy * gy;
      if (d2 < best) best = d2;
    }
  }
  if (best == ~0ull) {
    radius[i] = 1 << 30;
    return;
  }
  uint64_t r = (uint64_t)sqrt((double)best);
  while (r * r > best) r--;
  while ((r + 1) * (r + 1) <= best) r++;
  radius[i] = (int32_t)(r < (1u << 30) ? r : (1u << 30));
}

hipError_t launch_ctx_radius(hipStream_t s, const uint8_t *mark, int W, int H, int32_t *scratch, int32_t *radius)
{
  hipLaunchKernelGGL(k_ring_rows, dim3((W * H + 255) / 256), dim3(256), 0, s, mark, W, H, scratch);
  hipLaunchKernelGGL(k_ctx_radius, dim3((W * H + 255) / 256), dim3(256), 0, s, scratch, W, H, radius);
  return hipGetLastError();
}

// How far any target's neighbour scan can reach, from the context at the start of a run.
// A tile is full when all kTile^2 of its pixels lie in the image and hold a value (context
// pixels: they keep their values for the whole run).  For a pixel p of a tile whose box lies
// rho (k_ctx_radius over the full tiles) from the nearest full tile's box, every pixel of
// that full tile lies within rho + 1 + 2 (kTile - 1) sqrt 2 < rho + 21 of p: it holds
// kTile^2 = 64 valued pixels, at least the patch's L - 1 <= 63 neighbours, so p's scan
// ends before any offset farther than rho + 21.  reach = the largest rho over the tiles
// that hold a target (0 when none does; 2^30 when no tile is full).
template <typename Word>
__global__ void k_tile_context(const Word *tword, const uint8_t *tflag, int tw, int th, int tilesW, int tilesH,
                               uint8_t *full, uint8_t *hasTarget)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tilesW * tilesH) return;
  const int x0 = (t % tilesW) * kTile, y0 = (t / tilesW) * kTile;
  bool all = x0 + kTile <= tw && y0 + kTile <= th, any = false;
  if (sizeof(Word) == 4 && all && (tw & 3) == 0) {
    // a whole tile with 16-byte aligned rows: each row's 8 words and 8 flags in four loads
    for (int dy = 0; dy < kTile; dy++) {
      const size_t p = (size_t)(y0 + dy) * (unsigned)tw + (unsigned)x0;
      const uint4 a = *(const uint4 *)(tword + p), b = *(const uint4 *)(tword + p + 4);
      all = all && (a.x & a.y & a.z & a.w & b.x & b.y & b.z & b.w & 1u);
      any = any || (*(const uint32_t *)(tflag + p) | *(const uint32_t *)(tflag + p + 4));
    }
  } else {
    for (int dy = 0; dy < kTile && y0 + dy < th; dy++)
      for (int dx = 0; dx < kTile && x0 + dx < tw; dx++) {
        const size_t p = (size_t)(y0 + dy) * (unsigned)tw + (unsigned)(x0 + dx);
        all = all && (tword[p] & 1u);
        any = any || tflag[p];
      }
  }
  full[t] = all ? 1 : 0;
  hasTarget[t] = any ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_reach_max(const int32_t *rho, const uint8_t *hasTarget, unsigned n,
                                                   uint32_t *reach)
{
  __shared__ uint32_t part[4];
  uint32_t m = 0;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (hasTarget[i]) m = max(m, (uint32_t)rho[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if (lane_id() == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = max(max(part[0], part[1]), max(part[2], part[3]));
    if (m) atomicMax(reach, m);
  }
}

hipError_t launch_reach_radius(hipStream_t s, bool wide, const void *tword, const uint8_t *tflag, int tw, int th,
                               uint8_t *full, uint8_t *hasTarget, int32_t *scratch, int32_t *rhoFull, uint32_t *reach)
{
  const int tilesW = (tw + kTile - 1) / kTile, tilesH = (th + kTile - 1) / kTile;
  const unsigned nTiles = (unsigned)(tilesW * tilesH), blocks = (nTiles + 255) / 256;
  hipError_t e = launch_fill_u32(s, reach, 0u, 1);
  if (e != hipSuccess || !nTiles) return e;
  if (wide)
    hipLaunchKernelGGL((k_tile_context<uint64_t>), dim3(blocks), dim3(256), 0, s, (const uint64_t *)tword, tflag, tw,
                       th, tilesW, tilesH, full, hasTarget);
  else
    hipLaunchKernelGGL((k_tile_context<uint32_t>), dim3(blocks), dim3(256), 0, s, (const uint32_t *)tword, tflag, tw,
                       th, tilesW, tilesH, full, hasTarget);
  e = launch_ctx_radius(s, full, tilesW, tilesH, scratch, rhoFull);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reach_max, dim3(std::min(blocks, 1024u)), dim3(256), 0, s, rhoFull, hasTarget, nTiles, reach);
  return hipGetLastError();
}

// Steps of the workgroup offset scan taken per round once the first step has not filled
// the patch.
constexpr unsigned kScanBatch = 4;
// Pass-0 workgroup scans while at most kNearTargets targets are visited: the near part of
// the patch comes from those targets directly (see gather_neighbours_wg).
constexpr unsigned kNearTargets = 4096, kNearMax = 64;

// Neighbour patch of target (tx,ty) (lib/synthesize.h:189-241): neighbour 0 is the
// target itself; then the sorted offsets are scanned 64 at a time, keeping points in
// the target image (or wrapped into it when tiling) that have a value.
template <typename Word>
__device__ unsigned gather_neighbours(const Geometry &g, const Word *tword, const int2 *sourceOf, const int2 *offsets,
                                      int tx, int ty, int2 *nbOff, Word *nbPix, int2 *nbSrc,
                                      const uint8_t *tileHas = nullptr, const uint32_t *firstIdx = nullptr,
                                      unsigned *scanned = nullptr)
{
  const unsigned lane = lane_id();
  const size_t tp = (size_t)ty * (unsigned)g.tw + (unsigned)tx;
  if (lane == 0) {
    nbOff[0] = make_int2(0, 0);
    nbPix[0] = tword[tp];
    nbSrc[0] = sourceOf[tp];
  }
  unsigned count = 1;
  const unsigned L = g.patchLimit;
  unsigned start = 1;
  if (g.skipEmpty && tileHas)
    start = scan_start(g, tileHas, firstIdx, tx, ty, lane, 64u, [](bool v) { return __ballot(v) != 0ull; });
  for (unsigned base = start; base < g.nO && count < L; base += 64) {
    // loads without branches: lanes past the table or outside the image read entry 0 / the
    // target's own word and do not count
    const unsigned j = base + lane;
    const int2 off = offsets[j < g.nO ? j : 0u];
    int px = tx + off.x, py = ty + off.y;
    bool in = j < g.nO;
    if (px < 0) { if (g.tileH) px += g.tw; else in = false; }
    else if (px >= g.tw) { if (g.tileH) px -= g.tw; else in = false; }
    if (py < 0) { if (g.tileV) py += g.th; else in = false; }
    else if (py >= g.th) { if (g.tileV) py -= g.th; else in = false; }
    const size_t p = in ? (size_t)(__umul24((unsigned)py, (unsigned)g.tw) + (unsigned)px) : tp;
    const Word w = tword[p];
    const bool v = in && (w & 1u) != 0;
    uint64_t m = __ballot(v);
    unsigned rank = __popcll(m & lanemask_lt());
    if (v && count + rank < L) {
      unsigned idx = count + rank;
      nbOff[idx] = off;
      nbPix[idx] = w;
      nbSrc[idx] = sourceOf[p];
    }
    unsigned c2 = count + (unsigned)__popcll(m);
    count = c2 < L ? c2 : L;
    if (scanned) *scanned += 64u;
  }
  wave_sync();
  return count;
}

static bool rgb_fast_path();

// Heuristic 1 (lib/synthesize.h:537-580): lane i proposes source(neighbour i) - offset i.
// Candidates clipped/masked in the corpus, already probed by this target index
// (heuristic 2) or repeating an earlier lane's candidate are skipped; the rest are scored
// in lane order.  Returns the lanes whose candidate must be recorded in the prober map.
template <typename Word, int NC = 0>
__device__ __forceinline__ void coherence_phase(const Geometry &g, const Word *__restrict__ cword,
                                                const uint32_t *prober, unsigned k, unsigned count,
                                                const int2 *nbOff, const Word *nbPix, const int2 *nbSrc,
                                                const uint16_t *T, const uint32_t *Tm, uint32_t &best, int &winX,
                                                int &winY, bool &perfect, bool &scored, bool &record, int &cx,
                                                int &cy, unsigned &evals)
{
  const unsigned lane = lane_id();
  bool valid = false;
  uint32_t lastProber = k;
  cx = 0;
  cy = 0;
  {
    // the candidate's word and its recent-prober entry, loaded together and without a
    // branch (lanes without a candidate read entry 0 and stay invalid)
    const int2 src = nbSrc[lane < count ? lane : 0u];
    const int2 o = nbOff[lane < count ? lane : 0u];
    cx = src.x - o.x;
    cy = src.y - o.y;
    const bool in = lane < count && src.x != -1 && (unsigned)cx < (unsigned)g.cw && (unsigned)cy < (unsigned)g.ch;
    const unsigned ci = in ? cidx(g.cw, (unsigned)cx, (unsigned)cy) : 0u;
    const Word q = cword[ci];
    const uint32_t pr = prober[ci];
    valid = in && (q & 0xFEu) == 0xFEu;
    lastProber = in ? pr : k;
    if (!in) {
      cx = 0;
      cy = 0;
    }
  }
  // earlier lane with the same candidate
  bool dup = false;
  uint64_t vmask = __ballot(valid);
  if (__popcll(vmask) > 1) {
    for (unsigned j = 0; j < count; j++) {
      int ox = __shfl(cx, (int)j), oy = __shfl(cy, (int)j);
      if (j < lane && ((vmask >> j) & 1ull) && ox == cx && oy == cy) dup = true;
    }
  }
  const bool active = valid && !dup && lastProber != k;
  uint32_t F = 0, M = 0;
  bool alive = false;
  if (active)
    score_candidate<Word, 8, 8, NC>(g, cword, cx, cy, count, nbOff, nbPix, T, Tm, 0xFFFFFFFFu, F, M, alive,
                                    g.nMaps == 0);
  uint64_t amask = __ballot(active);
  evals += (unsigned)__popcll(amask);
  int win = -1, stop = 64;
  uint32_t b = best;
  resolve(active && alive, M, F, b, win, stop);
  if (win >= 0) {
    best = b;
    winX = __shfl(cx, win);
    winY = __shfl(cy, win);
    scored = true;
  }
  perfect = stop < 64;
  record = active && (int)lane < stop;
  // The evaluation count above includes lanes after a perfect match, which the
  // reference never reaches; correct it.
  if (perfect) evals -= (unsigned)__popcll(amask & ~((stop == 63) ? ~0ull : ((2ull << stop) - 1ull)));
}

// ------------------------------------------------------------------ workgroup visits

// Band of corpus rows that this workgroup's random probes come from: workgroups are
// dispatched to the XCDs in turn, so band x is probed from XCD x only and stays in that
// XCD's L2.  Returns the row range [r0, r1).
__device__ __forceinline__ void probe_band(const Geometry &g, int &r0, int &r1)
{
  if (g.bands) {
    const int x = (int)(blockIdx.x % (unsigned)kBands);
    r0 = x * g.ch / kBands;
    r1 = (x + 1) * g.ch / kBands;
  } else {
    r0 = 0;
    r1 = g.ch;
  }
}

// Index into the corpus point list of a random probe drawn from hash h (batched modes),
// restricted to the workgroup's band when banding is on and the band holds points.
__device__ __forceinline__ unsigned probe_index(const Geometry &g, const SynthBuffers &b, uint64_t h)
{
  unsigned lo = 0, n = g.nC;
  if (g.bands) {
    const unsigned x = blockIdx.x % (unsigned)kBands;
    const unsigned a = b.bandStart[x], e = b.bandStart[x + 1];
    if (e > a) {
      lo = a;
      n = e - a;
    }
  }
  return lo + (unsigned)(((h >> 32) * (uint64_t)n) >> 32);
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Maps one raw PRNG output to a corpus index as the engine's int_range(0, nC) does:
// GLib g_rand_int_range rejects outputs above maxvalue; the stand-alone build's
// lo + rand()/(RAND_MAX/n + 1) (lib/glibProxy.c:37-49) accepts every output.
__device__ __forceinline__ bool draw_index(int prngKind, uint32_t raw, uint32_t nC, uint32_t maxvalue, uint32_t &idx)
{
  if (prngKind == 1) {
    idx = raw / (2147483647u / nC + 1u);
    return true;
  }
  if (raw > maxvalue) return false;
  idx = raw % nC;
  return true;
}

// Shared state of one target visit by a workgroup of WAVES wavefronts.
template <typename Word, int WAVES>
struct VisitLDS {
  uint16_t T[512];
  uint32_t Tm[512];
  int2 off[64];
  Word pix[64];
  int2 src[64];
  uint32_t waveCnt[kScanBatch * WAVES];
  // visited targets nearer than any context pixel (pass 0, early chunks)
  uint32_t nearCnt;
  uint32_t nearD2[kNearMax], nearIdx[kNearMax];
  int2 nearPt[kNearMax];
  uint32_t candM[64 * WAVES], candF[64 * WAVES];
  int2 cand[64 * WAVES];
  uint8_t candOk[64 * WAVES];
  uint8_t candUsed[64 * WAVES];
  // visit result
  uint32_t best;
  int winX, winY;
  int perfect, scored, aborted;
  unsigned evals;
  uint32_t streamPos;
  uint32_t lastUsed;
  const uint8_t *tileHas;
  const int32_t *ctxRadius;
  const uint32_t *firstIdx;
};

// Neighbour patch of target (tx,ty) (lib/synthesize.h:189-241) gathered by the whole
// workgroup: neighbour 0 is the target itself; then the sorted offsets are scanned
// 64*WAVES at a time, keeping in-image (or wrapped, when tiling) points that have a value.
// A scan that the first step does not finish (pass-0 targets deep in a hole scan tens of
// thousands of offsets) goes on kScanBatch steps per round: their loads are in flight
// together, and the hits are placed in offset order by per-(step, wave) counts, so the
// patch is the same.
//
// Early in pass 0 (nVisited = targets[0, nVisited) visited, at most kNearTargets), a
// target deep in a hole would scan everything between a few scattered visited pixels and
// the hole's border.  Instead the context-only tile map (a snapshot taken before any
// visit) gives a radius rho inside which no context pixel lies (k_ctx_radius); inside it
// the only valued pixels are visited targets, which are found directly and ordered like
// the sorted offsets (x^2 + y^2 ascending, row-major offset index descending).  When at
// least L - 1 of them lie inside rho, the nearest L - 1 are the patch and nothing is
// scanned; else the scan resumes at rho.  The patch is the same as the plain scan's
// (targets of the current chunk that land early are the only difference a batched visit
// may see).
template <typename Word, int WAVES>
__device__ unsigned gather_neighbours_wg(const Geometry &g, const Word *tword, const int2 *sourceOf,
                                         const int2 *offsets, int tx, int ty, VisitLDS<Word, WAVES> &S,
                                         const int2 *targets, unsigned nVisited)
{
  constexpr unsigned LANES = 64u * WAVES;
  const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  const size_t tp = (size_t)ty * (unsigned)g.tw + (unsigned)tx;
  if (tid == 0) {
    S.off[0] = make_int2(0, 0);
    S.pix[0] = tword[tp];
    S.src[0] = sourceOf[tp];
    S.nearCnt = 0;
  }
  unsigned count = 1;
  const unsigned L = g.patchLimit;
  unsigned start = 1;
  auto anyTile = [](bool v) { return __syncthreads_or(v) != 0; };
  bool near = false;
  if (g.skipEmpty && S.ctxRadius && nVisited <= kNearTargets) {
    const unsigned rho = (unsigned)S.ctxRadius[(size_t)(ty / kTile) * g.tilesW + tx / kTile];
    // visited targets with d2 < lim count as near (no context pixel lies that near)
    uint64_t lim = (uint64_t)rho * rho;
    if (lim > 0xFFFFFFFFull) lim = 0xFFFFFFFFull;
    const int ow = min(g.cw, g.tw), oh = min(g.ch, g.th);
    // squared distance of visited target j, or ~0 for the target itself and points the
    // offset table does not reach
    auto dist2 = [&](unsigned j, int2 &q) -> uint64_t {
      q = targets[j];
      const int dx = q.x - tx, dy = q.y - ty;
      const uint64_t d2 = (uint64_t)((int64_t)dx * dx) + (uint64_t)((int64_t)dy * dy);
      return (d2 > 0 && dx > -ow && dx < ow && dy > -oh && dy < oh) ? d2 : ~0ull;
    };
    // how many visited targets lie nearer than sqrt(bound) (uniform over the workgroup)
    auto count_within = [&](uint64_t bound) -> unsigned {
      unsigned c = 0;
      int2 q;
      for (unsigned j = tid; j < nVisited; j += LANES) c += dist2(j, q) < bound ? 1u : 0u;
      __syncthreads();
      if (tid == 0) S.nearCnt = 0;
      __syncthreads();
      if (c) atomicAdd(&S.nearCnt, c);
      __syncthreads();
      return S.nearCnt;
    };
    unsigned n = nVisited ? count_within(lim) : 0u;
    bool ok = true;
    if (n > kNearMax) {
      // more than the slots hold: the patch takes the nearest L - 1 of them, so look for a
      // bound that keeps between L - 1 and kNearMax (a binary search on the squared
      // distance; none exists only when many of them share one distance)
      ok = false;
      uint64_t lo = 0, hi = lim;
      while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        const unsigned c = count_within(mid);
        if (c > kNearMax) {
          hi = mid;
        } else if (c + 1 < L) {
          lo = mid;
        } else {
          lim = mid;
          n = c;
          ok = true;
          break;
        }
      }
    }
    if (ok) {
      near = true;
      __syncthreads();
      if (tid == 0) S.nearCnt = 0;
      __syncthreads();
      for (unsigned j = tid; j < nVisited; j += LANES) {
        int2 q;
        const uint64_t d2 = dist2(j, q);
        if (d2 < lim) {
          const unsigned slot = atomicAdd(&S.nearCnt, 1u);
          if (slot < kNearMax) {
            S.nearD2[slot] = (uint32_t)d2;
            S.nearIdx[slot] = (uint32_t)(q.y - ty + oh - 1) * (uint32_t)(2 * ow - 1) + (uint32_t)(q.x - tx + ow - 1);
            S.nearPt[slot] = q;
          }
        }
      }
      __syncthreads();
      // ordered like the sorted offsets; the first L - 1 complete the patch
      if (tid < n) {
        const uint32_t d2 = S.nearD2[tid], id = S.nearIdx[tid];
        unsigned rank = 0;
        for (unsigned s = 0; s < n; s++) {
          const uint32_t e2 = S.nearD2[s], ie = S.nearIdx[s];
          rank += (e2 < d2 || (e2 == d2 && ie > id)) ? 1u : 0u;
        }
        if (rank + 1 < L) {
          const int2 q = S.nearPt[tid];
          const size_t pq = (size_t)q.y * (unsigned)g.tw + (unsigned)q.x;
          S.off[1 + rank] = make_int2(q.x - tx, q.y - ty);
          S.pix[1 + rank] = tword[pq];
          S.src[1 + rank] = sourceOf[pq];
        }
      }
      count = 1 + min(n, L - 1);
      start = rho == 0 ? 1u : offset_index(g, S.firstIdx, rho);
      __syncthreads();
    }
  }
  if (g.skipEmpty && S.tileHas && !near) start = scan_start(g, S.tileHas, S.firstIdx, tx, ty, tid, LANES, anyTile);
  unsigned steps = 1;
  for (unsigned base = start; base < g.nO && count < L; base += steps * LANES, steps = kScanBatch) {
    int2 off[kScanBatch];
    Word w[kScanBatch];
    size_t p[kScanBatch];
    bool v[kScanBatch];
#pragma unroll
    for (unsigned st = 0; st < kScanBatch; st++) {
      const unsigned j = base + st * LANES + tid;
      v[st] = false;
      p[st] = tp;
      w[st] = 0;
      off[st] = make_int2(0, 0);
      if (st < steps) {
        // loads without branches, as in the wave scan
        off[st] = offsets[j < g.nO ? j : 0u];
        int px = tx + off[st].x, py = ty + off[st].y;
        bool in = j < g.nO;
        if (px < 0) { if (g.tileH) px += g.tw; else in = false; }
        else if (px >= g.tw) { if (g.tileH) px -= g.tw; else in = false; }
        if (py < 0) { if (g.tileV) py += g.th; else in = false; }
        else if (py >= g.th) { if (g.tileV) py -= g.th; else in = false; }
        p[st] = in ? (size_t)(__umul24((unsigned)py, (unsigned)g.tw) + (unsigned)px) : tp;
        w[st] = tword[p[st]];
        v[st] = in && (w[st] & 1u) != 0;
      }
    }
    uint64_t m[kScanBatch];
#pragma unroll
    for (unsigned st = 0; st < kScanBatch; st++) {
      m[st] = __ballot(v[st]);
      if (st < steps && lane == 0) S.waveCnt[st * WAVES + wave] = (uint32_t)__popcll(m[st]);
    }
    __syncthreads();
    unsigned before = 0;
#pragma unroll
    for (unsigned st = 0; st < kScanBatch; st++) {
      if (st >= steps) break;
      unsigned mine = before;
#pragma unroll
      for (int wv = 0; wv < WAVES; wv++) {
        const unsigned c = S.waveCnt[st * WAVES + wv];
        mine += (wv < (int)wave) ? c : 0;
        before += c;
      }
      const unsigned idx = count + mine + (unsigned)__popcll(m[st] & lanemask_lt());
      if (v[st] && idx < L) {
        S.off[idx] = off[st];
        S.pix[idx] = w[st];
        S.src[idx] = sourceOf[p[st]];
      }
    }
    count = (count + before < L) ? count + before : L;
    __syncthreads();
  }
  return count;
}

// Sequential strict-less resolution, by wave 0, of the candidate slots S.cand*[0..nSlots)
// in slot order.  Slots marked in candUsed were drawn; candOk marks survivors of the early
// out.  On return candUsed keeps only the slots the reference would have scored (every
// drawn slot up to and including a perfect match), S.best/winX/winY/scored/perfect are
// updated, and lastUsed is the last such slot (-1 if none).
template <typename Word, int WAVES>
__device__ void resolve_slots_wave0(VisitLDS<Word, WAVES> &S, unsigned nSlots, int &lastUsed)
{
  const unsigned lane = lane_id();
  uint32_t best = S.best;
  bool perfect = false;
  lastUsed = -1;
  for (unsigned w = 0; w * 64 < nSlots; w++) {
    unsigned slot = w * 64 + lane;
    bool act = slot < nSlots && S.candUsed[slot] && !perfect;
    bool ok = act && S.candOk[slot];
    uint32_t M = ok ? S.candM[slot] : 0u, F = ok ? S.candF[slot] : 0u;
    int win = -1, stop = 64;
    resolve(ok, M, F, best, win, stop);
    if (win >= 0 && lane == 0) {
      int2 c = S.cand[w * 64 + win];
      S.winX = c.x;
      S.winY = c.y;
      S.scored = 1;
    }
    uint64_t am = __ballot(act);
    if (stop < 64) {
      perfect = true;
      am &= (stop == 63) ? ~0ull : ((2ull << stop) - 1ull);
    }
    if (slot < nSlots && !((am >> lane) & 1ull)) S.candUsed[slot] = 0;
    if (am) lastUsed = (int)(w * 64 + 63 - __clzll((long long)am));
  }
  if (lane == 0) {
    S.best = best;
    S.perfect = perfect ? 1 : 0;
  }
}

// One target visit (lib/synthesize.h:480-640) by a workgroup.  SERIAL draws the random
// probes from the exact PRNG stream (prngKind, stream, streamLen, S.streamPos) and aborts
// without side effects when the stream runs dry; otherwise probes come from a per-target
// counter-based stream.  Global writes happen only in the commit at the end.
// nVisited: targets[0, nVisited) are known to hold values (0: unknown).
template <typename Word, int WAVES, bool SERIAL, int NC = 0>
__device__ void visit_wg(const Geometry &g, const SynthBuffers &b, VisitLDS<Word, WAVES> &S, unsigned k,
                         unsigned seed, unsigned pass, int prngKind, const uint32_t *stream, uint32_t streamLen,
                         uint32_t maxvalue, unsigned long long &evalTot, unsigned long long &termTot,
                         unsigned &betTot, unsigned nVisited)
{
  const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid >> 6;
  Word *tword = (Word *)b.tword;
  const Word *cword = (const Word *)b.cword;
  const int2 t = b.targets[k];
  const unsigned count =
      gather_neighbours_wg<Word, WAVES>(g, tword, b.sourceOf, b.offsets, t.x, t.y, S, b.targets, nVisited);

  // heuristic 1 on wave 0
  bool record = false;
  int cx = 0, cy = 0;
  if (wave == 0) {
    uint32_t best = 0xFFFFFFFFu;
    int winX = 0, winY = 0;
    bool perfect = false, scored = false;
    unsigned evals = 0;
    coherence_phase<Word, NC>(g, cword, b.prober, k, count, S.off, S.pix, S.src, S.T, S.Tm, best, winX, winY, perfect,
                          scored, record, cx, cy, evals);
    if (lane == 0) {
      S.best = best;
      S.winX = winX;
      S.winY = winY;
      S.perfect = perfect;
      S.scored = scored;
      S.evals = evals;
      S.aborted = 0;
    }
  }
  __syncthreads();

  // random probes, 64*WAVES per round
  unsigned drawn = 0;
  const uint64_t key = ((uint64_t)seed << 40) ^ ((uint64_t)pass << 36) ^ ((uint64_t)k << 8);
  while (!S.perfect && drawn < g.maxProbe) {
    bool active = false;
    uint32_t idx = 0;
    const uint32_t pos = SERIAL ? S.streamPos : 0u;
    if (SERIAL) {
      if (pos + 64u * WAVES > streamLen) {
        __syncthreads();
        if (tid == 0) S.aborted = 1;
        __syncthreads();
        break;
      }
      uint32_t raw = stream[pos + tid];
      bool acc = draw_index(prngKind, raw, g.nC, maxvalue, idx);
      uint64_t am = __ballot(acc);
      if (lane == 0) S.waveCnt[wave] = (uint32_t)__popcll(am);
      __syncthreads();
      unsigned before = 0;
      for (int wv = 0; wv < (int)wave; wv++) before += S.waveCnt[wv];
      active = acc && drawn + before + (unsigned)__popcll(am & lanemask_lt()) < g.maxProbe;
    } else {
      unsigned j = drawn + tid;
      active = j < g.maxProbe;
      if (active) {
        uint64_t h = mix64(key * 0x100000001B3ull + j);
        idx = probe_index(g, b, h);
      }
    }
    uint32_t F = 0, M = 0;
    bool alive = false;
    int2 c = make_int2(0, 0);
    if (active) {
      c = b.corpusPts[idx];
      score_candidate<Word, 8, 8, NC>(g, cword, c.x, c.y, count, S.off, S.pix, S.T, S.Tm, S.best, F, M, alive,
                                      g.nMaps == 0);
    }
    S.cand[tid] = c;
    S.candM[tid] = M;
    S.candF[tid] = F;
    S.candOk[tid] = alive ? 1 : 0;
    S.candUsed[tid] = active ? 1 : 0;
    __syncthreads();
    if (wave == 0) {
      int lastUsed;
      resolve_slots_wave0<Word, WAVES>(S, 64u * WAVES, lastUsed);
      if (lane == 0) S.lastUsed = (uint32_t)(lastUsed + 1);
    }
    __syncthreads();
    const unsigned nUsed = (unsigned)__syncthreads_count(S.candUsed[tid] != 0);
    if (SERIAL && tid == 0) {
      // raw outputs consumed: through the last one used when the probes end in this
      // window, else the whole window
      bool done = S.perfect || drawn + nUsed >= g.maxProbe;
      S.streamPos = done ? pos + S.lastUsed : pos + 64u * WAVES;
    }
    drawn += nUsed;
    if (tid == 0) S.evals += nUsed;
    __syncthreads();
  }
  if (SERIAL && S.aborted) return;

  // commit (lib/synthesize.h:620-639)
  if (record) b.prober[cidx(g.cw, (unsigned)cx, (unsigned)cy)] = k;
  if (tid == 0) {
    size_t tp = (size_t)t.y * (unsigned)g.tw + (unsigned)t.x;
    Word old = tword[tp];
    Word nw = old | (Word)1;
    if (S.scored) {
      int2 cur = b.sourceOf[tp];
      if (cur.x != S.winX || cur.y != S.winY) {
        Word cmask = 0;
        for (int ch = 0; ch < g.nColors; ch++) cmask |= (Word)0xFF << (8 * (1 + ch));
        Word src = cword[cidx(g.cw, (unsigned)S.winX, (unsigned)S.winY)];
        nw = (old & ~cmask) | (src & cmask) | (Word)1;
        b.sourceOf[tp] = make_int2(S.winX, S.winY);
        betTot++;
      }
    }
    tword[tp] = nw;
    b.lastBest[tp] = S.best;
    if (g.skipEmpty) b.tileHas[(size_t)(t.y / kTile) * g.tilesW + t.x / kTile] = 1;
    evalTot += S.evals;
    termTot += (unsigned long long)S.evals * count;
  }
}

template <typename Word, int WAVES>
__device__ __forceinline__ void load_tables(const SynthBuffers &b, VisitLDS<Word, WAVES> &S)
{
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    S.T[i] = b.tables->T[i];
    S.Tm[i] = b.tables->Tm[i];
  }
  if (threadIdx.x == 0) {
    S.tileHas = b.tileHas;
    S.ctxRadius = b.ctxRadius;
    S.firstIdx = b.firstIdx;
  }
  __syncthreads();
}

// Batched visits: one workgroup per target of [k0, k1).
template <typename Word, int WAVES, int NC = 0>
__global__ __launch_bounds__(64 * WAVES) void k_synth_batched(Geometry g, SynthBuffers b, unsigned k0, unsigned k1,
                                                              unsigned seed, unsigned pass)
{
  __shared__ VisitLDS<Word, WAVES> S;
  load_tables<Word, WAVES>(b, S);
  unsigned long long ev = 0, tm = 0;
  unsigned bt = 0;
  for (unsigned k = k0 + blockIdx.x; k < k1; k += gridDim.x) {
    visit_wg<Word, WAVES, false, NC>(g, b, S, k, seed, pass, 0, nullptr, 0, 0, ev, tm, bt, g.nVisited);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    unsigned slot = blockIdx.x & (CNT_SLOTS - 1);
    if (ev) atomicAdd(&b.counters->candidates[slot], ev);
    if (tm) atomicAdd(&b.counters->terms[slot], tm);
    if (bt) atomicAdd(&b.counters->betters[slot], bt);
  }
}

// Stores of a visit's result that other visits of the same launch read.  With PUBLISH
// they are device-coherent stores (written through to memory), so that visits running on
// other XCDs of a long persistent launch see them as soon as their own cached copy of the
// line ages out, instead of when this XCD's L2 happens to evict the line.
template <bool PUBLISH, typename T>
__device__ __forceinline__ void put(T *p, T v)
{
  if constexpr (PUBLISH)
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    *p = v;
}

// One batched visit of target k by one wavefront (lib/synthesize.h:426-642).
// Per-wave tallies of the large-chunk visit kernels, kept in LDS: candidates, patch terms,
// betters, and the visits' offsets read, neighbours staged and count.
struct VisitTally {
  unsigned long long eval, term, scan, stage, visits;
  unsigned bet;
};

template <typename Word, int G0, int GL, bool PUBLISH, int NC = 0>
__device__ __forceinline__ void visit_wave(const Geometry &g, const SynthBuffers &b, unsigned k, unsigned seed,
                                           unsigned pass, int2 *sOff, Word *sPix, int2 *sSrc, const uint16_t *sT,
                                           const uint32_t *sTm, unsigned long long &myEval,
                                           unsigned long long &myTerm, unsigned &myBet, VisitTally *tally = nullptr)
{
  const unsigned lane = lane_id();
  Word *tword = (Word *)b.tword;
  const Word *cword = (const Word *)b.cword;
  const bool skip0 = g.nMaps == 0;
  const int2 t = b.targets[k];
  // the target's learned proposals do not depend on its neighbourhood: their load overlaps
  // the neighbour scan
  int2 prop = make_int2(-1, -1);
  if (b.props && lane < 16) prop = b.props[(size_t)(k - b.propsK0) * 16 + lane];
  unsigned scanned = 0;
  const unsigned count = gather_neighbours<Word>(g, tword, b.sourceOf, b.offsets, t.x, t.y, sOff, sPix, sSrc,
                                                 b.tileHas, b.firstIdx, &scanned);
  if (tally && lane == 0) {
    tally->scan += scanned;
    tally->stage += count;
    tally->visits++;
  }
  uint32_t best = 0xFFFFFFFFu;
  int winX = 0, winY = 0;
  bool perfect = false, scored = false;
  unsigned evals = 0;
  bool record = false;
  int cx, cy;
  coherence_phase<Word, NC>(g, cword, b.prober, k, count, sOff, sPix, sSrc, sT, sTm, best, winX, winY, perfect,
                            scored, record, cx, cy, evals);
  if (record) b.prober[cidx(g.cw, (unsigned)cx, (unsigned)cy)] = k;
  // learned proposals: up to 16 corpus points, scored in rank order like probes (groups of
  // 8 loads would save round trips but take 8 more VGPRs, one wave per SIMD less)
  if (b.props && !perfect) {
    const int2 c = prop;
    const bool valid = c.x >= 0;
    uint32_t F = 0, M = 0;
    bool alive = false;
    if (valid)
      score_candidate<Word, G0, GL, NC>(g, cword, c.x, c.y, count, sOff, sPix, sT, sTm, best, F, M, alive, skip0);
    int win = -1, stop = 64;
    uint32_t bb = best;
    resolve(valid && alive, M, F, bb, win, stop);
    if (win >= 0) {
      best = bb;
      winX = __shfl(c.x, win);
      winY = __shfl(c.y, win);
      scored = true;
    }
    uint64_t used = __ballot(valid);
    if (stop < 64) {
      perfect = true;
      used &= (stop == 63) ? ~0ull : ((2ull << stop) - 1ull);
    }
    evals += (unsigned)__popcll(used);
  }
  // random probes from a per-target counter-based stream (seed, pass, target index,
  // attempt).  Dense corpora draw a uniform pixel of the corpus image and keep it when
  // its word marks a corpus point (0xFF: selected and opaque), which is uniform over
  // corpus points and lets the validity test ride in the first group of loads; sparse
  // corpora index the corpus point list.  Draws are resolved in attempt order.
  const uint64_t key = ((uint64_t)seed << 40) ^ ((uint64_t)pass << 36) ^ ((uint64_t)k << 8);
  const bool byPixel = g.probeByPixel != 0;
  int r0, r1;
  probe_band(g, r0, r1);
  const uint64_t nPix = (uint64_t)g.cw * (uint64_t)(r1 - r0);
  const uint64_t pix0 = (uint64_t)g.cw * (uint64_t)r0;
  unsigned drawn = 0;
  for (unsigned attempt = 0; drawn < g.maxProbe && !perfect; attempt += 64) {
    const uint64_t h = mix64(key * 0x100000001B3ull + attempt + lane);
    int2 c;
    bool valid;
    uint32_t F = 0, M = 0;
    bool alive = false;
    if (byPixel) {
      const uint64_t p = pix0 + (((h >> 32) * nPix) >> 32);
      c = make_int2((int)(p % (unsigned)g.cw), (int)(p / (unsigned)g.cw));
      // the candidate's own word (neighbour 0, loaded with the first group) carries the
      // corpus-point flag
      const Word w0 = cword[cidx(g.cw, (unsigned)c.x, (unsigned)c.y)];
      score_candidate<Word, G0 + 1, GL, NC>(g, cword, c.x, c.y, count, sOff, sPix, sT, sTm, best, F, M, alive, false);
      valid = (w0 & 0xFFu) == 0xFFu;
    } else {
      c = b.corpusPts[probe_index(g, b, h)];
      valid = true;
      score_candidate<Word, G0, GL, NC>(g, cword, c.x, c.y, count, sOff, sPix, sT, sTm, best, F, M, alive, skip0);
    }
    const uint64_t vm = __ballot(valid);
    const bool active = valid && drawn + (unsigned)__popcll(vm & lanemask_lt()) < g.maxProbe;
    int win = -1, stop = 64;
    uint32_t bb = best;
    resolve(active && alive, M, F, bb, win, stop);
    if (win >= 0) {
      best = bb;
      winX = __shfl(c.x, win);
      winY = __shfl(c.y, win);
      scored = true;
    }
    uint64_t used = __ballot(active);
    if (stop < 64) {
      perfect = true;
      used &= (stop == 63) ? ~0ull : ((2ull << stop) - 1ull);
    }
    evals += (unsigned)__popcll(used);
    drawn += (unsigned)__popcll(__ballot(active));
  }
  // store (lib/synthesize.h:620-639)
  if (lane == 0) {
    const size_t tp = (size_t)t.y * (unsigned)g.tw + (unsigned)t.x;
    const Word old = tword[tp];
    Word nw = old | (Word)1;
    if (scored) {
      const int2 cur = b.sourceOf[tp];
      if (cur.x != winX || cur.y != winY) {
        Word cmask = 0;
        for (int ch = 0; ch < g.nColors; ch++) cmask |= (Word)0xFF << (8 * (1 + ch));
        const Word src = cword[cidx(g.cw, (unsigned)winX, (unsigned)winY)];
        nw = (old & ~cmask) | (src & cmask) | (Word)1;
        const uint64_t packed = (uint64_t)(uint32_t)winX | ((uint64_t)(uint32_t)winY << 32);
        put<PUBLISH>((uint64_t *)&b.sourceOf[tp], packed);
        myBet++;
      }
    }
    put<PUBLISH>(&tword[tp], nw);
    b.lastBest[tp] = best;
    if (g.skipEmpty) b.tileHas[(size_t)(t.y / kTile) * g.tilesW + t.x / kTile] = 1;
  }
  myEval += evals;
  myTerm += (unsigned long long)evals * count;
}

// ------------------------------------------------------------------ segmented visits

// A wavefront split into 64/LPT segments of LPT lanes, one target each.  Phases with few
// candidates (a patch of heuristic-1 candidates, 16 learned proposals, a handful of random
// probes) then fill the wave with several targets instead of leaving most lanes idle, and
// each instruction of the visit serves 64/LPT targets.  bits() is a ballot restricted to
// the lane's own segment, in segment-local lane order.
template <int LPT>
struct Seg {
  static_assert(LPT == 16 || LPT == 32 || LPT == 64, "segments of 16, 32 or 64 lanes");
  unsigned lane, seg, sl, base;
  __device__ __forceinline__ Seg()
  {
    lane = lane_id();
    seg = lane / LPT;
    sl = lane % LPT;
    base = seg * LPT;
  }
  __device__ __forceinline__ uint64_t bits(uint64_t m) const
  {
    return LPT == 64 ? m : (m >> base) & ((1ull << LPT) - 1ull);
  }
  __device__ __forceinline__ uint64_t lt() const { return (1ull << sl) - 1ull; }
};

// Sequential strict-less resolution within each segment, in segment lane order (the
// segmented form of resolve).  best/win/stop are the lane's segment's values.
template <int LPT>
__device__ __forceinline__ void resolve_seg(const Seg<LPT> &S, bool cand, uint32_t M, uint32_t F, uint32_t &best,
                                            int &win, int &stop)
{
  uint64_t rem = __ballot(cand);
  stop = LPT;
  for (;;) {
    const uint64_t m = __ballot(cand && M < best) & rem;
    if (!m) break;
    const uint64_t mine = S.bits(m);
    const int i = mine ? (int)__ffsll((unsigned long long)mine) - 1 : 0;
    const uint32_t Fi = __shfl(F, (int)S.base + i);
    uint64_t segRem = S.bits(rem);
    if (mine) {
      best = Fi;
      win = i;
      segRem &= (i == LPT - 1) ? 0ull : ~((2ull << i) - 1ull);
      if (Fi == 0) {
        stop = i;
        segRem = 0;
      }
    } else {
      segRem = 0;   // nothing left in this segment can beat its best
    }
    rem = __ballot((segRem >> S.sl) & 1ull);
  }
}

// Neighbour patch (lib/synthesize.h:189-241) of each segment's target, gathered by its
// LPT lanes: neighbour 0 is the target itself, then the sorted offsets LPT at a time.
// Returns the segment's patch size.
template <typename Word, int LPT>
__device__ unsigned gather_neighbours_seg(const Seg<LPT> &S, const Geometry &g, const Word *tword,
                                          const int2 *sourceOf, const int2 *offsets, int tx, int ty, bool active,
                                          int2 *nbOff, Word *nbPix, int2 *nbSrc, const uint8_t *tileHas,
                                          const uint32_t *firstIdx, unsigned &scanned)
{
  const size_t tp = (size_t)ty * (unsigned)g.tw + (unsigned)tx;
  if (active && S.sl == 0) {
    nbOff[0] = make_int2(0, 0);
    nbPix[0] = tword[tp];
    nbSrc[0] = sourceOf[tp];
  }
  unsigned count = 1;
  const unsigned L = g.patchLimit;
  unsigned base = 1;
  if (g.skipEmpty && tileHas && active)
    base = scan_start(g, tileHas, firstIdx, tx, ty, S.sl, (unsigned)LPT,
                      [&S](bool v) { return S.bits(__ballot(v)) != 0ull; });
  // a round's offsets depend only on base: each round loads the next round's, so only the
  // pixel gather waits on memory (offNext always holds entry base + sl, or entry 0 past nO)
  int2 offNext = offsets[base + S.sl < g.nO ? base + S.sl : 0u];
  for (;;) {
    const bool need = active && count < L && base < g.nO;
    if (__ballot(need) == 0ull) break;
    if (need) {
      const unsigned j = base + S.sl;
      const int2 off = offNext;
      offNext = offsets[j + LPT < g.nO ? j + LPT : 0u];
      int px = tx + off.x, py = ty + off.y;
      bool in = j < g.nO;
      if (px < 0) { if (g.tileH) px += g.tw; else in = false; }
      else if (px >= g.tw) { if (g.tileH) px -= g.tw; else in = false; }
      if (py < 0) { if (g.tileV) py += g.th; else in = false; }
      else if (py >= g.th) { if (g.tileV) py -= g.th; else in = false; }
      const size_t p = in ? (size_t)(__umul24((unsigned)py, (unsigned)g.tw) + (unsigned)px) : tp;
      const Word w = tword[p];
      const bool v = in && (w & 1u) != 0;
      const uint64_t m = S.bits(__ballot(v));
      const unsigned rank = (unsigned)__popcll(m & S.lt());
      if (v && count + rank < L) {
        const unsigned idx = count + rank;
        nbOff[idx] = off;
        nbPix[idx] = w;
        nbSrc[idx] = sourceOf[p];
      }
      const unsigned c2 = count + (unsigned)__popcll(m);
      count = c2 < L ? c2 : L;
      base += LPT;
      scanned += LPT;
    }
  }
  wave_sync();
  return count;
}

// One batched visit (lib/synthesize.h:426-642) of each of the wave's 64/LPT targets
// kb + seg (those below k1).  Per target it is visit_wave's visit: the same patch, the same
// heuristic-1 candidates in neighbour order (deduplicated, prober-checked), the same
// proposals, and the same per-target counter-based probe stream resolved in attempt order,
// so a target gets the same result whatever LPT is.  A probe round draws only the attempts
// the remaining budget needs (for probes by pixel, scaled by the corpus density).
template <typename Word, int LPT, int G0, int GL, bool PUBLISH, int NC = 0>
__device__ __forceinline__ void visit_seg(const Geometry &g, const SynthBuffers &b, unsigned kb, unsigned k1,
                                          unsigned seed, unsigned pass, int2 *sOff, Word *sPix, int2 *sSrc,
                                          const uint16_t *sT, const uint32_t *sTm, VisitTally *tally)
{
  const Seg<LPT> S;
  Word *tword = (Word *)b.tword;
  const Word *cword = (const Word *)b.cword;
  const bool skip0 = g.nMaps == 0;
  const unsigned k = kb + S.seg;
  const bool active = k < k1;
  const int2 t = b.targets[active ? k : kb];
  // the proposals do not depend on the neighbourhood: their load overlaps the scan
  int2 prop = make_int2(-1, -1);
  if (b.props && active && S.sl < 16) prop = b.props[(size_t)(k - b.propsK0) * 16 + S.sl];
  unsigned scanned = 0;
  const unsigned count = gather_neighbours_seg<Word, LPT>(S, g, tword, b.sourceOf, b.offsets, t.x, t.y, active, sOff,
                                                          sPix, sSrc, b.tileHas, b.firstIdx, scanned);
  uint32_t best = 0xFFFFFFFFu;
  int winX = 0, winY = 0;
  bool perfect = !active, scored = false;
  unsigned evals = 0;

  // heuristic 1 (lib/synthesize.h:537-580): neighbour i proposes source(i) - offset(i).
  // First each neighbour's candidate replaces its source in sSrc ((INT_MIN, INT_MIN) when
  // the neighbour has no source), so that a candidate can be compared with every earlier one.
  if (active)
    for (unsigned i = S.sl; i < count; i += LPT) {
      const int2 src = sSrc[i], o = sOff[i];
      sSrc[i] = src.x != -1 ? make_int2(src.x - o.x, src.y - o.y) : make_int2(INT_MIN, INT_MIN);
    }
  wave_sync();
  for (unsigned r0 = 0;; r0 += LPT) {
    const bool go = !perfect && r0 < count;
    if (__ballot(go) == 0ull) break;
    const unsigned i = r0 + S.sl;
    const bool mine = go && i < count;
    const int2 c = sSrc[mine ? i : 0u];
    const bool in = mine && c.x != INT_MIN && (unsigned)c.x < (unsigned)g.cw && (unsigned)c.y < (unsigned)g.ch;
    const unsigned ci = in ? cidx(g.cw, (unsigned)c.x, (unsigned)c.y) : 0u;
    const Word q = cword[ci];
    const uint32_t pr = b.prober[ci];
    bool valid = in && (q & 0xFEu) == 0xFEu && pr != k;
    // an earlier neighbour proposing the same point was scored (and recorded) already.
    // Entries [0, i) are read in pairs (entry i at most is read past them, and i < count <= 64)
    // and compared without an early exit: one LDS round trip per two entries and no divergent
    // exit (unrolling further costs registers and, on some instances, a wave per SIMD).
    if (valid) {
      bool dup = false;
#pragma unroll 1
      for (unsigned j = 0; j < i; j += 2) {
        const int2 e0 = sSrc[j], e1 = sSrc[j + 1];
        dup |= (e0.x == c.x) & (e0.y == c.y);
        dup |= (j + 1 < i) & (e1.x == c.x) & (e1.y == c.y);
      }
      valid = !dup;
    }
    uint32_t F = 0, M = 0;
    bool alive = false;
    if (valid) score_candidate<Word, 8, 8, NC>(g, cword, c.x, c.y, count, sOff, sPix, sT, sTm, best, F, M, alive, skip0);
    const uint64_t used0 = S.bits(__ballot(valid));
    int win = -1, stop = LPT;
    uint32_t bb = best;
    resolve_seg<LPT>(S, valid && alive, M, F, bb, win, stop);
    const int wx = __shfl(c.x, (int)S.base + (win >= 0 ? win : 0));
    const int wy = __shfl(c.y, (int)S.base + (win >= 0 ? win : 0));
    if (win >= 0) {
      best = bb;
      winX = wx;
      winY = wy;
      scored = true;
    }
    uint64_t used = used0, rec = used0;
    if (stop < LPT) {
      perfect = true;
      used &= (stop == LPT - 1) ? ~0ull : ((2ull << stop) - 1ull);
      rec &= (1ull << stop) - 1ull;
    }
    if (go) evals += (unsigned)__popcll(used);
    // recentProber (lib/synthesize.h:562-575): every candidate scored before the stop; the
    // loop breaks on a perfect match before recording it
    if (valid && ((rec >> S.sl) & 1ull)) b.prober[ci] = k;
  }

  // learned proposals: up to 16 corpus points, scored in rank order like probes
  if (b.props) {
    const bool go = !perfect;
    const bool valid = go && S.sl < 16 && prop.x >= 0;
    uint32_t F = 0, M = 0;
    bool alive = false;
    if (valid)
      score_candidate<Word, G0, GL, NC>(g, cword, prop.x, prop.y, count, sOff, sPix, sT, sTm, best, F, M, alive,
                                        skip0);
    int win = -1, stop = LPT;
    uint32_t bb = best;
    resolve_seg<LPT>(S, valid && alive, M, F, bb, win, stop);
    const int wx = __shfl(prop.x, (int)S.base + (win >= 0 ? win : 0));
    const int wy = __shfl(prop.y, (int)S.base + (win >= 0 ? win : 0));
    if (win >= 0) {
      best = bb;
      winX = wx;
      winY = wy;
      scored = true;
    }
    uint64_t used = S.bits(__ballot(valid));
    if (stop < LPT) {
      perfect = true;
      used &= (stop == LPT - 1) ? ~0ull : ((2ull << stop) - 1ull);
    }
    if (go) evals += (unsigned)__popcll(used);
  }

  // random probes: the per-target counter-based stream (seed, pass, target, attempt) of
  // visit_wave, taken in rounds of at most LPT attempts
  const uint64_t key = ((uint64_t)seed << 40) ^ ((uint64_t)pass << 36) ^ ((uint64_t)k << 8);
  const bool byPixel = g.probeByPixel != 0;
  int r0, r1;
  probe_band(g, r0, r1);
  const uint64_t nPix = (uint64_t)g.cw * (uint64_t)(r1 - r0);
  const uint64_t pix0 = (uint64_t)g.cw * (uint64_t)r0;
  unsigned drawn = 0, attempt = 0;
  for (;;) {
    const bool go = !perfect && drawn < g.maxProbe;
    if (__ballot(go) == 0ull) break;
    const unsigned remaining = go ? g.maxProbe - drawn : 0u;
    unsigned m = byPixel ? (unsigned)(((uint64_t)remaining * g.drawScale + 0xFFFFu) >> 16) + 1u : remaining;
    m = m < (unsigned)LPT ? m : (unsigned)LPT;
    const bool draw = go && S.sl < m;
    const uint64_t h = mix64(key * 0x100000001B3ull + attempt + S.sl);
    int2 c = make_int2(0, 0);
    bool valid = false;
    if (draw) {
      if (byPixel) {
        const uint64_t p = pix0 + (((h >> 32) * nPix) >> 32);
        c = make_int2((int)(p % (unsigned)g.cw), (int)(p / (unsigned)g.cw));
        // a corpus point: selected and opaque
        valid = (cword[cidx(g.cw, (unsigned)c.x, (unsigned)c.y)] & 0xFFu) == 0xFFu;
      } else {
        c = b.corpusPts[probe_index(g, b, h)];
        valid = true;
      }
    }
    const uint64_t vm = S.bits(__ballot(valid));
    const bool act = valid && (unsigned)__popcll(vm & S.lt()) < remaining;
    uint32_t F = 0, M = 0;
    bool alive = false;
    if (act) score_candidate<Word, G0, GL, NC>(g, cword, c.x, c.y, count, sOff, sPix, sT, sTm, best, F, M, alive, skip0);
    int win = -1, stop = LPT;
    uint32_t bb = best;
    resolve_seg<LPT>(S, act && alive, M, F, bb, win, stop);
    const int wx = __shfl(c.x, (int)S.base + (win >= 0 ? win : 0));
    const int wy = __shfl(c.y, (int)S.base + (win >= 0 ? win : 0));
    if (win >= 0) {
      best = bb;
      winX = wx;
      winY = wy;
      scored = true;
    }
    uint64_t used = S.bits(__ballot(act));
    drawn += (unsigned)__popcll(used);
    if (stop < LPT) {
      perfect = true;
      used &= (stop == LPT - 1) ? ~0ull : ((2ull << stop) - 1ull);
    }
    if (go) evals += (unsigned)__popcll(used);
    attempt += m;
  }

  // store (lib/synthesize.h:620-639)
  if (active && S.sl == 0) {
    const size_t tp = (size_t)t.y * (unsigned)g.tw + (unsigned)t.x;
    const Word old = tword[tp];
    Word nw = old | (Word)1;
    if (scored) {
      const int2 cur = b.sourceOf[tp];
      if (cur.x != winX || cur.y != winY) {
        Word cmask = 0;
        for (int ch = 0; ch < g.nColors; ch++) cmask |= (Word)0xFF << (8 * (1 + ch));
        const Word src = cword[cidx(g.cw, (unsigned)winX, (unsigned)winY)];
        nw = (old & ~cmask) | (src & cmask) | (Word)1;
        const uint64_t packed = (uint64_t)(uint32_t)winX | ((uint64_t)(uint32_t)winY << 32);
        put<PUBLISH>((uint64_t *)&b.sourceOf[tp], packed);
        atomicAdd(&tally->bet, 1u);
      }
    }
    put<PUBLISH>(&tword[tp], nw);
    b.lastBest[tp] = best;
    if (g.skipEmpty) b.tileHas[(size_t)(t.y / kTile) * g.tilesW + t.x / kTile] = 1;
    // segment leaders of one wave add at once: LDS atomics
    atomicAdd(&tally->eval, (unsigned long long)evals);
    atomicAdd(&tally->term, (unsigned long long)evals * count);
    atomicAdd(&tally->scan, (unsigned long long)scanned);
    atomicAdd(&tally->stage, (unsigned long long)count);
    atomicAdd(&tally->visits, 1ull);
  }
}

// The large-chunk kernels' tallies: one LDS record per wave, cleared before the block's
// first barrier, summed over the block's waves at the end, then one atomic per counter
// into the block's slot.
template <int WAVES>
__device__ __forceinline__ void tally_clear(VisitTally *sT)
{
  if (threadIdx.x < WAVES) sT[threadIdx.x] = VisitTally{};
}

template <int WAVES>
__device__ __forceinline__ void flush_tally(const SynthBuffers &b, const VisitTally *sT)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    VisitTally a = {};
    for (int w = 0; w < WAVES; w++) {
      a.eval += sT[w].eval;
      a.term += sT[w].term;
      a.scan += sT[w].scan;
      a.stage += sT[w].stage;
      a.visits += sT[w].visits;
      a.bet += sT[w].bet;
    }
    const unsigned slot = blockIdx.x & (CNT_SLOTS - 1);
    if (a.eval) atomicAdd(&b.counters->candidates[slot], a.eval);
    if (a.term) {
      atomicAdd(&b.counters->terms[slot], a.term);
      atomicAdd(&b.counters->termsWave[slot], a.term);
    }
    if (a.scan) atomicAdd(&b.counters->scanWave[slot], a.scan);
    if (a.stage) atomicAdd(&b.counters->stageWave[slot], a.stage);
    if (a.visits) atomicAdd(&b.counters->visitsWave[slot], a.visits);
    if (a.bet) atomicAdd(&b.counters->betters[slot], a.bet);
  }
}

template <typename Word, int WAVES, bool WAVE_KERNEL>
__device__ __forceinline__ void flush_counters(const SynthBuffers &b, unsigned long long myEval,
                                               unsigned long long myTerm, unsigned myBet,
                                               unsigned long long *sEval, unsigned long long *sTerm, unsigned *sBet)
{
  const int wave = threadIdx.x >> 6;
  if (lane_id() == 0) {
    sEval[wave] = myEval;
    sTerm[wave] = myTerm;
    sBet[wave] = myBet;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long e = 0, tm = 0;
    unsigned bt = 0;
    for (int w = 0; w < WAVES; w++) {
      e += sEval[w];
      tm += sTerm[w];
      bt += sBet[w];
    }
    const unsigned slot = blockIdx.x & (CNT_SLOTS - 1);
    if (e) atomicAdd(&b.counters->candidates[slot], e);
    if (tm) atomicAdd(&b.counters->terms[slot], tm);
    if (WAVE_KERNEL && tm) atomicAdd(&b.counters->termsWave[slot], tm);
    if (bt) atomicAdd(&b.counters->betters[slot], bt);
  }
}

// Batched visits, one wavefront per target (large chunks: every wave stays busy on its
// own target instead of idling through another wave's gather and heuristic-1 phases).
// The kernel is bound by VALU issue and latency, so the small-group shapes ask for 7 waves
// per SIMD (at most 72 VGPRs; they fit without spilling); the larger groups would spill.
template <typename Word, int WAVES, int G0, int GL, int NC = 0>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu((GL == 4 && G0 <= 4) ? 7 : 1)))
void k_synth_wave(Geometry g, SynthBuffers b, unsigned k0, unsigned k1, unsigned seed, unsigned pass)
{
  __shared__ uint16_t sT[512];
  __shared__ uint32_t sTm[512];
  __shared__ int2 sOff[WAVES][64];
  __shared__ Word sPix[WAVES][64];
  __shared__ int2 sSrc[WAVES][64];
  __shared__ VisitTally sTally[WAVES];
  tally_clear<WAVES>(sTally);
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    sT[i] = b.tables->T[i];
    sTm[i] = b.tables->Tm[i];
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  unsigned long long myEval = 0, myTerm = 0;
  unsigned myBet = 0;
  for (unsigned k = k0 + blockIdx.x * WAVES + wave; k < k1; k += gridDim.x * WAVES)
    visit_wave<Word, G0, GL, false, NC>(g, b, k, seed, pass, sOff[wave], sPix[wave], sSrc[wave], sT, sTm, myEval,
                                        myTerm, myBet, &sTally[wave]);
  if (lane_id() == 0) {
    sTally[wave].eval = myEval;
    sTally[wave].term = myTerm;
    sTally[wave].bet = myBet;
  }
  flush_tally<WAVES>(b, sTally);
}

// Batched visits, 64/LPT targets per wavefront (segmented visits; see visit_seg).
template <typename Word, int WAVES, int LPT, int G0, int GL, int NC = 0>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(6)))
void k_synth_seg(Geometry g, SynthBuffers b, unsigned k0, unsigned k1, unsigned seed, unsigned pass)
{
  constexpr int TPW = 64 / LPT;
  __shared__ uint16_t sT[512];
  __shared__ uint32_t sTm[512];
  __shared__ int2 sOff[WAVES][TPW][64];
  __shared__ Word sPix[WAVES][TPW][64];
  __shared__ int2 sSrc[WAVES][TPW][64];
  __shared__ VisitTally sTally[WAVES];
  tally_clear<WAVES>(sTally);
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    sT[i] = b.tables->T[i];
    sTm[i] = b.tables->Tm[i];
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const unsigned seg = lane_id() / LPT;
  for (unsigned kb = k0 + (blockIdx.x * WAVES + wave) * TPW; kb < k1; kb += gridDim.x * WAVES * TPW)
    visit_seg<Word, LPT, G0, GL, false, NC>(g, b, kb, k1, seed, pass, sOff[wave][seg], sPix[wave][seg],
                                            sSrc[wave][seg], sT, sTm, &sTally[wave]);
  flush_tally<WAVES>(b, sTally);
}

// Sliding-window batched visits in one persistent launch.  Waves take targets in
// synthesis order from a ticket counter; target k starts once all but W of the targets
// before it in the launch have finished, where W = max(minW, k * rampPermille / 1000)
// (capped at cap) when rampPermille > 0, else cap.  A visit therefore misses the results
// of at most W earlier visits, as a chunk of about 2W would, but no wave idles at chunk
// ends.  Every wait is on lower tickets, held by waves that are already running, so the
// lowest unfinished ticket never waits and the launch always drains; a spin limit bounds
// each wait all the same, and after one expiry (flagged in ctl) waits are skipped.
constexpr unsigned kSpinLimit = 1u << 18;

__device__ __forceinline__ unsigned load_agent(const unsigned *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename Word, int WAVES, int G0, int GL>
__global__ __launch_bounds__(64 * WAVES) void k_synth_window(Geometry g, SynthBuffers b, unsigned k0, unsigned k1,
                                                             unsigned seed, unsigned pass, unsigned minW,
                                                             unsigned rampPermille, unsigned cap)
{
  __shared__ uint16_t sT[512];
  __shared__ uint32_t sTm[512];
  __shared__ int2 sOff[WAVES][64];
  __shared__ Word sPix[WAVES][64];
  __shared__ int2 sSrc[WAVES][64];
  __shared__ unsigned long long sEval[WAVES], sTerm[WAVES];
  __shared__ unsigned sBet[WAVES];
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    sT[i] = b.tables->T[i];
    sTm[i] = b.tables->Tm[i];
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const unsigned lane = lane_id();
  unsigned *ticket = b.ctl + WCTL_TICKET, *done = b.ctl + WCTL_DONE, *expired = b.ctl + WCTL_EXPIRED;
  const unsigned n = k1 - k0;
  unsigned long long myEval = 0, myTerm = 0;
  unsigned myBet = 0;
  unsigned seen = 0;  // a lower bound of *done, which only grows
  for (;;) {
    unsigned tk = 0;
    if (lane == 0) tk = atomicAdd(ticket, 1u);
    tk = __builtin_amdgcn_readfirstlane(tk);
    if (tk >= n) break;
    const unsigned k = k0 + tk;
    unsigned W = cap;
    if (rampPermille) {
      const unsigned r = (unsigned)((uint64_t)k * rampPermille / 1000u);
      W = r > minW ? r : minW;
      if (W > cap) W = cap;
    }
    if (tk > W && seen < tk - W) {
      const unsigned need = tk - W;
      if (lane == 0) {
        seen = load_agent(done);
        unsigned spins = 0;
        while (seen < need) {
          if (++spins > kSpinLimit || load_agent(expired)) {
            atomicOr(expired, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
          seen = load_agent(done);
        }
      }
      seen = __builtin_amdgcn_readfirstlane(seen);
    }
    visit_wave<Word, G0, GL, true>(g, b, k, seed, pass, sOff[wave], sPix[wave], sSrc[wave], sT, sTm, myEval, myTerm,
                                   myBet);
    if (lane == 0) atomicAdd(done, 1u);
  }
  flush_counters<Word, WAVES, false>(b, myEval, myTerm, myBet, sEval, sTerm, sBet);
}

// Chunk sizes at or above this use one wavefront per target (env RESYNTH_WAVE_MIN), and
// the size of the first group of random-probe loads (env RESYNTH_FIRST_GROUP: 2, 4 or 8).
static unsigned wave_min()
{
  static unsigned v = [] {
    const char *e = std::getenv("RESYNTH_WAVE_MIN");
    return e ? (unsigned)std::atoi(e) : 2048u;
  }();
  return v;
}
unsigned synth_wave_min() { return wave_min(); }

// RESYNTH_RGB_FAST=0 keeps the generic scorer everywhere (A/B comparisons).
static bool rgb_fast_path()
{
  const char *e = std::getenv("RESYNTH_RGB_FAST");
  return e ? std::atoi(e) != 0 : true;
}

static int first_group()
{
  static int v = [] {
    const char *e = std::getenv("RESYNTH_FIRST_GROUP");
    int x = e ? std::atoi(e) : 2;
    return (x == 2 || x == 4 || x == 8) ? x : 2;
  }();
  return v;
}

template <typename Word, int G0>
static void launch_wave_g(hipStream_t s, dim3 grid, const Geometry &g, const SynthBuffers &b, unsigned k0,
                          unsigned k1, unsigned seed, unsigned pass)
{
  constexpr int WAVES = 4;
  // size of the later groups of random-probe loads (env RESYNTH_LATER_GROUP: 4 or 8)
  static const int later = [] {
    const char *e = std::getenv("RESYNTH_LATER_GROUP");
    return (e && std::atoi(e) == 8) ? 8 : 4;
  }();
  // RGB without maps (32-bit words) takes the branch-free scorer
  const bool rgb = sizeof(Word) == 4 && g.nColors == 3 && g.nMaps == 0 && rgb_fast_path();
  if (later == 8) {
    if (rgb)
      hipLaunchKernelGGL((k_synth_wave<Word, WAVES, G0, 8, 3>), grid, dim3(64 * WAVES), 0, s, g, b, k0, k1, seed, pass);
    else
      hipLaunchKernelGGL((k_synth_wave<Word, WAVES, G0, 8>), grid, dim3(64 * WAVES), 0, s, g, b, k0, k1, seed, pass);
  } else {
    if (rgb)
      hipLaunchKernelGGL((k_synth_wave<Word, WAVES, G0, 4, 3>), grid, dim3(64 * WAVES), 0, s, g, b, k0, k1, seed, pass);
    else
      hipLaunchKernelGGL((k_synth_wave<Word, WAVES, G0, 4>), grid, dim3(64 * WAVES), 0, s, g, b, k0, k1, seed, pass);
  }
}

// Lanes per target of the large-chunk visits: as many as a visit's random probes can use
// (a wave of 64 lanes per target wastes most of its issue slots on a visit that scores 16
// proposals and one probe).  Pass 0 keeps at least 32: its scans are long.  Measured on one
// MI355X: a 4096x4096 learned heal (16 probes in pass 0, 1 probe + 16 proposals later)
// took 16.9 ms with 32 then 16 lanes against 19.6 ms with one wavefront per target; a
// 2048x2048 plain heal (200 probes) 27.3 ms with 64 lanes against 29.2 ms.
// RESYNTH_LPT0 / RESYNTH_LPT override pass 0 / later passes: 16, 32, 64, or 1 (the
// wavefront-per-target kernel).
static int lanes_per_target(const Geometry &g, unsigned pass)
{
  static const int v0 = [] {
    const char *e = std::getenv("RESYNTH_LPT0");
    const int x = e ? std::atoi(e) : 0;
    return (x == 1 || x == 16 || x == 32 || x == 64) ? x : 0;
  }();
  static const int v1 = [] {
    const char *e = std::getenv("RESYNTH_LPT");
    const int x = e ? std::atoi(e) : 0;
    return (x == 1 || x == 16 || x == 32 || x == 64) ? x : 0;
  }();
  const int forced = pass == 0 ? v0 : v1;
  if (forced) return forced;
  int l = g.maxProbe >= 48 ? 64 : g.maxProbe >= 12 ? 32 : 16;
  if (pass == 0 && l < 32) l = 32;
  return l;
}

template <typename Word, int LPT>
static void launch_seg_l(hipStream_t s, const Geometry &g, const SynthBuffers &b, unsigned k0, unsigned k1,
                         unsigned seed, unsigned pass)
{
  constexpr int WAVES = 4, TPW = 64 / LPT;
  unsigned blocks = (k1 - k0 + WAVES * TPW - 1) / (WAVES * TPW);
  if (blocks > 16384) blocks = 16384;
  const bool rgb = sizeof(Word) == 4 && g.nColors == 3 && g.nMaps == 0 && rgb_fast_path();
  if (rgb)
    hipLaunchKernelGGL((k_synth_seg<Word, WAVES, LPT, 2, 4, 3>), dim3(blocks), dim3(64 * WAVES), 0, s, g, b, k0, k1,
                       seed, pass);
  else
    hipLaunchKernelGGL((k_synth_seg<Word, WAVES, LPT, 2, 4>), dim3(blocks), dim3(64 * WAVES), 0, s, g, b, k0, k1,
                       seed, pass);
}

template <typename Word>
static void launch_wave(hipStream_t s, const Geometry &g, const SynthBuffers &b, unsigned k0, unsigned k1,
                        unsigned seed, unsigned pass, int lpt)
{
  constexpr int WAVES = 4;
  switch (lpt ? lpt : lanes_per_target(g, pass)) {
  case 16: launch_seg_l<Word, 16>(s, g, b, k0, k1, seed, pass); return;
  case 32: launch_seg_l<Word, 32>(s, g, b, k0, k1, seed, pass); return;
  case 64: launch_seg_l<Word, 64>(s, g, b, k0, k1, seed, pass); return;
  default: break;
  }
  unsigned blocks = (k1 - k0 + WAVES - 1) / WAVES;
  if (blocks > 16384) blocks = 16384;
  switch (first_group()) {
  case 4: launch_wave_g<Word, 4>(s, dim3(blocks), g, b, k0, k1, seed, pass); break;
  case 8: launch_wave_g<Word, 8>(s, dim3(blocks), g, b, k0, k1, seed, pass); break;
  default: launch_wave_g<Word, 2>(s, dim3(blocks), g, b, k0, k1, seed, pass); break;
  }
}

hipError_t launch_synth_batched(hipStream_t s, bool wide, const Geometry &g, const SynthBuffers &b, unsigned k0,
                                unsigned k1, unsigned seed, unsigned pass, int lpt)
{
  constexpr int WAVES = 4;
  if (k1 <= k0) return hipSuccess;
  // learned proposals are scored by the wavefront kernels only
  if (lpt || k1 - k0 >= wave_min() || b.props) {
    if (wide) launch_wave<uint64_t>(s, g, b, k0, k1, seed, pass, lpt);
    else launch_wave<uint32_t>(s, g, b, k0, k1, seed, pass, lpt);
    return hipGetLastError();
  }
  unsigned blocks = k1 - k0;
  const bool rgb = !wide && g.nColors == 3 && g.nMaps == 0 && rgb_fast_path();
  if (blocks <= 256) {
    // few targets: a 1024-lane workgroup per target shortens each visit's latency
    if (wide)
      hipLaunchKernelGGL((k_synth_batched<uint64_t, 16>), dim3(blocks), dim3(1024), 0, s, g, b, k0, k1, seed, pass);
    else if (rgb)
      hipLaunchKernelGGL((k_synth_batched<uint32_t, 16, 3>), dim3(blocks), dim3(1024), 0, s, g, b, k0, k1, seed, pass);
    else
      hipLaunchKernelGGL((k_synth_batched<uint32_t, 16>), dim3(blocks), dim3(1024), 0, s, g, b, k0, k1, seed, pass);
    return hipGetLastError();
  }
  if (wide)
    hipLaunchKernelGGL((k_synth_batched<uint64_t, WAVES>), dim3(blocks), dim3(64 * WAVES), 0, s, g, b, k0, k1, seed,
                       pass);
  else if (rgb)
    hipLaunchKernelGGL((k_synth_batched<uint32_t, WAVES, 3>), dim3(blocks), dim3(64 * WAVES), 0, s, g, b, k0, k1,
                       seed, pass);
  else
    hipLaunchKernelGGL((k_synth_batched<uint32_t, WAVES>), dim3(blocks), dim3(64 * WAVES), 0, s, g, b, k0, k1, seed,
                       pass);
  return hipGetLastError();
}

// Many requests in one persistent launch.  Each request runs its passes as sliding
// windows (k_synth_window) with its own ticket counter; a wave keeps visiting targets of
// its current request and moves on to the next request when that one has no ticket to
// give.  The wave that finishes a pass applies the refiner's stop rule for its request
// (lib/refiner.h:111-115) and opens the next pass or retires the request.  Waits are on
// lower tickets of the same pass only, so every pass completes; waves leave when no
// request is active.  Spin limits bound every wait and the idle polling all the same
// (an expiry is flagged in the global words).
constexpr unsigned kIdleLimit = 1u << 24;

__device__ __forceinline__ unsigned long long load_agent64(const unsigned long long *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long first_lane64(unsigned long long v)
{
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Tickets a wave takes at once: one counter update per group instead of per visit.
constexpr unsigned kTicketGroup = 4;
// How far past its window a ticket may be handed out (it then waits for the visits before
// it): a few groups, so that waves queue on a request instead of leaving it.
constexpr unsigned kAdmitSlack = 4 * kTicketGroup;

// Visits of a pass that ticket t may overlap: the pass-0 ramp, or the cap.
__device__ __forceinline__ unsigned many_window(const ReqDesc &d, unsigned pass, unsigned t)
{
  if (pass != 0) return d.capW;
  const unsigned rr = (unsigned)((uint64_t)t * d.ramp / 1000u);
  const unsigned W = rr > d.minW ? rr : d.minW;
  return W < d.capW ? W : d.capW;
}

// A wave visits its ticket group 64/LPT targets at a time (segmented visits, visit_seg);
// the resident grid below follows from the occupancy query.  32-bit words ask for 6 waves
// per SIMD (at most 80 VGPRs), which they reach without spilling.
template <typename Word, int WAVES, int LPT, int NC = 0>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(sizeof(Word) == 4 ? 6 : 1)))
void k_synth_many(const ReqDesc *__restrict__ descs, ReqCtl *ctls, unsigned nReq, uint32_t *global)
{
  constexpr int TPW = 64 / LPT;
  constexpr int G0 = 2, GL = 4;
  __shared__ uint16_t sT[WAVES][512];
  __shared__ uint32_t sTm[WAVES][512];
  __shared__ int2 sOff[WAVES][TPW][64];
  __shared__ Word sPix[WAVES][TPW][64];
  __shared__ int2 sSrc[WAVES][TPW][64];
  __shared__ VisitTally sTally[WAVES];
  const int wave = threadIdx.x >> 6;
  const unsigned lane = lane_id();
  const unsigned seg = lane / LPT;
  unsigned *active = global + MG_ACTIVE, *expired = global + MG_EXPIRED;
  unsigned r = (blockIdx.x * WAVES + wave) % nReq;
  const Tables *loaded = nullptr;
  unsigned idleRound = 0, idleTotal = 0;
  for (;;) {
    r = __builtin_amdgcn_readfirstlane(r);
    const ReqDesc &d = descs[r];
    ReqCtl &c = ctls[r];
    unsigned long long t64 = 0;
    bool got = false;
    if (lane == 0) {
      const unsigned long long cur = load_agent64(&c.ctr);
      unsigned p = (unsigned)(cur >> 32);
      // admission: take tickets of this request only while they are within reach of its
      // window (a ticket far ahead would only spin); otherwise serve another request
      if (p < 6 && (unsigned)cur < d.sizes[p] &&
          (unsigned)cur <= (unsigned)load_agent64(&c.pass[p].word) + many_window(d, p, (unsigned)cur) + kAdmitSlack) {
        t64 = atomicAdd(&c.ctr, (unsigned long long)kTicketGroup);
        p = (unsigned)(t64 >> 32);
        got = p < 6 && (unsigned)t64 < d.sizes[p];
      }
    }
    if ((__ballot(got) & 1ull) == 0) {
      // no ticket here: leave when every request has finished, else try the next one
      unsigned act = 0;
      if (lane == 0) act = load_agent(active);
      if (__builtin_amdgcn_readfirstlane(act) == 0) break;
      r = (r + 1 == nReq) ? 0u : r + 1;
      if (++idleRound >= nReq) {
        idleRound = 0;
        __builtin_amdgcn_s_sleep(8);
      }
      if (++idleTotal > kIdleLimit) {
        if (lane == 0) atomicOr(expired, 1u);
        break;
      }
      continue;
    }
    idleRound = 0;
    idleTotal = 0;
    t64 = first_lane64(t64);
    const unsigned pass = (unsigned)(t64 >> 32), tk0 = (unsigned)t64;
    const unsigned nGot = min(kTicketGroup, d.sizes[pass] - tk0);
    const unsigned W = many_window(d, pass, tk0);
    if (tk0 > W) {
      const unsigned need = tk0 - W;
      if (lane == 0) {
        unsigned seen = (unsigned)load_agent64(&c.pass[pass].word);
        unsigned spins = 0;
        while (seen < need) {
          if (++spins > kSpinLimit || load_agent(expired)) {
            atomicOr(expired, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(2);
          seen = (unsigned)load_agent64(&c.pass[pass].word);
        }
      }
    }
    if (d.b.tables != loaded) {
      for (int i = lane; i < 512; i += 64) {
        sT[wave][i] = d.b.tables->T[i];
        sTm[wave][i] = d.b.tables->Tm[i];
      }
      wave_sync();
      loaded = d.b.tables;
    }
    Geometry g = d.g;
    g.skipEmpty = (pass == 0 && d.g.skipEmpty) ? 1 : 0;
    if (lane == 0) sTally[wave] = VisitTally{};
    wave_sync();
    for (unsigned j = 0; j < nGot; j += TPW)
      visit_seg<Word, LPT, G0, GL, true, NC>(g, d.b, tk0 + j, tk0 + nGot, d.seed, pass, sOff[wave][seg],
                                             sPix[wave][seg], sSrc[wave][seg], sT[wave], sTm[wave], &sTally[wave]);
    wave_sync();
    const unsigned long long ev = sTally[wave].eval, tm = sTally[wave].term;
    const unsigned bet = sTally[wave].bet;
    if (lane == 0) {
      const unsigned slot = tk0 & (CNT_SLOTS - 1);
      if (ev) atomicAdd(&d.b.counters->candidates[slot], ev);
      if (tm) atomicAdd(&d.b.counters->terms[slot], tm);
      const unsigned long long old = atomicAdd(&c.pass[pass].word, ((unsigned long long)bet << 32) | nGot);
      if ((unsigned)old + nGot == d.sizes[pass]) {
        // the pass is complete: lib/refiner.h:111-115 decides whether another follows
        const unsigned betters = (unsigned)(old >> 32) + bet;
        const bool capped = d.maxPasses && pass + 1 >= d.maxPasses;
        const bool more = pass < 5 && d.sizes[pass + 1] > 0 && !capped &&
                          !((double)((float)betters / (float)d.g.nT) < 0.1);
        if (more) {
          atomicExch(&c.ctr, (unsigned long long)(pass + 1) << 32);
        } else {
          // an empty next pass counts as run, as in the host loop
          c.passesRun = pass + 1 + ((pass < 5 && d.sizes[pass + 1] == 0 && !capped) ? 1 : 0);
          atomicExch(&c.ctr, 6ull << 32);
          atomicSub(active, 1u);
        }
      }
    }
  }
}

bool many_config_valid(const ManyConfig &c)
{
  return (c.waves == 2 || c.waves == 4 || c.waves == 8) && c.firstGroup == 2 && c.laterGroup == 4 &&
         (c.lanes == 16 || c.lanes == 32 || c.lanes == 64);
}

// The kernel of a shape and its resident grid (workgroups per CU from the occupancy query).
struct ManyKernel {
  const void *fn;
  unsigned grid, block;
};

template <typename Word, int WAVES, int LPT, int NC>
static ManyKernel many_kernel_nc()
{
  static const unsigned resident = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024u;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024u;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_synth_many<Word, WAVES, LPT, NC>, 64 * WAVES, 0) !=
            hipSuccess || per < 1)
      per = 1;
    return (unsigned)(cus * per);
  }();
  return {(const void *)k_synth_many<Word, WAVES, LPT, NC>, resident, 64u * WAVES};
}

// RGB without maps (32-bit words) in every request of the launch: the branch-free scorer.
template <typename Word, int WAVES, int LPT>
static ManyKernel many_kernel_t(bool rgb)
{
  if (sizeof(Word) == 4 && rgb) return many_kernel_nc<Word, WAVES, LPT, 3>();
  return many_kernel_nc<Word, WAVES, LPT, 0>();
}

template <typename Word, int WAVES>
static ManyKernel many_kernel_w(const ManyConfig &c, bool rgb)
{
  switch (c.lanes) {
  case 16: return many_kernel_t<Word, WAVES, 16>(rgb);
  case 32: return many_kernel_t<Word, WAVES, 32>(rgb);
  default: return many_kernel_t<Word, WAVES, 64>(rgb);
  }
}

template <typename Word>
static ManyKernel many_kernel_word(const ManyConfig &c, bool rgb)
{
  switch (c.waves) {
  case 2: return many_kernel_w<Word, 2>(c, rgb);
  case 8: return many_kernel_w<Word, 8>(c, rgb);
  default: return many_kernel_w<Word, 4>(c, rgb);
  }
}

// Clears the requests' control words and the launch's global words; global[MG_ACTIVE] =
// nReq.  (A kernel, not a memset node: fills in graphs here are kernels.)
__global__ void k_many_init(uint32_t *ctl, size_t ctlWords, uint32_t *global, unsigned nReq)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ctlWords; i += (size_t)gridDim.x * blockDim.x)
    ctl[i] = 0;
  if (blockIdx.x == 0)
    for (unsigned i = threadIdx.x; i < MG_WORDS; i += blockDim.x) global[i] = i == MG_ACTIVE ? nReq : 0u;
}

void many_graph_release(ManyGraph &G)
{
  if (G.exec) (void)hipGraphExecDestroy(G.exec);
  if (G.graph) (void)hipGraphDestroy(G.graph);
  G = ManyGraph{};
}

// The multi-request graph, two kernel nodes (init, then the persistent launch), built with
// the graph API so that later calls with the same kernel shape update the nodes'
// parameters (request count, buffers) in the instantiated graph instead of capturing and
// instantiating it again.
hipError_t many_graph_prepare(ManyGraph &G, bool wide, const ManyConfig &cfg, bool allRgb, const ReqDesc *descs,
                              ReqCtl *ctls, unsigned nReq, uint32_t *global, bool &instantiated)
{
  instantiated = false;
  if (!nReq || !many_config_valid(cfg)) return hipErrorInvalidValue;
  const bool rgb = allRgb && rgb_fast_path();
  const ManyKernel K = wide ? many_kernel_word<uint64_t>(cfg, rgb) : many_kernel_word<uint32_t>(cfg, rgb);
  uint32_t *ctlWords = (uint32_t *)ctls;
  size_t nCtl = (size_t)nReq * sizeof(ReqCtl) / sizeof(uint32_t);
  const ReqDesc *d = descs;
  ReqCtl *c = ctls;
  unsigned n = nReq;
  uint32_t *gl = global;
  void *initArgs[] = {&ctlWords, &nCtl, &gl, &n};
  void *manyArgs[] = {&d, &c, &n, &gl};
  hipKernelNodeParams pi = {};
  pi.func = (void *)k_many_init;
  pi.gridDim = dim3((unsigned)std::min<size_t>(1024, (nCtl + 255) / 256));
  pi.blockDim = dim3(256);
  pi.kernelParams = initArgs;
  hipKernelNodeParams pm = {};
  pm.func = const_cast<void *>(K.fn);
  pm.gridDim = dim3(K.grid);
  pm.blockDim = dim3(K.block);
  pm.kernelParams = manyArgs;
  if (G.exec && G.fn == K.fn) {
    hipError_t e = hipGraphExecKernelNodeSetParams(G.exec, G.init, &pi);
    if (e == hipSuccess) e = hipGraphExecKernelNodeSetParams(G.exec, G.many, &pm);
    if (e == hipSuccess) return hipSuccess;
    (void)hipGetLastError();   // build it again below
  }
  many_graph_release(G);
  hipError_t e = hipGraphCreate(&G.graph, 0);
  if (e == hipSuccess) e = hipGraphAddKernelNode(&G.init, G.graph, nullptr, 0, &pi);
  if (e == hipSuccess) e = hipGraphAddKernelNode(&G.many, G.graph, &G.init, 1, &pm);
  if (e == hipSuccess) e = hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    many_graph_release(G);
    return e;
  }
  G.fn = K.fn;
  instantiated = true;
  return hipSuccess;
}

template <typename Word, int GL>
static hipError_t launch_window_t(hipStream_t s, const Geometry &g, const SynthBuffers &b, unsigned k0, unsigned k1,
                                  unsigned seed, unsigned pass, unsigned minW, unsigned rampPermille, unsigned cap)
{
  constexpr int WAVES = 4;
  constexpr int G0 = 2;
  // one resident grid: blocks beyond what the device holds at once would only start
  // after the tickets have run out
  static int resident = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_synth_window<Word, WAVES, G0, GL>, 64 * WAVES, 0) !=
            hipSuccess || per < 1)
      per = 1;
    return cus * per;
  }();
  unsigned blocks = (k1 - k0 + WAVES - 1) / WAVES;
  if (blocks > (unsigned)resident) blocks = (unsigned)resident;
  hipLaunchKernelGGL((k_synth_window<Word, WAVES, G0, GL>), dim3(blocks), dim3(64 * WAVES), 0, s, g, b, k0, k1, seed,
                     pass, minW, rampPermille, cap);
  return hipGetLastError();
}

hipError_t launch_synth_window(hipStream_t s, bool wide, const Geometry &g, const SynthBuffers &b, unsigned k0,
                               unsigned k1, unsigned seed, unsigned pass, unsigned minW, unsigned rampPermille,
                               unsigned cap)
{
  if (k1 <= k0) return hipSuccess;
  if (!b.ctl) return hipErrorInvalidValue;
  static const int later = [] {
    const char *e = std::getenv("RESYNTH_LATER_GROUP");
    return (e && std::atoi(e) == 8) ? 8 : 4;
  }();
  hipError_t e = launch_fill_u32(s, b.ctl, 0u, WCTL_WORDS);
  if (e != hipSuccess) return e;
  if (wide)
    return later == 8 ? launch_window_t<uint64_t, 8>(s, g, b, k0, k1, seed, pass, minW, rampPermille, cap)
                      : launch_window_t<uint64_t, 4>(s, g, b, k0, k1, seed, pass, minW, rampPermille, cap);
  return later == 8 ? launch_window_t<uint32_t, 8>(s, g, b, k0, k1, seed, pass, minW, rampPermille, cap)
                    : launch_window_t<uint32_t, 4>(s, g, b, k0, k1, seed, pass, minW, rampPermille, cap);
}

// Serial visits: one workgroup visits targets k0..k1-1 in order (lib/refiner.h:75-100),
// drawing probes from the exact PRNG stream.
template <typename Word, int WAVES, int NC = 0>
__global__ __launch_bounds__(64 * WAVES) void k_synth_serial(Geometry g, SynthBuffers b, unsigned k0, unsigned k1,
                                                             int prngKind, const uint32_t *__restrict__ stream,
                                                             uint32_t streamLen, StreamState *state)
{
  __shared__ VisitLDS<Word, WAVES> S;
  load_tables<Word, WAVES>(b, S);
  uint32_t maxvalue;   // of g_rand_int_range(0, nC), GLib >= 2.2
  if (g.nC <= 0x80000000u) {
    uint32_t leftover = (0x80000000u % g.nC) * 2u;
    if (leftover >= g.nC) leftover -= g.nC;
    maxvalue = 0xFFFFFFFFu - leftover;
  } else {
    maxvalue = g.nC - 1u;
  }
  if (threadIdx.x == 0) S.streamPos = state->pos;
  __syncthreads();
  unsigned long long ev = 0, tm = 0;
  unsigned bt = 0;
  unsigned k = k0;
  bool aborted = false;
  for (; k < k1; k++) {
    const uint32_t posBefore = S.streamPos;
    __syncthreads();
    visit_wg<Word, WAVES, true, NC>(g, b, S, k, 0, 0, prngKind, stream, streamLen, maxvalue, ev, tm, bt,
                                    g.skipEmpty ? k : 0u);
    __syncthreads();
    if (S.aborted) {
      __syncthreads();
      if (threadIdx.x == 0) S.streamPos = posBefore;
      aborted = true;
      __syncthreads();
      break;
    }
  }
  if (threadIdx.x == 0) {
    state->pos = S.streamPos;
    state->kDone = k;
    state->aborted = aborted ? 1u : 0u;
    b.counters->candidates[0] += ev;
    b.counters->terms[0] += tm;
    b.counters->betters[0] += bt;
  }
}

hipError_t launch_synth_serial(hipStream_t s, bool wide, const Geometry &g, const SynthBuffers &b, unsigned k0,
                               unsigned k1, int prngKind, const uint32_t *stream, uint32_t streamLen,
                               StreamState *state)
{
  constexpr int WAVES = 4;
  if (wide)
    hipLaunchKernelGGL((k_synth_serial<uint64_t, WAVES>), dim3(1), dim3(64 * WAVES), 0, s, g, b, k0, k1, prngKind,
                       stream, streamLen, state);
  else if (g.nColors == 3 && g.nMaps == 0 && rgb_fast_path())
    hipLaunchKernelGGL((k_synth_serial<uint32_t, WAVES, 3>), dim3(1), dim3(64 * WAVES), 0, s, g, b, k0, k1, prngKind,
                       stream, streamLen, state);
  else
    hipLaunchKernelGGL((k_synth_serial<uint32_t, WAVES>), dim3(1), dim3(64 * WAVES), 0, s, g, b, k0, k1, prngKind,
                       stream, streamLen, state);
  return hipGetLastError();
}

// ------------------------------------------------------------------ adaptation

// SimpleAPI adaptation (lib/adaptSimple.h:226-307 + lib/engine.c:338-431): the target word
// (hasValue = use_context && unselected && opaque), the corpus word (mask byte 0xFF <=>
// selected in the corpus mask) and the point flags of every pixel.  A thread takes a column
// of four rows: their corpus words are adjacent in the tiled layout and go out in one store
// (rows past the image get zero words, in the layout's padding).
__device__ __forceinline__ void store4(uint32_t *dst, const uint32_t (&v)[4])
{
  *(uint4 *)dst = make_uint4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void store4(uint64_t *dst, const uint64_t (&v)[4])
{
  ((ulonglong2 *)dst)[0] = make_ulonglong2(v[0], v[1]);
  ((ulonglong2 *)dst)[1] = make_ulonglong2(v[2], v[3]);
}

template <typename Word>
__global__ void k_adapt_simple(const unsigned char *__restrict__ image, size_t rowBytes, unsigned bpp,
                               const unsigned char *__restrict__ mask, size_t maskRowBytes,
                               const unsigned char *__restrict__ mask2, size_t mask2RowBytes, int w, int h,
                               int nColors, int alphaIdx, int useContext, Word *tword, Word *cword, uint8_t *tflag,
                               uint8_t *cflag)
{
  const unsigned quads = ((unsigned)h + 3) / 4;
  for (unsigned q = blockIdx.y; q < quads; q += gridDim.y)
    for (unsigned x = blockIdx.x * blockDim.x + threadIdx.x; x < (unsigned)w; x += gridDim.x * blockDim.x) {
      Word c4[4];
      for (int r = 0; r < 4; r++) {
        const unsigned y = q * 4 + r;
        c4[r] = 0;
        if (y >= (unsigned)h) continue;
        const size_t p = (size_t)y * (unsigned)w + x;
        const unsigned char *px = image + y * rowBytes + (size_t)x * bpp;
        Word chans = 0;
        for (int c = 0; c < nColors; c++) chans |= (Word)px[c] << (8 * (1 + c));
        const bool opaque = alphaIdx < 0 ? true : px[alphaIdx] != 0;
        const unsigned char m = mask[y * maskRowBytes + x];
        const unsigned char cm = mask2 ? mask2[y * mask2RowBytes + x] : (unsigned char)~m;
        const bool hasValue = useContext && m == 0 && opaque;
        tword[p] = chans | (Word)(hasValue ? 1u : 0u);
        c4[r] = chans | (Word)(cm == 0xFF ? (opaque ? 0xFFu : 0xFEu) : 0u);
        tflag[p] = m != 0;
        cflag[p] = (cm == 0xFF) && opaque;
      }
      store4(cword + cidx(w, x, q * 4), c4);
    }
}

hipError_t launch_adapt_simple(hipStream_t s, bool wide, const unsigned char *image, size_t rowBytes, unsigned bpp,
                               const unsigned char *mask, size_t maskRowBytes, const unsigned char *mask2,
                               size_t mask2RowBytes, int w, int h, int nColors, int alphaIdx, int useContext,
                               void *tword, void *cword, uint8_t *tflag, uint8_t *cflag)
{
  const dim3 blocks(std::max(1u, std::min(64u, ((unsigned)w + 255) / 256)),
                    std::max(1u, std::min(65535u, ((unsigned)h + 3) / 4)));
  if (wide)
    hipLaunchKernelGGL((k_adapt_simple<uint64_t>), dim3(blocks), dim3(256), 0, s, image, rowBytes, bpp, mask,
                       maskRowBytes, mask2, mask2RowBytes, w, h, nColors, alphaIdx, useContext, (uint64_t *)tword,
                       (uint64_t *)cword, tflag, cflag);
  else
    hipLaunchKernelGGL((k_adapt_simple<uint32_t>), dim3(blocks), dim3(256), 0, s, image, rowBytes, bpp, mask,
                       maskRowBytes, mask2, mask2RowBytes, w, h, nColors, alphaIdx, useContext, (uint32_t *)tword,
                       (uint32_t *)cword, tflag, cflag);
  return hipGetLastError();
}

// Engine-layout adaptation: pixmap [mask][colours][alpha?][maps] -> words and flags.
// Target (isTarget): flag = mask != 0, word bit 0 = hasValue.  Corpus: flag = selected
// and opaque, word byte 0 = 0xFF when selected.
template <typename Word>
__global__ void k_adapt_maps(const unsigned char *__restrict__ pix, int w, int h, int depth, int colorEnd,
                             int mapStart, int mapEnd, int alphaIdx, int useContext, int isTarget, Word *word,
                             uint8_t *flag)
{
  size_t n = (size_t)w * h;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    const unsigned char *px = pix + p * depth;
    Word chans = 0;
    int c = 0;
    for (int j = 1; j < colorEnd; j++, c++) chans |= (Word)px[j] << (8 * (1 + c));
    for (int j = mapStart; j < mapEnd; j++, c++) chans |= (Word)px[j] << (8 * (1 + c));
    bool opaque = alphaIdx < 0 ? true : px[alphaIdx] != 0;
    unsigned char m = px[0];
    if (isTarget) {
      bool hasValue = useContext && m == 0 && opaque;
      word[p] = chans | (Word)(hasValue ? 1u : 0u);
      flag[p] = m != 0;
    } else {
      word[cidx(w, (unsigned)(p % (unsigned)w), (unsigned)(p / (unsigned)w))] =
          chans | (Word)(m == 0xFF ? (opaque ? 0xFFu : 0xFEu) : 0u);
      flag[p] = (m == 0xFF) && opaque;
    }
  }
}

hipError_t launch_adapt_maps(hipStream_t s, bool wide, const unsigned char *pix, int w, int h, int depth,
                             int colorEnd, int mapStart, int mapEnd, int alphaIdx, int useContext, int isTarget,
                             void *word, uint8_t *flag)
{
  size_t n = (size_t)w * h;
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  if (blocks == 0) blocks = 1;
  if (wide)
    hipLaunchKernelGGL((k_adapt_maps<uint64_t>), dim3(blocks), dim3(256), 0, s, pix, w, h, depth, colorEnd, mapStart,
                       mapEnd, alphaIdx, useContext, isTarget, (uint64_t *)word, flag);
  else
    hipLaunchKernelGGL((k_adapt_maps<uint32_t>), dim3(blocks), dim3(256), 0, s, pix, w, h, depth, colorEnd, mapStart,
                       mapEnd, alphaIdx, useContext, isTarget, (uint32_t *)word, flag);
  return hipGetLastError();
}

// Packs engine-layout pixels (frozen-state scorer input) into words.
// Corpus words (corpusFlags) go to the tiled corpus layout of a w-wide corpus.
template <typename Word>
__global__ void k_pack_words(const unsigned char *pix, int n, int w, int depth, int colorEnd, int mapStart, int mapEnd,
                             int corpusFlags, Word *out)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned char *px = pix + (size_t)p * depth;
  Word chans = 0;
  int c = 0;
  for (int j = 1; j < colorEnd; j++, c++) chans |= (Word)px[j] << (8 * (1 + c));
  for (int j = mapStart; j < mapEnd; j++, c++) chans |= (Word)px[j] << (8 * (1 + c));
  const Word v = chans | (Word)(corpusFlags ? (px[0] == 0xFF ? 0xFFu : 0u) : (px[0] & 1u));
  if (corpusFlags) out[cidx(w, (unsigned)p % (unsigned)w, (unsigned)p / (unsigned)w)] = v;
  else out[p] = v;
}

hipError_t launch_pack_words(hipStream_t s, bool wide, const unsigned char *pix, int n, int w, int depth,
                             int colorEnd, int mapStart, int mapEnd, int corpusFlags, void *out)
{
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (blocks == 0) return hipSuccess;
  if (wide)
    hipLaunchKernelGGL((k_pack_words<uint64_t>), dim3(blocks), dim3(256), 0, s, pix, n, w, depth, colorEnd, mapStart,
                       mapEnd, corpusFlags, (uint64_t *)out);
  else
    hipLaunchKernelGGL((k_pack_words<uint32_t>), dim3(blocks), dim3(256), 0, s, pix, n, w, depth, colorEnd, mapStart,
                       mapEnd, corpusFlags, (uint32_t *)out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ compaction

// Row-major list of the flagged pixels (lib/engine.c:338-431 build these lists by a
// row-major scan): per-tile counts, one exclusive scan of the tile counts, scatter.
constexpr int CT_THREADS = 256;
constexpr int CT_ITEMS = 16;
constexpr int CT_TILE = CT_THREADS * CT_ITEMS;

// A thread's CT_ITEMS (16) flags as 0/1 bytes: one 16-byte load where the run is whole and
// aligned, byte loads at the end of the buffer.
__device__ __forceinline__ void load_flags(const uint8_t *flags, size_t n, size_t base, uint8_t (&f)[CT_ITEMS])
{
  if (base + CT_ITEMS <= n && ((uintptr_t)(flags + base) & 15u) == 0) {
    const uint4 v = *(const uint4 *)(flags + base);
    const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
    for (int q = 0; q < 4; q++)
      for (int b = 0; b < 4; b++) f[q * 4 + b] = ((w4[q] >> (8 * b)) & 0xFFu) != 0;
  } else {
    for (int i = 0; i < CT_ITEMS; i++) f[i] = (base + i < n) ? (flags[base + i] != 0) : 0;
  }
}

__global__ __launch_bounds__(CT_THREADS) void k_tile_counts(const uint8_t *flags, size_t n, unsigned *tileCount)
{
  __shared__ unsigned red[CT_THREADS / 64];
  size_t base = (size_t)blockIdx.x * CT_TILE + (size_t)threadIdx.x * CT_ITEMS;
  uint8_t f[CT_ITEMS];
  load_flags(flags, n, base, f);
  unsigned c = 0;
  for (int i = 0; i < CT_ITEMS; i++) c += f[i];
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (lane_id() == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned s = 0;
    for (int w = 0; w < CT_THREADS / 64; w++) s += red[w];
    tileCount[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(1024) void k_scan_tiles(unsigned *tileCount, unsigned nTiles, unsigned *total)
{
  __shared__ unsigned carry;
  __shared__ unsigned waveSum[16];
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (unsigned base = 0; base < nTiles; base += 1024) {
    unsigned i = base + threadIdx.x;
    unsigned v = i < nTiles ? tileCount[i] : 0;
    // inclusive wave scan
    unsigned x = v;
    for (int o = 1; o < 64; o <<= 1) {
      unsigned y = __shfl_up(x, o);
      if ((int)lane_id() >= o) x += y;
    }
    if (lane_id() == 63) waveSum[threadIdx.x >> 6] = x;
    __syncthreads();
    unsigned waveOff = 0;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) waveOff += waveSum[w];
    unsigned excl = carry + waveOff + x - v;
    if (i < nTiles) tileCount[i] = excl;
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// The tile's points are gathered in LDS in order, then written out by consecutive threads
// (a thread's own points would go out as 64 scattered runs per store).
__global__ __launch_bounds__(CT_THREADS) void k_tile_scatter(const uint8_t *flags, size_t n, int w,
                                                             const unsigned *tileOffset, int2 *out)
{
  __shared__ unsigned waveSum[CT_THREADS / 64];
  __shared__ int2 stage[CT_TILE];
  size_t base = (size_t)blockIdx.x * CT_TILE + (size_t)threadIdx.x * CT_ITEMS;
  uint8_t f[CT_ITEMS];
  load_flags(flags, n, base, f);
  unsigned c = 0;
  for (int i = 0; i < CT_ITEMS; i++) c += f[i];
  unsigned x = c;
  for (int o = 1; o < 64; o <<= 1) {
    unsigned y = __shfl_up(x, o);
    if ((int)lane_id() >= o) x += y;
  }
  if (lane_id() == 63) waveSum[threadIdx.x >> 6] = x;
  __syncthreads();
  unsigned off = 0, total = 0;
  for (int wv = 0; wv < CT_THREADS / 64; wv++) {
    if (wv < (int)(threadIdx.x >> 6)) off += waveSum[wv];
    total += waveSum[wv];
  }
  off += x - c;
  if (c) {
    // the run's first pixel, then step along the row (one division per thread)
    unsigned px = (unsigned)(base % (unsigned)w), py = (unsigned)(base / (unsigned)w);
    for (int i = 0; i < CT_ITEMS; i++) {
      if (f[i]) stage[off++] = make_int2((int)px, (int)py);
      if (++px == (unsigned)w) {
        px = 0;
        py++;
      }
    }
  }
  __syncthreads();
  int2 *dst = out + tileOffset[blockIdx.x];
  for (unsigned i = threadIdx.x; i < total; i += CT_THREADS) dst[i] = stage[i];
}

size_t compact_temp_bytes(size_t n) { return (((n + CT_TILE - 1) / CT_TILE) + 64) * sizeof(unsigned); }

hipError_t launch_compact(hipStream_t s, const uint8_t *flags, int w, int h, int2 *out, unsigned *count, void *temp)
{
  size_t n = (size_t)w * h;
  unsigned tiles = (unsigned)((n + CT_TILE - 1) / CT_TILE);
  if (tiles == 0) tiles = 1;
  unsigned *tc = (unsigned *)temp;
  hipLaunchKernelGGL(k_tile_counts, dim3(tiles), dim3(CT_THREADS), 0, s, flags, n, tc);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, tc, tiles, count);
  hipLaunchKernelGGL(k_tile_scatter, dim3(tiles), dim3(CT_THREADS), 0, s, flags, n, w, tc, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ sorted offsets

// lib/engine.c:465-497 sorts the row-major offsets of [-w+1,w-1]x[-h+1,h-1] by x^2+y^2
// with a comparator that never returns 0 under GLib's merge sort, which orders equal
// distances by descending row-major index.  The same order is a stable LSD radix sort
// by x^2+y^2 of the offsets listed in descending row-major order.
// The sorted offsets (lib/engine.c:465-497): every (x, y) in [-w+1, w-1] x [-h+1, h-1]
// by x^2 + y^2 ascending, ties in descending row-major index (y descending, then x
// descending), the order GLib's merge sort gives the reference's never-equal comparator.
// The table is symmetric under x -> -x and y -> -y, so only the quadrant x, y >= 0 is
// sorted (a quarter of the entries); each quadrant entry (a, b) then writes its 1, 2 or 4
// signed offsets straight to their places.  Within a ring of equal x^2 + y^2 the quadrant
// entries come in descending b, and the ring's offsets are:
//   y = +b rows in that order, each (a, b) then (-a, b);
//   the y = 0 row, (a, 0) then (-a, 0);
//   y = -b rows in the reverse order, each (a, -b) then (-a, -b).

// Quadrant entry q = b*qw + a of the quadrant's [0, qw) x [0, qh) corner, generated in
// descending q so that a stable sort by the key leaves ties in descending q, i.e. descending
// b.  The sorted payload is (b << shift) | a (shift = bits of qw - 1; it fits, as w*h <
// 2^31), so the later kernels need no division.  Keys above lim2 (a table cut at x^2 + y^2
// <= lim2) become lim2 + 1 and sort after every kept entry.
__global__ void k_quad_keys(int qw, unsigned shift, uint32_t lim2, uint32_t *keys, uint32_t *vals, size_t nQ)
{
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nQ; e += (size_t)gridDim.x * blockDim.x) {
    const uint32_t q = (uint32_t)(nQ - 1 - e);
    const uint32_t a = q % (unsigned)qw, b = q / (unsigned)qw;
    const uint32_t d2 = a * a + b * b;
    keys[e] = d2 > lim2 ? lim2 + 1u : d2;
    vals[e] = (b << shift) | a;
  }
}

// Offsets a sorted quadrant entry stands for.
__global__ void k_quad_members(unsigned shift, const uint32_t *vals, uint32_t *mem, size_t nQ)
{
  const uint32_t amask = (1u << shift) - 1u;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nQ; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t v = vals[i];
    mem[i] = ((v & amask) > 0 ? 2u : 1u) * ((v >> shift) > 0 ? 2u : 1u);
  }
}

// first: exclusive prefix sums of the member counts, so that first[lo] is where the ring
// of entries lo.. begins in the full table.
__global__ void k_quad_expand(unsigned shift, const uint32_t *keys, const uint32_t *vals, const uint32_t *first,
                              size_t nQ, int2 *out)
{
  const uint32_t amask = (1u << shift) - 1u;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nQ; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t D = keys[i];
    size_t lo = i, hi = i + 1;
    while (lo > 0 && keys[lo - 1] == D) lo--;
    while (hi < nQ && keys[hi] == D) hi++;
    uint32_t before = 0, rowsAbove = 0, axis = 0;   // y > 0 offsets before entry i; all of them; the y = 0 ones
    for (size_t j = lo; j < hi; j++) {
      const uint32_t vj = vals[j];
      const uint32_t m = (vj & amask) > 0 ? 2u : 1u;
      if ((vj >> shift) > 0) {
        rowsAbove += m;
        if (j < i) before += m;
      } else {
        axis = m;
      }
    }
    const uint32_t v = vals[i];
    const int a = (int)(v & amask), b = (int)(v >> shift);
    const uint32_t m = a > 0 ? 2u : 1u;
    const size_t start = first[lo];
    if (b > 0) {
      const size_t up = start + before, down = start + rowsAbove + axis + (rowsAbove - before - m);
      out[up] = make_int2(a, b);
      out[down] = make_int2(a, -b);
      if (a > 0) {
        out[up + 1] = make_int2(-a, b);
        out[down + 1] = make_int2(-a, -b);
      }
    } else {
      const size_t mid = start + rowsAbove;
      out[mid] = make_int2(a, 0);
      if (a > 0) out[mid + 1] = make_int2(-a, 0);
    }
  }
}

static uint64_t isqrt_u64(uint64_t v)
{
  uint64_t r = (uint64_t)std::sqrt((double)v);
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

// The sorted offsets of [-w+1, w-1] x [-h+1, h-1] with x^2 + y^2 <= maxD2: the table's first
// nO entries (the sort key is x^2 + y^2), built from the quadrant's [0, qw) x [0, qh) corner
// and its nKeep entries inside the disc.  maxD2 at or above (w-1)^2 + (h-1)^2 is the whole
// table.
OffsetPlan plan_offsets(int w, int h, uint64_t maxD2)
{
  OffsetPlan p;
  p.w = w;
  p.h = h;
  const uint64_t far2 = (uint64_t)(w - 1) * (w - 1) + (uint64_t)(h - 1) * (h - 1);
  if (maxD2 >= far2) {
    p.qw = w;
    p.qh = h;
    p.lim2 = (uint32_t)far2;
    p.nQ = p.nKeep = (size_t)w * h;
    p.nO = (size_t)(2 * w - 1) * (size_t)(2 * h - 1);
    p.whole = true;
    return p;
  }
  const uint64_t R = isqrt_u64(maxD2);
  p.qw = (int)std::min<uint64_t>((uint64_t)w, R + 1);
  p.qh = (int)std::min<uint64_t>((uint64_t)h, R + 1);
  p.lim2 = (uint32_t)maxD2;
  p.nQ = (size_t)p.qw * p.qh;
  p.nKeep = 0;
  p.nO = 0;
  for (int a = 0; a < p.qw; a++) {
    const uint64_t bmax = std::min<uint64_t>((uint64_t)(h - 1), isqrt_u64(maxD2 - (uint64_t)a * a));
    p.nKeep += bmax + 1;
    p.nO += (a ? 2 : 1) * (2 * bmax + 1);
  }
  p.whole = false;
  return p;
}

static unsigned bits_of(uint64_t mx)
{
  unsigned b = 1;
  while (b < 32 && (mx >> b)) b++;
  return b;
}

static size_t quad_scan_bytes(size_t nQ)
{
  size_t scanBytes = 0;
  (void)rocprim::exclusive_scan(nullptr, scanBytes, (uint32_t *)nullptr, (uint32_t *)nullptr, 0u, nQ,
                                rocprim::plus<uint32_t>());
  return scanBytes;
}

// The five u32 arrays and rocPRIM's work area each start on a 256-byte boundary of the
// (256-aligned) temporary buffer: the multi-block sort and scan keep 64-bit look-back
// state in the work area and need it aligned, whatever w*h is.
static size_t quad_array_bytes(size_t nQ) { return (nQ * sizeof(uint32_t) + 255) & ~(size_t)255; }

size_t offsets_temp_bytes(const OffsetPlan &p)
{
  size_t sortBytes = 0;
  (void)rocprim::radix_sort_pairs<rocprim::default_config>(nullptr, sortBytes, (uint32_t *)nullptr,
                                                           (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                           (uint32_t *)nullptr, p.nQ, 0, 32);
  return 5 * quad_array_bytes(p.nQ) + std::max(sortBytes, quad_scan_bytes(p.nKeep)) + 1024;
}

hipError_t launch_sorted_offsets(hipStream_t s, const OffsetPlan &p, int2 *out, void *temp, size_t tempBytes)
{
  const size_t nQ = p.nQ, A = quad_array_bytes(nQ);
  char *base = (char *)temp;
  uint32_t *k0 = (uint32_t *)base, *v0 = (uint32_t *)(base + A), *k1 = (uint32_t *)(base + 2 * A),
           *v1 = (uint32_t *)(base + 3 * A), *first = (uint32_t *)(base + 4 * A);
  char *work = base + 5 * A;
  size_t workBytes = tempBytes - 5 * A;
  unsigned blocks = (unsigned)((nQ + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  unsigned shift = 1;
  while (shift < 31 && ((unsigned)(p.qw - 1) >> shift)) shift++;   // bits of qw - 1 (at least 1)
  hipLaunchKernelGGL(k_quad_keys, dim3(blocks), dim3(256), 0, s, p.qw, shift, p.lim2, k0, v0, nQ);
  hipError_t e = rocprim::radix_sort_pairs<rocprim::default_config>(work, workBytes, k0, k1, v0, v1, nQ, 0,
                                                                    bits_of(p.whole ? p.lim2 : (uint64_t)p.lim2 + 1), s);
  if (e != hipSuccess) return e;
  // the kept entries are the sorted prefix [0, nKeep): whole rings, as a ring's entries share one key
  const size_t nK = p.nKeep;
  unsigned kblocks = (unsigned)std::min<size_t>((nK + 255) / 256, 65536);
  if (!kblocks) return hipGetLastError();
  hipLaunchKernelGGL(k_quad_members, dim3(kblocks), dim3(256), 0, s, shift, v1, k0, nK);
  e = rocprim::exclusive_scan(work, workBytes, k0, first, 0u, nK, rocprim::plus<uint32_t>(), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_quad_expand, dim3(kblocks), dim3(256), 0, s, shift, k1, v1, first, nK, out);
  return hipGetLastError();
}

// Radix sort of (key, value) u32 pairs on the low `bits` bits of the keys.
size_t sort_pairs_temp_bytes(size_t n)
{
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs<rocprim::default_config>(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                           (uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, 32);
  return bytes + 256;
}

hipError_t launch_sort_pairs(hipStream_t s, const uint32_t *kIn, uint32_t *kOut, const uint32_t *vIn, uint32_t *vOut,
                             size_t n, unsigned bits, void *temp, size_t tempBytes)
{
  if (!n) return hipSuccess;
  return rocprim::radix_sort_pairs<rocprim::default_config>(temp, tempBytes, const_cast<uint32_t *>(kIn), kOut,
                                                            const_cast<uint32_t *>(vIn), vOut, n, 0, bits, s);
}

// ------------------------------------------------------------------ small kernels

// First index in the row-major corpus point list of every band's first row.
__global__ void k_band_starts(const int2 *corpusPts, const unsigned *nC, int ch, uint32_t *out)
{
  const int x = (int)threadIdx.x;
  if (x > kBands) return;
  const int row = x * ch / kBands;
  unsigned lo = 0, hi = *nC;
  while (lo < hi) {
    const unsigned mid = (lo + hi) / 2;
    if (corpusPts[mid].y < row) lo = mid + 1;
    else hi = mid;
  }
  out[x] = lo;
}

hipError_t launch_band_starts(hipStream_t s, const int2 *corpusPts, const unsigned *nC, int ch, uint32_t *out)
{
  hipLaunchKernelGGL(k_band_starts, dim3(1), dim3(64), 0, s, corpusPts, nC, ch, out);
  return hipGetLastError();
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, size_t n)
{
  if (!n) return hipSuccess;
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

// Tiles holding a pixel that has a value (target word bit 0).
template <typename Word>
__global__ void k_tile_init(const Word *tword, int tw, int th, uint8_t *tileHas)
{
  const unsigned tilesW = (unsigned)(tw + kTile - 1) / kTile;
  for (unsigned y = blockIdx.y; y < (unsigned)th; y += gridDim.y)
    for (unsigned x = blockIdx.x * blockDim.x + threadIdx.x; x < (unsigned)tw; x += gridDim.x * blockDim.x)
      if (tword[(size_t)y * (unsigned)tw + x] & 1u) tileHas[(size_t)(y / kTile) * tilesW + x / kTile] = 1;
}

hipError_t launch_tile_init(hipStream_t s, bool wide, const void *tword, int tw, int th, uint8_t *tileHas)
{
  if (tw <= 0 || th <= 0) return hipSuccess;
  const dim3 blocks(std::min(64u, ((unsigned)tw + 255) / 256), std::min(65535u, (unsigned)th));
  if (wide) hipLaunchKernelGGL((k_tile_init<uint64_t>), dim3(blocks), dim3(256), 0, s, (const uint64_t *)tword, tw, th, tileHas);
  else hipLaunchKernelGGL((k_tile_init<uint32_t>), dim3(blocks), dim3(256), 0, s, (const uint32_t *)tword, tw, th, tileHas);
  return hipGetLastError();
}

// firstIdx[rho] = #{(x,y) : |x| <= w-1, |y| <= h-1, x^2 + y^2 < rho^2}: the position in the
// sorted offsets where distance rho begins (the sort key is x^2 + y^2).
__global__ __launch_bounds__(256) void k_first_idx(int w, int h, unsigned rhoMax, uint32_t *firstIdx)
{
  __shared__ unsigned long long part[4];
  const unsigned rho = blockIdx.x;
  const long long r2 = (long long)rho * rho;
  unsigned long long cnt = 0;
  for (int x = -(w - 1) + (int)threadIdx.x; x <= w - 1; x += blockDim.x) {
    long long rem = r2 - (long long)x * x - 1;   // y^2 <= rem
    if (rem < 0) continue;
    long long ys = (long long)sqrt((double)rem);
    while (ys * ys > rem) ys--;
    while ((ys + 1) * (ys + 1) <= rem) ys++;
    if (ys > h - 1) ys = h - 1;
    cnt += 2 * ys + 1;
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane_id() == 0) part[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) firstIdx[rho] = (uint32_t)(part[0] + part[1] + part[2] + part[3]);
}

hipError_t launch_first_idx(hipStream_t s, int w, int h, unsigned rhoMax, uint32_t *firstIdx)
{
  hipLaunchKernelGGL(k_first_idx, dim3(rhoMax + 1), dim3(256), 0, s, w, h, rhoMax, firstIdx);
  return hipGetLastError();
}

__global__ void k_gather_targets(const int2 *rm, const uint32_t *perm, int2 *out, unsigned n)
{
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = rm[perm[i]];
}

hipError_t launch_gather_targets(hipStream_t s, const int2 *rm, const uint32_t *perm, int2 *out, unsigned n)
{
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_gather_targets, dim3((n + 255) / 256), dim3(256), 0, s, rm, perm, out, n);
  return hipGetLastError();
}

// The batched mode's random target order: out[i] = rm[perm(i)] for a keyed pseudo-random
// permutation of [0, n), drawn on the device in one pass.  perm is a 4-round Feistel
// network on 2*hb-bit words (2^(2 hb) >= n, at most 4n), each round function a 64-bit
// mixer of (seed, round, half); it is a bijection of the words, and cycle-walking (apply it
// again while the result is n or more) restricts it to a bijection of [0, n) -- the walk
// from i returns below n within its own cycle, about 2-4 steps on average.
__device__ __forceinline__ uint32_t feistel_round(uint32_t seed, uint32_t round, uint32_t half)
{
  uint64_t z = (((uint64_t)seed << 32) | (((uint64_t)round << 28) ^ half)) + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (uint32_t)((z ^ (z >> 31)) >> 32);
}

__global__ void k_random_order(uint32_t seed, unsigned n, unsigned hb, const int2 *rm, int2 *out)
{
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t mask = (1u << hb) - 1u;
  uint32_t x = i;
  do {
    uint32_t L = x >> hb, R = x & mask;
#pragma unroll
    for (uint32_t r = 0; r < 4; r++) {
      const uint32_t t = L ^ (feistel_round(seed, r, R) & mask);
      L = R;
      R = t;
    }
    x = (L << hb) | R;
  } while (x >= n);
  out[i] = rm[x];
}

hipError_t launch_random_order(hipStream_t s, uint32_t seed, unsigned n, const int2 *rm, int2 *out)
{
  if (!n) return hipSuccess;
  unsigned hb = 1;
  while ((1ull << (2 * hb)) < (unsigned long long)n) hb++;
  hipLaunchKernelGGL(k_random_order, dim3((n + 255) / 256), dim3(256), 0, s, seed, n, hb, rm, out);
  return hipGetLastError();
}

// lib/adaptSimple.h:111-168 writes every pixel back; only target pixels can differ, so
// only their colour channels are written (alpha and context are left as they are).
template <typename Word>
__global__ void k_writeback_simple(const Word *tword, const int2 *targets, unsigned nT, unsigned char *image,
                                   size_t rowBytes, unsigned bpp, int nColors, int w)
{
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nT) return;
  int2 t = targets[i];
  Word v = tword[(size_t)t.y * (unsigned)w + (unsigned)t.x];
  unsigned char *px = image + (size_t)t.y * rowBytes + (size_t)t.x * bpp;
  for (int c = 0; c < nColors; c++) px[c] = (unsigned char)chan(v, c);
}

hipError_t launch_writeback_simple(hipStream_t s, bool wide, const void *tword, const int2 *targets, unsigned nT,
                                   unsigned char *image, size_t rowBytes, unsigned bpp, int nColors, int w)
{
  if (!nT) return hipSuccess;
  if (wide)
    hipLaunchKernelGGL((k_writeback_simple<uint64_t>), dim3((nT + 255) / 256), dim3(256), 0, s,
                       (const uint64_t *)tword, targets, nT, image, rowBytes, bpp, nColors, w);
  else
    hipLaunchKernelGGL((k_writeback_simple<uint32_t>), dim3((nT + 255) / 256), dim3(256), 0, s,
                       (const uint32_t *)tword, targets, nT, image, rowBytes, bpp, nColors, w);
  return hipGetLastError();
}

// lib/synthesize.h:403-419 setColor: colour pixelels [1, colorEndBip) of target pixels.
template <typename Word>
__global__ void k_writeback_pixmap(const Word *tword, const int2 *targets, unsigned nT, unsigned char *pix, int depth,
                                   int nColors, int w)
{
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nT) return;
  int2 t = targets[i];
  size_t p = (size_t)t.y * (unsigned)w + (unsigned)t.x;
  Word v = tword[p];
  for (int c = 0; c < nColors; c++) pix[p * depth + 1 + c] = (unsigned char)chan(v, c);
}

hipError_t launch_writeback_pixmap(hipStream_t s, bool wide, const void *tword, const int2 *targets, unsigned nT,
                                   unsigned char *pix, int depth, int nColors, int w)
{
  if (!nT) return hipSuccess;
  if (wide)
    hipLaunchKernelGGL((k_writeback_pixmap<uint64_t>), dim3((nT + 255) / 256), dim3(256), 0, s,
                       (const uint64_t *)tword, targets, nT, pix, depth, nColors, w);
  else
    hipLaunchKernelGGL((k_writeback_pixmap<uint32_t>), dim3((nT + 255) / 256), dim3(256), 0, s,
                       (const uint32_t *)tword, targets, nT, pix, depth, nColors, w);
  return hipGetLastError();
}

// Sum and count of the final match errors over the targets.  Callers pass the targets in
// row-major order, so that the loads of a wavefront fall on a few rows; each workgroup
// reduces in LDS and adds once (one pair of 64-bit atomics per workgroup, not per wavefront).
__global__ __launch_bounds__(256) void k_mean_error(const uint32_t *lastBest, const int2 *targets, unsigned nT, int w,
                                                    unsigned long long *sumCount)
{
  __shared__ unsigned long long part[2][4];
  unsigned long long s = 0, c = 0;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < nT; i += gridDim.x * blockDim.x) {
    int2 t = targets[i];
    uint32_t v = lastBest[(size_t)t.y * (unsigned)w + (unsigned)t.x];
    if (v != 0xFFFFFFFFu) {
      s += v;
      c++;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o);
    c += __shfl_xor(c, o);
  }
  const unsigned wave = threadIdx.x >> 6;
  if (lane_id() == 0) {
    part[0][wave] = s;
    part[1][wave] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned nw = blockDim.x >> 6;
    unsigned long long bs = 0, bc = 0;
    for (unsigned k = 0; k < nw; k++) {
      bs += part[0][k];
      bc += part[1][k];
    }
    if (bc) {
      atomicAdd(&sumCount[0], bs);
      atomicAdd(&sumCount[1], bc);
    }
  }
}

hipError_t launch_mean_error(hipStream_t s, const uint32_t *lastBest, const int2 *targets, unsigned nT, int w,
                             unsigned long long *sumCount)
{
  unsigned blocks = (nT + 1023) / 1024;
  if (blocks > 512) blocks = 512;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_mean_error, dim3(blocks), dim3(256), 0, s, lastBest, targets, nT, w, sumCount);
  return hipGetLastError();
}

// ------------------------------------------------------------------ frozen-state scorer

template <typename Word, int NC = 0>
__global__ __launch_bounds__(64) void k_score_frozen(Geometry g, const Word *cword, const Tables *tables,
                                                     const unsigned *nbCount, const int2 *nbOffset,
                                                     const Word *nbWord, const unsigned *candCount, unsigned maxCands,
                                                     const int2 *cands, int32_t *winner, uint32_t *bestSum,
                                                     int32_t *stopped)
{
  __shared__ uint16_t sT[512];
  __shared__ uint32_t sTm[512];
  __shared__ int2 sOff[64];
  __shared__ Word sPix[64];
  for (int i = threadIdx.x; i < 512; i += blockDim.x) {
    sT[i] = tables->T[i];
    sTm[i] = tables->Tm[i];
  }
  const unsigned prob = blockIdx.x;
  const unsigned lane = lane_id();
  unsigned count = nbCount[prob];
  if (count > 64) count = 64;
  if (lane < count) {
    sOff[lane] = nbOffset[(size_t)prob * 64 + lane];
    sPix[lane] = nbWord[(size_t)prob * 64 + lane];
  }
  __syncthreads();
  unsigned nc = candCount[prob];
  uint32_t best = 0xFFFFFFFFu;
  int winIdx = -1;
  unsigned consumed = nc;
  for (unsigned base = 0; base < nc; base += 64) {
    unsigned j = base + lane;
    bool active = j < nc;
    uint32_t F = 0, M = 0;
    bool alive = false;
    if (active) {
      int2 c = cands[(size_t)prob * maxCands + j];
      score_candidate<Word, 8, 8, NC>(g, cword, c.x, c.y, count, sOff, sPix, sT, sTm, best, F, M, alive);
    }
    int win = -1, stop = 64;
    uint32_t bb = best;
    resolve(active && alive, M, F, bb, win, stop);
    if (win >= 0) {
      best = bb;
      winIdx = (int)(base + win);
    }
    if (stop < 64) {
      consumed = base + stop + 1;
      break;
    }
  }
  if (lane == 0) {
    winner[prob] = winIdx;
    bestSum[prob] = best;
    stopped[prob] = (int32_t)consumed;
  }
}

hipError_t launch_score_frozen(hipStream_t s, bool wide, int cw, int ch, int nColors, int nMaps, uint32_t penalty,
                               const void *cword, const Tables *tables, unsigned nProblems, const unsigned *nbCount,
                               const int2 *nbOffset, const void *nbWord, const unsigned *candCount,
                               unsigned maxCands, const int2 *cands, int32_t *winner, uint32_t *bestSum,
                               int32_t *stopped)
{
  Geometry g = {};
  g.cw = cw;
  g.ch = ch;
  g.nColors = nColors;
  g.nMaps = nMaps;
  g.penalty = penalty;
  if (!nProblems) return hipSuccess;
  if (wide)
    hipLaunchKernelGGL((k_score_frozen<uint64_t>), dim3(nProblems), dim3(64), 0, s, g, (const uint64_t *)cword,
                       tables, nbCount, nbOffset, (const uint64_t *)nbWord, candCount, maxCands, cands, winner,
                       bestSum, stopped);
  else if (nColors == 3 && nMaps == 0 && rgb_fast_path())
    hipLaunchKernelGGL((k_score_frozen<uint32_t, 3>), dim3(nProblems), dim3(64), 0, s, g, (const uint32_t *)cword,
                       tables, nbCount, nbOffset, (const uint32_t *)nbWord, candCount, maxCands, cands, winner,
                       bestSum, stopped);
  else
    hipLaunchKernelGGL((k_score_frozen<uint32_t>), dim3(nProblems), dim3(64), 0, s, g, (const uint32_t *)cword,
                       tables, nbCount, nbOffset, (const uint32_t *)nbWord, candCount, maxCands, cands, winner,
                       bestSum, stopped);
  return hipGetLastError();
}

}  // namespace rs
