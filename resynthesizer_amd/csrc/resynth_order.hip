// Target ordering of match-context types 2..8 on the device (lib/orderTarget.h:154-343,
// lib/brushfire.h:63-148): the sort of the targets by their direction key.  The banded
// shuffle that follows it (lib/orderTarget.h:58-79) depends only on the target count and
// the engine's PRNG, so the host draws it as an index permutation while these kernels run;
// launch_order_compose applies it.
//
// The reference sorts with g_array_sort, a top-down merge sort (left half n/2, take the
// left element iff cmp <= 0) over comparators that never return 0.  For a comparator that
// is consistent (a strict weak order) this is a stable sort with a fixed tie rule:
//   "less" comparators (take left iff key(l) < key(r)): ascending key, ties by original
//     index descending;
//   "more" comparators (take left iff key(l) >= key(r)): descending key, ties by original
//     index ascending.
// Both are one stable radix sort of a mapped key over the points in reverse or forward
// original order.
//
// The inward/outward keys of types 2, 5 and 8 are a float ratio that is NaN (0/0) for the
// target's centre pixel when no other target pixel lies on its ray; a NaN compares false
// both ways, so the merge sort is no longer a sort.  Its result is still exact here: with
// one NaN element every merged run containing it has the form A ++ [NaN] ++ B, A and B
// sorted, and which of a sibling run's elements join A is decided by one threshold per
// merge level (see k_order_nan_plan).  The final order is the stable sort with the A
// elements first, then the NaN element, then the B elements.
#include "resynth_internal.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <climits>

namespace rs {

namespace {

constexpr int kMaxLevels = 40;

struct OrderPlan {
  int nLevels;                 // merge levels above the NaN element's leaf (0: no NaN)
  unsigned nanIdx;             // original index of the NaN element
  int lo[kMaxLevels], hi[kMaxLevels];   // sibling range of each level, original indices
  int mode[kMaxLevels];        // 0: all join A; 1: A iff key < T; 2: A iff key > T
  float T[kMaxLevels];
};

struct OrderState {
  int bounds[4];               // ulx, uly, lrx, lry (lib/engineTypes.h:187-206)
  unsigned nanCount;
  unsigned nanIdx;
  unsigned maxc[401];          // largest x^2 + y^2 along each ray (lib/brushfire.h:72-98)
  OrderPlan plan;
};

__device__ __forceinline__ bool is_less_type(int type) { return type == 5 || type == 6 || type == 7; }
__device__ __forceinline__ bool is_ratio_type(int type) { return type == 2 || type == 5 || type == 8; }

__global__ void k_order_init(OrderState *st)
{
  const int t = threadIdx.x;
  if (t == 0) {
    // get_bounds starts from {G_MAXINT, G_MAXINT, 0, 0}
    st->bounds[0] = INT_MAX;
    st->bounds[1] = INT_MAX;
    st->bounds[2] = 0;
    st->bounds[3] = 0;
    st->nanCount = 0;
    st->nanIdx = UINT_MAX;
    st->plan.nLevels = 0;
  }
  for (int i = t; i < 401; i += blockDim.x) st->maxc[i] = 0;
}

__global__ void k_order_bounds(const int2 *pts, unsigned n, OrderState *st)
{
  int mnx = INT_MAX, mny = INT_MAX, mxx = 0, mxy = 0;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int2 p = pts[i];
    mnx = min(mnx, p.x);
    mny = min(mny, p.y);
    mxx = max(mxx, p.x);
    mxy = max(mxy, p.y);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mnx = min(mnx, __shfl_xor(mnx, o));
    mny = min(mny, __shfl_xor(mny, o));
    mxx = max(mxx, __shfl_xor(mxx, o));
    mxy = max(mxy, __shfl_xor(mxy, o));
  }
  if (__lane_id() == 0) {
    atomicMin(&st->bounds[0], mnx);
    atomicMin(&st->bounds[1], mny);
    atomicMax(&st->bounds[2], mxx);
    atomicMax(&st->bounds[3], mxy);
  }
}

// get_center (lib/engineTypes.h:210-226), then the point as an offset from it.
__device__ __forceinline__ int2 order_offset(const OrderState *st, int2 p)
{
  const int cx = (st->bounds[2] - st->bounds[0]) / 2 + st->bounds[0];
  const int cy = (st->bounds[3] - st->bounds[1]) / 2 + st->bounds[1];
  return make_int2(p.x - cx, p.y - cy);
}

// grad (lib/brushfire.h:63-67): (guint)(atan2((gfloat)y, (gfloat)x) * 200 / G_PI + 200).
// On the axes and diagonals the product lands exactly on a ray boundary, where one ulp of
// atan2 decides the truncation; those nine directions take the host C library's values
// (axisRays: centre, +x, -x, +y, -y, (+,+), (-,+), (-,-), (+,-)).  Any other integer
// direction lies far (in ulps) from every boundary.
__device__ __forceinline__ unsigned order_ray(int x, int y, const uint32_t *axisRays)
{
  if (x == 0 && y == 0) return axisRays[0];
  if (y == 0) return axisRays[x > 0 ? 1 : 2];
  if (x == 0) return axisRays[y > 0 ? 3 : 4];
  if (x == y) return axisRays[x > 0 ? 5 : 7];
  if (x == -y) return axisRays[x < 0 ? 6 : 8];
  const double a = atan2((double)(float)y, (double)(float)x);
  return (unsigned)(a * 200.0 / 3.14159265358979323846 + 200.0);
}

// x*x + y*y in the reference's int arithmetic (two's complement wrap), as unsigned bits.
__device__ __forceinline__ uint32_t order_cart(int2 o)
{
  return (uint32_t)o.x * (uint32_t)o.x + (uint32_t)o.y * (uint32_t)o.y;
}

__global__ void k_order_rays(const int2 *pts, unsigned n, const uint32_t *axisRays, OrderState *st)
{
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int2 o = order_offset(st, pts[i]);
    atomicMax(&st->maxc[order_ray(o.x, o.y, axisRays)], order_cart(o));
  }
}

// Order-preserving map of a float (not NaN) to unsigned; -0 is +0.
__device__ __forceinline__ uint32_t float_key(float f)
{
  uint32_t b = __float_as_uint(f == 0.0f ? 0.0f : f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Sort keys and payloads: the point at original index i goes to slot n-1-i for the "less"
// comparators (so that the stable sort leaves ties in descending index order) and to slot
// i for the "more" ones, whose keys are complemented.  Ratio types also keep the ratio by
// original index (for the NaN plan); the NaN element gets the largest key and sorts last.
__global__ void k_order_keys(const int2 *pts, unsigned n, int type, const uint32_t *axisRays, OrderState *st,
                             uint32_t *keys, uint32_t *vals, float *ratio)
{
  const bool less = is_less_type(type);
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int2 o = order_offset(st, pts[i]);
    uint32_t key;
    if (is_ratio_type(type)) {
      // proportion_inward (lib/brushfire.h:117-128): (gfloat)(int)(y*y + x*x) / max along the ray
      const float num = __int2float_rn((int)order_cart(o));
      const float den = __uint2float_rn(st->maxc[order_ray(o.x, o.y, axisRays)]);
      const float r = __fdiv_rn(num, den);
      ratio[i] = r;
      if (r != r) {
        key = 0xFFFFFFFFu;
        atomicAdd(&st->nanCount, 1u);
        atomicMin(&st->nanIdx, i);
      } else {
        key = less ? float_key(r) : ~float_key(r);
      }
    } else {
      // lessHorizontal / lessVertical (lib/engineTypes.h:113-141): x*x or y*y as int
      const bool horiz = type == 3 || type == 6;
      const int v = horiz ? (int)((uint32_t)o.x * (uint32_t)o.x) : (int)((uint32_t)o.y * (uint32_t)o.y);
      const uint32_t u = (uint32_t)v ^ 0x80000000u;
      key = less ? u : ~u;
    }
    const unsigned slot = less ? n - 1 - i : i;
    keys[slot] = key;
    vals[slot] = i;
  }
}

// Block-wide max (sign > 0) or min (sign < 0) of ratio[lo, hi).
__device__ float block_extreme(const float *ratio, int lo, int hi, bool wantMax, float *red)
{
  float v = wantMax ? -INFINITY : INFINITY;
  for (int i = lo + (int)threadIdx.x; i < hi; i += blockDim.x) v = wantMax ? fmaxf(v, ratio[i]) : fminf(v, ratio[i]);
  for (int o = 32; o > 0; o >>= 1) {
    const float w = __shfl_xor(v, o);
    v = wantMax ? fmaxf(v, w) : fminf(v, w);
  }
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  float r = red[0];
  for (unsigned w = 1; w < blockDim.x / 64; w++) r = wantMax ? fmaxf(r, red[w]) : fminf(r, red[w]);
  __syncthreads();
  return r;
}

// The merge levels above the NaN element's leaf, bottom-up, and for each the rule by which
// its sibling run's elements join A (the part of the merged run before the NaN element).
// With T the extreme of A so far (the last element A contributes to a merge):
//   less comparator, NaN in the left run: the whole right run is taken before the NaN
//     reaches the front, so all of it joins A, and T = max(T, its max);
//   less, NaN in the right run: left elements taken before A is exhausted, i.e. key < T,
//     join A (none while A is empty: T = -inf);
//   more comparator, NaN in the left run: right elements with key > T join A (T = min(A),
//     +inf while A is empty);
//   more, NaN in the right run: the whole left run is taken first and joins A, and
//     T = min(T, its min).
// One workgroup; the sibling ranges partition the other n-1 elements, so the reductions
// read each ratio once.
__global__ __launch_bounds__(1024) void k_order_nan_plan(unsigned n, int type, const float *ratio, OrderState *st)
{
  __shared__ float red[16];
  OrderPlan &P = st->plan;
  if (st->nanCount != 1) {
    // no NaN (nothing to split); more than one cannot occur for target sides up to 65536
    if (threadIdx.x == 0) P.nLevels = 0;
    return;
  }
  const unsigned nan = st->nanIdx;
  const bool less = is_less_type(type);
  // the path from the root: level d splits [lo, hi) at lo + (hi - lo) / 2
  int los[kMaxLevels], his[kMaxLevels];
  bool nanLeft[kMaxLevels];
  int levels = 0;
  int lo = 0, hi = (int)n;
  while (hi - lo > 1 && levels < kMaxLevels) {
    const int mid = lo + (hi - lo) / 2;
    if ((int)nan < mid) {
      los[levels] = mid;
      his[levels] = hi;
      nanLeft[levels] = true;
      hi = mid;
    } else {
      los[levels] = lo;
      his[levels] = mid;
      nanLeft[levels] = false;
      lo = mid;
    }
    levels++;
  }
  float T = less ? -INFINITY : INFINITY;
  for (int d = levels - 1; d >= 0; d--) {
    int mode;
    if (less) {
      if (nanLeft[d]) {
        mode = 0;
        T = fmaxf(T, block_extreme(ratio, los[d], his[d], true, red));
      } else {
        mode = 1;
      }
    } else {
      if (nanLeft[d]) {
        mode = 2;
      } else {
        mode = 0;
        T = fminf(T, block_extreme(ratio, los[d], his[d], false, red));
      }
    }
    if (threadIdx.x == 0) {
      const int k = levels - 1 - d;   // bottom-up slot
      P.lo[k] = los[d];
      P.hi[k] = his[d];
      P.mode[k] = mode;
      P.T[k] = (mode == 0) ? 0.0f : T;
    }
  }
  if (threadIdx.x == 0) {
    P.nLevels = levels;
    P.nanIdx = nan;
  }
}

// flag[p] = 1 when the element at sorted position p joins A (every element when there is
// no NaN; the NaN element itself gets 0 and sorts last).
__global__ void k_order_flags(const uint32_t *sortedVals, unsigned n, const float *ratio, const OrderState *st,
                              uint32_t *flags)
{
  const OrderPlan &P = st->plan;
  const int levels = P.nLevels;
  for (unsigned p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const unsigned i = sortedVals[p];
    uint32_t f = 1;
    if (levels > 0) {
      if (i == P.nanIdx) {
        f = 0;
      } else {
        for (int k = 0; k < levels; k++) {
          if ((int)i >= P.lo[k] && (int)i < P.hi[k]) {
            const float r = ratio[i];
            f = P.mode[k] == 0 ? 1u : P.mode[k] == 1 ? (r < P.T[k] ? 1u : 0u) : (r > P.T[k] ? 1u : 0u);
            break;
          }
        }
      }
    }
    flags[p] = f;
  }
}

// Stable partition of the sorted payloads: A elements (in sorted order), the NaN element,
// then the B elements.  aBefore is the exclusive prefix sum of flags.
__global__ void k_order_scatter(const uint32_t *sortedVals, const uint32_t *flags, const uint32_t *aBefore, unsigned n,
                                const OrderState *st, uint32_t *out)
{
  const bool hasNan = st->plan.nLevels > 0;
  const unsigned nA = aBefore[n - 1] + flags[n - 1];
  for (unsigned p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const unsigned i = sortedVals[p];
    unsigned dst;
    if (flags[p]) dst = aBefore[p];
    else if (hasNan && i == st->nanIdx) dst = nA;
    else dst = nA + (hasNan ? 1u : 0u) + (p - aBefore[p]);
    out[dst] = i;
  }
}

__global__ void k_order_compose(const int2 *rm, const uint32_t *order, const uint32_t *perm, int2 *out, unsigned n)
{
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = rm[order[perm[i]]];
}

size_t scan_temp_bytes(size_t n)
{
  size_t bytes = 0;
  (void)rocprim::exclusive_scan(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr, 0u, n,
                                rocprim::plus<uint32_t>());
  return bytes;
}

size_t sort_temp_bytes(size_t n)
{
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs<rocprim::default_config>(nullptr, bytes, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                           (uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, 32);
  return bytes;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

size_t order_temp_bytes(unsigned n)
{
  const size_t m = n ? n : 1;
  return align256(sizeof(OrderState)) + 6 * align256(m * 4) + align256(m * 4) +
         align256(std::max(sort_temp_bytes(m), scan_temp_bytes(m))) + 256;
}

// Sorts the row-major targets rm[0..n) for match-context type 2..8 and writes the original
// indices in the reference's sorted order to order[0..n) (inside temp).  axisRays: the
// host's ray values of the nine exact directions (see order_ray).
hipError_t launch_order_sort(hipStream_t s, const int2 *rm, unsigned n, int type, const uint32_t *axisRays,
                             void *temp, size_t tempBytes, const uint32_t **order)
{
  char *base = (char *)(((uintptr_t)temp + 255) & ~(uintptr_t)255);
  const size_t m = n ? n : 1, a = align256(m * 4);
  OrderState *st = (OrderState *)base;
  char *q = base + align256(sizeof(OrderState));
  uint32_t *k0 = (uint32_t *)q, *v0 = (uint32_t *)(q + a), *k1 = (uint32_t *)(q + 2 * a),
           *v1 = (uint32_t *)(q + 3 * a), *flags = (uint32_t *)(q + 4 * a), *aBefore = (uint32_t *)(q + 5 * a);
  float *ratio = (float *)(q + 6 * a);
  void *work = q + 7 * a;
  size_t workBytes = std::max(sort_temp_bytes(m), scan_temp_bytes(m));
  if ((char *)work + workBytes > (char *)temp + tempBytes) return hipErrorInvalidValue;
  *order = v1;
  if (!n) return hipSuccess;
  const unsigned blocks = std::min(4096u, (n + 255) / 256);
  hipLaunchKernelGGL(k_order_init, dim3(1), dim3(256), 0, s, st);
  hipLaunchKernelGGL(k_order_bounds, dim3(blocks), dim3(256), 0, s, rm, n, st);
  const bool ratioType = type == 2 || type == 5 || type == 8;
  if (ratioType) hipLaunchKernelGGL(k_order_rays, dim3(blocks), dim3(256), 0, s, rm, n, axisRays, st);
  hipLaunchKernelGGL(k_order_keys, dim3(blocks), dim3(256), 0, s, rm, n, type, axisRays, st, k0, v0, ratio);
  hipError_t e = rocprim::radix_sort_pairs<rocprim::default_config>(work, workBytes, k0, k1, v0, v1, n, 0, 32, s);
  if (e != hipSuccess) return e;
  if (!ratioType) return hipGetLastError();
  hipLaunchKernelGGL(k_order_nan_plan, dim3(1), dim3(1024), 0, s, n, type, (const float *)ratio, st);
  hipLaunchKernelGGL(k_order_flags, dim3(blocks), dim3(256), 0, s, (const uint32_t *)v1, n, (const float *)ratio,
                     (const OrderState *)st, flags);
  e = rocprim::exclusive_scan(work, workBytes, flags, aBefore, 0u, n, rocprim::plus<uint32_t>(), s);
  if (e != hipSuccess) return e;
  // the partitioned order overwrites the sort keys, which are no longer needed
  hipLaunchKernelGGL(k_order_scatter, dim3(blocks), dim3(256), 0, s, (const uint32_t *)v1, (const uint32_t *)flags,
                     (const uint32_t *)aBefore, n, (const OrderState *)st, k1);
  *order = k1;
  return hipGetLastError();
}

// out[i] = rm[order[perm[i]]]: the sorted targets after the banded shuffle perm.
hipError_t launch_order_compose(hipStream_t s, const int2 *rm, const uint32_t *order, const uint32_t *perm, int2 *out,
                                unsigned n)
{
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_order_compose, dim3((n + 255) / 256), dim3(256), 0, s, rm, order, perm, out, n);
  return hipGetLastError();
}

}  // namespace rs
