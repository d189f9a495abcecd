// Learned candidate proposals for the batched engine: a small BF16 MLP (FP32 accumulate)
// that embeds a pixel's neighbourhood, trained with InfoNCE on neighbourhood -> best-match
// pairs logged from the plain engine, and an inverted-file index over corpus embeddings
// whose searches are MFMA GEMMs with fused top-k.  The exact integer scorer still judges
// every proposal; proposals only replace most of the random probes.
//
// Network: x (120 inputs + 8 zero pads: for each of the 30 nearest offsets, R, G, B / 255
// and a presence flag) -> 256 -> ReLU -> 256 -> ReLU -> 64 -> L2 normalisation.
#include "resynth_learn.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

namespace rs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
  // round to nearest even (inputs here are finite)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Word index of corpus pixel (x, y) in the engine's corpus layout: rows of four pixel rows,
// each pixel's four words side by side (8 x 4-pixel tiles per 128-byte line).
__device__ __forceinline__ size_t corpus_word_index(int cw, unsigned x, unsigned y)
{
  const size_t rowOf4 = (size_t)(((unsigned)cw + 7u) & ~7u) * 4u;
  return (size_t)(y >> 2) * rowOf4 + ((x << 2) | (y & 3u));
}

// ------------------------------------------------------------------ GEMM
//
// C[M x N] = op(A)[M x K] * op(B)[K x N], A and B bf16, FP32 accumulation on
// v_mfma_f32_16x16x32_bf16.  op(A) is A ([M x K] row-major, lda) or, with TA, the
// transpose of A stored [K x M] (lda).  op(B) is B ([K x N], ldb) or, with TB, the
// transpose of B stored [N x K] (ldb).  K must be a multiple of 32; M and N are bounded.
// Workgroup tile 64 x 64, four waves of 32 x 32 (2 x 2 MFMA tiles), K step 32; both
// tiles are staged in LDS k-contiguous, so each lane's MFMA operand is one 16-byte read.
constexpr int GT = 64, GK = 32, GPAD = 8;

template <bool TA, bool TB>
__device__ __forceinline__ void gemm_load_tiles(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B,
                                                int ldb, int m0, int n0, int k0, uint16_t (*As)[GK + GPAD],
                                                uint16_t (*Bs)[GK + GPAD])
{
  const int t = threadIdx.x;  // 256 threads
  // A tile: 64 rows (m) x 32 k
  if (!TA) {
    const int r = t >> 2, c = (t & 3) * 8;  // 64 rows x 4 groups of 8
    const int m = m0 + r;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (m < M) v = *(const uint4 *)(A + (size_t)m * lda + k0 + c);
    *(uint4 *)&As[r][c] = v;
  } else {
    // stored [K x M]: 32 rows of k, 64 m each; thread reads 8 consecutive m of one k
    const int kr = t >> 3, c = (t & 7) * 8;
    const int m = m0 + c;
    uint16_t tmp[8];
    if (m + 7 < M) {
      uint4 v = *(const uint4 *)(A + (size_t)(k0 + kr) * lda + m);
      __builtin_memcpy(tmp, &v, 16);
    } else {
      for (int i = 0; i < 8; i++) tmp[i] = (m + i < M) ? A[(size_t)(k0 + kr) * lda + m + i] : (uint16_t)0;
    }
    for (int i = 0; i < 8; i++) As[c + i][kr] = tmp[i];
  }
  // B tile into Bs[n][k]
  if (TB) {
    const int r = t >> 2, c = (t & 3) * 8;
    const int n = n0 + r;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (n < N) v = *(const uint4 *)(B + (size_t)n * ldb + k0 + c);
    *(uint4 *)&Bs[r][c] = v;
  } else {
    const int kr = t >> 3, c = (t & 7) * 8;
    const int n = n0 + c;
    uint16_t tmp[8];
    if (n + 7 < N) {
      uint4 v = *(const uint4 *)(B + (size_t)(k0 + kr) * ldb + n);
      __builtin_memcpy(tmp, &v, 16);
    } else {
      for (int i = 0; i < 8; i++) tmp[i] = (n + i < N) ? B[(size_t)(k0 + kr) * ldb + n + i] : (uint16_t)0;
    }
    for (int i = 0; i < 8; i++) Bs[c + i][kr] = tmp[i];
  }
}

// Accumulates the workgroup's 64 x 64 tile; acc[a][b] is the 16 x 16 MFMA tile (a, b) of
// this wave's 32 x 32 quadrant.  Lane l of a 16 x 16 tile holds rows 4*(l/16)+r, column l%16.
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_tile(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                                          int m0, int n0, f32x4 (&acc)[2][2])
{
  __shared__ __attribute__((aligned(16))) uint16_t As[GT][GK + GPAD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[GT][GK + GPAD];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += GK) {
    gemm_load_tiles<TA, TB>(M, N, K, A, lda, B, ldb, m0, n0, k0, As, Bs);
    __syncthreads();
    const int r = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 fa[2], fb[2];
    for (int a = 0; a < 2; a++) fa[a] = *(const bf16x8 *)&As[wm + a * 16 + r][kq];
    for (int b = 0; b < 2; b++) fb[b] = *(const bf16x8 *)&Bs[wn + b * 16 + r][kq];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
}

// Epilogues
enum { EPI_F32 = 0, EPI_BIAS_RELU_BF16 = 1, EPI_BIAS_F32 = 2, EPI_BF16 = 3 };

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B,
                                              int ldb, float *C, uint16_t *Cb, int ldc, const float *bias,
                                              float scale)
{
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  f32x4 acc[2][2];
  gemm_tile<TA, TB>(M, N, K, A, lda, B, ldb, m0, n0, acc);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + a * 16 + (lane >> 4) * 4 + r;
        const int n = n0 + wn + b * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        float v = acc[a][b][r] * scale;
        if (EPI == EPI_BIAS_RELU_BF16 || EPI == EPI_BIAS_F32) v += bias[n];
        if (EPI == EPI_BIAS_RELU_BF16) {
          v = v > 0.f ? v : 0.f;
          Cb[(size_t)m * ldc + n] = f2bf(v);
        } else if (EPI == EPI_BF16) {
          Cb[(size_t)m * ldc + n] = f2bf(v);
        } else {
          C[(size_t)m * ldc + n] = v;
        }
      }
}

template <bool TA, bool TB, int EPI>
static hipError_t gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                       float *C, uint16_t *Cb, int ldc, const float *bias, float scale)
{
  if (M <= 0 || N <= 0) return hipSuccess;
  if (K % GK) return hipErrorInvalidValue;
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
  hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, Cb, ldc, bias, scale);
  return hipGetLastError();
}

hipError_t learn_gemm(hipStream_t s, int ta, int tb, int M, int N, int K, const uint16_t *A, int lda,
                      const uint16_t *B, int ldb, float *C, int ldc)
{
  if (!ta && !tb) return gemm<false, false, EPI_F32>(s, M, N, K, A, lda, B, ldb, C, nullptr, ldc, nullptr, 1.f);
  if (!ta && tb) return gemm<false, true, EPI_F32>(s, M, N, K, A, lda, B, ldb, C, nullptr, ldc, nullptr, 1.f);
  if (ta && !tb) return gemm<true, false, EPI_F32>(s, M, N, K, A, lda, B, ldb, C, nullptr, ldc, nullptr, 1.f);
  return gemm<true, true, EPI_F32>(s, M, N, K, A, lda, B, ldb, C, nullptr, ldc, nullptr, 1.f);
}

// ------------------------------------------------------------------ elementwise helpers

__global__ void k_f32_to_bf16(const float *x, uint16_t *y, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

__global__ void k_zero(float *x, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.f;
}

// y += x
__global__ void k_axpy(const float *x, float *y, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

__global__ void k_scale1(float *x, float a) { x[0] *= a; }

static hipError_t to_bf16(hipStream_t s, const float *x, uint16_t *y, size_t n)
{
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}

// One wavefront per row of 64: z = e / |e|, also returning |e| (guarded against 0).
__global__ void k_normalize64(const float *e, float *z, float *norm, uint16_t *zb, int rows)
{
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  const float v = e[(size_t)row * 64 + lane];
  float s = v * v;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  const float n = sqrtf(fmaxf(s, 1e-12f));
  const float zz = v / n;
  if (z) z[(size_t)row * 64 + lane] = zz;
  if (zb) zb[(size_t)row * 64 + lane] = f2bf(zz);
  if (norm && lane == 0) norm[row] = n;
}

// Backward of the normalisation: de = (dz - z (z . dz)) / |e|.
__global__ void k_normalize64_bwd(const float *z, const float *norm, const float *dz, uint16_t *deb, float *de,
                                  int rows)
{
  const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  const float zz = z[(size_t)row * 64 + lane], g = dz[(size_t)row * 64 + lane];
  float d = zz * g;
  for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
  const float v = (g - zz * d) / norm[row];
  if (deb) deb[(size_t)row * 64 + lane] = f2bf(v);
  if (de) de[(size_t)row * 64 + lane] = v;
}

// InfoNCE over a B x B logit matrix L (row i: query i against every key; the positive is
// the diagonal).  One workgroup per row: loss_i = logsumexp(L_i) - L_ii, and
// dL_i = (softmax(L_i) - e_i) / B written as bf16 (and transposed into dLt).
__global__ __launch_bounds__(256) void k_infonce_rows(const float *L, int B, float *rowLoss, uint16_t *dL,
                                                      uint16_t *dLt)
{
  const int i = blockIdx.x;
  __shared__ float red[256];
  const float *row = L + (size_t)i * B;
  float mx = -3.4e38f;
  for (int j = threadIdx.x; j < B; j += 256) mx = fmaxf(mx, row[j]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  mx = red[0];
  __syncthreads();
  float sum = 0.f;
  for (int j = threadIdx.x; j < B; j += 256) sum += expf(row[j] - mx);
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  sum = red[0];
  const float lse = mx + logf(sum);
  if (threadIdx.x == 0) rowLoss[i] = lse - row[i];
  const float inv = 1.f / (float)B;
  for (int j = threadIdx.x; j < B; j += 256) {
    float g = expf(row[j] - lse);
    if (j == i) g -= 1.f;
    g *= inv;
    dL[(size_t)i * B + j] = f2bf(g);
    dLt[(size_t)j * B + i] = f2bf(g);
  }
}

// Backward of ReLU through a bf16 activation: g_pre = g * (act > 0), as bf16.
__global__ void k_relu_bwd(const float *g, const uint16_t *act, uint16_t *out, size_t n)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (act[i] != 0 && !(act[i] & 0x8000u)) ? f2bf(g[i]) : (uint16_t)0;
}

__global__ void k_sum(const float *x, int n, float *out)
{
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Adam on FP32 master weights, refreshing the bf16 copies used by the GEMMs.
__global__ void k_adam(float *w, const float *g, float *m, float *v, uint16_t *wb, size_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

// ------------------------------------------------------------------ model

// Parameter offsets in the flat FP32 vector (weights [in x out] row-major).
constexpr size_t OW1 = 0, OB1 = OW1 + (size_t)kIn * kH1, OW2 = OB1 + kH1, OB2 = OW2 + (size_t)kH1 * kH2,
                 OW3 = OB2 + kH2, OB3 = OW3 + (size_t)kH2 * kEmb, NPARAM = OB3 + kEmb;
static_assert(NPARAM == kParamCount, "parameter layout");

// Forward of n rows of x (bf16 [n x kIn]) into activations (bf16) and e (fp32 [n x 64]).
hipError_t mlp_forward(hipStream_t s, const float *w, const uint16_t *wb, const uint16_t *x, int n, uint16_t *h1,
                       uint16_t *h2, float *e)
{
  hipError_t r;
  if ((r = gemm<false, false, EPI_BIAS_RELU_BF16>(s, n, kH1, kIn, x, kIn, wb + OW1, kH1, nullptr, h1, kH1, w + OB1,
                                                  1.f)))
    return r;
  if ((r = gemm<false, false, EPI_BIAS_RELU_BF16>(s, n, kH2, kH1, h1, kH1, wb + OW2, kH2, nullptr, h2, kH2, w + OB2,
                                                  1.f)))
    return r;
  return gemm<false, false, EPI_BIAS_F32>(s, n, kEmb, kH2, h2, kH2, wb + OW3, kEmb, e, nullptr, kEmb, w + OB3, 1.f);
}

hipError_t mlp_embed(hipStream_t s, const float *w, const uint16_t *wb, const uint16_t *x, int n, LearnScratch &t,
                     uint16_t *zb, float *z)
{
  hipError_t r;
  if ((r = mlp_forward(s, w, wb, x, n, t.h1, t.h2, t.e))) return r;
  hipLaunchKernelGGL(k_normalize64, dim3((n + 3) / 4), dim3(256), 0, s, t.e, z, nullptr, zb, n);
  return hipGetLastError();
}

__global__ void k_colsum_bf16_par(const uint16_t *g, int rows, int cols, float *out);

// Backward of one tower given dE (fp32 [n x 64]); accumulates parameter gradients into
// grad (which the caller zeroes).  x, h1, h2 are the forward's inputs and activations.
static hipError_t mlp_backward(hipStream_t s, const uint16_t *wb, const uint16_t *x, const uint16_t *h1,
                               const uint16_t *h2, const uint16_t *dEb, int n, LearnScratch &t, float *grad)
{
  hipError_t r;
  // dW3 = h2^T dE, db3 = colsum dE
  if ((r = gemm<true, false, EPI_F32>(s, kH2, kEmb, n, h2, kH2, dEb, kEmb, t.gw, nullptr, kEmb, nullptr, 1.f))) return r;
  hipLaunchKernelGGL(k_axpy, dim3((kH2 * kEmb + 255) / 256), dim3(256), 0, s, t.gw, grad + OW3, (size_t)kH2 * kEmb);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s, t.gb, (size_t)kEmb);
  hipLaunchKernelGGL(k_colsum_bf16_par, dim3((kEmb + 63) / 64, (n + 255) / 256), dim3(256), 0, s, dEb, n, kEmb, t.gb);
  hipLaunchKernelGGL(k_axpy, dim3(1), dim3(256), 0, s, t.gb, grad + OB3, (size_t)kEmb);
  // dH2 = dE W3^T -> through ReLU
  if ((r = gemm<false, true, EPI_F32>(s, n, kH2, kEmb, dEb, kEmb, wb + OW3, kEmb, t.g, nullptr, kH2, nullptr, 1.f)))
    return r;
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)(((size_t)n * kH2 + 255) / 256)), dim3(256), 0, s, t.g, h2, t.gb16,
                     (size_t)n * kH2);
  if ((r = gemm<true, false, EPI_F32>(s, kH1, kH2, n, h1, kH1, t.gb16, kH2, t.gw, nullptr, kH2, nullptr, 1.f))) return r;
  hipLaunchKernelGGL(k_axpy, dim3((kH1 * kH2 + 255) / 256), dim3(256), 0, s, t.gw, grad + OW2, (size_t)kH1 * kH2);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s, t.gb, (size_t)kH2);
  hipLaunchKernelGGL(k_colsum_bf16_par, dim3((kH2 + 63) / 64, (n + 255) / 256), dim3(256), 0, s, t.gb16, n, kH2, t.gb);
  hipLaunchKernelGGL(k_axpy, dim3(1), dim3(256), 0, s, t.gb, grad + OB2, (size_t)kH2);
  // dH1 = dH2pre W2^T -> through ReLU
  if ((r = gemm<false, true, EPI_F32>(s, n, kH1, kH2, t.gb16, kH2, wb + OW2, kH2, t.g, nullptr, kH1, nullptr, 1.f)))
    return r;
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)(((size_t)n * kH1 + 255) / 256)), dim3(256), 0, s, t.g, h1, t.gb16,
                     (size_t)n * kH1);
  if ((r = gemm<true, false, EPI_F32>(s, kIn, kH1, n, x, kIn, t.gb16, kH1, t.gw, nullptr, kH1, nullptr, 1.f))) return r;
  hipLaunchKernelGGL(k_axpy, dim3((kIn * kH1 + 255) / 256), dim3(256), 0, s, t.gw, grad + OW1, (size_t)kIn * kH1);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, s, t.gb, (size_t)kH1);
  hipLaunchKernelGGL(k_colsum_bf16_par, dim3((kH1 + 63) / 64, (n + 255) / 256), dim3(256), 0, s, t.gb16, n, kH1, t.gb);
  hipLaunchKernelGGL(k_axpy, dim3(1), dim3(256), 0, s, t.gb, grad + OB1, (size_t)kH1);
  return hipGetLastError();
}

// InfoNCE loss and gradients for n pairs (xq[i], xk[i]) with temperature tau; both towers
// share the weights.  grad (fp32, kParamCount) is overwritten; *loss (device) gets the mean.
hipError_t mlp_infonce(hipStream_t s, const float *w, const uint16_t *wb, const uint16_t *xq, const uint16_t *xk, int n,
                       float tau, LearnScratch &q, LearnScratch &k, float *grad, float *loss)
{
  hipError_t r;
  hipLaunchKernelGGL(k_zero, dim3((kParamCount + 255) / 256), dim3(256), 0, s, grad, (size_t)kParamCount);
  if ((r = mlp_forward(s, w, wb, xq, n, q.h1, q.h2, q.e))) return r;
  if ((r = mlp_forward(s, w, wb, xk, n, k.h1, k.h2, k.e))) return r;
  hipLaunchKernelGGL(k_normalize64, dim3((n + 3) / 4), dim3(256), 0, s, q.e, q.z, q.norm, q.zb, n);
  hipLaunchKernelGGL(k_normalize64, dim3((n + 3) / 4), dim3(256), 0, s, k.e, k.z, k.norm, k.zb, n);
  // L = zq zk^T / tau
  if ((r = gemm<false, true, EPI_F32>(s, n, n, kEmb, q.zb, kEmb, k.zb, kEmb, q.L, nullptr, n, nullptr, 1.f / tau)))
    return r;
  hipLaunchKernelGGL(k_infonce_rows, dim3(n), dim3(256), 0, s, q.L, n, q.rowLoss, q.dL, q.dLt);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, q.rowLoss, n, loss);
  hipLaunchKernelGGL(k_scale1, dim3(1), dim3(1), 0, s, loss, 1.f / (float)n);
  // dzq = dL zk / tau ; dzk = dL^T zq / tau
  if ((r = gemm<false, false, EPI_F32>(s, n, kEmb, n, q.dL, n, k.zb, kEmb, q.dz, nullptr, kEmb, nullptr, 1.f / tau)))
    return r;
  if ((r = gemm<false, false, EPI_F32>(s, n, kEmb, n, q.dLt, n, q.zb, kEmb, k.dz, nullptr, kEmb, nullptr, 1.f / tau)))
    return r;
  hipLaunchKernelGGL(k_normalize64_bwd, dim3((n + 3) / 4), dim3(256), 0, s, q.z, q.norm, q.dz, q.dEb, nullptr, n);
  hipLaunchKernelGGL(k_normalize64_bwd, dim3((n + 3) / 4), dim3(256), 0, s, k.z, k.norm, k.dz, k.dEb, nullptr, n);
  if ((r = mlp_backward(s, wb, xq, q.h1, q.h2, q.dEb, n, q, grad))) return r;
  return mlp_backward(s, wb, xk, k.h1, k.h2, k.dEb, n, k, grad);
}

hipError_t adam_step(hipStream_t s, float *w, const float *g, float *m, float *v, uint16_t *wb, float lr, int step)
{
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((kParamCount + 255) / 256), dim3(256), 0, s, w, g, m, v, wb, (size_t)kParamCount, lr,
                     b1, b2, 1e-8f, c1, c2);
  return hipGetLastError();
}

hipError_t params_to_bf16(hipStream_t s, const float *w, uint16_t *wb) { return to_bf16(s, w, wb, kParamCount); }

// dst row i = src row idx[i]; cols is a multiple of 8 (16-byte rows pieces).
__global__ void k_gather_rows(const uint16_t *src, int cols, const uint32_t *idx, int n, uint16_t *dst)
{
  const int per = cols / 8;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * per) return;
  const int i = (int)(t / per), c = (int)(t % per);
  ((uint4 *)(dst + (size_t)i * cols))[c] = ((const uint4 *)(src + (size_t)idx[i] * cols))[c];
}

hipError_t gather_rows_bf16(hipStream_t s, const uint16_t *src, int cols, const uint32_t *idx, int n, uint16_t *dst)
{
  if (n <= 0) return hipSuccess;
  const size_t work = (size_t)n * (cols / 8);
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, src, cols, idx, n, dst);
  return hipGetLastError();
}

// dst row i = src row floor(i * nSrc / n): n rows spread evenly over the nSrc rows.
__global__ void k_gather_spread_rows(const uint16_t *src, int cols, uint32_t nSrc, int n, uint16_t *dst)
{
  const int per = cols / 8;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * per) return;
  const int i = (int)(t / per), c = (int)(t % per);
  const uint32_t r = (uint32_t)((uint64_t)i * nSrc / (uint32_t)n);
  ((uint4 *)(dst + (size_t)i * cols))[c] = ((const uint4 *)(src + (size_t)r * cols))[c];
}

hipError_t gather_spread_rows_bf16(hipStream_t s, const uint16_t *src, int cols, uint32_t nSrc, int n, uint16_t *dst)
{
  if (n <= 0) return hipSuccess;
  const size_t work = (size_t)n * (cols / 8);
  hipLaunchKernelGGL(k_gather_spread_rows, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, src, cols, nSrc, n,
                     dst);
  return hipGetLastError();
}

__device__ __forceinline__ uint64_t lmix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// idx[i] = uniform draw in [0, range) from a counter-based hash of (key, i).
__global__ void k_hash_idx(uint32_t *idx, int n, uint32_t range, uint64_t key)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h = lmix64(key * 0x100000001B3ull + (uint64_t)i);
  idx[i] = (uint32_t)(((h >> 32) * (uint64_t)range) >> 32);
}

hipError_t hash_indices(hipStream_t s, uint32_t *idx, int n, uint32_t range, uint64_t key)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_hash_idx, dim3((n + 255) / 256), dim3(256), 0, s, idx, n, range, key);
  return hipGetLastError();
}

}  // namespace rs

namespace rs {

// ------------------------------------------------------------------ features

template <typename Word>
__global__ __launch_bounds__(256) void k_features(const Word *words, int w, int h, int tileH, int tileV, int nColors,
                                                  const int2 *offsets, unsigned nOffsets, const int2 *pts, int n,
                                                  int isCorpus, int scale2, uint64_t dropKey, uint16_t *out)
{
  const int i = blockIdx.x * 8 + (threadIdx.x >> 5);
  const int j = threadIdx.x & 31;
  if (i >= n) return;
  uint16_t v[4] = {0, 0, 0, 0};
  if (j < kNbr && j > 0 && (unsigned)j < nOffsets) {
    int2 p = pts[i];
    if (scale2) {
      p.x *= 2;
      p.y *= 2;
    }
    const int2 o = offsets[j];
    int x = p.x + o.x, y = p.y + o.y;
    bool in = true;
    if (x < 0) { if (tileH) x += w; else in = false; }
    else if (x >= w) { if (tileH) x -= w; else in = false; }
    if (y < 0) { if (tileV) y += h; else in = false; }
    else if (y >= h) { if (tileV) y -= h; else in = false; }
    if (in && x >= 0 && x < w && y >= 0 && y < h) {
      const size_t at =
          isCorpus ? corpus_word_index(w, (unsigned)x, (unsigned)y) : (size_t)y * (unsigned)w + (unsigned)x;
      const Word wd = words[at];
      bool present = isCorpus ? ((wd & 0xFEu) == 0xFEu) : ((wd & 1u) != 0);
      if (present && dropKey) {
        // keep each neighbour with a per-point probability in [0.2, 1)
        const uint64_t hp = lmix64(dropKey * 0x100000001B3ull + (uint64_t)i);
        const float keep = 0.2f + 0.8f * (float)(hp >> 40) * (1.0f / 16777216.0f);
        const uint64_t hj = lmix64(hp + (uint64_t)j * 0x9E3779B97F4A7C15ull);
        present = (float)(hj >> 40) * (1.0f / 16777216.0f) < keep;
      }
      if (present) {
        for (int c = 0; c < 3; c++)
          if (c < nColors) v[c] = f2bf((float)((uint32_t)(wd >> (8 * (1 + c))) & 0xFFu) * (1.0f / 255.0f));
        v[3] = 0x3F80;  // 1.0
      }
    }
  }
  // threads 30 and 31 write the 8 zero pads (inputs 120..127); thread 0 writes zeros too
  uint2 packed = make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
  *(uint2 *)(out + (size_t)i * kIn + j * 4) = packed;
}

hipError_t learn_features(hipStream_t s, bool wide, const void *words, int w, int h, int tileH, int tileV, int nColors,
                          const int2 *offsets, unsigned nOffsets, const int2 *pts, int n, bool isCorpus, bool scale2,
                          uint64_t dropKey, uint16_t *out)
{
  if (n <= 0) return hipSuccess;
  dim3 grid((n + 7) / 8);
  if (wide)
    hipLaunchKernelGGL(k_features<uint64_t>, grid, dim3(256), 0, s, (const uint64_t *)words, w, h, tileH, tileV,
                       nColors, offsets, nOffsets, pts, n, (int)isCorpus, (int)scale2, dropKey, out);
  else
    hipLaunchKernelGGL(k_features<uint32_t>, grid, dim3(256), 0, s, (const uint32_t *)words, w, h, tileH, tileV,
                       nColors, offsets, nOffsets, pts, n, (int)isCorpus, (int)scale2, dropKey, out);
  return hipGetLastError();
}

template <typename Word>
__global__ void k_key_flags(const Word *cword, int cw, int ch, uint8_t *flags)
{
  const int w2 = (cw + 1) / 2, h2 = (ch + 1) / 2;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)w2 * h2) return;
  const int x = (int)(i % w2) * 2, y = (int)(i / w2) * 2;
  flags[i] = ((cword[corpus_word_index(cw, (unsigned)x, (unsigned)y)] & 0xFFu) == 0xFFu) ? 1 : 0;
}

hipError_t learn_key_flags(hipStream_t s, bool wide, const void *cword, int cw, int ch, uint8_t *flags)
{
  const size_t n = (size_t)((cw + 1) / 2) * ((ch + 1) / 2);
  if (!n) return hipSuccess;
  dim3 grid((unsigned)((n + 255) / 256));
  if (wide)
    hipLaunchKernelGGL(k_key_flags<uint64_t>, grid, dim3(256), 0, s, (const uint64_t *)cword, cw, ch, flags);
  else
    hipLaunchKernelGGL(k_key_flags<uint32_t>, grid, dim3(256), 0, s, (const uint32_t *)cword, cw, ch, flags);
  return hipGetLastError();
}

// ------------------------------------------------------------------ GEMM with fused top-P

template <int P>
__device__ __forceinline__ void topk_insert(float (&v)[P], int (&ix)[P], float val, int idx)
{
  if (P == 1) {
    if (val > v[0]) {
      v[0] = val;
      ix[0] = idx;
    }
  } else {
    if (val > v[0]) {
      v[1] = v[0];
      ix[1] = ix[0];
      v[0] = val;
      ix[0] = idx;
    } else if (val > v[1]) {
      v[1] = val;
      ix[1] = idx;
    }
  }
}

template <int P>
__global__ __launch_bounds__(256) void k_topk_cols(const uint16_t *rows, int M, const uint16_t *cols, int N,
                                                   int32_t *out)
{
  __shared__ float sV[4][32][P];
  __shared__ int sI[4][32][P];
  const int m0 = blockIdx.x * GT;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wn = (wave & 1) * 32;   // waves 0,1 hold rows 0..31 of the tile, waves 2,3 rows 32..63
  float bv[8][P];
  int bi[8][P];
  for (int q = 0; q < 8; q++)
    for (int p = 0; p < P; p++) {
      bv[q][p] = -3.4e38f;
      bi[q][p] = -1;
    }
  for (int n0 = 0; n0 < N; n0 += GT) {
    f32x4 acc[2][2];
    gemm_tile<false, true>(M, N, kEmb, rows, kEmb, cols, kEmb, m0, n0, acc);
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        const int n = n0 + wn + b * 16 + (lane & 15);
        if (n >= N) continue;
        for (int r = 0; r < 4; r++) topk_insert<P>(bv[a * 4 + r], bi[a * 4 + r], acc[a][b][r], n);
      }
  }
  // merge across the 16 lanes that share rows
  for (int o = 1; o < 16; o <<= 1)
    for (int q = 0; q < 8; q++) {
      float ov[P];
      int oi[P];
      for (int p = 0; p < P; p++) {
        ov[p] = __shfl_xor(bv[q][p], o);
        oi[p] = __shfl_xor(bi[q][p], o);
      }
      for (int p = 0; p < P; p++)
        if (oi[p] >= 0) topk_insert<P>(bv[q], bi[q], ov[p], oi[p]);
    }
  // remove duplicates a lane may have taken from its partner (insert keeps both copies
  // of the same column); keep the list distinct
  if (P == 2)
    for (int q = 0; q < 8; q++)
      if (bi[q][1] == bi[q][0]) {
        bv[q][1] = -3.4e38f;
        bi[q][1] = -1;
      }
  if ((lane & 15) == 0)
    for (int a = 0; a < 2; a++)
      for (int r = 0; r < 4; r++) {
        const int rl = a * 16 + (lane >> 4) * 4 + r;
        for (int p = 0; p < P; p++) {
          sV[wave][rl][p] = bv[a * 4 + r][p];
          sI[wave][rl][p] = bi[a * 4 + r][p];
        }
      }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int ml = threadIdx.x, w0 = (ml >= 32) ? 2 : 0, rl = ml & 31;
    float v[P];
    int ix[P];
    for (int p = 0; p < P; p++) {
      v[p] = sV[w0][rl][p];
      ix[p] = sI[w0][rl][p];
    }
    for (int p = 0; p < P; p++)
      if (sI[w0 + 1][rl][p] >= 0 && sI[w0 + 1][rl][p] != ix[0]) topk_insert<P>(v, ix, sV[w0 + 1][rl][p], sI[w0 + 1][rl][p]);
    const int m = m0 + ml;
    if (m < M)
      for (int p = 0; p < P; p++) out[(size_t)m * P + p] = ix[p];
  }
}

hipError_t learn_topk_cols(hipStream_t s, const uint16_t *rows, int M, const uint16_t *cols, int N, int P,
                           int32_t *out)
{
  if (M <= 0 || N <= 0) return hipSuccess;
  dim3 grid((M + GT - 1) / GT);
  if (P == 1)
    hipLaunchKernelGGL(k_topk_cols<1>, grid, dim3(256), 0, s, rows, M, cols, N, out);
  else if (P == 2)
    hipLaunchKernelGGL(k_topk_cols<2>, grid, dim3(256), 0, s, rows, M, cols, N, out);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// ------------------------------------------------------------------ fine search

__global__ void k_sample_pairs(const int2 *targets, unsigned nT, const int2 *sourceOf, int tw, int n, uint64_t key,
                               int2 *pts, int2 *src)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h = lmix64(key * 0x100000001B3ull + (uint64_t)i);
  const int2 t = targets[(unsigned)(((h >> 32) * (uint64_t)nT) >> 32)];
  pts[i] = t;
  src[i] = sourceOf[(size_t)t.y * (unsigned)tw + (unsigned)t.x];
}

hipError_t learn_sample_pairs(hipStream_t s, const int2 *targets, unsigned nT, const int2 *sourceOf, int tw, int n,
                              uint64_t key, int2 *pts, int2 *src)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample_pairs, dim3((n + 255) / 256), dim3(256), 0, s, targets, nT, sourceOf, tw, n, key, pts,
                     src);
  return hipGetLastError();
}


}  // namespace rs

namespace rs {

// ------------------------------------------------------------------ two-level index
//
// S super-centroids (every (nK/S)-th key) split the keys; inside each super-cluster, Q
// sub-centroids (evenly spaced members of its bucket) split them again, giving C = S*Q
// clusters.  Assignment and query routing cost S + Q dot products per vector instead of C.

// Block per super-cluster s: sub-centroid q copies the key at position q*size/Q of s's
// bucket (key 0 when the bucket is empty).
__global__ void k_pick_subcentroids(const uint32_t *startS, const uint32_t *membersS, int Q, const uint16_t *keyZ,
                                    uint16_t *subCent)
{
  const int s = blockIdx.x;
  const uint32_t b = startS[s], size = startS[s + 1] - b;
  for (int q = threadIdx.x; q < Q; q += blockDim.x) {
    const uint32_t key = size ? membersS[b + (uint32_t)(((uint64_t)q * size) / (uint64_t)Q)] : 0u;
    const uint4 *src = (const uint4 *)(keyZ + (size_t)key * kEmb);
    uint4 *dst = (uint4 *)(subCent + ((size_t)s * Q + q) * kEmb);
    for (int g = 0; g < 8; g++) dst[g] = src[g];
  }
}

// Tiles of 64 keys that do not straddle super-clusters: tileStart[s] = first tile of s.
__global__ void k_super_tiles(const uint32_t *startS, int S, uint32_t *tileStart)
{
  if (threadIdx.x != 0) return;
  uint32_t t = 0;
  for (int s = 0; s < S; s++) {
    tileStart[s] = t;
    t += (startS[s + 1] - startS[s] + 63) / 64;
  }
  tileStart[S] = t;
}

// Block per tile of 64 keys of one super-cluster: scores against its Q sub-centroids on
// MFMA (the key rows gathered through the bucket), keeps each key's best sub-centroid and
// writes cluster = s*Q + q.
__global__ __launch_bounds__(256) void k_sub_assign(const uint32_t *startS, const uint32_t *membersS,
                                                    const uint32_t *tileStart, int S, int Q, const uint16_t *keyZ,
                                                    const uint16_t *subCent, int32_t *cluster)
{
  __shared__ __attribute__((aligned(16))) uint16_t As[GT][kEmb + GPAD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[GT][kEmb + GPAD];
  __shared__ float sV[4][32];
  __shared__ int sI[4][32];
  __shared__ uint32_t sKey[GT];
  const uint32_t t = blockIdx.x;
  if (t >= tileStart[S]) return;
  int lo = 0, hi = S;  // the super-cluster holding tile t
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (tileStart[mid] <= t) lo = mid;
    else hi = mid;
  }
  const int s = lo;
  const uint32_t base = startS[s] + (t - tileStart[s]) * 64u, end = startS[s + 1];
  const int rows = (int)min(64u, end - base);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  if (tid < GT) sKey[tid] = tid < rows ? membersS[base + tid] : 0u;
  __syncthreads();
  // A: 64 gathered key rows x 64 dims (two k-steps kept in LDS)
  {
    const int r = tid >> 2, c = (tid & 3) * 16;
    uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0;
    if (r < rows) {
      const uint4 *src = (const uint4 *)(keyZ + (size_t)sKey[r] * kEmb + c);
      v0 = src[0];
      v1 = src[1];
    }
    *(uint4 *)&As[r][c] = v0;
    *(uint4 *)&As[r][c + 8] = v1;
  }
  float bv[8];
  int bi[8];
  for (int q = 0; q < 8; q++) {
    bv[q] = -3.4e38f;
    bi[q] = -1;
  }
  const uint16_t *cent = subCent + (size_t)s * Q * kEmb;
  for (int n0 = 0; n0 < Q; n0 += GT) {
    __syncthreads();
    {
      const int r = tid >> 2, c = (tid & 3) * 16;
      uint4 v0 = make_uint4(0, 0, 0, 0), v1 = v0;
      if (n0 + r < Q) {
        const uint4 *src = (const uint4 *)(cent + (size_t)(n0 + r) * kEmb + c);
        v0 = src[0];
        v1 = src[1];
      }
      *(uint4 *)&Bs[r][c] = v0;
      *(uint4 *)&Bs[r][c + 8] = v1;
    }
    __syncthreads();
    f32x4 acc[2][2];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int rr = lane & 15, kq = (lane >> 4) * 8;
    for (int k0 = 0; k0 < kEmb; k0 += 32) {
      bf16x8 fa[2], fb[2];
      for (int a = 0; a < 2; a++) fa[a] = *(const bf16x8 *)&As[wm + a * 16 + rr][k0 + kq];
      for (int b = 0; b < 2; b++) fb[b] = *(const bf16x8 *)&Bs[wn + b * 16 + rr][k0 + kq];
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
    }
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        const int n = n0 + wn + b * 16 + (lane & 15);
        if (n >= Q) continue;
        for (int r = 0; r < 4; r++)
          if (acc[a][b][r] > bv[a * 4 + r]) {
            bv[a * 4 + r] = acc[a][b][r];
            bi[a * 4 + r] = n;
          }
      }
  }
  for (int o = 1; o < 16; o <<= 1)
    for (int q = 0; q < 8; q++) {
      const float ov = __shfl_xor(bv[q], o);
      const int oi = __shfl_xor(bi[q], o);
      if (ov > bv[q] || (ov == bv[q] && oi >= 0 && (bi[q] < 0 || oi < bi[q]))) {
        bv[q] = ov;
        bi[q] = oi;
      }
    }
  if ((lane & 15) == 0)
    for (int a = 0; a < 2; a++)
      for (int r = 0; r < 4; r++) {
        const int rl = a * 16 + (lane >> 4) * 4 + r;
        sV[wave][rl] = bv[a * 4 + r];
        sI[wave][rl] = bi[a * 4 + r];
      }
  __syncthreads();
  if (tid < GT && tid < rows) {
    const int w0 = (tid >= 32) ? 2 : 0, rl = tid & 31;
    const int q = (sV[w0 + 1][rl] > sV[w0][rl]) ? sI[w0 + 1][rl] : sI[w0][rl];
    cluster[sKey[tid]] = s * Q + (q < 0 ? 0 : q);
  }
}

hipError_t learn_sub_index(hipStream_t s, const uint32_t *startS, const uint32_t *membersS, int S, int Q, int nK,
                           const uint16_t *keyZ, uint16_t *subCent, uint32_t *tileStart, int32_t *cluster)
{
  hipLaunchKernelGGL(k_pick_subcentroids, dim3(S), dim3(256), 0, s, startS, membersS, Q, keyZ, subCent);
  hipLaunchKernelGGL(k_super_tiles, dim3(1), dim3(64), 0, s, startS, S, tileStart);
  const unsigned tiles = (unsigned)((nK + 63) / 64 + S);
  hipLaunchKernelGGL(k_sub_assign, dim3(tiles), dim3(256), 0, s, startS, membersS, tileStart, S, Q, keyZ, subCent,
                     cluster);
  return hipGetLastError();
}


// Bias gradients: column sums of a bf16 [rows x cols] matrix into out (zeroed by the caller),
// 64 columns x a chunk of rows per block.
__global__ __launch_bounds__(256) void k_colsum_bf16_par(const uint16_t *g, int rows, int cols, float *out)
{
  __shared__ float part[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), rg = threadIdx.x >> 6;
  const int r0 = blockIdx.y * 256, r1 = min(rows, r0 + 256);
  float s = 0.f;
  if (c < cols)
    for (int r = r0 + rg; r < r1; r += 4) s += bf2f(g[(size_t)r * cols + c]);
  part[rg][threadIdx.x & 63] = s;
  __syncthreads();
  if (rg == 0 && c < cols) atomicAdd(&out[c], part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                                                  part[3][threadIdx.x]);
}

}  // namespace rs

namespace rs {

// ------------------------------------------------------------------ fused embedding
//
// The fused embedding takes 64-row blocks through all three layers with the activations in
// LDS; the weights come as transposed bf16 copies ([out x in], k-contiguous) so that every
// MFMA B operand is one 16-byte load.  Four waves split each layer's output columns.
constexpr int kRowsE = 64, kLdsW = 264;   // rows per block, LDS row stride (bf16)

__global__ void k_transpose_w(const float *w, int in, int out, uint16_t *wt)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= in * out) return;
  const int r = i / out, c = i % out;   // w[r][c] -> wt[c][r]
  wt[(size_t)c * in + r] = f2bf(w[i]);
}

hipError_t transpose_weights(hipStream_t s, const float *w, uint16_t *wt)
{
  const size_t OW1 = 0, OB1 = OW1 + (size_t)kIn * kH1, OW2 = OB1 + kH1, OB2 = OW2 + (size_t)kH1 * kH2,
               OW3 = OB2 + kH2;
  hipLaunchKernelGGL(k_transpose_w, dim3((kIn * kH1 + 255) / 256), dim3(256), 0, s, w + OW1, kIn, kH1, wt);
  hipLaunchKernelGGL(k_transpose_w, dim3((kH1 * kH2 + 255) / 256), dim3(256), 0, s, w + OW2, kH1, kH2,
                     wt + (size_t)kIn * kH1);
  hipLaunchKernelGGL(k_transpose_w, dim3((kH2 * kEmb + 255) / 256), dim3(256), 0, s, w + OW3, kH2, kEmb,
                     wt + (size_t)kIn * kH1 + (size_t)kH1 * kH2);
  return hipGetLastError();
}

// Two floats rounded to bf16 (nearest even) and packed, low half first.
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi)
{
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  bf16x2 v;
  v.x = (__bf16)lo;
  v.y = (__bf16)hi;
  return __builtin_bit_cast(uint32_t, v);
}

// One layer on 64 rows: out[64 x N] = act(in[64 x K] W + b); this wave computes output
// features [wave*N/4, (wave+1)*N/4), CB = N/64 blocks of 16, from its slice of the weights:
// weight(ks, b, feature, kq) returns a lane's A operand for k-step ks and feature block b,
// from registers or from the layer's transposed weights in LDS.  The weights are the MFMA A
// operand and the rows the B operand, so a lane ends up with 4 consecutive features of one
// row and writes them with one 8-byte (16-byte for the fp32 last layer) LDS store.  The
// row fragments are read from LDS kPipe steps ahead of the MFMAs that use them, so that LDS
// latency hides behind the matrix cores (one wave per SIMD has no other wave to switch to).
constexpr int kPipe = 3;

template <int K, int N, bool RELU, typename Weight>
__device__ __forceinline__ void mlp_layer(const uint16_t (*in)[kLdsW], Weight weight, const float *bias,
                                          uint16_t (*out)[kLdsW], float (*outF)[kEmb + 4])
{
  constexpr int CB = N / 64, KS = K / 32, STEPS = KS * 4;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c0 = wave * (N / 4);
  const int rr = lane & 15, kq = (lane >> 4) * 8, f4 = (lane >> 4) * 4;
  f32x4 acc[4][CB];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < CB; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  // step st = ks * 4 + a: rows a*16.., k-step ks
  bf16x8 fx[kPipe];
#pragma unroll
  for (int st = 0; st < kPipe; st++) fx[st] = *(const bf16x8 *)&in[(st & 3) * 16 + rr][(st >> 2) * 32 + kq];
#pragma unroll
  for (int st = 0; st < STEPS; st++) {
    const int ks = st >> 2, a = st & 3;
    const bf16x8 cur = fx[st % kPipe];
    if (st + kPipe < STEPS)
      fx[st % kPipe] = *(const bf16x8 *)&in[((st + kPipe) & 3) * 16 + rr][((st + kPipe) >> 2) * 32 + kq];
#pragma unroll
    for (int b = 0; b < CB; b++)
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(weight(ks, b, c0 + b * 16 + rr, kq), cur, acc[a][b], 0, 0, 0);
  }
  // acc[a][b][r]: feature c0 + b*16 + f4 + r of row a*16 + rr
#pragma unroll
  for (int b = 0; b < CB; b++) {
    const int col = c0 + b * 16 + f4;
    const float4 bb = *(const float4 *)(bias + col);
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const int row = a * 16 + rr;
      float v0 = acc[a][b][0] + bb.x, v1 = acc[a][b][1] + bb.y, v2 = acc[a][b][2] + bb.z, v3 = acc[a][b][3] + bb.w;
      if (RELU) {
        v0 = v0 > 0.f ? v0 : 0.f;
        v1 = v1 > 0.f ? v1 : 0.f;
        v2 = v2 > 0.f ? v2 : 0.f;
        v3 = v3 > 0.f ? v3 : 0.f;
        *(uint2 *)&out[row][col] = make_uint2(pack_bf16(v0, v1), pack_bf16(v2, v3));
      } else {
        *(float4 *)&outF[row][col] = make_float4(v0, v1, v2, v3);
      }
    }
  }
}

// Loads this wave's slice of one layer's transposed weights (wt: [N x K] bf16) into fw.
template <int K, int N>
__device__ __forceinline__ void load_layer(const uint16_t *wt, bf16x8 (&fw)[K / 32][N / 64])
{
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int c0 = wave * (N / 4), rr = lane & 15, kq = (lane >> 4) * 8;
#pragma unroll
  for (int ks = 0; ks < K / 32; ks++)
#pragma unroll
    for (int cb = 0; cb < N / 64; cb++) fw[ks][cb] = *(const bf16x8 *)(wt + (size_t)(c0 + cb * 16 + rr) * K + ks * 32 + kq);
}

// k_mlp_embed: persistent, one 4-wave workgroup per compute unit.  Each wave loads its
// column slices of the first two layers' weights into registers once (192 VGPRs of bf16
// operands; the third layer's 32 KB sit in LDS, which leaves registers to read the row
// fragments ahead) and then takes 64-row blocks in turn: the rows go through the three
// layers and the normalisation with the activations in LDS (two ping-pong buffers), and the
// next block's input rows are fetched into registers while this one computes.  Re-reading
// the 229 KB of weights from L2 for every 64 rows left the matrix cores idle 88% of the time.
__global__ __launch_bounds__(256, 1) void k_mlp_embed(const uint16_t *x, int n, const uint16_t *wt, const float *w,
                                                      uint16_t *zb, float *z)
{
  __shared__ __attribute__((aligned(16))) uint16_t bufA[kRowsE][kLdsW];
  __shared__ __attribute__((aligned(16))) uint16_t bufB[kRowsE][kLdsW];
  __shared__ __attribute__((aligned(16))) uint16_t w3s[kEmb][kLdsW];
  const size_t OB1 = (size_t)kIn * kH1, OW2 = OB1 + kH1, OB2 = OW2 + (size_t)kH1 * kH2, OW3 = OB2 + kH2,
               OB3 = OW3 + (size_t)kH2 * kEmb;
  const uint16_t *wt1 = wt, *wt2 = wt + (size_t)kIn * kH1, *wt3 = wt2 + (size_t)kH1 * kH2;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  bf16x8 w1[kIn / 32][kH1 / 64], w2[kH1 / 32][kH2 / 64];
  load_layer<kIn, kH1>(wt1, w1);
  load_layer<kH1, kH2>(wt2, w2);
  // third layer [kEmb x kH2] transposed, 16 bytes a thread per step
  for (int i = tid; i < kEmb * kH2 / 8; i += 256) {
    const int r = i / (kH2 / 8), c = (i % (kH2 / 8)) * 8;
    *(uint4 *)&w3s[r][c] = *(const uint4 *)(wt3 + (size_t)r * kH2 + c);
  }
  const int nBlocks = (n + kRowsE - 1) / kRowsE;
  // input rows of a block: 64 x 128 bf16 = 1024 uint4, four per thread
  uint4 xin[4];
  auto fetch = [&](int blk) {
    for (int u = 0; u < 4; u++) {
      const int i = tid + u * 256, r = i >> 4, c = (i & 15) * 8;
      const int row = blk * kRowsE + r;
      xin[u] = (blk < nBlocks && row < n) ? *(const uint4 *)(x + (size_t)row * kIn + c) : make_uint4(0, 0, 0, 0);
    }
  };
  auto stage = [&]() {
    for (int u = 0; u < 4; u++) {
      const int i = tid + u * 256, r = i >> 4, c = (i & 15) * 8;
      *(uint4 *)&bufA[r][c] = xin[u];
    }
  };
  int blk = blockIdx.x;
  fetch(blk);
  stage();
  __syncthreads();
  float (*eF)[kEmb + 4] = (float (*)[kEmb + 4])&bufB[0][0];   // 64 x 68 floats fit in bufB
  for (; blk < nBlocks; blk += gridDim.x) {
    fetch(blk + (int)gridDim.x);
    mlp_layer<kIn, kH1, true>(bufA, [&](int ks, int b, int, int) { return w1[ks][b]; }, w + OB1, bufB, nullptr);
    __syncthreads();
    mlp_layer<kH1, kH2, true>(bufB, [&](int ks, int b, int, int) { return w2[ks][b]; }, w + OB2, bufA, nullptr);
    __syncthreads();
    mlp_layer<kH2, kEmb, false>(
        bufA, [&](int ks, int, int f, int kq) { return *(const bf16x8 *)&w3s[f][ks * 32 + kq]; }, w + OB3, nullptr,
        eF);
    __syncthreads();
    stage();   // bufA is free: the next block's rows go in while eF is normalised
    // normalise: wave w takes rows w*16 .. w*16+15 all at once, four lanes per row with 16
    // columns each (a lane per column and a 64-lane reduction per row would serialise 16
    // reductions at one wave per SIMD)
    {
      const int r = wave * 16 + (lane >> 2), seg = (lane & 3) * 16;
      const int row = blk * kRowsE + r;
      float v[16];
      float s2 = 0.f;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const float4 f = *(const float4 *)&eF[r][seg + 4 * q];
        v[4 * q] = f.x;
        v[4 * q + 1] = f.y;
        v[4 * q + 2] = f.z;
        v[4 * q + 3] = f.w;
      }
#pragma unroll
      for (int c = 0; c < 16; c++) s2 += v[c] * v[c];
      s2 += __shfl_xor(s2, 1);
      s2 += __shfl_xor(s2, 2);
      const float nrm = sqrtf(fmaxf(s2, 1e-12f));
      if (row < n) {
#pragma unroll
        for (int c = 0; c < 16; c++) v[c] = v[c] / nrm;
        if (zb) {
          uint32_t pk[8];
#pragma unroll
          for (int c = 0; c < 8; c++) pk[c] = (uint32_t)f2bf(v[2 * c]) | ((uint32_t)f2bf(v[2 * c + 1]) << 16);
          uint4 *dst = (uint4 *)(zb + (size_t)row * kEmb + seg);
          dst[0] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
          dst[1] = make_uint4(pk[4], pk[5], pk[6], pk[7]);
        }
        if (z) {
          float4 *dst = (float4 *)(z + (size_t)row * kEmb + seg);
#pragma unroll
          for (int q = 0; q < 4; q++) dst[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
        }
      }
    }
    __syncthreads();
  }
}

static int compute_units()
{
  static const int cus = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
      return 256;
    return v;
  }();
  return cus;
}

hipError_t mlp_embed_fused(hipStream_t s, const float *w, const uint16_t *wt, const uint16_t *x, int n, uint16_t *zb,
                           float *z, int cuShare)
{
  if (n <= 0) return hipSuccess;
  const int cus = compute_units();
  const int share = cuShare > 0 && cuShare < 1000 ? std::max(1, cus * cuShare / 1000) : cus;
  const int blocks = std::min((n + kRowsE - 1) / kRowsE, share);
  hipLaunchKernelGGL(k_mlp_embed, dim3(blocks), dim3(256), 0, s, x, n, wt, w, zb, z);
  return hipGetLastError();
}

}  // namespace rs

namespace rs {

// ------------------------------------------------------------------ sorted buckets

// keys[i] = (u32) cluster[i * stride] (negative ids sort last), vals[i] = i.
__global__ void k_sort_keys(const int32_t *cluster, int stride, int n, uint32_t *keys, uint32_t *vals)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint32_t)cluster[(size_t)i * stride];
  vals[i] = (uint32_t)i;
}

hipError_t learn_sort_keys(hipStream_t s, const int32_t *cluster, int stride, int n, uint32_t *keys, uint32_t *vals)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sort_keys, dim3((n + 255) / 256), dim3(256), 0, s, cluster, stride, n, keys, vals);
  return hipGetLastError();
}

// start[c] = first position of cluster c in the sorted ids (n for c = C).
__global__ void k_bucket_starts(const uint32_t *sorted, int n, int C, uint32_t *start)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > C) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (sorted[mid] < (uint32_t)c) lo = mid + 1;
    else hi = mid;
  }
  start[c] = (uint32_t)lo;
}

hipError_t learn_bucket_starts(hipStream_t s, const uint32_t *sorted, int n, int C, uint32_t *start)
{
  hipLaunchKernelGGL(k_bucket_starts, dim3((C + 256) / 256), dim3(256), 0, s, sorted, n, C, start);
  return hipGetLastError();
}

// Keys and their corpus points copied into cluster order.
__global__ void k_permute_keys(const uint32_t *members, int n, const uint16_t *keyZ, const int2 *keyPts,
                               uint16_t *keyZs, int2 *keyPtsS)
{
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)n * 8) return;
  const int i = (int)(t >> 3), g = (int)(t & 7);
  const uint32_t m = members[i];
  ((uint4 *)(keyZs + (size_t)i * kEmb))[g] = ((const uint4 *)(keyZ + (size_t)m * kEmb))[g];
  if (g == 0) keyPtsS[i] = keyPts[m];
}

hipError_t learn_permute_keys(hipStream_t s, const uint32_t *members, int n, const uint16_t *keyZ, const int2 *keyPts,
                              uint16_t *keyZs, int2 *keyPtsS)
{
  if (n <= 0) return hipSuccess;
  const size_t work = (size_t)n * 8;
  hipLaunchKernelGGL(k_permute_keys, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, members, n, keyZ, keyPts,
                     keyZs, keyPtsS);
  return hipGetLastError();
}

// Top-N search on the matrix cores, shared by the query routing (a query's best sub-clusters
// inside its super-cluster) and the fine search (a query's best keys inside its clusters).
// One wavefront takes 16 consecutive queries of `order`, which is sorted by the cluster
// searched, so the tile's queries mostly share one cluster and a cluster's rows are read
// by consecutive waves from L2.  For each distinct cluster among the tile's queries it
// walks the cluster's rows 16 at a time: two v_mfma_f32_16x16x32_bf16 steps (64 dimensions)
// give the 16 x 16 dot products, and lane l holds those of queries 4*(l/16)+r against row
// column l%16.  Each lane keeps, per query, the best LT rows of its column slot (a sorted
// list, branch-free insertion); the 16 lanes of a query's row group then merge their lists
// into the query's top N.  The merge is exact unless more than LT of a query's top N fall
// into one column slot (never for N <= LT; probability about 3e-6 for N = 16, LT = 8).
template <int LT>
__device__ __forceinline__ void lane_top_insert(float (&v)[LT], int (&id)[LT], float d, int key)
{
  const bool c = d > v[LT - 1];
  v[LT - 1] = c ? d : v[LT - 1];
  id[LT - 1] = c ? key : id[LT - 1];
  for (int j = LT - 1; j > 0; j--) {
    const bool up = v[j] > v[j - 1];
    const float tv = v[j - 1];
    const int ti = id[j - 1];
    v[j - 1] = up ? v[j] : tv;
    id[j - 1] = up ? id[j] : ti;
    v[j] = up ? tv : v[j];
    id[j] = up ? ti : id[j];
  }
}

// Cluster c holds rows [start[c], start[c+1]), or [c*fixedLen, (c+1)*fixedLen) when start
// is null.  POINTS: writes the top 16 rows' points (keyPts * 2, or (-1, -1)) to
// props[q*16 ..]; otherwise writes the top `topn` row indices (or -1) to ids[q*topn ..].
// rowCl[q*P + p] is the p-th cluster of query q (negative: none; in the routing mode a
// negative super-cluster counts as 0).
template <int LT, bool POINTS>
__global__ __launch_bounds__(256) void k_topn_mfma(const uint16_t *qz, int nq, const uint32_t *order,
                                                   const int32_t *rowCl, int P, const uint16_t *rows,
                                                   const uint32_t *start, int fixedLen, int topn,
                                                   const int2 *keyPts, int2 *props, int32_t *ids)
{
  constexpr float kNone = -3.4e38f;
  const int lane = threadIdx.x & 63;
  const int t0 = (int)(blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (t0 >= nq) return;  // uniform per wavefront: the MFMA steps below run with every lane
  const int col = lane & 15, grp = lane >> 4, kq = grp * 8;
  const uint32_t qa = order[min(t0 + col, nq - 1)];
  const bf16x8 a0 = *(const bf16x8 *)(qz + (size_t)qa * kEmb + kq);
  const bf16x8 a1 = *(const bf16x8 *)(qz + (size_t)qa * kEmb + 32 + kq);
  int rq[4];
  for (int r = 0; r < 4; r++) {
    const int i = t0 + grp * 4 + r;
    rq[r] = i < nq ? (int)order[i] : -1;
  }
  float v[4][LT];
  int id[4][LT];
  for (int r = 0; r < 4; r++)
    for (int j = 0; j < LT; j++) {
      v[r][j] = kNone;
      id[r][j] = -1;
    }
  for (int p = 0; p < P; p++) {
    int rc[4];
    unsigned done = 0;
    for (int r = 0; r < 4; r++) {
      int c = rq[r] >= 0 ? rowCl[(size_t)rq[r] * P + p] : -1;
      if (!POINTS && rq[r] >= 0 && c < 0) c = 0;
      rc[r] = c;
      if (c < 0) done |= 1u << r;
    }
    for (;;) {
      int c = 0x7FFFFFFF;
      for (int r = 0; r < 4; r++)
        if (!((done >> r) & 1u)) c = min(c, rc[r]);
      c = min(c, __shfl_xor(c, 16));
      c = min(c, __shfl_xor(c, 32));
      if (c == 0x7FFFFFFF) break;
      bool match[4];
      for (int r = 0; r < 4; r++) {
        match[r] = !((done >> r) & 1u) && rc[r] == c;
        if (match[r]) done |= 1u << r;
      }
      const uint32_t b = start ? start[c] : (uint32_t)c * (uint32_t)fixedLen;
      const uint32_t e = start ? start[c + 1] : b + (uint32_t)fixedLen;
      if (b >= e) continue;
      uint32_t kk = min(b + col, e - 1);
      bf16x8 b0 = *(const bf16x8 *)(rows + (size_t)kk * kEmb + kq);
      bf16x8 b1 = *(const bf16x8 *)(rows + (size_t)kk * kEmb + 32 + kq);
      for (uint32_t m0 = b; m0 < e; m0 += 16) {
        // the next tile's operands (clamped into the cluster, so always in bounds)
        kk = min(m0 + 16 + col, e - 1);
        const bf16x8 n0 = *(const bf16x8 *)(rows + (size_t)kk * kEmb + kq);
        const bf16x8 n1 = *(const bf16x8 *)(rows + (size_t)kk * kEmb + 32 + kq);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc, 0, 0, 0);
        const uint32_t key = m0 + col;
        const bool inside = key < e;
        for (int r = 0; r < 4; r++) lane_top_insert<LT>(v[r], id[r], (inside && match[r]) ? acc[r] : kNone, (int)key);
        b0 = n0;
        b1 = n1;
      }
    }
  }
  // Merge: topn rounds per query; the lane whose list head is the row group's maximum (the
  // lowest such lane on a tie) hands it over and pops it.
  const uint64_t groupMask = 0xFFFFull << (grp * 16);
  int out[4] = {-1, -1, -1, -1};
  for (int r = 0; r < 4; r++) {
    for (int k = 0; k < topn; k++) {
      float best = v[r][0];
      best = fmaxf(best, __shfl_xor(best, 1));
      best = fmaxf(best, __shfl_xor(best, 2));
      best = fmaxf(best, __shfl_xor(best, 4));
      best = fmaxf(best, __shfl_xor(best, 8));
      const uint64_t hit = __ballot(v[r][0] == best) & groupMask;
      const int winner = (int)__builtin_ctzll(hit);
      const int key = __shfl(id[r][0], winner);
      if (col == k) out[r] = key;
      if (lane == winner) {
        for (int j = 0; j + 1 < LT; j++) {
          v[r][j] = v[r][j + 1];
          id[r][j] = id[r][j + 1];
        }
        v[r][LT - 1] = kNone;
        id[r][LT - 1] = -1;
      }
    }
  }
  for (int r = 0; r < 4; r++) {
    if (rq[r] < 0) continue;
    const int k = out[r];
    if (POINTS)
      props[(size_t)rq[r] * 16 + col] = k >= 0 ? make_int2(keyPts[k].x * 2, keyPts[k].y * 2) : make_int2(-1, -1);
    else if (col < topn)
      ids[(size_t)rq[r] * topn + col] = k;
  }
}

// Grid-stride, so that at most one atomic per workgroup reaches the single total.
__global__ __launch_bounds__(256) void k_count_scored(const int32_t *qcl, int n, int P, const uint32_t *start,
                                                      unsigned long long *total)
{
  __shared__ unsigned long long part[4];
  unsigned long long v = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    for (int p = 0; p < P; p++) {
      const int c = qcl[(size_t)i * P + p];
      if (c >= 0) v += start[c + 1] - start[c];
    }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(total, t);
  }
}

hipError_t learn_count_scored(hipStream_t s, const int32_t *qcl, int n, int P, const uint32_t *start,
                              unsigned long long *total)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_scored, dim3(std::min((n + 255) / 256, 512)), dim3(256), 0, s, qcl, n, P, start, total);
  return hipGetLastError();
}

hipError_t learn_query_sub(hipStream_t s, const uint16_t *qz, int nq, const uint32_t *order, const int32_t *qsuper,
                           int Q, const uint16_t *subCent, int P, int32_t *qcl)
{
  if (nq <= 0) return hipSuccess;
  if (P < 1 || P > 2) return hipErrorInvalidValue;
  const unsigned tiles = (unsigned)((nq + 15) / 16);
  hipLaunchKernelGGL((k_topn_mfma<2, false>), dim3((tiles + 3) / 4), dim3(256), 0, s, qz, nq, order, qsuper, 1,
                     subCent, nullptr, Q, P, nullptr, nullptr, qcl);
  return hipGetLastError();
}

hipError_t learn_fine_top16_sorted(hipStream_t s, const uint16_t *qz, int nq, const uint32_t *order,
                                   const int32_t *qcl, int P, const uint16_t *keyZs, const uint32_t *start,
                                   const int2 *keyPtsS, int2 *props)
{
  if (nq <= 0) return hipSuccess;
  const unsigned tiles = (unsigned)((nq + 15) / 16);
  hipLaunchKernelGGL((k_topn_mfma<8, true>), dim3((tiles + 3) / 4), dim3(256), 0, s, qz, nq, order, qcl, P, keyZs,
                     start, 0, 16, keyPtsS, props, nullptr);
  return hipGetLastError();
}

}  // namespace rs
