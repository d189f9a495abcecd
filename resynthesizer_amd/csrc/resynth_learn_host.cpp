// Host side of the learned proposals: the model object (FP32 master weights, BF16 copies,
// Adam state) and its C ABI (include/resynth_mi355x.h), training loops and test entries.
#include "../../include/resynth_mi355x.h"
#include "resynth_internal.h"
#include "resynth_learn.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

using namespace rs;

// A failed HIP call unwinds to the C entry point, which returns RESYNTH_ERROR_DEVICE (or
// NULL) instead of ending the caller's process.
#define LHIP(expr)                                                                                    \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) hip_fail(e_, #expr, __FILE__, __LINE__);                                    \
  } while (0)

struct ResynthModel {
  hipStream_t s = nullptr;
  float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
  uint16_t *wb = nullptr;
  uint16_t *wt = nullptr;   // transposed bf16 weights for the fused embedding kernel
  int step = 0;
  int version = 0, wtVersion = -1;   // parameter version and the version wt was made from
  // training scratch for batches of up to cap rows
  int cap = 0;
  LearnScratch q, k;
  uint16_t *bq = nullptr, *bk = nullptr;   // gathered batch inputs
  uint32_t *idx = nullptr;
  // logged training pairs (bf16 rows of kIn), grow-only
  uint16_t *pq = nullptr, *pk = nullptr;
  unsigned nPairs = 0, capPairs = 0;
  std::vector<void *> owned;
  std::mutex mu;

  void *alloc(size_t bytes)
  {
    void *p = nullptr;
    LHIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
    owned.push_back(p);
    return p;
  }
};

namespace {

// Device memory owned by one call.
struct DevTmp {
  void *p = nullptr;
  void *alloc(size_t bytes)
  {
    LHIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
    return p;
  }
  DevTmp() = default;
  DevTmp(const DevTmp &) = delete;
  DevTmp &operator=(const DevTmp &) = delete;
  ~DevTmp()
  {
    if (p) (void)hipFree(p);
  }
};

// Runs a model entry; a HIP failure becomes RESYNTH_ERROR_DEVICE after the model's stream
// has drained.
template <typename F>
int model_guarded(ResynthModel *M, F &&body)
{
  try {
    return body();
  } catch (const HipFailure &) {
    if (M && M->s) (void)hipStreamSynchronize(M->s);
    (void)hipGetLastError();
    return RESYNTH_ERROR_DEVICE;
  } catch (const std::bad_alloc &) {
    return RESYNTH_ERROR_DEVICE;
  }
}

void alloc_scratch(ResynthModel *M, LearnScratch &t, int n, bool query)
{
  t.h1 = (uint16_t *)M->alloc((size_t)n * kH1 * 2);
  t.h2 = (uint16_t *)M->alloc((size_t)n * kH2 * 2);
  t.e = (float *)M->alloc((size_t)n * kEmb * 4);
  t.z = (float *)M->alloc((size_t)n * kEmb * 4);
  t.zb = (uint16_t *)M->alloc((size_t)n * kEmb * 2);
  t.norm = (float *)M->alloc((size_t)n * 4);
  if (query) {
    t.L = (float *)M->alloc((size_t)n * n * 4);
    t.rowLoss = (float *)M->alloc((size_t)n * 4);
    t.dL = (uint16_t *)M->alloc((size_t)n * n * 2);
    t.dLt = (uint16_t *)M->alloc((size_t)n * n * 2);
  }
  t.dz = (float *)M->alloc((size_t)n * kEmb * 4);
  t.dEb = (uint16_t *)M->alloc((size_t)n * kEmb * 2);
  t.g = (float *)M->alloc((size_t)n * kH1 * 4);
  t.gb16 = (uint16_t *)M->alloc((size_t)n * kH1 * 2);
  t.gw = (float *)M->alloc((size_t)kH1 * kH2 * 4);
  t.gb = (float *)M->alloc(kH1 * 4);
}

// Training scratch for batches of n rows (grow-only; the old buffers stay owned).
void ensure_scratch(ResynthModel *M, int n)
{
  if (n <= M->cap) return;
  alloc_scratch(M, M->q, n, true);
  alloc_scratch(M, M->k, n, false);
  M->bq = (uint16_t *)M->alloc((size_t)n * kIn * 2);
  M->bk = (uint16_t *)M->alloc((size_t)n * kIn * 2);
  M->idx = (uint32_t *)M->alloc((size_t)n * 4);
  M->cap = n;
}

}  // namespace

namespace rs {

// He-uniform weights and zero biases from a seeded generator (deterministic per seed).
void model_init_params(unsigned seed, std::vector<float> &p)
{
  p.assign(kParamCount, 0.f);
  std::mt19937 gen(seed);
  size_t o = 0;
  const int dims[4] = {kIn, kH1, kH2, kEmb};
  for (int l = 0; l < 3; l++) {
    const int fin = dims[l], fout = dims[l + 1];
    const float a = std::sqrt(6.f / (float)(l == 0 ? 120 : fin));
    std::uniform_real_distribution<float> U(-a, a);
    for (int i = 0; i < fin; i++)
      for (int j = 0; j < fout; j++) p[o + (size_t)i * fout + j] = (l == 0 && i >= 120) ? 0.f : U(gen);
    o += (size_t)fin * fout;
    o += fout;  // biases stay zero
  }
}

// Device-resident training on pairs (xq[i], xk[i]), i < n, in bf16 [n x kIn].  Each step
// draws `batch` pairs by a counter-based hash of (seed, step).  losses (host) may be NULL.
int model_train_device(ResynthModel *M, const uint16_t *xq, const uint16_t *xk, unsigned n, unsigned steps,
                       unsigned batch, float lr, float tau, unsigned seed, float *losses)
{
  if (!batch || batch % 64 || !n) return -1;
  ensure_scratch(M, (int)batch);
  hipStream_t s = M->s;
  std::vector<float> hl(steps, 0.f);
  for (unsigned t = 0; t < steps; t++) {
    LHIP(hash_indices(s, M->idx, (int)batch, n, ((uint64_t)seed << 32) ^ (uint64_t)(M->step + 1)));
    LHIP(gather_rows_bf16(s, xq, kIn, M->idx, (int)batch, M->bq));
    LHIP(gather_rows_bf16(s, xk, kIn, M->idx, (int)batch, M->bk));
    LHIP(mlp_infonce(s, M->w, M->wb, M->bq, M->bk, (int)batch, tau, M->q, M->k, M->grad, M->loss + t % 1024));
    M->step++;
    M->version++;
    LHIP(adam_step(s, M->w, M->grad, M->m, M->v, M->wb, lr, M->step));
    if (losses && (t % 1024 == 1023 || t + 1 == steps)) {
      const unsigned base = t - t % 1024, cnt = t - base + 1;
      LHIP(hipMemcpyAsync(hl.data() + base, M->loss, cnt * sizeof(float), hipMemcpyDeviceToHost, s));
      LHIP(hipStreamSynchronize(s));
    }
  }
  LHIP(hipStreamSynchronize(s));
  if (losses) std::copy(hl.begin(), hl.end(), losses);
  return 0;
}

// Room for n more pairs; returns where to write them (device).  The store grows by copying.
void model_reserve_pairs(ResynthModel *M, unsigned n, uint16_t **pq, uint16_t **pk)
{
  if (M->nPairs + n > M->capPairs) {
    const unsigned cap = std::max(M->nPairs + n, M->capPairs * 2);
    uint16_t *q = nullptr, *k = nullptr;
    LHIP(hipMalloc(&q, (size_t)cap * kIn * 2));
    if (hipMalloc(&k, (size_t)cap * kIn * 2) != hipSuccess) {
      (void)hipFree(q);
      hip_fail(hipErrorOutOfMemory, "hipMalloc (pair store)", __FILE__, __LINE__);
    }
    if (M->nPairs) {
      LHIP(hipMemcpyAsync(q, M->pq, (size_t)M->nPairs * kIn * 2, hipMemcpyDeviceToDevice, M->s));
      LHIP(hipMemcpyAsync(k, M->pk, (size_t)M->nPairs * kIn * 2, hipMemcpyDeviceToDevice, M->s));
    }
    LHIP(hipStreamSynchronize(M->s));
    if (M->pq) LHIP(hipFree(M->pq));
    if (M->pk) LHIP(hipFree(M->pk));
    M->pq = q;
    M->pk = k;
    M->capPairs = cap;
  }
  *pq = M->pq + (size_t)M->nPairs * kIn;
  *pk = M->pk + (size_t)M->nPairs * kIn;
}

void model_commit_pairs(ResynthModel *M, unsigned n) { M->nPairs += n; }

hipStream_t model_stream(const ResynthModel *M) { return M->s; }
const float *model_weights(const ResynthModel *M) { return M->w; }
const uint16_t *model_weights_bf16(const ResynthModel *M) { return M->wb; }

// Transposed weights, refreshed on stream s when the parameters changed since last time.
const uint16_t *model_weights_t(const ResynthModel *Mc, hipStream_t s)
{
  ResynthModel *M = const_cast<ResynthModel *>(Mc);
  if (M->wtVersion != M->version) {
    LHIP(hipStreamSynchronize(M->s));
    LHIP(transpose_weights(s, M->w, M->wt));
    M->wtVersion = M->version;
  }
  return M->wt;
}

}  // namespace rs

extern "C" {

unsigned int resynth_model_param_count(void) { return (unsigned)kParamCount; }

ResynthModel *resynth_model_create(unsigned int seed)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return nullptr;
  ResynthModel *M = new ResynthModel();
  try {
    LHIP(hipStreamCreateWithFlags(&M->s, hipStreamNonBlocking));
    M->w = (float *)M->alloc(kParamCount * 4);
    M->m = (float *)M->alloc(kParamCount * 4);
    M->v = (float *)M->alloc(kParamCount * 4);
    M->grad = (float *)M->alloc(kParamCount * 4);
    M->loss = (float *)M->alloc(1024 * 4);
    M->wb = (uint16_t *)M->alloc(kParamCount * 2);
    M->wt = (uint16_t *)M->alloc(kParamCount * 2);
    std::vector<float> p;
    model_init_params(seed, p);
    LHIP(hipMemcpyAsync(M->w, p.data(), kParamCount * 4, hipMemcpyHostToDevice, M->s));
    LHIP(hipMemsetAsync(M->m, 0, kParamCount * 4, M->s));
    LHIP(hipMemsetAsync(M->v, 0, kParamCount * 4, M->s));
    LHIP(params_to_bf16(M->s, M->w, M->wb));
    LHIP(hipStreamSynchronize(M->s));
  } catch (const HipFailure &) {
    resynth_model_free(M);
    return nullptr;
  }
  return M;
}

void resynth_model_free(ResynthModel *M)
{
  if (!M) return;
  if (M->s) (void)hipStreamSynchronize(M->s);
  if (M->pq) (void)hipFree(M->pq);
  if (M->pk) (void)hipFree(M->pk);
  for (void *p : M->owned) (void)hipFree(p);
  if (M->s) (void)hipStreamDestroy(M->s);
  delete M;
}

int resynth_model_get_params(const ResynthModel *Mc, float *out)
{
  ResynthModel *M = const_cast<ResynthModel *>(Mc);
  if (!M || !out) return -1;
  return model_guarded(M, [&]() -> int {
    LHIP(hipMemcpyAsync(out, M->w, kParamCount * 4, hipMemcpyDeviceToHost, M->s));
    LHIP(hipStreamSynchronize(M->s));
    return 0;
  });
}

int resynth_model_set_params(ResynthModel *M, const float *in)
{
  if (!M || !in) return -1;
  std::lock_guard<std::mutex> lock(M->mu);
  return model_guarded(M, [&]() -> int {
    LHIP(hipMemcpyAsync(M->w, in, kParamCount * 4, hipMemcpyHostToDevice, M->s));
    LHIP(params_to_bf16(M->s, M->w, M->wb));
    LHIP(hipMemsetAsync(M->m, 0, kParamCount * 4, M->s));
    LHIP(hipMemsetAsync(M->v, 0, kParamCount * 4, M->s));
    M->step = 0;
    M->version++;
    LHIP(hipStreamSynchronize(M->s));
    return 0;
  });
}

int resynth_model_embed(const ResynthModel *Mc, const uint16_t *x, unsigned int n, float *out)
{
  ResynthModel *M = const_cast<ResynthModel *>(Mc);
  if (!M || !x || !out) return -1;
  if (!n) return 0;
  std::lock_guard<std::mutex> lock(M->mu);
  return model_guarded(M, [&]() -> int {
    hipStream_t s = M->s;
    DevTmp tx, tz;
    uint16_t *dx = (uint16_t *)tx.alloc((size_t)n * kIn * 2);
    float *dz = (float *)tz.alloc((size_t)n * kEmb * 4);
    LHIP(hipMemcpyAsync(dx, x, (size_t)n * kIn * 2, hipMemcpyHostToDevice, s));
    // the fused kernel the engine embeds with
    LHIP(mlp_embed_fused(s, M->w, model_weights_t(M, s), dx, (int)n, nullptr, dz));
    LHIP(hipMemcpyAsync(out, dz, (size_t)n * kEmb * 4, hipMemcpyDeviceToHost, s));
    LHIP(hipStreamSynchronize(s));
    return 0;
  });
}

int resynth_model_infonce(ResynthModel *M, const uint16_t *xq, const uint16_t *xk, unsigned int n, float tau,
                          float *loss, float *grad)
{
  if (!M || !xq || !xk || !n || n % 64) return -1;
  std::lock_guard<std::mutex> lock(M->mu);
  return model_guarded(M, [&]() -> int {
    ensure_scratch(M, (int)n);
    hipStream_t s = M->s;
    LHIP(hipMemcpyAsync(M->bq, xq, (size_t)n * kIn * 2, hipMemcpyHostToDevice, s));
    LHIP(hipMemcpyAsync(M->bk, xk, (size_t)n * kIn * 2, hipMemcpyHostToDevice, s));
    LHIP(mlp_infonce(s, M->w, M->wb, M->bq, M->bk, (int)n, tau, M->q, M->k, M->grad, M->loss));
    if (loss) LHIP(hipMemcpyAsync(loss, M->loss, 4, hipMemcpyDeviceToHost, s));
    if (grad) LHIP(hipMemcpyAsync(grad, M->grad, kParamCount * 4, hipMemcpyDeviceToHost, s));
    LHIP(hipStreamSynchronize(s));
    return 0;
  });
}

int resynth_model_train(ResynthModel *M, const uint16_t *xq, const uint16_t *xk, unsigned int n, unsigned int steps,
                        unsigned int batch, float lr, float tau, unsigned int seed, float *losses)
{
  if (!M || !xq || !xk || !n || !batch || batch % 64) return -1;
  std::lock_guard<std::mutex> lock(M->mu);
  return model_guarded(M, [&]() -> int {
    hipStream_t s = M->s;
    DevTmp tq, tk;
    uint16_t *dq = (uint16_t *)tq.alloc((size_t)n * kIn * 2);
    uint16_t *dk = (uint16_t *)tk.alloc((size_t)n * kIn * 2);
    LHIP(hipMemcpyAsync(dq, xq, (size_t)n * kIn * 2, hipMemcpyHostToDevice, s));
    LHIP(hipMemcpyAsync(dk, xk, (size_t)n * kIn * 2, hipMemcpyHostToDevice, s));
    const int rc = model_train_device(M, dq, dk, n, steps, batch, lr, tau, seed, losses);
    LHIP(hipStreamSynchronize(s));
    return rc;
  });
}

unsigned int resynth_model_pair_count(const ResynthModel *M) { return M ? M->nPairs : 0u; }

void resynth_model_clear_pairs(ResynthModel *M)
{
  if (M) M->nPairs = 0;
}

int resynth_model_fit(ResynthModel *M, unsigned int steps, unsigned int batch, float lr, float tau, unsigned int seed,
                      float *losses)
{
  if (!M || !M->nPairs) return -1;
  std::lock_guard<std::mutex> lock(M->mu);
  return model_guarded(M, [&]() -> int {
    return model_train_device(M, M->pq, M->pk, M->nPairs, steps, batch, lr, tau, seed, losses);
  });
}

int resynth_gemm_bf16(int ta, int tb, int M, int N, int K, const uint16_t *A, const uint16_t *B, float *C)
{
  if (M <= 0 || N <= 0 || K <= 0 || K % 32 || M % 8 || N % 8 || !A || !B || !C) return -1;
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return RESYNTH_ERROR_DEVICE;
  return model_guarded(nullptr, [&]() -> int {
    DevTmp ta_, tb_, tc_;
    uint16_t *dA = (uint16_t *)ta_.alloc((size_t)M * K * 2);
    uint16_t *dB = (uint16_t *)tb_.alloc((size_t)K * N * 2);
    float *dC = (float *)tc_.alloc((size_t)M * N * 4);
    LHIP(hipMemcpy(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice));
    LHIP(hipMemcpy(dB, B, (size_t)K * N * 2, hipMemcpyHostToDevice));
    const int lda = ta ? M : K, ldb = tb ? K : N;
    LHIP(learn_gemm(nullptr, ta, tb, M, N, K, dA, lda, dB, ldb, dC, N));
    LHIP(hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost));
    return 0;
  });
}

}  // extern "C"
