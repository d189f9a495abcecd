// Declarations shared by the learned-proposal kernels (resynth_learn.hip) and their host
// side (resynth_learn_host.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace rs {

constexpr int kNbr = 30;         // neighbourhood: the 30 nearest offsets (offset 0 first)
constexpr int kIn = 128;         // 30 x (R, G, B, present) = 120 inputs, zero-padded to 128
constexpr int kH1 = 256, kH2 = 256, kEmb = 64;
constexpr int kParamCount = kIn * kH1 + kH1 + kH1 * kH2 + kH2 + kH2 * kEmb + kEmb;

// Device scratch of one tower for n rows (training batch or inference block).
struct LearnScratch {
  uint16_t *h1 = nullptr, *h2 = nullptr;    // bf16 [n x 256] activations
  float *e = nullptr, *z = nullptr;         // [n x 64] embedding, normalised
  uint16_t *zb = nullptr;                   // bf16 [n x 64]
  float *norm = nullptr;                    // [n]
  float *L = nullptr;                       // [n x n] logits (query tower only)
  float *rowLoss = nullptr;                 // [n]
  uint16_t *dL = nullptr, *dLt = nullptr;   // bf16 [n x n]
  float *dz = nullptr;                      // [n x 64]
  uint16_t *dEb = nullptr;                  // bf16 [n x 64]
  float *g = nullptr;                       // [n x 256]
  uint16_t *gb16 = nullptr;                 // bf16 [n x 256]
  float *gw = nullptr;                      // [256 x 256]
  float *gb = nullptr;                      // [256]
};

hipError_t learn_gemm(hipStream_t s, int ta, int tb, int M, int N, int K, const uint16_t *A, int lda,
                      const uint16_t *B, int ldb, float *C, int ldc);
hipError_t mlp_forward(hipStream_t s, const float *w, const uint16_t *wb, const uint16_t *x, int n, uint16_t *h1,
                       uint16_t *h2, float *e);
hipError_t mlp_embed(hipStream_t s, const float *w, const uint16_t *wb, const uint16_t *x, int n, LearnScratch &t,
                     uint16_t *zb, float *z);
hipError_t mlp_infonce(hipStream_t s, const float *w, const uint16_t *wb, const uint16_t *xq, const uint16_t *xk, int n,
                       float tau, LearnScratch &q, LearnScratch &k, float *grad, float *loss);
hipError_t adam_step(hipStream_t s, float *w, const float *g, float *m, float *v, uint16_t *wb, float lr, int step);
hipError_t params_to_bf16(hipStream_t s, const float *w, uint16_t *wb);

}  // namespace rs
