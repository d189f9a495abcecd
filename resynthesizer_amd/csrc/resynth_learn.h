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
// Transposed bf16 copies of the three weight matrices ([out x in] each, kParamCount-sized
// scratch minus biases), for the fused embedding kernel.
hipError_t transpose_weights(hipStream_t s, const float *w, uint16_t *wt);
// Fused embedding of n rows (bf16 x [n x kIn]) into normalised z (bf16 and/or fp32).  The
// persistent grid takes every compute unit (each of its workgroups fills one), or
// cuShare permille of them, which leaves the rest to concurrent work on other streams.
hipError_t mlp_embed_fused(hipStream_t s, const float *w, const uint16_t *wt, const uint16_t *x, int n, uint16_t *zb,
                           float *z, int cuShare = 0);
hipError_t gather_rows_bf16(hipStream_t s, const uint16_t *src, int cols, const uint32_t *idx, int n, uint16_t *dst);
hipError_t gather_spread_rows_bf16(hipStream_t s, const uint16_t *src, int cols, uint32_t nSrc, int n, uint16_t *dst);
hipError_t hash_indices(hipStream_t s, uint32_t *idx, int n, uint32_t range, uint64_t key);

// ---- the model object (resynth_learn_host.cpp)
}  // namespace rs
struct ResynthModel;
namespace rs {
const float *model_weights(const ResynthModel *M);
const uint16_t *model_weights_bf16(const ResynthModel *M);
const uint16_t *model_weights_t(const ResynthModel *M, hipStream_t s);
void model_reserve_pairs(ResynthModel *M, unsigned n, uint16_t **pq, uint16_t **pk);
void model_commit_pairs(ResynthModel *M, unsigned n);
hipStream_t model_stream(const ResynthModel *M);

// ---- neighbourhood features, index and search (resynth_learn.hip)

// Features of n points (x, y) of a pixel-word image (32- or 64-bit words, flag byte 0,
// colours in bytes 1..nColors) into bf16 rows [n x kIn].  offsets: the engine's sorted
// offsets (the first kNbr are used; offset 0 is the point itself and is never present).
// isCorpus: a neighbour is present when it is selected in the corpus (flag 0xFE/0xFF);
// otherwise (target) when its hasValue bit is set.  scale2: point coordinates are doubled
// (half-resolution key grid).  dropKey != 0: each neighbour of a target point is dropped
// with a per-point probability drawn from the hash of (dropKey, point).
hipError_t learn_features(hipStream_t s, bool wide, const void *words, int w, int h, int tileH, int tileV, int nColors,
                          const int2 *offsets, unsigned nOffsets, const int2 *pts, int n, bool isCorpus, bool scale2,
                          uint64_t dropKey, uint16_t *out);
// Flags of the half-resolution key grid: 1 where (2x, 2y) is a corpus point.
hipError_t learn_key_flags(hipStream_t s, bool wide, const void *cword, int cw, int ch, uint8_t *flags);
// Top-P (P = 1 or 2) columns by dot product for every row: rows [M x 64] and cols [N x 64]
// bf16 (MFMA GEMM with the selection fused in), out [M x P] column indices.
hipError_t learn_topk_cols(hipStream_t s, const uint16_t *rows, int M, const uint16_t *cols, int N, int P,
                           int32_t *out);
// Pairs for training: pts[i] = targets[hash(key, i) % nT] and src[i] = sourceOf of that
// target (row-major target image of width tw).
hipError_t learn_sample_pairs(hipStream_t s, const int2 *targets, unsigned nT, const int2 *sourceOf, int tw, int n,
                              uint64_t key, int2 *pts, int2 *src);
// Second level of the index: sub-centroids of every super-cluster (bucketed by startS /
// membersS) and each key's cluster s*Q + q (tileStart: S+1 scratch words).
hipError_t learn_sub_index(hipStream_t s, const uint32_t *startS, const uint32_t *membersS, int S, int Q, int nK,
                           const uint16_t *keyZ, uint16_t *subCent, uint32_t *tileStart, int32_t *cluster);
// Query routing inside the query's super-cluster: the best P (1 or 2) of its Q
// sub-centroids, as cluster ids s*Q + c.  order: the queries sorted by super-cluster.
hipError_t learn_query_sub(hipStream_t s, const uint16_t *qz, int nq, const uint32_t *order, const int32_t *qsuper,
                           int Q, const uint16_t *subCent, int P, int32_t *qcl);
// Adds to *total the keys the fine search will score: the summed sizes of every query's
// clusters.
hipError_t learn_count_scored(hipStream_t s, const int32_t *qcl, int n, int P, const uint32_t *start,
                              unsigned long long *total);
// Sort keys from cluster ids (every stride-th entry) with values 0..n-1.
hipError_t learn_sort_keys(hipStream_t s, const int32_t *cluster, int stride, int n, uint32_t *keys, uint32_t *vals);
// start[c] = first position of cluster c in the sorted ids (start[C] = n).
hipError_t learn_bucket_starts(hipStream_t s, const uint32_t *sorted, int n, int C, uint32_t *start);
// Key embeddings and their corpus points copied into cluster order.
hipError_t learn_permute_keys(hipStream_t s, const uint32_t *members, int n, const uint16_t *keyZ, const int2 *keyPts,
                              uint16_t *keyZs, int2 *keyPtsS);
// Fine search: for each query, the top-16 keys (by dot product) among the keys of its P
// clusters (keys in cluster order); writes their corpus coordinates (key grid point * 2)
// or (-1, -1).  order: the queries sorted by their first cluster.
hipError_t learn_fine_top16_sorted(hipStream_t s, const uint16_t *qz, int nq, const uint32_t *order,
                                   const int32_t *qcl, int P, const uint16_t *keyZs, const uint32_t *start,
                                   const int2 *keyPtsS, int2 *props);

}  // namespace rs
