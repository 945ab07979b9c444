// gemm.hpp - batched FP64 matrix products on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// One launch evaluates a list of independent products (column-major, BLAS conventions)
//   C = alpha * (op(A) op(B) + op(A2) op(B2)) + beta * C        (the A2/B2 part only if K2 > 0)
// tiled in 64x64 blocks of C; work items are (descriptor, tile) pairs.
#pragma once
#include "common.hpp"

namespace nrg {

// One K-part of a grouped product: scale * op(A) op(B), op(A) M x K, op(B) K x N.
template <typename T>
struct GemmPartT {
  const T *A, *B;
  int lda, ldb, tA, tB, K;
  double scale;
};

template <typename T>
struct GemmDescT {
  const T *A, *B, *A2, *B2;
  T *C;
  int M, N, K, K2;
  int lda, ldb, lda2, ldb2, ldc;
  int tA, tB, tA2, tB2;        // 1: use the transpose
  double alpha, beta;          // beta == 0: C is not read
  const GemmPartT<T> *parts;   // if nparts > 0: C = alpha * sum_p scale_p op(A_p) op(B_p) + beta * C
  int nparts;
};
using GemmDesc = GemmDescT<double>;
using GemmPart = GemmPartT<double>;
using GemmDesc32 = GemmDescT<float>;

constexpr int GEMM_TILE = 64;

template <typename T>
inline int gemm_tiles(const GemmDescT<T> &g) {
  return ((g.M + GEMM_TILE - 1) / GEMM_TILE) * ((g.N + GEMM_TILE - 1) / GEMM_TILE);
}

// appends (desc index, tile) items of descriptors [d0, d1) to `items`
template <typename T>
inline void gemm_items(const std::vector<GemmDescT<T>> &d, size_t d0, size_t d1, std::vector<int> &items) {
  for (size_t i = d0; i < d1; i++) {
    const int t = gemm_tiles(d[i]);
    for (int k = 0; k < t; k++) {
      items.push_back((int)i);
      items.push_back(k);
    }
  }
}

// The two stages of a grouped triple-product table (nrg_gemm3_grouped): gd[0, nd1) form the
// intermediates, gd[nd1, ...) accumulate into each distinct C with K-parts
// parts[prange[g - nd1].first, .second) (the descriptors' parts pointers are set by the executor).
struct Gemm3Plan {
  std::vector<GemmDesc> gd;
  size_t nd1 = 0;
  std::vector<GemmPart> parts;
  std::vector<std::pair<size_t, size_t>> prange;
};
int gemm3_plan(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, Gemm3Plan &pl);

__global__ void k_gemm_f64(const GemmDesc *descs, const int2 *items);
__global__ void k_gemm_f32(const GemmDesc32 *descs, const int2 *items);

}  // namespace nrg
