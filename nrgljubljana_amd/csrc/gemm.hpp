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

constexpr int GEMM_TILE = 64;         // tile of C per workgroup for small products
constexpr int GEMM_BIG = 128;         // tile for products with M, N >= GEMM_BIG
constexpr int GEMM_BIG_FLAG = 1 << 30;   // marks a work item's tile index as a GEMM_BIG tile

// Big tiles pay off when the tile has enough K to amortise its staging: M, N >= 128 and total
// K >= 256 (descriptors with a part list carry their total K in K as a hint).
template <typename T>
inline bool gemm_is_big(const GemmDescT<T> &g) {
  const int ktot = g.nparts > 0 ? g.K : g.K + (g.K2 > 0 ? g.K2 : 0);
  return g.M >= GEMM_BIG && g.N >= GEMM_BIG && ktot >= 256;
}

template <typename T>
inline int gemm_tiles(const GemmDescT<T> &g) {
  const int t = gemm_is_big(g) ? GEMM_BIG : GEMM_TILE;
  return ((g.M + t - 1) / t) * ((g.N + t - 1) / t);
}

// Fewer big tiles than this in one launch leave most compute units idle: the launch then uses
// the small tiles only (four per big tile).
constexpr int GEMM_BIG_MIN_TILES = 160;

// appends (desc index, tile) items of descriptors [d0, d1) to `items`: first every big tile (tile
// index with GEMM_BIG_FLAG), then every small one, so that gemm_launch can split the list
template <typename T>
inline void gemm_items(const std::vector<GemmDescT<T>> &d, size_t d0, size_t d1, std::vector<int> &items) {
  size_t nbig = 0;
  for (size_t i = d0; i < d1; i++)
    if (gemm_is_big(d[i])) nbig += gemm_tiles(d[i]);
  const bool use_big = nbig >= (size_t)GEMM_BIG_MIN_TILES;
  for (int big = 1; big >= 0; big--)
    for (size_t i = d0; i < d1; i++) {
      const bool b = use_big && gemm_is_big(d[i]);
      if ((int)b != big) continue;
      const int ts = b ? GEMM_BIG : GEMM_TILE;
      const int t = ((d[i].M + ts - 1) / ts) * ((d[i].N + ts - 1) / ts);
      for (int k = 0; k < t; k++) {
        items.push_back((int)i);
        items.push_back(big ? (k | GEMM_BIG_FLAG) : k);
      }
    }
}

// Launches the `count` work items of a gemm_items list: the leading big tiles on k_gemm_*_big,
// the rest on k_gemm_* (items_host: the same list in host memory).
void gemm_launch(nrg_ctx *ctx, hipStream_t st, const GemmDesc *descs, const int *items_dev, const int *items_host,
                 size_t count);
void gemm_launch(nrg_ctx *ctx, hipStream_t st, const GemmDesc32 *descs, const int *items_dev, const int *items_host,
                 size_t count);

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
__global__ void k_gemm_f64_big(const GemmDesc *descs, const int2 *items);
__global__ void k_gemm_f32_big(const GemmDesc32 *descs, const int2 *items);

}  // namespace nrg
