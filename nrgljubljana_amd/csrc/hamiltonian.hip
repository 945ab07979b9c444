// hamiltonian.hip - assembly of the Hamiltonian blocks of one NRG step (nrg_hamiltonian).
//
// Replaces hamiltonian() (c++/core.hpp:72-88) and the building steps the symmetry tables call,
// offdiag_function_impl, diag_function_impl and diag_offdiag_function (c++/matrix.hpp:32-111).
// Every block of the arena is cut into 64 x 64 tiles; one workgroup owns one tile, keeps it in
// registers, applies the terms that touch it in the caller's order and writes it once, so no
// element is written twice, no atomics are needed and zero regions cost one store. Transposed
// sub-blocks (offdiag terms with i > j) are staged through LDS so that both the read and the write
// coalesce. Products and sums are rounded separately (floating-point contraction is off in the
// kernel), as in the reference's `hsub += factor * mat`, so the result is the reference's to the
// last bit.
#include "common.hpp"
#include <algorithm>
#include <cstring>

namespace nrg {

constexpr int HT = 64;   // tile edge

struct HamBlk {
  int64_t n, off;
};
struct HamTile {
  int32_t blk, tr, tc;   // tile (tr, tc) of block blk
  int32_t t0, t1;        // its terms: term_list[t0, t1)
};

__global__ __launch_bounds__(256) void k_ham_tile(const HamBlk *blks, const HamTile *tiles, const int *term_list,
                                                  const nrg_ham_term *terms, double *H) {
#pragma clang fp contract(off)   // products and sums rounded separately, never fused
  const HamTile t = tiles[blockIdx.x];
  const HamBlk b = blks[t.blk];
  const int64_t n = b.n;
  const int64_t r0 = (int64_t)t.tr * HT, c0 = (int64_t)t.tc * HT;
  const int tid = threadIdx.x, cl = tid & (HT - 1), rw = tid >> 6;   // thread: column cl, rows rw + 4q
  __shared__ double S[HT][HT + 1];
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; q++) acc[q] = 0.0;
  const int64_t c = c0 + cl;
  for (int k = t.t0; k < t.t1; k++) {
    const nrg_ham_term T = terms[term_list[k]];
    if (T.kind == NRG_HAM_BLOCK) {
      const bool cin = c >= T.col && c < T.col + T.cols;
      if (!T.trans) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int64_t r = r0 + rw + 4 * q;
          if (cin && r >= T.row && r < T.row + T.rows)
            acc[q] = acc[q] + T.factor * T.src[(r - T.row) * T.ld + (c - T.col)];
        }
      } else {
        // element (r, c) = S(c - col, r - row): stage rows [c0 - col, +64) x columns [r0 - row, +64) of S
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int64_t sr = c0 - T.col + rw + 4 * q, sc = r0 - T.row + cl;
          S[rw + 4 * q][cl] = (sr >= 0 && sr < T.cols && sc >= 0 && sc < T.rows) ? T.src[sr * T.ld + sc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int64_t r = r0 + rw + 4 * q;
          if (cin && r >= T.row && r < T.row + T.rows)
            acc[q] = acc[q] + T.factor * S[cl][rw + 4 * q];
        }
      }
    } else if (T.kind == NRG_HAM_DIAG) {
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int64_t r = r0 + rw + 4 * q;
        if (r == c && r >= T.row && r < T.row + T.rows) {
          if (T.src) acc[q] = acc[q] + T.factor * T.src[r - T.row];
          acc[q] = acc[q] + T.shift;
        }
      }
    } else {   // NRG_HAM_IDENT
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int64_t r = r0 + rw + 4 * q;
        if (r - T.row == c - T.col && r >= T.row && r < T.row + T.rows) acc[q] = acc[q] + T.factor;
      }
    }
  }
  if (c < n) {
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int64_t r = r0 + rw + 4 * q;
      if (r < n) H[b.off + r * n + c] = acc[q];
    }
  }
}

int hamiltonian_run(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H, const nrg_ham_term *terms,
                    int nterms, hipStream_t st) {
  if (nblk < 0 || nterms < 0 || (nblk > 0 && (!n || !off_H || !H)) || (nterms > 0 && !terms)) {
    set_error("nrg_hamiltonian: invalid arguments");
    return NRG_E_ARG;
  }
  std::vector<HamBlk> blks(nblk);
  std::vector<int64_t> tile0(nblk + 1, 0), nt(nblk);
  for (int b = 0; b < nblk; b++) {
    if (n[b] < 0 || n[b] > INT32_MAX || off_H[b] < 0) {
      set_error("nrg_hamiltonian: block " + std::to_string(b) + " has an invalid order or offset");
      return NRG_E_ARG;
    }
    blks[b] = HamBlk{n[b], off_H[b]};
    nt[b] = (n[b] + HT - 1) / HT;
    tile0[b + 1] = tile0[b] + nt[b] * nt[b];
  }
  if (tile0[nblk] > INT32_MAX) {
    set_error("nrg_hamiltonian: too many tiles");
    return NRG_E_ARG;
  }
  // validate, then the terms of every tile in the caller's order as one CSR list (count, then fill)
  for (int k = 0; k < nterms; k++) {
    const nrg_ham_term &T = terms[k];
    const int64_t b = T.block;
    const bool kind_ok = T.kind == NRG_HAM_BLOCK || T.kind == NRG_HAM_DIAG || T.kind == NRG_HAM_IDENT;
    bool ok = kind_ok && b >= 0 && b < nblk && T.row >= 0 && T.rows >= 0;
    if (ok) {
      const int64_t N = n[b];
      if (T.kind == NRG_HAM_BLOCK)
        ok = T.col >= 0 && T.cols >= 0 && T.row + T.rows <= N && T.col + T.cols <= N &&
             (T.rows == 0 || T.cols == 0 || (T.src && T.ld >= (T.trans ? T.rows : T.cols)));
      else if (T.kind == NRG_HAM_DIAG)
        ok = T.row + T.rows <= N;
      else
        ok = T.col >= 0 && T.row + T.rows <= N && T.col + T.rows <= N;
    }
    if (!ok) {
      set_error("nrg_hamiltonian: term " + std::to_string(k) + " is inconsistent with its block");
      return NRG_E_ARG;
    }
  }
  // calls f(tile id) for every tile term k touches
  auto tiles_of = [&](int k, auto &&f) {
    const nrg_ham_term &T = terms[k];
    if (T.rows == 0 || (T.kind == NRG_HAM_BLOCK && T.cols == 0)) return;
    const int64_t base = tile0[T.block], m = nt[T.block];
    if (T.kind == NRG_HAM_BLOCK) {
      for (int64_t tr = T.row / HT; tr <= (T.row + T.rows - 1) / HT; tr++)
        for (int64_t tc = T.col / HT; tc <= (T.col + T.cols - 1) / HT; tc++) f(base + tr * m + tc);
    } else if (T.kind == NRG_HAM_DIAG) {
      for (int64_t tr = T.row / HT; tr <= (T.row + T.rows - 1) / HT; tr++) f(base + tr * m + tr);
    } else {   // the tiles the shifted diagonal (row + l, col + l) crosses, in order
      int64_t last = -1;
      for (int64_t l = 0; l < T.rows;) {
        const int64_t id = base + (T.row + l) / HT * m + (T.col + l) / HT;
        if (id != last) f(id);
        last = id;
        l += std::min(HT - (T.row + l) % HT, HT - (T.col + l) % HT);   // to the next tile edge
      }
    }
  };
  const int64_t ntiles = tile0[nblk];
  std::vector<int64_t> start(ntiles + 1, 0);
  for (int k = 0; k < nterms; k++) tiles_of(k, [&](int64_t id) { start[id + 1]++; });
  for (int64_t i = 0; i < ntiles; i++) start[i + 1] += start[i];
  if (start[ntiles] > INT32_MAX) {
    set_error("nrg_hamiltonian: too many (term, tile) pairs");
    return NRG_E_ARG;
  }
  std::vector<int> tl(start[ntiles]);
  {
    std::vector<int64_t> cur(start.begin(), start.end() - 1);
    for (int k = 0; k < nterms; k++) tiles_of(k, [&](int64_t id) { tl[cur[id]++] = k; });
  }
  std::vector<HamTile> tiles((size_t)ntiles);
  for (int b = 0; b < nblk; b++)
    for (int64_t i = 0; i < nt[b] * nt[b]; i++) {
      const int64_t id = tile0[b] + i;
      tiles[id] = HamTile{b, (int32_t)(i / nt[b]), (int32_t)(i % nt[b]), (int32_t)start[id], (int32_t)start[id + 1]};
    }
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t bB = sizeof(HamBlk) * blks.size(), bT = sizeof(HamTile) * tiles.size(), bL = sizeof(int) * tl.size();
  const size_t bX = sizeof(nrg_ham_term) * (size_t)nterms;
  const size_t oB = 0, oT = al(bB), oL = oT + al(bT), oX = oL + al(bL);
  std::vector<char> blob(oX + bX + 8);
  if (bB) memcpy(blob.data() + oB, blks.data(), bB);
  if (bT) memcpy(blob.data() + oT, tiles.data(), bT);
  if (bL) memcpy(blob.data() + oL, tl.data(), bL);
  if (bX) memcpy(blob.data() + oX, terms, bX);
  NRG_HIP(hipEventRecord(ctx->ev[0], st));
  if (int r = upload_desc(ctx, blob.data(), blob.size(), st)) return r;
  if (!tiles.empty())
    NRG_TIMED(ctx, st, "k_ham_tile", hipLaunchKernelGGL(k_ham_tile, dim3((unsigned)tiles.size()), dim3(256), 0, st,
                                                        (const HamBlk *)ctx->desc.as<char>(oB),
                                                        (const HamTile *)ctx->desc.as<char>(oT), ctx->desc.as<int>(oL),
                                                        (const nrg_ham_term *)ctx->desc.as<char>(oX), H));
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[1], st));
  NRG_HIP(hipStreamSynchronize(st));
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
  ctx->total_ms = ms;
  if (ctx->kt.on) ctx->kt.collect();
  return NRG_OK;
}

}  // namespace nrg
