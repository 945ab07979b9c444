// common.hpp - shared device helpers and host-side context of the gfx950 NRG core.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <map>
#include <memory>
#include "../../include/nrgcore.h"

namespace nrg {

constexpr int WAVE = 64;
constexpr int NMAX = 8192;   // largest block order accepted by the eigensolver

// ---------------------------------------------------------------- error handling
void set_error(const std::string &msg);
#define NRG_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) {                                                                        \
      ::nrg::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                         \
      return NRG_E_HIP;                                                                            \
    }                                                                                              \
  } while (0)

// ---------------------------------------------------------------- device reductions
// Cross-lane moves that do not go through the LDS: DPP permutations within rows of 16 lanes and the
// permlane swaps between rows and between wave halves (gfx950). Lane semantics, as checked on the GPU:
//   DPP_MIRROR: lane i of a row reads lane 15 - i; DPP_HALF_MIRROR: lane i of 8 reads lane 7 - i;
//   DPP_QXOR2 / DPP_QXOR1: lane i of a quad reads lane i ^ 2 / i ^ 1;
//   swap16(a, b): a <- rows (a0, b0, a2, b2), b <- rows (a1, b1, a3, b3);
//   swap32(a, b): a <- halves (a_lo, b_lo), b <- halves (a_hi, b_hi).
constexpr int DPP_MIRROR = 0x140, DPP_HALF_MIRROR = 0x141, DPP_QXOR2 = 0x4E, DPP_QXOR1 = 0xB1;
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
template <int CTRL>
__device__ __forceinline__ unsigned dpp32(unsigned v) { return __builtin_amdgcn_update_dpp(0u, v, CTRL, 0xF, 0xF, true); }
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  u32x2_t u = __builtin_bit_cast(u32x2_t, v);
  u[0] = dpp32<CTRL>(u[0]);
  u[1] = dpp32<CTRL>(u[1]);
  return __builtin_bit_cast(double, u);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) { return __builtin_bit_cast(float, dpp32<CTRL>(__builtin_bit_cast(unsigned, v))); }
__device__ __forceinline__ void swap16(unsigned &a, unsigned &b) {
  const u32x2_t r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
  a = r[0];
  b = r[1];
}
__device__ __forceinline__ void swap32(unsigned &a, unsigned &b) {
  const u32x2_t r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
  a = r[0];
  b = r[1];
}
template <bool HALF32>
__device__ __forceinline__ void swap_rows(double &a, double &b) {
  u32x2_t ua = __builtin_bit_cast(u32x2_t, a), ub = __builtin_bit_cast(u32x2_t, b);
  unsigned a0 = ua[0], a1 = ua[1], b0 = ub[0], b1 = ub[1];
  if (HALF32) { swap32(a0, b0); swap32(a1, b1); } else { swap16(a0, b0); swap16(a1, b1); }
  ua[0] = a0; ua[1] = a1; ub[0] = b0; ub[1] = b1;
  a = __builtin_bit_cast(double, ua);
  b = __builtin_bit_cast(double, ub);
}
template <bool HALF32>
__device__ __forceinline__ void swap_rows(float &a, float &b) {
  unsigned ua = __builtin_bit_cast(unsigned, a), ub = __builtin_bit_cast(unsigned, b);
  if (HALF32) swap32(ua, ub); else swap16(ua, ub);
  a = __builtin_bit_cast(float, ua);
  b = __builtin_bit_cast(float, ub);
}
// Sum over each row of 16 lanes, in every lane of the row. Every step adds a lane and its partner in
// both, so all lanes of a row get the same bits.
template <typename T>
__device__ __forceinline__ T row16_sum(T v) {
  v += dpp<DPP_QXOR1>(v);
  v += dpp<DPP_QXOR2>(v);
  v += dpp<DPP_HALF_MIRROR>(v);
  v += dpp<DPP_MIRROR>(v);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum_t(T v) {
  v = row16_sum(v);
  T a = v, b = v;
  swap_rows<false>(a, b);   // a + b: rows 0 + 1 in rows 0, 1; rows 2 + 3 in rows 2, 3
  v = a + b;
  a = v;
  b = v;
  swap_rows<true>(a, b);    // the two halves
  return a + b;
}
__device__ __forceinline__ double wave_sum(double v) { return wave_sum_t(v); }
__device__ __forceinline__ float wave_sum(float v) { return wave_sum_t(v); }
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp<DPP_QXOR1>(v));
  v = fmax(v, dpp<DPP_QXOR2>(v));
  v = fmax(v, dpp<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp<DPP_MIRROR>(v));
  double a = v, b = v;
  swap_rows<false>(a, b);
  v = fmax(a, b);
  a = v;
  b = v;
  swap_rows<true>(a, b);
  return fmax(a, b);
}

// Sum over a workgroup of NT threads. `red` holds NT/64 entries of LDS. All threads get the sum.
// 1/x to about an ulp: the hardware reciprocal estimate (v_rcp_f64) refined by two Newton-Raphson
// steps (e = 1 - x r, r += r e), about half the instructions of an IEEE division. For the O(k^2)
// terms of the divide-and-conquer kernels, whose sums and products tolerate an ulp per term; x must
// be a finite normal number (their denominators are separations of distinct poles and roots).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// TRAIL = false leaves out the closing barrier, for callers that pass another barrier before `red`
// is written again.
template <int NT, bool TRAIL = true, typename T>
__device__ __forceinline__ T block_sum(T v, T *red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  if (lane == 0) red[w] = v;
  __syncthreads();
  T s = 0;
#pragma unroll
  for (int i = 0; i < NT / WAVE; i++) s += red[i];
  if (TRAIL) __syncthreads();
  return s;
}
template <int NT>
__device__ __forceinline__ double block_max(double v, double *red) {
  v = wave_max(v);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = red[0];
#pragma unroll
  for (int i = 1; i < NT / WAVE; i++) s = fmax(s, red[i]);
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------- host-side device buffer
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);   // grow-only
  void release();
  template <typename T> T *as(size_t byte_off = 0) const { return (T *)((char *)p + byte_off); }
};

struct PinnedBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
};

constexpr int NPHASE = 8;

// Optional per-launch HIP-event timing of named kernels (off by default).
struct KTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  struct Mark { const char *name; size_t b, e; };
  std::vector<Mark> marks;
  std::map<std::string, std::pair<double, long>> acc;
  size_t begin(hipStream_t st);
  void end(hipStream_t st, size_t b, const char *name);
  void collect();  // call after the stream has been synchronised
  void reset();
  void release();
};

}  // namespace nrg

// Launch `launch` on `st`, bracketed by timing events when ctx->kt.on.
#define NRG_TIMED(ctx, st, name, launch)                     \
  do {                                                        \
    size_t kb_ = (ctx)->kt.on ? (ctx)->kt.begin(st) : 0;      \
    launch;                                                   \
    if ((ctx)->kt.on) (ctx)->kt.end(st, kb_, name);           \
  } while (0)

struct nrg_ctx {
  int device = 0;
  hipEvent_t ev[nrg::NPHASE + 1];
  float phase_ms[nrg::NPHASE];
  double total_ms = 0;
  nrg::DevBuf ws;        // general workspace
  nrg::DevBuf ws2;       // second workspace (typed arenas of the mixed form)
  nrg::DevBuf desc;      // descriptor tables
  nrg::DevBuf flags;     // per-block status words
  nrg::DevBuf dcbuf;     // deflation arrays of divide-and-conquer merges too large for LDS
  nrg::PinnedBuf hdesc;  // pinned staging of descriptor tables
  nrg::PinnedBuf hflags;
  nrg::KTimer kt;
  // extra streams for independent work of one call (default priority, created on first use),
  // forked from and joined back into the caller's stream with the events
  static constexpr int NAUX = 4;   // the last one carries the eigensolver's two-stage chain
  hipStream_t aux[NAUX] = {};
  hipEvent_t fork_ev = nullptr, join_ev[NAUX] = {};
  // stream on which the eigensolver captures and replays its graphs (the caller's stream may be the
  // null stream, which cannot be captured), with its fork and join events
  hipStream_t gstream = nullptr;
  hipEvent_t gfork = nullptr, gjoin = nullptr;
  // the eigensolver's host-side work lists, kept between calls so that their memory is reused
  std::shared_ptr<void> syev_cache;
};

namespace nrg {
// looks up the device's address of the no-allocation scratch array (noalloc.hip), once per device
int scratch_setup(int device);
// creates ctx->aux / fork_ev / join_ev once
int ensure_aux_streams(nrg_ctx *ctx);
int syev_run(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H, const int64_t *off_w,
             double *w, int mode, int *info, hipStream_t st);
std::string syev_tuning_json(nrg_ctx *ctx);
int gemm3_run(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, int precision, hipStream_t st);
int fdm_run(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double wnf,
            int sign, double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg, hipStream_t st);
int dmnrg_run(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double emin, double emax,
              int sign, double *out_pos, double *out_neg, hipStream_t st);
int hamiltonian_run(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H, const nrg_ham_term *terms,
                    int nterms, hipStream_t st);
int cfs_run(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z, int last,
            int sign, double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg, hipStream_t st);
int mats_run(nrg_ctx *ctx, int mode, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, double wnf,
             double Z, int sign, double *G, hipStream_t st);
int temp_run(nrg_ctx *ctx, int kind, const nrg_fdm_pair *pairs, int npairs, double temperature, double Z, double *out,
             hipStream_t st);
int broaden_run(nrg_ctx *ctx, const double *pos_e, const double *pos_w, int64_t npos, const double *neg_e,
                const double *neg_w, int64_t nneg, const double *mesh, int64_t nmesh, double alpha, double omega0,
                double *out_pos, double *out_neg, hipStream_t st);
// Uploads `bytes` of host data through the pinned staging buffer into ctx->desc at byte offset 0.
int upload_desc(nrg_ctx *ctx, const void *src, size_t bytes, hipStream_t st);
// The same in two steps, for tables written straight into the staging buffer: stage_desc returns
// the buffer (`bytes` long, once the previous upload has left it), upload_staged copies its first
// `bytes` into ctx->desc.
int stage_desc(nrg_ctx *ctx, size_t bytes, hipStream_t st, char **host);
int upload_staged(nrg_ctx *ctx, size_t bytes, hipStream_t st);
}  // namespace nrg
