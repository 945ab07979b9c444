// gemm.hip - batched products on the matrix cores (see gemm.hpp).
//
// A 256-thread workgroup computes one TM x TM tile of C: TM = 64 for small products (wave w owns
// the 32 x 32 quarter (w % 2, w / 2) as 2 x 2 MFMA tiles) and TM = 128 when M, N >= 128 (each wave
// 64 x 64 as 4 x 4 MFMA tiles: twice the flops per byte staged and per LDS fragment read). K
// advances in chunks (32 for the small tile, 16 for the big one) staged through LDS; the next chunk is fetched into registers while the
// matrix cores work on the current one. The MFMA computes D = op(B)^T op(A)^T = C^T so that in the
// result layout of the instruction (FP64: lane l, register i holds D[l/16 + 4 i][l % 16]; FP32:
// D[4 (l/16) + i][l % 16]) consecutive lanes hold consecutive rows of the column-major C, and the
// stores coalesce.
#include "gemm.hpp"

namespace nrg {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T ld_op(const T *X, int ld, int t, int r, int c, int R, int Cn) {
  // element (r, c) of op(X) (R x Cn), zero outside
  if (r >= R || c >= Cn) return (T)0;
  return t ? X[(size_t)r * ld + c] : X[(size_t)c * ld + r];
}

template <typename T> struct Acc;
template <> struct Acc<double> { typedef v4d type; };
template <> struct Acc<float> { typedef v4f type; };

template <typename T> __device__ __forceinline__ int out_row(int lane, int r);
template <> __device__ __forceinline__ int out_row<double>(int lane, int r) { return (lane >> 4) + 4 * r; }
template <> __device__ __forceinline__ int out_row<float>(int lane, int r) { return 4 * (lane >> 4) + r; }

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v4f mfma(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// TM x TM tile of C with K chunks of KC staged through LDS
template <typename T, int TM, int KC>
__device__ __forceinline__ void gemm_tile(const GemmDescT<T> &G, int tile, T *lds) {
  constexpr int LD = TM + 1;          // LDS row stride
  constexpr int NF = TM / 32;         // MFMA tiles per wave and dimension
  constexpr int PT = TM * KC / 256;   // elements of each operand staged per thread and chunk
  const int tm = (G.M + TM - 1) / TM;
  const int m0 = (tile % tm) * TM, n0 = (tile / tm) * TM;
  T *As = lds;             // op(A)(m0 + m, k0 + k) at As[k * LD + m]
  T *Bs = lds + KC * LD;   // op(B)(k0 + k, n0 + n) at Bs[k * LD + n]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv & 1) * (TM / 2), wn = (wv >> 1) * (TM / 2);
  typename Acc<T>::type acc[NF][NF];
#pragma unroll
  for (int i = 0; i < NF; i++)
#pragma unroll
    for (int j = 0; j < NF; j++) acc[i][j] = (typename Acc<T>::type){0, 0, 0, 0};
  // C's old values (beta != 0) are loaded before the K loop on the small tile, so their latency hides
  // behind it (the big tile has no registers to spare and a K loop long enough to amortise them)
  constexpr bool PRE = TM == 64;
  const bool rmw = G.beta != 0.0;
  typename Acc<T>::type cold[NF][NF];
#pragma unroll
  for (int i = 0; i < NF; i++)
#pragma unroll
    for (int j = 0; j < NF; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + i * 16 + (lane & 15);
        const int n = n0 + wn + j * 16 + out_row<T>(lane, r);
        cold[i][j][r] = (PRE && rmw && m < G.M && n < G.N) ? G.C[(size_t)n * G.ldc + m] : (T)0;
      }
  auto fetch = [&](const T *A, const T *B, int lda, int ldb, int tA, int tB, int Kp, int kb, T sc, T (&ra)[PT],
                   T (&rb)[PT]) {
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int idx = tid + 256 * q;
      if (tA) {   // contiguous along k
        const int k = idx % KC, m = idx / KC;
        ra[q] = ld_op(A, lda, 1, m0 + m, kb + k, G.M, Kp);
      } else {    // contiguous along m
        const int m = idx % TM, k = idx / TM;
        ra[q] = ld_op(A, lda, 0, m0 + m, kb + k, G.M, Kp);
      }
      if (tB) {   // contiguous along n
        const int n = idx % TM, k = idx / TM;
        rb[q] = sc * ld_op(B, ldb, 1, kb + k, n0 + n, Kp, G.N);
      } else {    // contiguous along k
        const int k = idx % KC, n = idx / KC;
        rb[q] = sc * ld_op(B, ldb, 0, kb + k, n0 + n, Kp, G.N);
      }
    }
  };
  // interior chunk (tile and K chunk inside the operands): no bounds checks, one base address per
  // operand and fixed strides between the elements a thread stages
  auto fetch_fast = [&](const T *A, const T *B, int lda, int ldb, int tA, int tB, int kb, T sc, T (&ra)[PT],
                        T (&rb)[PT]) {
    if (tA) {
      const T *Ab = A + (size_t)(m0 + tid / KC) * lda + kb + tid % KC;
#pragma unroll
      for (int q = 0; q < PT; q++) ra[q] = Ab[(size_t)q * (256 / KC) * lda];
    } else {
      const T *Ab = A + (size_t)(kb + tid / TM) * lda + m0 + tid % TM;
#pragma unroll
      for (int q = 0; q < PT; q++) ra[q] = Ab[(size_t)q * (256 / TM) * lda];
    }
    if (tB) {
      const T *Bb = B + (size_t)(kb + tid / TM) * ldb + n0 + tid % TM;
#pragma unroll
      for (int q = 0; q < PT; q++) rb[q] = sc * Bb[(size_t)q * (256 / TM) * ldb];
    } else {
      const T *Bb = B + (size_t)(n0 + tid / KC) * ldb + kb + tid % KC;
#pragma unroll
      for (int q = 0; q < PT; q++) rb[q] = sc * Bb[(size_t)q * (256 / KC) * ldb];
    }
  };
  const bool mn_inside = m0 + TM <= G.M && n0 + TM <= G.N;
  auto stash = [&](int tA, int tB, const T (&ra)[PT], const T (&rb)[PT]) {
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int idx = tid + 256 * q;
      if (tA) As[(idx % KC) * LD + idx / KC] = ra[q];
      else As[(idx / TM) * LD + idx % TM] = ra[q];
      if (tB) Bs[(idx / TM) * LD + idx % TM] = rb[q];
      else Bs[(idx % KC) * LD + idx / KC] = rb[q];
    }
  };
  auto compute = [&]() {
#pragma unroll
    for (int kk = 0; kk < KC / 4; kk++) {
      const int k = kk * 4 + (lane >> 4);
      T fa[NF], fb[NF];
#pragma unroll
      for (int j = 0; j < NF; j++) fa[j] = Bs[k * LD + wn + j * 16 + (lane & 15)];   // MFMA A operand: op(B)^T
#pragma unroll
      for (int i = 0; i < NF; i++) fb[i] = As[k * LD + wm + i * 16 + (lane & 15)];   // MFMA B operand: op(A)^T
#pragma unroll
      for (int i = 0; i < NF; i++)
#pragma unroll
        for (int j = 0; j < NF; j++) acc[i][j] = mfma(fa[j], fb[i], acc[i][j]);
    }
  };
  // K-parts: the inline pair (A, B) and (A2, B2), or the descriptor's part list
  const int np = G.nparts > 0 ? G.nparts : (G.K2 > 0 ? 2 : 1);
  auto part_of = [&](int p, const T *&A, const T *&B, int &lda, int &ldb, int &tA, int &tB, int &Kp, T &sc) {
    if (G.nparts > 0) {
      const GemmPartT<T> P = G.parts[p];
      A = P.A; B = P.B; lda = P.lda; ldb = P.ldb; tA = P.tA; tB = P.tB; Kp = P.K; sc = (T)P.scale;
    } else if (p == 0) {
      A = G.A; B = G.B; lda = G.lda; ldb = G.ldb; tA = G.tA; tB = G.tB; Kp = G.K; sc = (T)1;
    } else {
      A = G.A2; B = G.B2; lda = G.lda2; ldb = G.ldb2; tA = G.tA2; tB = G.tB2; Kp = G.K2; sc = (T)1;
    }
  };
  // software pipeline over the chunks of all parts: fetch chunk c+1 while computing chunk c
  const T *A, *B;
  int lda, ldb, tA, tB, Kp;
  T sc;
  int p = 0, kb = 0;
  part_of(0, A, B, lda, ldb, tA, tB, Kp, sc);
  while (p < np && Kp <= 0) { if (++p < np) part_of(p, A, B, lda, ldb, tA, tB, Kp, sc); }
  if (p < np) {
    T ra[PT], rb[PT];
    if (mn_inside && KC <= Kp) fetch_fast(A, B, lda, ldb, tA, tB, 0, sc, ra, rb);
    else fetch(A, B, lda, ldb, tA, tB, Kp, 0, sc, ra, rb);
    int ctA = tA, ctB = tB;
    for (;;) {
      stash(ctA, ctB, ra, rb);
      __syncthreads();
      // advance to the next chunk
      kb += KC;
      if (kb >= Kp) {
        kb = 0;
        do { if (++p < np) part_of(p, A, B, lda, ldb, tA, tB, Kp, sc); } while (p < np && Kp <= 0);
      }
      const bool more = p < np;
      if (more) {
        if (mn_inside && kb + KC <= Kp) fetch_fast(A, B, lda, ldb, tA, tB, kb, sc, ra, rb);
        else fetch(A, B, lda, ldb, tA, tB, Kp, kb, sc, ra, rb);
      }
      compute();
      __syncthreads();
      if (!more) break;
      ctA = tA;
      ctB = tB;
    }
  }
  // acc[i][j] holds D = C^T for rows n0+wn+16j+out_row(lane, r) (of D) and column m0+wm+16i+(lane&15)
#pragma unroll
  for (int i = 0; i < NF; i++)
#pragma unroll
    for (int j = 0; j < NF; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + i * 16 + (lane & 15);
        const int n = n0 + wn + j * 16 + out_row<T>(lane, r);
        if (m < G.M && n < G.N) {
          const T v = (T)G.alpha * acc[i][j][r];
          T *c = G.C + (size_t)n * G.ldc + m;
          *c = rmw ? v + (T)G.beta * (PRE ? cold[i][j][r] : *c) : v;
        }
      }
}

// three waves per SIMD (at most 170 registers per lane, no spills): 1.2x on 1000x1000x128 products
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_gemm_f64(const GemmDesc *descs, const int2 *items) {
  __shared__ double lds[2 * 32 * (GEMM_TILE + 1)];
  const int2 it = items[blockIdx.x];
  gemm_tile<double, GEMM_TILE, 32>(descs[it.x], it.y, lds);
}

__global__ __launch_bounds__(256) void k_gemm_f32(const GemmDesc32 *descs, const int2 *items) {
  __shared__ float lds[2 * 32 * (GEMM_TILE + 1)];
  const int2 it = items[blockIdx.x];
  gemm_tile<float, GEMM_TILE, 32>(descs[it.x], it.y, lds);
}

// two waves per SIMD (at most 256 registers per lane): one workgroup's LDS staging and barriers
// then hide behind the other's matrix-core work (1.4x on 1024^3 products over the unconstrained
// allocation, which takes 364 registers and leaves one wave per SIMD)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_gemm_f64_big(const GemmDesc *descs, const int2 *items) {
  __shared__ double lds[2 * 16 * (GEMM_BIG + 1)];
  const int2 it = items[blockIdx.x];
  gemm_tile<double, GEMM_BIG, 16>(descs[it.x], it.y & ~GEMM_BIG_FLAG, lds);
}

__global__ __launch_bounds__(256) void k_gemm_f32_big(const GemmDesc32 *descs, const int2 *items) {
  __shared__ float lds[2 * 16 * (GEMM_BIG + 1)];
  const int2 it = items[blockIdx.x];
  gemm_tile<float, GEMM_BIG, 16>(descs[it.x], it.y & ~GEMM_BIG_FLAG, lds);
}

static size_t leading_big(const int *items_host, size_t count) {
  size_t nb = 0;
  while (nb < count && (items_host[2 * nb + 1] & GEMM_BIG_FLAG)) nb++;
  return nb;
}

void gemm_launch(nrg_ctx *ctx, hipStream_t st, const GemmDesc *descs, const int *items_dev, const int *items_host,
                 size_t count) {
  const size_t nb = leading_big(items_host, count);
  if (nb)
    NRG_TIMED(ctx, st, "k_gemm_f64_big", hipLaunchKernelGGL(k_gemm_f64_big, dim3((unsigned)nb), dim3(256), 0, st, descs,
                                                            (const int2 *)items_dev));
  if (count > nb)
    NRG_TIMED(ctx, st, "k_gemm_f64", hipLaunchKernelGGL(k_gemm_f64, dim3((unsigned)(count - nb)), dim3(256), 0, st, descs,
                                                        (const int2 *)items_dev + nb));
}

void gemm_launch(nrg_ctx *ctx, hipStream_t st, const GemmDesc32 *descs, const int *items_dev, const int *items_host,
                 size_t count) {
  const size_t nb = leading_big(items_host, count);
  if (nb)
    NRG_TIMED(ctx, st, "k_gemm_f32_big", hipLaunchKernelGGL(k_gemm_f32_big, dim3((unsigned)nb), dim3(256), 0, st, descs,
                                                            (const int2 *)items_dev));
  if (count > nb)
    NRG_TIMED(ctx, st, "k_gemm_f32", hipLaunchKernelGGL(k_gemm_f32, dim3((unsigned)(count - nb)), dim3(256), 0, st, descs,
                                                        (const int2 *)items_dev + nb));
}

}  // namespace nrg
