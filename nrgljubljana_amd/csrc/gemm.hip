// gemm.hip - batched FP64 products on v_mfma_f64_16x16x4f64 (see gemm.hpp).
//
// A 256-thread workgroup computes a 64x64 tile of C; wave w owns the 32x32 quarter
// (w % 2, w / 2) as 2x2 MFMA tiles. K advances in chunks of 32 staged through LDS; the next
// chunk is fetched into registers while the matrix cores work on the current one.
// The MFMA computes D = op(B)^T op(A)^T = C^T so that in the result layout of the
// instruction (lane l, register i holds D[l/16 + 4 i][l % 16]) consecutive lanes hold
// consecutive rows of the column-major C, and the stores coalesce.
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

// FP64: lane l, register r holds D[l/16 + 4 r][l % 16]; FP32: D[4 (l/16) + r][l % 16]
template <typename T> __device__ __forceinline__ int out_row(int lane, int r);
template <> __device__ __forceinline__ int out_row<double>(int lane, int r) { return (lane >> 4) + 4 * r; }
template <> __device__ __forceinline__ int out_row<float>(int lane, int r) { return 4 * (lane >> 4) + r; }

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v4f mfma(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

constexpr int KC = 32;   // K chunk staged through LDS

template <typename T>
__device__ __forceinline__ void gemm_tile(const GemmDescT<T> G, const int2 it) {
  const int tm = (G.M + GEMM_TILE - 1) / GEMM_TILE;
  const int m0 = (it.y % tm) * GEMM_TILE, n0 = (it.y / tm) * GEMM_TILE;
  __shared__ T As[KC][GEMM_TILE + 1];   // op(A)(m0 + m, k0 + k) at As[k][m]
  __shared__ T Bs[KC][GEMM_TILE + 1];   // op(B)(k0 + k, n0 + n) at Bs[k][n]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
  typename Acc<T>::type acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (typename Acc<T>::type){0, 0, 0, 0};
  // each thread stages 8 elements of op(A) and 8 of op(B) per chunk
  auto fetch = [&](const T *A, const T *B, int lda, int ldb, int tA, int tB, int Kp, int kb, T sc, T (&ra)[8],
                   T (&rb)[8]) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q;
      if (tA) {   // contiguous along k
        const int k = idx & (KC - 1), m = idx / KC;
        ra[q] = ld_op(A, lda, 1, m0 + m, kb + k, G.M, Kp);
      } else {    // contiguous along m
        const int m = idx & 63, k = idx >> 6;
        ra[q] = ld_op(A, lda, 0, m0 + m, kb + k, G.M, Kp);
      }
      if (tB) {   // contiguous along n
        const int n = idx & 63, k = idx >> 6;
        rb[q] = sc * ld_op(B, ldb, 1, kb + k, n0 + n, Kp, G.N);
      } else {    // contiguous along k
        const int k = idx & (KC - 1), n = idx / KC;
        rb[q] = sc * ld_op(B, ldb, 0, kb + k, n0 + n, Kp, G.N);
      }
    }
  };
  auto stash = [&](int tA, int tB, const T (&ra)[8], const T (&rb)[8]) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int idx = tid + 256 * q;
      if (tA) As[idx & (KC - 1)][idx / KC] = ra[q];
      else As[idx >> 6][idx & 63] = ra[q];
      if (tB) Bs[idx >> 6][idx & 63] = rb[q];
      else Bs[idx & (KC - 1)][idx / KC] = rb[q];
    }
  };
  auto compute = [&]() {
#pragma unroll
    for (int kk = 0; kk < KC / 4; kk++) {
      const int k = kk * 4 + (lane >> 4);
      T fa[2], fb[2];
#pragma unroll
      for (int j = 0; j < 2; j++) fa[j] = Bs[k][wn + j * 16 + (lane & 15)];   // MFMA A operand: op(B)^T
#pragma unroll
      for (int i = 0; i < 2; i++) fb[i] = As[k][wm + i * 16 + (lane & 15)];   // MFMA B operand: op(A)^T
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = mfma(fa[j], fb[i], acc[i][j]);
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
    T ra[8], rb[8];
    fetch(A, B, lda, ldb, tA, tB, Kp, 0, sc, ra, rb);
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
      if (more) fetch(A, B, lda, ldb, tA, tB, Kp, kb, sc, ra, rb);
      compute();
      __syncthreads();
      if (!more) break;
      ctA = tA;
      ctB = tB;
    }
  }
  // acc[i][j] holds D = C^T for rows n0+wn+16j+out_row(lane, r) (of D) and column m0+wm+16i+(lane&15)
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + i * 16 + (lane & 15);
        const int n = n0 + wn + j * 16 + out_row<T>(lane, r);
        if (m < G.M && n < G.N) {
          T *c = G.C + (size_t)n * G.ldc + m;
          const T v = (T)G.alpha * acc[i][j][r];
          *c = G.beta == 0.0 ? v : v + (T)G.beta * *c;
        }
      }
}

__global__ __launch_bounds__(256) void k_gemm_f64(const GemmDesc *descs, const int2 *items) {
  const int2 it = items[blockIdx.x];
  gemm_tile<double>(descs[it.x], it);
}

__global__ __launch_bounds__(256) void k_gemm_f32(const GemmDesc32 *descs, const int2 *items) {
  const int2 it = items[blockIdx.x];
  gemm_tile<float>(descs[it.x], it);
}

}  // namespace nrg
