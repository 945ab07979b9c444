// gemm.hip - batched FP64 products on v_mfma_f64_16x16x4f64 (see gemm.hpp).
//
// A 256-thread workgroup computes a 64x64 tile of C; wave w owns the 32x32 quarter
// (w % 2, w / 2) as 2x2 MFMA tiles. K advances in chunks of 16 staged through LDS.
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

template <typename T>
__device__ __forceinline__ void gemm_tile(const GemmDescT<T> G, const int2 it) {
  const int tm = (G.M + GEMM_TILE - 1) / GEMM_TILE;
  const int m0 = (it.y % tm) * GEMM_TILE, n0 = (it.y / tm) * GEMM_TILE;
  __shared__ T As[16][GEMM_TILE + 1];   // op(A)(m0 + m, k0 + k) at As[k][m]
  __shared__ T Bs[16][GEMM_TILE + 1];   // op(B)(k0 + k, n0 + n) at Bs[k][n]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
  typename Acc<T>::type acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (typename Acc<T>::type){0, 0, 0, 0};
  auto part = [&](const T *A, const T *B, int lda, int ldb, int tA, int tB, int Kp, double scd) {
    const T sc = (T)scd;
    for (int kb = 0; kb < Kp; kb += 16) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int idx = tid + 256 * q;
        if (tA) {   // contiguous along k
          const int k = idx & 15, m = idx >> 4;
          As[k][m] = ld_op(A, lda, 1, m0 + m, kb + k, G.M, Kp);
        } else {    // contiguous along m
          const int m = idx & 63, k = idx >> 6;
          As[k][m] = ld_op(A, lda, 0, m0 + m, kb + k, G.M, Kp);
        }
        if (tB) {   // contiguous along n
          const int n = idx & 63, k = idx >> 6;
          Bs[k][n] = sc * ld_op(B, ldb, 1, kb + k, n0 + n, Kp, G.N);
        } else {    // contiguous along k
          const int k = idx & 15, n = idx >> 4;
          Bs[k][n] = sc * ld_op(B, ldb, 0, kb + k, n0 + n, Kp, G.N);
        }
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < 4; kk++) {
        const int k = kk * 4 + (lane >> 4);
        T fa[2], fb[2];
#pragma unroll
        for (int j = 0; j < 2; j++) fa[j] = Bs[k][wn + j * 16 + (lane & 15)];   // MFMA A operand: op(B)^T
#pragma unroll
        for (int i = 0; i < 2; i++) fb[i] = As[k][wm + i * 16 + (lane & 15)];   // MFMA B operand: op(A)^T
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++)
            acc[i][j] = mfma(fa[j], fb[i], acc[i][j]);
      }
      __syncthreads();
    }
  };
  if (G.nparts > 0) {
    for (int p = 0; p < G.nparts; p++) {
      const GemmPartT<T> P = G.parts[p];
      part(P.A, P.B, P.lda, P.ldb, P.tA, P.tB, P.K, P.scale);
    }
  } else {
    part(G.A, G.B, G.lda, G.ldb, G.tA, G.tB, G.K, 1.0);
    if (G.K2 > 0) part(G.A2, G.B2, G.lda2, G.ldb2, G.tA2, G.tB2, G.K2, 1.0);
  }
  // acc[i][j] holds D = C^T for rows n0+wn+16j+(lane>>4)+4r (of D) and column m0+wm+16i+(lane&15)
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
