// checkdiag.hip - check_diag of a whole step on the device (nrg_check_diag_batched).
//
// The reference checks every diagonalisation (c++/diag.hpp:290 -> check_diag, c++/numerics.hpp:340-404):
// the eigenvalues must be finite and the eigenvector rows orthonormal, |<v_r, v_r> - 1| within
// NORMALIZATION_EPSILON and |<v_r, v_s>| within ORTHOGONALITY_EPSILON, by O(M^2 n) scalar loops on the
// host. Here the Gram matrices G_b = V_b V_b^T of all blocks are one grouped FP64 product on the matrix
// cores (the k_gemm_f64 kernels), and one reduction launch takes, per block, max |G_rr - 1|,
// max_{r != s} |G_rs| and whether an eigenvalue is not finite. Maxima of non-negative doubles are
// combined with 64-bit integer atomicMax on their bit patterns (the order of non-negative IEEE doubles,
// +inf included, is the order of their bit patterns); a NaN is counted as +inf.
#include <cstring>
#include <limits>

#include "gemm.hpp"

namespace nrg {

namespace {

struct ChkBlk {
  int64_t goff;   // Gram matrix of the block in the workspace (row-major M x M)
  int64_t woff;   // its eigenvalues
  int M;          // checked eigenvectors (rows)
  int rpi;        // Gram rows per work item
};

constexpr int CHK_THREADS = 256;
constexpr int CHK_ELEMS = 8192;   // Gram elements per work item (about)

__device__ inline unsigned long long nonneg_bits(double x) {
  if (!(x <= std::numeric_limits<double>::max())) x = std::numeric_limits<double>::infinity();   // NaN, inf
  return (unsigned long long)__double_as_longlong(x);
}

__device__ inline unsigned long long wave_max(unsigned long long v) {
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    const unsigned long long u = __shfl_xor(v, o);
    v = u > v ? u : v;
  }
  return v;
}

__global__ void k_chk_clear(unsigned long long *res, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) res[i] = 0ull;
}

// item (b, r0): Gram rows [r0, r0 + rpi) of block b; item r0 == 0 also scans the eigenvalues
__global__ __launch_bounds__(CHK_THREADS) void k_chk_reduce(const ChkBlk *blk, const int2 *items, const double *G,
                                                            const double *w, unsigned long long *res) {
  const int2 it = items[blockIdx.x];
  const ChkBlk B = blk[it.x];
  const int M = B.M, r0 = it.y, r1 = min(M, it.y + B.rpi);
  unsigned long long dn = 0, dof = 0, wf = 0;
  const double *g = G + B.goff;
  const int tot = (r1 - r0) * M;   // at most max(M, 2 * CHK_ELEMS)
  for (int e = threadIdx.x; e < tot; e += CHK_THREADS) {
    const int r = r0 + e / M, c = e % M;
    const double x = g[(int64_t)r * M + c];
    if (r == c) {
      const unsigned long long v = nonneg_bits(fabs(x - 1.0));
      dn = v > dn ? v : dn;
    } else {
      const unsigned long long v = nonneg_bits(fabs(x));
      dof = v > dof ? v : dof;
    }
  }
  if (r0 == 0)
    for (int i = threadIdx.x; i < M; i += CHK_THREADS)
      if (!isfinite(w[B.woff + i])) wf = nonneg_bits(std::numeric_limits<double>::infinity());
  __shared__ unsigned long long red[3][CHK_THREADS / WAVE];
  dn = wave_max(dn);
  dof = wave_max(dof);
  wf = wave_max(wf);
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  if (lane == 0) {
    red[0][wv] = dn;
    red[1][wv] = dof;
    red[2][wv] = wf;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long m = 0;
    for (int k = 0; k < CHK_THREADS / WAVE; k++) m = red[threadIdx.x][k] > m ? red[threadIdx.x][k] : m;
    if (m) atomicMax(res + 3 * (int64_t)it.x + threadIdx.x, m);
  }
}

}  // namespace

}  // namespace nrg

using namespace nrg;

extern "C" int nrg_check_diag_batched(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *rows,
                                      const int64_t *off_V, const double *V_arena, const int64_t *off_w,
                                      const double *w_arena, double eps_norm, double eps_orth, double *dev_norm,
                                      double *dev_orth, int *info, void *stream) {
  try {
    if (!ctx || nblk < 0 || (nblk > 0 && (!n || !rows || !off_V || !V_arena || !off_w || !w_arena))) {
      set_error("nrg_check_diag_batched: invalid arguments");
      return NRG_E_ARG;
    }
    if (nblk == 0) return NRG_OK;
    for (int b = 0; b < nblk; b++)
      if (n[b] < 0 || rows[b] < 0 || rows[b] > n[b] || n[b] > (int64_t)1 << 30 || off_V[b] < 0 || off_w[b] < 0) {
        set_error("nrg_check_diag_batched: block " + std::to_string(b) + " has an invalid shape or offset");
        return NRG_E_ARG;
      }
    NRG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    // Gram matrices: G_b = V_b V_b^T, V_b row-major rows x n (= column-major n x rows, ld n), so the
    // column-major product is C = op(A) op(B) with A = B = V_b, op(A) = A^T
    std::vector<GemmDesc> gd;
    std::vector<ChkBlk> cb(nblk);
    std::vector<int> citems;
    int64_t gtot = 0;
    for (int b = 0; b < nblk; b++) {
      const int M = (int)rows[b];
      cb[b] = ChkBlk{gtot, off_w[b], M, std::max(1, CHK_ELEMS / std::max(M, 1))};
      if (M == 0) continue;
      GemmDesc g{};
      g.A = V_arena + off_V[b]; g.lda = (int)n[b]; g.tA = 1;
      g.B = V_arena + off_V[b]; g.ldb = (int)n[b]; g.tB = 0;
      g.C = nullptr; g.ldc = M;   // C set below, once the workspace is allocated
      g.M = M; g.N = M; g.K = (int)n[b]; g.alpha = 1.0; g.beta = 0.0;
      g.lda2 = g.ldb2 = 1;
      gd.push_back(g);
      for (int r0 = 0; r0 < M; r0 += cb[b].rpi) {
        citems.push_back(b);
        citems.push_back(r0);
      }
      gtot += (int64_t)M * M;
    }
    if (gd.empty()) {
      for (int b = 0; b < nblk; b++) {
        if (dev_norm) dev_norm[b] = 0.0;
        if (dev_orth) dev_orth[b] = 0.0;
        if (info) info[b] = 0;
      }
      return NRG_OK;
    }
    const size_t resb = sizeof(unsigned long long) * 3 * (size_t)nblk;
    const size_t oG = 0, oR = (sizeof(double) * (size_t)gtot + 255) & ~(size_t)255;
    if (int r = ctx->ws2.ensure(oR + resb)) return r;
    double *G = ctx->ws2.as<double>();
    unsigned long long *res = (unsigned long long *)(ctx->ws2.as<char>(oR));
    {
      size_t k = 0;
      for (int b = 0; b < nblk; b++)
        if (rows[b] > 0) gd[k++].C = G + oG / sizeof(double) + cb[b].goff;
    }
    std::vector<int> gitems;
    gemm_items(gd, 0, gd.size(), gitems);
    // descriptor blob: [GemmDesc][ChkBlk][gemm items][check items]
    const size_t bD = sizeof(GemmDesc) * gd.size(), bB = sizeof(ChkBlk) * cb.size();
    const size_t oB = (bD + 255) & ~(size_t)255, oI1 = oB + ((bB + 255) & ~(size_t)255);
    const size_t oI2 = oI1 + sizeof(int) * gitems.size();
    const size_t total = oI2 + sizeof(int) * citems.size();
    char *h = nullptr;
    if (int r = stage_desc(ctx, total, st, &h)) return r;
    std::memcpy(h, gd.data(), bD);
    std::memcpy(h + oB, cb.data(), bB);
    std::memcpy(h + oI1, gitems.data(), sizeof(int) * gitems.size());
    std::memcpy(h + oI2, citems.data(), sizeof(int) * citems.size());
    if (int r = upload_staged(ctx, total, st)) return r;
    NRG_HIP(hipEventRecord(ctx->ev[0], st));
    const int nres = 3 * nblk;
    hipLaunchKernelGGL(k_chk_clear, dim3((unsigned)((nres + 255) / 256)), dim3(256), 0, st, res, nres);
    gemm_launch(ctx, st, (const GemmDesc *)ctx->desc.as<char>(0), ctx->desc.as<int>(oI1), gitems.data(),
                gitems.size() / 2);
    hipLaunchKernelGGL(k_chk_reduce, dim3((unsigned)(citems.size() / 2)), dim3(CHK_THREADS), 0, st,
                       (const ChkBlk *)ctx->desc.as<char>(oB), (const int2 *)ctx->desc.as<int>(oI2), G, w_arena, res);
    NRG_HIP(hipGetLastError());
    NRG_HIP(hipEventRecord(ctx->ev[1], st));
    if (int r = ctx->hflags.ensure(resb)) return r;
    // the kernels complete before the copy is issued (a DMA copy was seen to overtake kernels queued
    // before it under rocprofv3's kernel tracing)
    NRG_HIP(hipStreamSynchronize(st));
    NRG_HIP(hipMemcpyAsync(ctx->hflags.p, res, resb, hipMemcpyDeviceToHost, st));
    NRG_HIP(hipStreamSynchronize(st));
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->total_ms = ms;
    const unsigned long long *hr = (const unsigned long long *)ctx->hflags.p;
    bool bad = false;
    for (int b = 0; b < nblk; b++) {
      double dn, dof, wf;
      std::memcpy(&dn, hr + 3 * b, 8);
      std::memcpy(&dof, hr + 3 * b + 1, 8);
      std::memcpy(&wf, hr + 3 * b + 2, 8);
      const int code = wf != 0.0 ? 3 : !(dn <= eps_norm) ? 1 : !(dof <= eps_orth) ? 2 : 0;
      if (dev_norm) dev_norm[b] = dn;
      if (dev_orth) dev_orth[b] = dof;
      if (info) info[b] = code;
      bad |= code != 0;
    }
    if (bad) {
      set_error("nrg_check_diag_batched: eigenvectors not orthonormal or eigenvalues not finite (see info)");
      return NRG_E_NOCONV;
    }
    return NRG_OK;
  } catch (const std::exception &e) {
    set_error(std::string("nrg_check_diag_batched: ") + e.what());
    return NRG_E_HIP;
  }
}
