// spectral.hpp - what the spectral-weight paths of one step share (fdm.hip: the binned forms,
// mats.hip: the Matsubara forms): device pair descriptors, the density-matrix products each
// algorithm forms on the matrix cores, argument checks, and the replicated accumulators.
#pragma once
#include "gemm.hpp"
#include <cstring>
#include <vector>

namespace nrg {

// A pair with the FDM products (prod_fit, c++/numerics.hpp:283-306).
struct FdmPairDev {
  int64_t dim_i, dim_j, K_i, K_j, ld_op;
  const double *E_i, *E_j, *op1, *op2;
  const double *op2rho;   // (dim_j - K_j) x K_i, row-major, ld K_i:  (op2 rho_i)(K_j + r, c)
  const double *rhoop2;   // K_j x (dim_i - K_i), row-major, ld dim_i - K_i:  (rho_j op2)(r, K_i + c)
  double factor;
};

// A pair with the DM-NRG products over the kept states (c++/algo_DMNRG.hpp:32-44).
struct DmnrgPairDev {
  int64_t K_i, K_j, ld_op;
  const double *E_i, *E_j, *op1, *op2;
  const double *pa;   // K_j x K_i row-major: (op2 rho_i^T)(j, m) over the kept states
  const double *pb;   // K_j x K_i row-major: (rho_j op1)(j, m)
  double factor;
};

// Accumulators are replicated FDM_COPIES times (workgroup w adds into replica w mod FDM_COPIES):
// hot bins and frequencies receive thousands of additions per step, and replicas spread the
// read-modify-writes on one address over many addresses.
constexpr int FDM_COPIES = 32;

// Which density matrices an algorithm reads, for the argument checks:
//   fdm    rho_i where K_i > 0 and dim_j > K_j, rho_j where K_j > 0 and dim_i > K_i, covering K
//   dmnrg  both wherever K_i, K_j > 0, covering K
//   cfs    as fdm, and inside the subspace (n_rho <= dim, prod_fit_left's bound)
enum class RhoUse { none, fdm, dmnrg, cfs };
int check_pairs(const char *who, const nrg_fdm_pair *pairs, int npairs, RhoUse use);

// The FDM products of pair q at buf (GEMM descriptors appended to gd); d gets every field.
// Returns the doubles the products take; buf == nullptr only counts them.
int64_t fdm_products(const nrg_fdm_pair &q, double *buf, FdmPairDev &d, std::vector<GemmDesc> &gd);
// The DM-NRG products of pair q at buf (2 K_i K_j doubles if both are kept), as fdm_products.
int64_t dmnrg_products(const nrg_fdm_pair &q, double *buf, DmnrgPairDev &d, std::vector<GemmDesc> &gd);

// p[0, n) = 0 on stream st
void replicas_zero(nrg_ctx *ctx, hipStream_t st, double *p, size_t n);
// out_h[b] += sum over the FDM_COPIES replicas of histogram h (< nout) at priv + (c * nh + h) * n
void replicas_reduce(nrg_ctx *ctx, hipStream_t st, const double *priv, int64_t n, int nh, int nout, double *o0,
                     double *o1, double *o2, double *o3);

// the same two launches without the per-kernel timers (they create events on first use)
void replicas_zero_untimed(hipStream_t st, double *p, size_t n);
void replicas_reduce_untimed(hipStream_t st, const double *priv, int64_t n, int nh, int nout, double *o0, double *o1,
                             double *o2, double *o3);
// k_temp_sum over nitems (pair, tile) items into FDM_COPIES partial sums at priv (mats.hip)
void temp_launch(hipStream_t st, int kind, const FdmPairDev *pdev, const int2 *idev, unsigned nitems, double beta,
                 double Z, double *priv);

// One step of a spectral path: uploads [GEMM descriptors | pair descriptors | GEMM items | element
// items] as one blob, runs the products on the matrix cores, then kernels(pairs_dev, items_dev) with
// the element items as int2 (pair, tile). Timing: phase 0 the products, phase 1 the rest.
template <class PD, class F>
int run_spectral_step(nrg_ctx *ctx, hipStream_t st, const std::vector<GemmDesc> &gd, const std::vector<PD> &pd,
                      const std::vector<int> &items, F &&kernels) {
  std::vector<int> gitems;
  gemm_items(gd, 0, gd.size(), gitems);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t bG = sizeof(GemmDesc) * gd.size(), bPd = sizeof(PD) * pd.size();
  const size_t oG = 0, oPd = al(bG), oGI = oPd + al(bPd), oI = oGI + al(sizeof(int) * gitems.size());
  std::vector<char> blob(oI + sizeof(int) * items.size() + 8);
  if (bG) memcpy(blob.data() + oG, gd.data(), bG);
  if (bPd) memcpy(blob.data() + oPd, pd.data(), bPd);
  if (!gitems.empty()) memcpy(blob.data() + oGI, gitems.data(), sizeof(int) * gitems.size());
  if (!items.empty()) memcpy(blob.data() + oI, items.data(), sizeof(int) * items.size());
  NRG_HIP(hipEventRecord(ctx->ev[0], st));
  if (int r = upload_desc(ctx, blob.data(), blob.size(), st)) return r;
  if (!gitems.empty())
    gemm_launch(ctx, st, (const GemmDesc *)ctx->desc.as<char>(oG), ctx->desc.as<int>(oGI), gitems.data(), gitems.size() / 2);
  NRG_HIP(hipEventRecord(ctx->ev[1], st));
  if (!items.empty()) kernels((const PD *)ctx->desc.as<char>(oPd), (const int2 *)ctx->desc.as<char>(oI));
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[2], st));
  NRG_HIP(hipStreamSynchronize(st));
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[2]);
  ctx->total_ms = ms;
  hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
  ctx->phase_ms[0] = ms;
  hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]);
  ctx->phase_ms[1] = ms;
  if (ctx->kt.on) ctx->kt.collect();
  return NRG_OK;
}

}  // namespace nrg
