// broaden.hip - broadening of a binned spectrum onto an energy mesh (nrg_broaden).
//
// Replaces the double loop of SpectrumRealFreq::continuous (c++/spectrum.hpp:180-209): every mesh
// energy E gets sum_peaks w BR_NEW(+-E, +-e), with BR_NEW the log-Gaussian kernel of width alpha
// blended into a Gaussian of width omega0 for |E| <= omega0 (c++/spectral.hpp:30-71). One
// workgroup per mesh point; the peaks are split over its threads and the two sums (A(+E), A(-E))
// are reduced in the workgroup.
#include "common.hpp"
#include <cmath>

namespace nrg {

constexpr double SQRTPI = 1.7724538509055160273;

// BR_L (c++/spectral.hpp:39-44): log-Gaussian, zero across the sign boundary, gamma = alpha / 4
__device__ __forceinline__ double br_l(double e, double ept, double alpha) {
  if ((e < 0.0 && ept > 0.0) || (e > 0.0 && ept < 0.0)) return 0.0;
  if (ept == 0.0) return 0.0;
  const double gamma = alpha / 4.0;
  const double u = log(e / ept) / alpha - gamma;
  return exp(-(u * u)) / (alpha * fabs(e) * SQRTPI);
}

// BR_NEW (c++/spectral.hpp:57-71); bh = exp(-(log(|e|/omega0)/alpha)^2) depends on e only
__device__ __forceinline__ double br_new(double e, double ept, double alpha, double omega0, double bh, bool low) {
  const double pl = br_l(e, ept, alpha);
  if (!low) return pl;
  const double v = (e - ept) / omega0;
  return pl * bh + exp(-(v * v)) / (omega0 * SQRTPI) * (1.0 - bh);
}

__global__ __launch_bounds__(256) void k_broaden(const double *pos_e, const double *pos_w, int64_t npos,
                                                 const double *neg_e, const double *neg_w, int64_t nneg,
                                                 const double *mesh, double alpha, double omega0, double *out_pos,
                                                 double *out_neg) {
  __shared__ double red[2][256 / WAVE];
  const double E = mesh[blockIdx.x];
  const bool low = fabs(E) <= omega0;   // |e| > omega0: BR_NEW is the log-Gaussian alone
  double bh = 0.0;
  if (low) {
    const double u = log(fabs(E) / omega0) / alpha;
    bh = exp(-(u * u));
  }
  double vp = 0.0, vn = 0.0;
  for (int64_t i = threadIdx.x; i < npos; i += 256) {
    const double e = pos_e[i], w = pos_w[i];
    vp += w * br_new(E, e, alpha, omega0, bh, low);
    vn += w * br_new(-E, e, alpha, omega0, bh, low);
  }
  for (int64_t i = threadIdx.x; i < nneg; i += 256) {
    const double e = neg_e[i], w = neg_w[i];
    vn += w * br_new(-E, -e, alpha, omega0, bh, low);
    vp += w * br_new(E, -e, alpha, omega0, bh, low);
  }
  vp = wave_sum(vp);
  vn = wave_sum(vn);
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  if (lane == 0) {
    red[0][wv] = vp;
    red[1][wv] = vn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sp = 0.0, sn = 0.0;
    for (int w = 0; w < 256 / WAVE; w++) {
      sp += red[0][w];
      sn += red[1][w];
    }
    out_pos[blockIdx.x] = sp;
    out_neg[blockIdx.x] = sn;
  }
}

int broaden_run(nrg_ctx *ctx, const double *pos_e, const double *pos_w, int64_t npos, const double *neg_e,
                const double *neg_w, int64_t nneg, const double *mesh, int64_t nmesh, double alpha, double omega0,
                double *out_pos, double *out_neg, hipStream_t st) {
  if (npos < 0 || nneg < 0 || nmesh < 0 || (npos && (!pos_e || !pos_w)) || (nneg && (!neg_e || !neg_w)) ||
      (nmesh && (!mesh || !out_pos || !out_neg)) || !(alpha > 0) || !(omega0 > 0) || nmesh > INT32_MAX) {
    set_error("nrg_broaden: invalid arguments");
    return NRG_E_ARG;
  }
  NRG_HIP(hipEventRecord(ctx->ev[0], st));
  if (nmesh)
    NRG_TIMED(ctx, st, "k_broaden", hipLaunchKernelGGL(k_broaden, dim3((unsigned)nmesh), dim3(256), 0, st, pos_e, pos_w,
                                                       npos, neg_e, neg_w, nneg, mesh, alpha, omega0, out_pos, out_neg));
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
