// mats.hip - Matsubara-axis Green's functions of one NRG step (nrg_fdm_mats, nrg_dmnrg_mats,
// nrg_ft_mats).
//
// Replaces, for one step, the pair loop of BaseSpectrum::calc (c++/algo.hpp:55-67) into
// Algo_FDMmats::calc (c++/algo_FDM.hpp:167-213), Algo_DMNRGmats::calc (c++/algo_DMNRG.hpp:74-100) and
// Algo_FTmats::calc (c++/algo_FT.hpp:59-84) with ChainMatsubara::add (c++/spectrum.hpp:34-42): every
// evaluated element (i, j) of every pair adds  factor * w / (i omega_n - E)  to G(i omega_n) for
// n < mats, with E = E_j - E_i, omega_n = T pi (2n + 1) for fermions and T pi 2n for bosons
// (ww, c++/matsubara.hpp:20-27), and, for bosons at n = 0 where |E| <= WEIGHT_TOL and the term has
// that limit, factor * (-s / T) instead.
//
// Launch 1 forms the density-matrix products of the algorithm (the same ones as the binned forms) on
// the FP64 matrix cores. Launch 2 takes a 64 x 64 element tile per workgroup in four 64 x 16 slabs:
// the 256 threads evaluate a slab's 1024 elements into LDS as (E^2, -w E, -w), then each thread sums
// those elements for its own frequencies,
//   Re += -w E / (E^2 + omega^2),   Im += -w / (E^2 + omega^2)   (times omega at the end),
// with the reciprocal from v_rcp_f64 and two Newton steps, and adds its sums into replicated
// accumulators once per tile. The bosonic n = 0 term is summed while the elements are evaluated.
//
// nrg_temp_sum (temp_run, k_temp_sum) is the one-number-per-step kind of the same loop: the
// conductance and moments G(T), I1(T), I2(T) and the susceptibility chi(T) of c++/algo_FT.hpp.
#include "spectral.hpp"
#include <cmath>

namespace nrg {

constexpr double WEIGHT_TOL = 1e-8;   // c++/numerics.hpp:318
constexpr int MATS_FDM = 0, MATS_DMNRG = 1, MATS_FT = 2;

struct MatsArgs {
  const void *pairs;   // FdmPairDev (FDM, FT) or DmnrgPairDev (DM-NRG)
  double T, wnf, Z, msign;
  int bosonic;
  int mats;
  int n_lo;    // first frequency of the slab sums: 1 for bosons (n = 0 is summed apart), else 0
  int ntn;     // threads per frequency group: 64, 128 or 256 (256 / ntn groups split the elements);
               // a group covers ntn * NPT frequencies of a chunk
  double *priv;   // FDM_COPIES replicas of 2 * mats doubles (Re, Im interleaved)
};

// 1 / d for d > 0 normal: v_rcp_f64 refined by two Newton-Raphson steps r <- r + r (1 - d r)
__device__ __forceinline__ double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

template <int MODE>
using MatsPair = typename std::conditional<MODE == MATS_DMNRG, DmnrgPairDev, FdmPairDev>::type;

// Element (i, j): returns false if the algorithm does not evaluate it; else E = E_j - E_i, the generic
// numerator w and the numerator s of the bosonic n = 0 limit (-s / T), both times the pair factor,
// and lim = whether that limit applies to this term.
template <int MODE>
__device__ __forceinline__ bool mats_element(const MatsArgs &a, const MatsPair<MODE> &P, int i, int j, double Ei,
                                             double Ej, double Bi, double Bj, double &w, double &s, bool &lim) {
  if constexpr (MODE == MATS_FDM) {
    if (i >= P.dim_i || j >= P.dim_j) return false;
    const bool ki = i < P.K_i, kj = j < P.K_j;
    if (ki && kj) return false;
    const size_t e = (size_t)j * P.ld_op + i;
    const double o1 = P.op1[e], o2 = P.op2[e];
    double wA, wB;
    if (!ki && !kj) {   // term1, D_i x D_j
      wA = o1 * o2 * a.wnf * Bi;
      wB = o1 * o2 * a.msign * a.wnf * Bj;
      lim = true;
    } else if (!ki) {   // term2, D_i x K_j: rho_j op2
      wA = o1 * o2 * a.wnf * Bi;
      wB = o1 * P.rhoop2[(size_t)j * (P.dim_i - P.K_i) + (i - P.K_i)] * a.msign;
      lim = false;
    } else {            // term3, K_i x D_j: op2 rho_i
      wA = o1 * P.op2rho[(size_t)(j - P.K_j) * P.K_i + i];
      wB = a.msign * o1 * o2 * a.wnf * Bj;
      lim = false;
    }
    w = (wA + wB) * P.factor;
    s = wA * P.factor;
    return true;
  } else if constexpr (MODE == MATS_DMNRG) {
    if (i >= P.K_i || j >= P.K_j) return false;
    const size_t e = (size_t)j * P.ld_op + i, q = (size_t)j * P.K_i + i;
    const double wA = P.pa[q] * P.op1[e];
    const double wB = P.pb[q] * P.op2[e];
    w = (wA + a.msign * wB) * P.factor;
    s = wA * P.factor;
    lim = true;
    return true;
  } else {   // MATS_FT
    if (i >= P.K_i || j >= P.K_j) return false;
    const size_t e = (size_t)j * P.ld_op + i;
    const double oo = P.op1[e] * P.op2[e];
    w = oo * ((a.msign * Bj + Bi) / a.Z) * P.factor;
    s = oo * (Bj / a.Z) * P.factor;
    lim = true;
    return true;
  }
}

template <int MODE, int NPT>
__global__ __launch_bounds__(256) void k_mats(MatsArgs a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const MatsPair<MODE> P = ((const MatsPair<MODE> *)a.pairs)[it.x];
  int64_t dim_i, dim_j;
  if constexpr (MODE == MATS_FDM) {
    dim_i = P.dim_i;
    dim_j = P.dim_j;
  } else {
    dim_i = P.K_i;
    dim_j = P.K_j;
  }
  const int ti = (int)((dim_i + 63) / 64);
  const int i0 = (it.y % ti) * 64, j0 = (it.y / ti) * 64;
  const int tid = threadIdx.x;
  __shared__ double Ei[64], Ej[64], Bi[64], Bj[64];
  __shared__ double sa[1024], sb[1024], sc[1024];
  __shared__ double red[4];
  if (tid < 64) {
    const int i = i0 + tid;
    const double e = i < dim_i ? P.E_i[i] : 0.0;
    Ei[tid] = e;
    Bi[tid] = MODE == MATS_DMNRG ? 0.0 : exp(-e / a.T);
  } else if (tid < 128) {
    const int j = j0 + tid - 64;
    const double e = j < dim_j ? P.E_j[j] : 0.0;
    Ej[tid - 64] = e;
    Bj[tid - 64] = MODE == MATS_DMNRG ? 0.0 : exp(-e / a.T);
  }
  // this thread's frequencies of chunk blockIdx.y: n = n_lo + chunk * ntn * NPT + lane + ntn * k
  const int lane = tid % a.ntn, grp = tid / a.ntn, ngrp = 256 / a.ntn;
  const int nbase = a.n_lo + (int)blockIdx.y * a.ntn * NPT + lane;
  double om[NPT], om2[NPT], re[NPT], im[NPT];
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int n = nbase + a.ntn * k;
    om[k] = a.T * M_PI * (double)(2 * n + (a.bosonic ? 0 : 1));
    om2[k] = om[k] * om[k];
    re[k] = 0.0;
    im[k] = 0.0;
  }
  double z0 = 0.0;   // bosonic n = 0, chunk 0 only
  const bool do_z0 = a.bosonic && blockIdx.y == 0;
  for (int slab = 0; slab < 4; slab++) {
    __syncthreads();   // energies loaded; the previous slab's sums done
    const int il = tid & 63;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int jl = slab * 16 + (tid >> 6) + 4 * q;   // 0..63
      double w = 0.0, sv = 0.0;
      bool lim = false;
      const bool ok = mats_element<MODE>(a, P, i0 + il, j0 + jl, Ei[il], Ej[jl], Bi[il], Bj[jl], w, sv, lim);
      const double E = Ej[jl] - Ei[il];
      const int e = ((tid >> 6) + 4 * q) * 64 + il;
      if (ok) {
        sa[e] = E * E;
        sb[e] = -w * E;
        sc[e] = -w;
        // n = 0 for bosons: w / (0 i - E), or the limit -s / T
        if (do_z0) z0 += (lim && fabs(E) <= WEIGHT_TOL) ? -sv / a.T : -w / E;
      } else {
        sa[e] = 1.0;
        sb[e] = 0.0;
        sc[e] = 0.0;
      }
    }
    __syncthreads();
    for (int e = grp; e < 1024; e += ngrp) {
      const double ea = sa[e], eb = sb[e], ec = sc[e];
#pragma unroll
      for (int k = 0; k < NPT; k++) {
        const double r = recip(ea + om2[k]);
        re[k] = fma(eb, r, re[k]);
        im[k] = fma(ec, r, im[k]);
      }
    }
  }
  double *acc = a.priv + (size_t)(blockIdx.x % FDM_COPIES) * 2 * a.mats;
#pragma unroll
  for (int k = 0; k < NPT; k++) {
    const int n = nbase + a.ntn * k;
    if (n < a.mats) {
      atomicAdd(acc + 2 * n, re[k]);
      atomicAdd(acc + 2 * n + 1, im[k] * om[k]);
    }
  }
  if (do_z0) {   // workgroup sum of z0: wave reduction, then the four waves
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) z0 += __shfl_xor(z0, off);
    if ((tid & 63) == 0) red[tid >> 6] = z0;
    __syncthreads();
    if (tid == 0) atomicAdd(acc, red[0] + red[1] + red[2] + red[3]);
  }
}

template <int MODE>
void launch_mats(int npt, dim3 grid, hipStream_t st, const MatsArgs &ma, const int2 *idev) {
  if (npt == 1) hipLaunchKernelGGL((k_mats<MODE, 1>), grid, dim3(256), 0, st, ma, idev);
  else if (npt == 2) hipLaunchKernelGGL((k_mats<MODE, 2>), grid, dim3(256), 0, st, ma, idev);
  else if (npt == 4) hipLaunchKernelGGL((k_mats<MODE, 4>), grid, dim3(256), 0, st, ma, idev);
  else hipLaunchKernelGGL((k_mats<MODE, 8>), grid, dim3(256), 0, st, ma, idev);
}

int mats_run(nrg_ctx *ctx, int mode, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, double wnf,
             double Z, int sign, double *G, hipStream_t st) {
  const char *who = mode == MATS_FDM ? "nrg_fdm_mats" : mode == MATS_DMNRG ? "nrg_dmnrg_mats" : "nrg_ft_mats";
  if (npairs < 0 || (npairs > 0 && !pairs) || mats < 0 || mats > 32767 || (mats > 0 && !G) || !(T > 0) ||
      (sign != 1 && sign != -1) || (mode == MATS_FT && !(Z > 0))) {
    set_error(std::string(who) + ": invalid arguments");
    return NRG_E_ARG;
  }
  const RhoUse use = mode == MATS_FDM ? RhoUse::fdm : mode == MATS_DMNRG ? RhoUse::dmnrg : RhoUse::none;
  if (int r = check_pairs(who, pairs, npairs, use)) return r;
  const bool bosonic = sign == 1;   // gf_sign: 1 for bosons, -1 for fermions
  const int n_lo = bosonic ? 1 : 0;
  const int64_t nsum = std::max<int64_t>(mats - n_lo, 0);
  // frequency slots per chunk ntn * npt close above nsum: up to 512 one wave per group (the four
  // waves split the elements), then 1024 in groups of two waves, then chunks of 2048 (all waves)
  int ntn = 64, npt = 1;
  while (npt < 8 && 64 * npt < nsum) npt *= 2;
  if (nsum > 512) ntn = nsum > 1024 ? 256 : 128;
  const int nchunks = (int)std::max<int64_t>((nsum + ntn * npt - 1) / (ntn * npt), 1);
  std::vector<GemmDesc> gd;
  std::vector<FdmPairDev> pf;
  std::vector<DmnrgPairDev> pm;
  std::vector<int> items;
  int64_t tot = 0;
  for (int p = 0; p < npairs && mode != MATS_FT; p++) {
    FdmPairDev f;
    DmnrgPairDev m;
    tot += mode == MATS_FDM ? fdm_products(pairs[p], nullptr, f, gd) : dmnrg_products(pairs[p], nullptr, m, gd);
  }
  if (int r = ctx->ws2.ensure(sizeof(double) * (size_t)(tot + 1))) return r;
  if (int r = ctx->ws.ensure(sizeof(double) * (size_t)FDM_COPIES * 2 * (size_t)std::max<int64_t>(mats, 1))) return r;
  double *prod = ctx->ws2.as<double>();
  for (int64_t p = 0, off = 0; p < npairs && mats > 0; p++) {
    const nrg_fdm_pair &q = pairs[p];
    int64_t di = q.dim_i, dj = q.dim_j;
    if (mode != MATS_FDM) {
      if (!(q.K_i > 0 && q.K_j > 0)) continue;
      di = q.K_i;
      dj = q.K_j;
    }
    const int64_t id = mode == MATS_DMNRG ? (int64_t)pm.size() : (int64_t)pf.size();
    if (mode == MATS_DMNRG) {
      DmnrgPairDev m;
      off += dmnrg_products(q, prod + off, m, gd);
      pm.push_back(m);
    } else {
      FdmPairDev f;
      off += fdm_products(q, mode == MATS_FDM ? prod + off : nullptr, f, gd);
      if (mode == MATS_FT) f.op2rho = f.rhoop2 = nullptr;
      pf.push_back(f);
    }
    // element tiles; FDM skips those entirely in K_i x K_j
    const int ti = (int)((di + 63) / 64), tj = (int)((dj + 63) / 64);
    for (int b = 0; b < tj; b++)
      for (int c = 0; c < ti; c++) {
        if (mode == MATS_FDM && (int64_t)(c + 1) * 64 <= q.K_i && (int64_t)(b + 1) * 64 <= q.K_j) continue;
        items.push_back((int)id);
        items.push_back(b * ti + c);
      }
  }
  if (items.size() / 2 > (size_t)INT32_MAX) {
    set_error(std::string(who) + ": too many element tiles");
    return NRG_E_ARG;
  }
  auto launch = [&](const void *pdev, const int2 *idev) {
    MatsArgs ma;
    ma.pairs = pdev;
    ma.T = T; ma.wnf = wnf; ma.Z = Z; ma.msign = -(double)sign;
    ma.bosonic = bosonic ? 1 : 0;
    ma.mats = (int)mats;
    ma.n_lo = n_lo;
    ma.ntn = ntn;
    ma.priv = ctx->ws.as<double>();
    replicas_zero(ctx, st, ma.priv, (size_t)FDM_COPIES * 2 * mats);
    const dim3 grid((unsigned)(items.size() / 2), (unsigned)nchunks);
    if (mode == MATS_FDM)
      NRG_TIMED(ctx, st, "k_mats_fdm", launch_mats<MATS_FDM>(npt, grid, st, ma, idev));
    else if (mode == MATS_DMNRG)
      NRG_TIMED(ctx, st, "k_mats_dmnrg", launch_mats<MATS_DMNRG>(npt, grid, st, ma, idev));
    else
      NRG_TIMED(ctx, st, "k_mats_ft", launch_mats<MATS_FT>(npt, grid, st, ma, idev));
    replicas_reduce(ctx, st, ma.priv, 2 * mats, 1, 1, G, G, G, G);
  };
  if (mode == MATS_DMNRG)
    return run_spectral_step(ctx, st, gd, pm, items, [&](const DmnrgPairDev *d, const int2 *i) { launch(d, i); });
  return run_spectral_step(ctx, st, gd, pf, items, [&](const FdmPairDev *d, const int2 *i) { launch(d, i); });
}

// ------------------------------------------------------------------------------------------
// Temperature-dependent sums of one step: Algo_GT (n = 0, 1, 2: GT, I1T, I2T, c++/algo_FT.hpp:95-134)
// and Algo_CHIT (c++/algo_FT.hpp:136-170): one number per step, the sum over the kept x kept
// elements of every pair of factor * op1 op2 * stat(E_1, E_p), E_1 = E_j, E_p = E_i.
// ------------------------------------------------------------------------------------------
constexpr int TEMP_GT = 0, TEMP_I1T = 1, TEMP_I2T = 2, TEMP_CHIT = 3;

struct TempArgs {
  const FdmPairDev *pairs;
  double beta, Z;
  double *priv;   // FDM_COPIES partial sums
};

// chit_weight (c++/numerics.hpp:323-336): (exp(-beta Em) - exp(-beta En)) / (beta En - beta Em), or
// exp(-beta Em) where the two are within WEIGHT_TOL
__device__ __forceinline__ double chit_weight(double En, double Em, double beta) {
  const double bn = beta * En, bm = beta * Em, x = bn - bm;
  if (fabs(x) > WEIGHT_TOL) return (exp(-bm) - exp(-bn)) / x;
  return exp(-bm);
}

template <int KIND>
__global__ __launch_bounds__(256) void k_temp_sum(TempArgs a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const FdmPairDev P = a.pairs[it.x];
  const int ti = (int)((P.K_i + 63) / 64);
  const int i = (it.y % ti) * 64 + (threadIdx.x & 63), j0 = (it.y / ti) * 64;
  double acc = 0.0;
  if (i < P.K_i) {
    const double Ep = P.E_i[i];
    for (int jl = threadIdx.x >> 6; jl < 64; jl += 4) {
      const int j = j0 + jl;
      if (j >= P.K_j) break;
      const double E1 = P.E_j[j];
      const size_t e = (size_t)j * P.ld_op + i;
      double stat;
      if constexpr (KIND == TEMP_CHIT) {
        stat = chit_weight(E1, Ep, a.beta) / a.Z;
      } else {
        const double d = E1 - Ep;
        const double pw = KIND == TEMP_GT ? 1.0 : KIND == TEMP_I1T ? d : d * d;
        stat = a.beta / (exp(a.beta * E1) + exp(a.beta * Ep)) * pw / a.Z;
      }
      acc += P.op1[e] * P.op2[e] * stat;
    }
  }
  acc *= P.factor;
  __shared__ double red[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(a.priv + blockIdx.x % FDM_COPIES, red[0] + red[1] + red[2] + red[3]);
}

// k_temp_sum over `nitems` (pair, tile) items, partial sums into priv[0, FDM_COPIES), no timing
void temp_launch(hipStream_t st, int kind, const FdmPairDev *pdev, const int2 *idev, unsigned nitems, double beta,
                 double Z, double *priv) {
  if (!nitems) return;
  const TempArgs ta{pdev, beta, Z, priv};
  if (kind == TEMP_GT) hipLaunchKernelGGL(k_temp_sum<TEMP_GT>, dim3(nitems), dim3(256), 0, st, ta, idev);
  else if (kind == TEMP_I1T) hipLaunchKernelGGL(k_temp_sum<TEMP_I1T>, dim3(nitems), dim3(256), 0, st, ta, idev);
  else if (kind == TEMP_I2T) hipLaunchKernelGGL(k_temp_sum<TEMP_I2T>, dim3(nitems), dim3(256), 0, st, ta, idev);
  else hipLaunchKernelGGL(k_temp_sum<TEMP_CHIT>, dim3(nitems), dim3(256), 0, st, ta, idev);
}

int temp_run(nrg_ctx *ctx, int kind, const nrg_fdm_pair *pairs, int npairs, double temperature, double Z, double *out,
             hipStream_t st) {
  if (kind < TEMP_GT || kind > TEMP_CHIT || npairs < 0 || (npairs > 0 && !pairs) || !out || !(temperature > 0) ||
      !(Z > 0)) {
    set_error("nrg_temp_sum: invalid arguments");
    return NRG_E_ARG;
  }
  if (int r = check_pairs("nrg_temp_sum", pairs, npairs, RhoUse::none)) return r;
  if (int r = ctx->ws.ensure(sizeof(double) * FDM_COPIES)) return r;
  std::vector<GemmDesc> gd;   // no products
  std::vector<FdmPairDev> pd;
  std::vector<int> items;
  for (int p = 0; p < npairs; p++) {
    const nrg_fdm_pair &q = pairs[p];
    if (!(q.K_i > 0 && q.K_j > 0)) continue;
    FdmPairDev d;
    fdm_products(q, nullptr, d, gd);
    const int ti = (int)((q.K_i + 63) / 64), tj = (int)((q.K_j + 63) / 64);
    for (int t = 0; t < ti * tj; t++) {
      items.push_back((int)pd.size());
      items.push_back(t);
    }
    pd.push_back(d);
  }
  return run_spectral_step(ctx, st, gd, pd, items, [&](const FdmPairDev *pdev, const int2 *idev) {
    TempArgs ta{pdev, 1.0 / temperature, Z, ctx->ws.as<double>()};
    replicas_zero(ctx, st, ta.priv, FDM_COPIES);
    const dim3 grid((unsigned)(items.size() / 2));
    if (kind == TEMP_GT)
      NRG_TIMED(ctx, st, "k_temp_sum", hipLaunchKernelGGL(k_temp_sum<TEMP_GT>, grid, dim3(256), 0, st, ta, idev));
    else if (kind == TEMP_I1T)
      NRG_TIMED(ctx, st, "k_temp_sum", hipLaunchKernelGGL(k_temp_sum<TEMP_I1T>, grid, dim3(256), 0, st, ta, idev));
    else if (kind == TEMP_I2T)
      NRG_TIMED(ctx, st, "k_temp_sum", hipLaunchKernelGGL(k_temp_sum<TEMP_I2T>, grid, dim3(256), 0, st, ta, idev));
    else
      NRG_TIMED(ctx, st, "k_temp_sum", hipLaunchKernelGGL(k_temp_sum<TEMP_CHIT>, grid, dim3(256), 0, st, ta, idev));
    replicas_reduce(ctx, st, ta.priv, 1, 1, 1, out, out, out, out);
  });
}

}  // namespace nrg
