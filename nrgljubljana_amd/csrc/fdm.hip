// fdm.hip - FDM spectral weights of one NRG step, binned on the logarithmic mesh (nrg_fdm_bin).
//
// Replaces, for one step, the pair loop of BaseSpectrum::calc (c++/algo.hpp:55-67) into
// Algo_FDM::calc = Algo_FDMgt::calc + Algo_FDMls::calc (c++/algo_FDM.hpp:30-60, 84-114, 136-142)
// and the binning ChainBinning::add -> Bins::add -> Bins::add_std (c++/spectrum.hpp:19-28,
// c++/bins.hpp:103-132). For a pair (Ii, Ij) with op(j, i) = <j|op|i> (dim_j x dim_i, row-major):
//
//   greater:  D_i x D_j, D_i x K_j:  op1 op2 exp(-E_i/T) wnf      K_i x D_j:  op1 (op2 rho_i)
//   lesser:   D_i x D_j, K_i x D_j:  -sign op1 op2 exp(-E_j/T) wnf   D_i x K_j:  -sign op1 (rho_j op2)
//
// each at energy E_j - E_i, times the pair factor. Launch 1 forms the two products (only the
// sub-blocks the three ranges use) on the FP64 matrix cores; launch 2 evaluates every element of
// every pair in 64 x 64 tiles (Boltzmann factors of the tile's states in LDS, one log10 per element
// shared by both halves) and scatters into replicated bins with FP64 atomics, which one more launch
// adds into the caller's arrays.
//
// nrg_dmnrg_bin (dmnrg_run, k_dmnrg_bin) does the same for Algo_DMNRG::calc (c++/algo_DMNRG.hpp:30-52):
// the kept x kept elements of every pair, with the two density-matrix products formed on the matrix
// cores and an energy window instead of Boltzmann factors.
//
// nrg_cfs_bin (cfs_run, k_cfs_bin) does it for Algo_CFS::calc (c++/algo_CFS.hpp:13-144): at the last
// step the kept x kept elements with Boltzmann factors, otherwise the discarded x kept (ls) and
// kept x discarded (gt) elements with the density-matrix products op2 rho_i and rho_j^T op1.
#include "spectral.hpp"
#include <cmath>
#include <cstring>

namespace nrg {

struct FdmArgs {
  const FdmPairDev *pairs;
  double T, wnf, msign;   // msign = -sign
  double log10emin, bpd, discard;
  int64_t nbins;
  double *gt_pos, *gt_neg, *ls_pos, *ls_neg;
};

// Bin position of a non-negative energy (Bins::add_std): below zero_epsilon = 1e-14 (c++/bins.hpp:46)
// or below the grid -> bin 0, above it -> the last bin (low = true: the whole weight in bin idx);
// otherwise split between idx and idx + 1 with weights 1 - rem and rem.
__device__ __forceinline__ void bin_pos(double energy, double log10emin, double bpd, int64_t nbins, int64_t &idx,
                                        double &rem, bool &low) {
  idx = 0;
  rem = 0.0;
  low = true;
  if (energy < 1e-14) return;
  const double x = (log10(energy) - log10emin) * bpd;
  const double ip = floor(x);
  if (ip < 0) return;
  if (ip >= (double)(nbins - 1)) {
    idx = nbins - 1;
    return;
  }
  idx = (int64_t)ip;
  rem = x - ip;
  low = false;
}

// A tile's window of the histograms in LDS. Along a row of a tile the energy difference E_j - E_i is
// monotone and the tile's differences span a narrow range of bins, except near zero: for each sign s
// of E_j - E_i (0: >= 0) and each of NH histograms, the window holds the WINB bins below the tile's
// largest |E_j - E_i| of that sign (one decade at 1000 bins per decade). An element whose bins fall
// inside adds there with an LDS atomic; the others add to the replica in global memory; the window is
// then added into the replica, one global atomic per non-zero bin.
constexpr int WINB = 1024;

template <int NH>
struct BinWindow {
  double *lds;     // [2][NH][WINB]
  int64_t lo[2];   // first bin of each sign's window
  __device__ double *at(int s, int h) const { return lds + (s * NH + h) * WINB; }
};

// lo of a sign's window from the largest difference dtop of that sign (none when dtop < 0)
__device__ __forceinline__ int64_t window_lo(double dtop, double log10emin, double bpd, int64_t nbins) {
  int64_t idx;
  double rem;
  bool low;
  bin_pos(dtop > 0.0 ? dtop : 0.0, log10emin, bpd, nbins, idx, rem, low);
  int64_t lo = idx + 2 - WINB;
  lo = lo < nbins - WINB ? lo : nbins - WINB;
  return lo > 0 ? lo : 0;
}

// Bins::add + Bins::add_std (c++/bins.hpp:103-132) with the energy already made non-negative: into the
// window when both bins of the element are inside it, else into the replica
__device__ __forceinline__ void win_add(double discard, double *win, int64_t lo, double *bins, double energy, double w,
                                        int64_t idx, double rem, bool low) {
  if (w == 0.0 || fabs(w) < discard * energy) return;
  const int64_t b = idx - lo;
  if (low) {
    if (b >= 0 && b < WINB) atomicAdd(win + b, w);
    else atomicAdd(bins + idx, w);
    return;
  }
  if (b >= 0 && b + 1 < WINB) {
    atomicAdd(win + b, (1.0 - rem) * w);
    atomicAdd(win + b + 1, rem * w);
  } else {
    atomicAdd(bins + idx, (1.0 - rem) * w);
    atomicAdd(bins + idx + 1, rem * w);
  }
}

// Window placement: the tile's energies are in LDS (e_i[0, ni), e_j[0, nj)); wave 0 reduces e_i, wave 1
// e_j, and every thread then reads the two windows' first bins. Also clears the window.
template <int NH>
__device__ __forceinline__ void window_setup(BinWindow<NH> &w, double *lds, const double *e_i, int ni, const double *e_j,
                                             int nj, double log10emin, double bpd, int64_t nbins, double *mm) {
  w.lds = lds;
  for (int k = threadIdx.x; k < 2 * NH * WINB; k += 256) lds[k] = 0.0;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (wv < 2) {
    const double *e = wv == 0 ? e_i : e_j;
    const bool in = l < (wv == 0 ? ni : nj);
    const double hi = wave_max(in ? e[l] : -INFINITY), lo = -wave_max(in ? -e[l] : -INFINITY);
    if (l == 0) {
      mm[2 * wv] = lo;
      mm[2 * wv + 1] = hi;
    }
  }
  __syncthreads();
  // mm: min E_i, max E_i, min E_j, max E_j
  w.lo[0] = window_lo(mm[3] - mm[0], log10emin, bpd, nbins);
  w.lo[1] = window_lo(mm[1] - mm[2], log10emin, bpd, nbins);
}

// adds the window into the replicas (hist(s, h) = the replica of histogram h for sign s)
template <int NH, typename F>
__device__ __forceinline__ void window_flush(const BinWindow<NH> &w, int64_t nbins, F hist) {
  __syncthreads();
  for (int k = threadIdx.x; k < 2 * NH * WINB; k += 256) {
    const double v = w.lds[k];
    const int s = k / (NH * WINB), h = (k / WINB) % NH;
    const int64_t b = w.lo[s] + k % WINB;
    if (v != 0.0 && b < nbins) atomicAdd(hist(s, h) + b, v);
  }
}

// The scatter goes to FDM_COPIES replicas of the histograms; k_fdm_reduce then adds the replicas
// into the caller's arrays.

__global__ __launch_bounds__(256) void k_fdm_zero(double *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0.0;
}

// replica c of histogram h (< nh) at priv + (c * nh + h) * nbins; FDM: h = 0 gt_pos, 1 gt_neg,
// 2 ls_pos, 3 ls_neg; DM-NRG: 0 pos, 1 neg
__global__ __launch_bounds__(256) void k_fdm_reduce(const double *priv, int64_t nbins, int nh, double *o0, double *o1,
                                                    double *o2, double *o3) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int h = blockIdx.y;
  if (b >= nbins) return;
  double s = 0.0;
#pragma unroll 8
  for (int c = 0; c < FDM_COPIES; c++) s += priv[((size_t)c * nh + h) * nbins + b];
  double *out = h == 0 ? o0 : h == 1 ? o1 : h == 2 ? o2 : o3;
  out[b] += s;
}

void replicas_zero(nrg_ctx *ctx, hipStream_t st, double *p, size_t n) {
  NRG_TIMED(ctx, st, "k_fdm_zero", hipLaunchKernelGGL(k_fdm_zero, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)),
                                                      dim3(256), 0, st, p, n));
}

void replicas_reduce(nrg_ctx *ctx, hipStream_t st, const double *priv, int64_t n, int nh, int nout, double *o0,
                     double *o1, double *o2, double *o3) {
  NRG_TIMED(ctx, st, "k_fdm_reduce", hipLaunchKernelGGL(k_fdm_reduce, dim3((unsigned)((n + 255) / 256), nout), dim3(256),
                                                        0, st, priv, n, nh, o0, o1, o2, o3));
}

void replicas_zero_untimed(hipStream_t st, double *p, size_t n) {
  hipLaunchKernelGGL(k_fdm_zero, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0, st, p, n);
}

void replicas_reduce_untimed(hipStream_t st, const double *priv, int64_t n, int nh, int nout, double *o0, double *o1,
                             double *o2, double *o3) {
  hipLaunchKernelGGL(k_fdm_reduce, dim3((unsigned)((n + 255) / 256), nout), dim3(256), 0, st, priv, n, nh, o0, o1, o2, o3);
}

int check_pairs(const char *who, const nrg_fdm_pair *pairs, int npairs, RhoUse use) {
  for (int p = 0; p < npairs; p++) {
    const nrg_fdm_pair &q = pairs[p];
    if (q.dim_i < 0 || q.dim_j < 0 || q.K_i < 0 || q.K_j < 0 || q.K_i > q.dim_i || q.K_j > q.dim_j ||
        (q.dim_i && q.dim_j && (!q.op1 || !q.op2 || !q.E_i || !q.E_j || q.ld_op < q.dim_i))) {
      set_error(std::string(who) + ": pair " + std::to_string(p) + " is inconsistent");
      return NRG_E_ARG;
    }
    if (q.dim_i > INT32_MAX / 2 || q.dim_j > INT32_MAX / 2 || q.ld_op > INT32_MAX || q.n_rho_i > INT32_MAX ||
        q.n_rho_j > INT32_MAX) {
      set_error(std::string(who) + ": dimension too large");
      return NRG_E_ARG;
    }
    bool ok = true;
    if (use == RhoUse::fdm || use == RhoUse::cfs) {
      // the products need rho on every kept state they touch
      const bool need_i = q.K_i > 0 && q.dim_j > q.K_j, need_j = q.K_j > 0 && q.dim_i > q.K_i;
      ok = !(need_i && (!q.rho_i || q.n_rho_i < q.K_i)) && !(need_j && (!q.rho_j || q.n_rho_j < q.K_j));
      if (use == RhoUse::cfs)   // prod_fit_left / prod_adj_fit_left: rho within the subspace
        ok = ok && !(need_i && q.n_rho_i > q.dim_i) && !(need_j && q.n_rho_j > q.dim_j);
    } else if (use == RhoUse::dmnrg) {
      ok = !(q.K_i && q.K_j && (!q.rho_i || q.n_rho_i < q.K_i || !q.rho_j || q.n_rho_j < q.K_j));
    }
    if (!ok) {
      set_error(std::string(who) + ": pair " + std::to_string(p) + " needs density matrices covering its kept states");
      return NRG_E_ARG;
    }
  }
  return NRG_OK;
}

int64_t fdm_products(const nrg_fdm_pair &q, double *buf, FdmPairDev &d, std::vector<GemmDesc> &gd) {
  d = FdmPairDev{};
  d.dim_i = q.dim_i; d.dim_j = q.dim_j; d.K_i = q.K_i; d.K_j = q.K_j; d.ld_op = q.ld_op;
  d.E_i = q.E_i; d.E_j = q.E_j; d.op1 = q.op1; d.op2 = q.op2; d.factor = q.factor;
  const int64_t n_or = q.K_i > 0 && q.dim_j > q.K_j ? (q.dim_j - q.K_j) * q.K_i : 0;
  const int64_t n_ro = q.K_j > 0 && q.dim_i > q.K_i ? q.K_j * (q.dim_i - q.K_i) : 0;
  if (!buf) return n_or + n_ro;
  if (n_or) {
    d.op2rho = buf;
    // (op2 rho_i)[K_j:, :K_i] = op2[K_j:, :c] rho_i[:c, :K_i], c = min(n_rho_i, dim_i) (prod_fit)
    const int64_t c = std::min(q.n_rho_i, q.dim_i);
    // row-major C (R x K_i) = X (R x c) * Y (c x K_i): column-major C^T = Y^T X^T
    GemmDesc g{};
    g.A = q.rho_i; g.lda = (int)q.n_rho_i; g.tA = 0;
    g.B = q.op2 + (size_t)q.K_j * q.ld_op; g.ldb = (int)q.ld_op; g.tB = 0;
    g.C = buf; g.ldc = (int)q.K_i;
    g.M = (int)q.K_i; g.N = (int)(q.dim_j - q.K_j); g.K = (int)c; g.alpha = 1.0; g.beta = 0.0;
    gd.push_back(g);
  }
  if (n_ro) {
    d.rhoop2 = buf + n_or;
    // (rho_j op2)[:K_j, K_i:] = rho_j[:K_j, :c] op2[:c, K_i:], c = min(n_rho_j, dim_j)
    const int64_t c = std::min(q.n_rho_j, q.dim_j);
    GemmDesc g{};
    g.A = q.op2 + q.K_i; g.lda = (int)q.ld_op; g.tA = 0;
    g.B = q.rho_j; g.ldb = (int)q.n_rho_j; g.tB = 0;
    g.C = buf + n_or; g.ldc = (int)(q.dim_i - q.K_i);
    g.M = (int)(q.dim_i - q.K_i); g.N = (int)q.K_j; g.K = (int)c; g.alpha = 1.0; g.beta = 0.0;
    gd.push_back(g);
  }
  return n_or + n_ro;
}

int64_t dmnrg_products(const nrg_fdm_pair &q, double *buf, DmnrgPairDev &d, std::vector<GemmDesc> &gd) {
  d = DmnrgPairDev{};
  d.K_i = q.K_i; d.K_j = q.K_j; d.ld_op = q.ld_op;
  d.E_i = q.E_i; d.E_j = q.E_j; d.op1 = q.op1; d.op2 = q.op2; d.factor = q.factor;
  const int64_t n = q.K_i > 0 && q.K_j > 0 ? 2 * q.K_i * q.K_j : 0;
  if (!buf || !n) return n;
  double *pa = buf, *pb = buf + q.K_i * q.K_j;
  d.pa = pa; d.pb = pb;
  // row-major PA (K_j x K_i) = op2[:K_j, :K_i] rho_i[:K_i, :K_i]^T: column-major PA^T = rho_i op2s^T
  GemmDesc g{};
  g.A = q.rho_i; g.lda = (int)q.n_rho_i; g.tA = 1;
  g.B = q.op2; g.ldb = (int)q.ld_op; g.tB = 0;
  g.C = pa; g.ldc = (int)q.K_i;
  g.M = (int)q.K_i; g.N = (int)q.K_j; g.K = (int)q.K_i; g.alpha = 1.0; g.beta = 0.0;
  gd.push_back(g);
  // row-major PB (K_j x K_i) = rho_j[:K_j, :K_j] op1[:K_j, :K_i]: column-major PB^T = op1s^T rho_j^T
  GemmDesc h{};
  h.A = q.op1; h.lda = (int)q.ld_op; h.tA = 0;
  h.B = q.rho_j; h.ldb = (int)q.n_rho_j; h.tB = 0;
  h.C = pb; h.ldc = (int)q.K_i;
  h.M = (int)q.K_i; h.N = (int)q.K_j; h.K = (int)q.K_j; h.alpha = 1.0; h.beta = 0.0;
  gd.push_back(h);
  return n;
}

__global__ __launch_bounds__(256) void k_fdm_bin(FdmArgs a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const FdmPairDev P = a.pairs[it.x];
  {
    const size_t c = (size_t)(blockIdx.x % FDM_COPIES) * 4 * a.nbins;
    a.gt_pos += c;
    a.gt_neg += c + a.nbins;
    a.ls_pos += c + 2 * a.nbins;
    a.ls_neg += c + 3 * a.nbins;
  }
  const int ti = (int)((P.dim_i + 63) / 64);
  const int i0 = (it.y % ti) * 64, j0 = (it.y / ti) * 64;
  __shared__ double Ei[64], Ej[64], Bi[64], Bj[64];
  if (threadIdx.x < 64) {
    const int i = i0 + threadIdx.x;
    const double e = i < P.dim_i ? P.E_i[i] : 0.0;
    Ei[threadIdx.x] = e;
    Bi[threadIdx.x] = exp(-e / a.T);
  } else if (threadIdx.x < 128) {
    const int j = j0 + threadIdx.x - 64;
    const double e = j < P.dim_j ? P.E_j[j] : 0.0;
    Ej[threadIdx.x - 64] = e;
    Bj[threadIdx.x - 64] = exp(-e / a.T);
  }
  __shared__ double wlds[2 * 2 * WINB], mm[4];
  __syncthreads();
  BinWindow<2> win;
  window_setup(win, wlds, Ei, (int)min<int64_t>(64, P.dim_i - i0), Ej, (int)min<int64_t>(64, P.dim_j - j0), a.log10emin,
               a.bpd, a.nbins, mm);
  const int il = threadIdx.x & 63;
  const int i = i0 + il;
  const bool Di = i >= P.K_i;
  for (int jl = threadIdx.x >> 6; jl < 64 && i < P.dim_i; jl += 4) {
    const int j = j0 + jl;
    if (j >= P.dim_j) break;
    const bool Dj = j >= P.K_j;
    if (!Di && !Dj) continue;
    const double o1 = P.op1[(size_t)j * P.ld_op + i];
    const double o2 = P.op2[(size_t)j * P.ld_op + i];
    double wgt, wls;
    if (Di) wgt = o1 * o2 * Bi[il] * a.wnf;   // evaluation order of the reference terms
    else wgt = o1 * P.op2rho[(size_t)(j - P.K_j) * P.K_i + i];
    if (Dj) wls = o1 * o2 * a.msign * Bj[jl] * a.wnf;
    else wls = o1 * P.rhoop2[(size_t)j * (P.dim_i - P.K_i) + (i - P.K_i)] * a.msign;
    wgt *= P.factor;
    wls *= P.factor;
    const double de = Ej[jl] - Ei[il];
    const bool pos = de >= 0.0;
    const double energy = pos ? de : -de;
    // bin position, shared by both halves
    int64_t idx;
    double rem;
    bool low;
    bin_pos(energy, a.log10emin, a.bpd, a.nbins, idx, rem, low);
    const int sg = pos ? 0 : 1;
    win_add(a.discard, win.at(sg, 0), win.lo[sg], pos ? a.gt_pos : a.gt_neg, energy, wgt, idx, rem, low);
    win_add(a.discard, win.at(sg, 1), win.lo[sg], pos ? a.ls_pos : a.ls_neg, energy, wls, idx, rem, low);
  }
  window_flush(win, a.nbins, [&](int sg, int h) {
    return h == 0 ? (sg == 0 ? a.gt_pos : a.gt_neg) : (sg == 0 ? a.ls_pos : a.ls_neg);
  });
}

int fdm_run(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double wnf,
            int sign, double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg, hipStream_t st) {
  if (npairs < 0 || (npairs > 0 && !pairs) || !grid || grid->nbins < 2 || !(T > 0) || (sign != 1 && sign != -1) ||
      !gt_pos || !gt_neg || !ls_pos || !ls_neg || !(grid->bins_per_decade > 0)) {
    set_error("nrg_fdm_bin: invalid arguments");
    return NRG_E_ARG;
  }
  if (grid->nbins > (1 << 20)) {   // the replicated histograms take 1 KiB per bin
    set_error("nrg_fdm_bin: more than 2^20 bins");
    return NRG_E_ARG;
  }
  if (int r = check_pairs("nrg_fdm_bin", pairs, npairs, RhoUse::fdm)) return r;
  // product buffers
  std::vector<GemmDesc> gd;
  std::vector<FdmPairDev> pd(npairs);
  int64_t tot = 0;
  for (int p = 0; p < npairs; p++) tot += fdm_products(pairs[p], nullptr, pd[p], gd);
  if (int r = ctx->ws2.ensure(sizeof(double) * (size_t)(tot + 1))) return r;
  double *prod = ctx->ws2.as<double>();
  std::vector<int> items;
  for (int64_t p = 0, off = 0; p < npairs; p++) {
    const nrg_fdm_pair &q = pairs[p];
    off += fdm_products(q, prod + off, pd[p], gd);
    // element tiles, skipping those that lie entirely in K_i x K_j
    const int ti = (int)((q.dim_i + 63) / 64), tj = (int)((q.dim_j + 63) / 64);
    for (int b = 0; b < tj; b++)
      for (int a = 0; a < ti; a++) {
        const bool allK = (int64_t)(a + 1) * 64 <= q.K_i && (int64_t)(b + 1) * 64 <= q.K_j;
        if (allK) continue;
        items.push_back((int)p);
        items.push_back(b * ti + a);
      }
  }
  if (int r = ctx->ws.ensure(sizeof(double) * (size_t)FDM_COPIES * 4 * (size_t)grid->nbins)) return r;
  return run_spectral_step(ctx, st, gd, pd, items, [&](const FdmPairDev *pdev, const int2 *idev) {
    FdmArgs fa;
    fa.pairs = pdev;
    fa.T = T; fa.wnf = wnf; fa.msign = -(double)sign;
    fa.log10emin = grid->log10emin; fa.bpd = grid->bins_per_decade; fa.discard = grid->discard_immediately;
    fa.nbins = grid->nbins;
    double *priv = ctx->ws.as<double>();   // FDM_COPIES x 4 histograms
    fa.gt_pos = priv; fa.gt_neg = priv; fa.ls_pos = priv; fa.ls_neg = priv;   // offset per replica in the kernel
    replicas_zero(ctx, st, priv, (size_t)FDM_COPIES * 4 * grid->nbins);
    NRG_TIMED(ctx, st, "k_fdm_bin", hipLaunchKernelGGL(k_fdm_bin, dim3((unsigned)(items.size() / 2)), dim3(256), 0, st, fa,
                                                       idev));
    replicas_reduce(ctx, st, priv, grid->nbins, 4, 4, gt_pos, gt_neg, ls_pos, ls_neg);
  });
}

// ------------------------------------------------------------------------------------------
// DM-NRG (Algo_DMNRG::calc, c++/algo_DMNRG.hpp:30-52)
// ------------------------------------------------------------------------------------------
struct DmnrgArgs {
  const DmnrgPairDev *pairs;
  double emin, emax, msign;
  double log10emin, bpd, discard;
  int64_t nbins;
  double *priv;   // FDM_COPIES replicas of (pos, neg)
};

__global__ __launch_bounds__(256) void k_dmnrg_bin(DmnrgArgs a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const DmnrgPairDev P = a.pairs[it.x];
  double *pos = a.priv + (size_t)(blockIdx.x % FDM_COPIES) * 2 * a.nbins, *neg = pos + a.nbins;
  const int ti = (int)((P.K_i + 63) / 64);
  const int m0 = (it.y % ti) * 64, j0 = (it.y / ti) * 64;
  __shared__ double Em[64], Ej[64];
  if (threadIdx.x < 64) {
    const int m = m0 + threadIdx.x;
    Em[threadIdx.x] = m < P.K_i ? P.E_i[m] : 0.0;
  } else if (threadIdx.x < 128) {
    const int j = j0 + threadIdx.x - 64;
    Ej[threadIdx.x - 64] = j < P.K_j ? P.E_j[j] : 0.0;
  }
  __shared__ double wlds[2 * WINB], mm[4];
  __syncthreads();
  BinWindow<1> win;
  window_setup(win, wlds, Em, (int)min<int64_t>(64, P.K_i - m0), Ej, (int)min<int64_t>(64, P.K_j - j0), a.log10emin,
               a.bpd, a.nbins, mm);
  const int ml = threadIdx.x & 63;
  const int m = m0 + ml;
  for (int jl = threadIdx.x >> 6; jl < 64 && m < P.K_i; jl += 4) {
    const int j = j0 + jl;
    if (j >= P.K_j) break;
    const double de = Ej[jl] - Em[ml];
    const double energy = fabs(de);
    if (energy < a.emin || energy > a.emax) continue;   // outside the step's window: no contribution
    const size_t e = (size_t)j * P.ld_op + m, q = (size_t)j * P.K_i + m;
    const double wA = P.pa[q] * P.op1[e];
    const double wB = P.pb[q] * P.op2[e];
    const double w = (wA + a.msign * wB) * P.factor;
    int64_t idx;
    double rem;
    bool low;
    bin_pos(energy, a.log10emin, a.bpd, a.nbins, idx, rem, low);
    const int sg = de >= 0.0 ? 0 : 1;
    win_add(a.discard, win.at(sg, 0), win.lo[sg], sg == 0 ? pos : neg, energy, w, idx, rem, low);
  }
  window_flush(win, a.nbins, [&](int sg, int) { return sg == 0 ? pos : neg; });
}

int dmnrg_run(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double emin, double emax,
              int sign, double *out_pos, double *out_neg, hipStream_t st) {
  if (npairs < 0 || (npairs > 0 && !pairs) || !grid || grid->nbins < 2 || (sign != 1 && sign != -1) || !out_pos ||
      !out_neg || !(grid->bins_per_decade > 0) || !(emin >= 0) || !(emax >= emin)) {
    set_error("nrg_dmnrg_bin: invalid arguments");
    return NRG_E_ARG;
  }
  if (grid->nbins > (1 << 20)) {
    set_error("nrg_dmnrg_bin: more than 2^20 bins");
    return NRG_E_ARG;
  }
  if (int r = check_pairs("nrg_dmnrg_bin", pairs, npairs, RhoUse::dmnrg)) return r;
  std::vector<GemmDesc> gd;
  std::vector<DmnrgPairDev> pd;
  DmnrgPairDev scratch;
  int64_t tot = 0;
  for (int p = 0; p < npairs; p++) tot += dmnrg_products(pairs[p], nullptr, scratch, gd);
  if (int r = ctx->ws2.ensure(sizeof(double) * (size_t)(tot + 1))) return r;
  if (int r = ctx->ws.ensure(sizeof(double) * (size_t)FDM_COPIES * 2 * (size_t)grid->nbins)) return r;
  double *prod = ctx->ws2.as<double>();
  std::vector<int> items;
  for (int64_t p = 0, off = 0; p < npairs; p++) {
    const nrg_fdm_pair &q = pairs[p];
    if (!(q.K_i > 0 && q.K_j > 0)) continue;
    DmnrgPairDev d;
    off += dmnrg_products(q, prod + off, d, gd);
    const int ti = (int)((q.K_i + 63) / 64), tj = (int)((q.K_j + 63) / 64);
    for (int t = 0; t < ti * tj; t++) {
      items.push_back((int)pd.size());
      items.push_back(t);
    }
    pd.push_back(d);
  }
  return run_spectral_step(ctx, st, gd, pd, items, [&](const DmnrgPairDev *pdev, const int2 *idev) {
    DmnrgArgs da;
    da.pairs = pdev;
    da.emin = emin; da.emax = emax; da.msign = -(double)sign;
    da.log10emin = grid->log10emin; da.bpd = grid->bins_per_decade; da.discard = grid->discard_immediately;
    da.nbins = grid->nbins;
    da.priv = ctx->ws.as<double>();
    replicas_zero(ctx, st, da.priv, (size_t)FDM_COPIES * 2 * grid->nbins);
    NRG_TIMED(ctx, st, "k_dmnrg_bin", hipLaunchKernelGGL(k_dmnrg_bin, dim3((unsigned)(items.size() / 2)), dim3(256), 0, st,
                                                         da, idev));
    replicas_reduce(ctx, st, da.priv, grid->nbins, 2, 2, out_pos, out_neg, out_pos, out_neg);
  });
}

// ------------------------------------------------------------------------------------------
// CFS (Algo_CFS::calc = Algo_CFSgt::calc + Algo_CFSls::calc, c++/algo_CFS.hpp:29-56, 79-106)
// ------------------------------------------------------------------------------------------
struct CfsArgs {
  const FdmPairDev *pairs;   // op2rho: (op2 rho_i)[K_j:, :K_i]; rhoop2: (rho_j^T op1)[:K_j, K_i:], ld dim_i - K_i
  double T, Z, msign;
  double log10emin, bpd, discard;
  int64_t nbins;
  int last;       // 0: below the last step, 1: the last step, 2: Algo_FT (kept x kept into the gt pair)
  double *priv;   // FDM_COPIES replicas of (gt_pos, gt_neg, ls_pos, ls_neg)
};

__global__ __launch_bounds__(256) void k_cfs_bin(CfsArgs a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const FdmPairDev P = a.pairs[it.x];
  double *gp = a.priv + (size_t)(blockIdx.x % FDM_COPIES) * 4 * a.nbins;
  double *gn = gp + a.nbins, *lp = gn + a.nbins, *ln = lp + a.nbins;
  const int ti = (int)((P.dim_i + 63) / 64);
  const int i0 = (it.y % ti) * 64, j0 = (it.y / ti) * 64;
  __shared__ double Ei[64], Ej[64], Bi[64], Bj[64];
  if (threadIdx.x < 64) {
    const int i = i0 + threadIdx.x;
    const double e = i < P.dim_i ? P.E_i[i] : 0.0;
    Ei[threadIdx.x] = e;
    Bi[threadIdx.x] = a.last ? exp(-e / a.T) : 0.0;
  } else if (threadIdx.x < 128) {
    const int j = j0 + threadIdx.x - 64;
    const double e = j < P.dim_j ? P.E_j[j] : 0.0;
    Ej[threadIdx.x - 64] = e;
    Bj[threadIdx.x - 64] = a.last ? exp(-e / a.T) : 0.0;
  }
  __shared__ double wlds[2 * 2 * WINB], mm[4];
  __syncthreads();
  BinWindow<2> win;
  window_setup(win, wlds, Ei, (int)min<int64_t>(64, P.dim_i - i0), Ej, (int)min<int64_t>(64, P.dim_j - j0), a.log10emin,
               a.bpd, a.nbins, mm);
  const int il = threadIdx.x & 63;
  const int i = i0 + il;
  const bool ki = i < P.K_i;
  for (int jl = threadIdx.x >> 6; jl < 64 && i < P.dim_i; jl += 4) {
    const int j = j0 + jl;
    if (j >= P.dim_j) break;
    const bool kj = j < P.K_j;
    const size_t e = (size_t)j * P.ld_op + i;
    double wg = 0.0, wl = 0.0;
    if (a.last == 2) {   // FT: kept x kept, ((-sign) exp(-E_j/T) + exp(-E_i/T)) / Z, one histogram pair
      if (!(ki && kj)) continue;
      wg = P.op1[e] * P.op2[e] * ((a.msign * Bj[jl] + Bi[il]) / a.Z);
    } else if (a.last) {   // Eq. (11), kept x kept
      if (!(ki && kj)) continue;
      const double oo = P.op1[e] * P.op2[e];
      wg = oo * Bi[il] / a.Z;
      wl = oo * Bj[jl] * a.msign / a.Z;
    } else if (ki && !kj) {   // Eq. (16), discarded j x kept i
      wl = P.op1[e] * P.op2rho[(size_t)(j - P.K_j) * P.K_i + i] * a.msign;
    } else if (!ki && kj) {   // Eq. (15), kept j x discarded i
      wg = P.rhoop2[(size_t)j * (P.dim_i - P.K_i) + (i - P.K_i)] * P.op2[e];
    } else {
      continue;
    }
    const double de = Ej[jl] - Ei[il];
    const double energy = fabs(de);
    int64_t idx;
    double rem;
    bool low;
    bin_pos(energy, a.log10emin, a.bpd, a.nbins, idx, rem, low);
    const int sg = de >= 0.0 ? 0 : 1;
    win_add(a.discard, win.at(sg, 0), win.lo[sg], sg == 0 ? gp : gn, energy, wg * P.factor, idx, rem, low);
    win_add(a.discard, win.at(sg, 1), win.lo[sg], sg == 0 ? lp : ln, energy, wl * P.factor, idx, rem, low);
  }
  window_flush(win, a.nbins, [&](int sg, int h) { return h == 0 ? (sg == 0 ? gp : gn) : (sg == 0 ? lp : ln); });
}

// last: 0 / 1 as step.last() for Algo_CFS; 2 for Algo_FT (gt_pos / gt_neg receive the weights, the
// lesser pair is not touched and may be null)
int cfs_run(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z, int last,
            int sign, double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg, hipStream_t st) {
  const char *who = last == 2 ? "nrg_ft_bin" : "nrg_cfs_bin";
  if (npairs < 0 || (npairs > 0 && !pairs) || !grid || grid->nbins < 2 || (sign != 1 && sign != -1) || !gt_pos ||
      !gt_neg || (last != 2 && (!ls_pos || !ls_neg)) || !(grid->bins_per_decade > 0) || (last && !(T > 0 && Z > 0))) {
    set_error(std::string(who) + ": invalid arguments");
    return NRG_E_ARG;
  }
  if (grid->nbins > (1 << 20)) {
    set_error(std::string(who) + ": more than 2^20 bins");
    return NRG_E_ARG;
  }
  if (int r = check_pairs(who, pairs, npairs, last ? RhoUse::none : RhoUse::cfs)) return r;
  std::vector<int64_t> o_or(npairs, -1), o_ro(npairs, -1);
  int64_t tot = 0;
  if (!last)
    for (int p = 0; p < npairs; p++) {
      const nrg_fdm_pair &q = pairs[p];
      if (q.K_i > 0 && q.dim_j > q.K_j) { o_or[p] = tot; tot += (q.dim_j - q.K_j) * q.K_i; }
      if (q.K_j > 0 && q.dim_i > q.K_i) { o_ro[p] = tot; tot += q.K_j * (q.dim_i - q.K_i); }
    }
  if (int r = ctx->ws2.ensure(sizeof(double) * (size_t)(tot + 1))) return r;
  if (int r = ctx->ws.ensure(sizeof(double) * (size_t)FDM_COPIES * 4 * (size_t)grid->nbins)) return r;
  double *prod = ctx->ws2.as<double>();
  std::vector<GemmDesc> gd;
  std::vector<FdmPairDev> pd;
  std::vector<int> items;
  for (int p = 0; p < npairs; p++) {
    const nrg_fdm_pair &q = pairs[p];
    FdmPairDev d{};
    d.dim_i = q.dim_i; d.dim_j = q.dim_j; d.K_i = q.K_i; d.K_j = q.K_j; d.ld_op = q.ld_op;
    d.E_i = q.E_i; d.E_j = q.E_j; d.op1 = q.op1; d.op2 = q.op2; d.factor = q.factor;
    d.op2rho = o_or[p] >= 0 ? prod + o_or[p] : nullptr;
    d.rhoop2 = o_ro[p] >= 0 ? prod + o_ro[p] : nullptr;
    if (o_or[p] >= 0) {
      // row-major (op2 rho_i)[K_j:, :K_i] = op2[K_j:, :n] rho_i[:n, :K_i]: column-major rho_i^T-view times op2^T-view
      GemmDesc g{};
      g.A = q.rho_i; g.lda = (int)q.n_rho_i; g.tA = 0;
      g.B = q.op2 + (size_t)q.K_j * q.ld_op; g.ldb = (int)q.ld_op; g.tB = 0;
      g.C = (double *)d.op2rho; g.ldc = (int)q.K_i;
      g.M = (int)q.K_i; g.N = (int)(q.dim_j - q.K_j); g.K = (int)q.n_rho_i; g.alpha = 1.0; g.beta = 0.0;
      gd.push_back(g);
    }
    if (o_ro[p] >= 0) {
      // row-major X = (rho_j^T op1)[:K_j, K_i:], X(r, c) = sum_m rho_j(m, r) op1(m, K_i + c), m < n_rho_j:
      // column-major X^T = op1[:n, K_i:]-view (M x n) times rho_j (n x K_j, stored row-major: transposed)
      GemmDesc g{};
      g.A = q.op1 + q.K_i; g.lda = (int)q.ld_op; g.tA = 0;
      g.B = q.rho_j; g.ldb = (int)q.n_rho_j; g.tB = 1;
      g.C = (double *)d.rhoop2; g.ldc = (int)(q.dim_i - q.K_i);
      g.M = (int)(q.dim_i - q.K_i); g.N = (int)q.K_j; g.K = (int)q.n_rho_j; g.alpha = 1.0; g.beta = 0.0;
      gd.push_back(g);
    }
    // element tiles of the full dim_i x dim_j grid that meet the elements this step evaluates
    const int64_t ti = (q.dim_i + 63) / 64, tj = (q.dim_j + 63) / 64;
    bool any = false;
    for (int64_t b = 0; b < tj; b++)
      for (int64_t c = 0; c < ti; c++) {
        const int64_t a0 = c * 64, a1 = std::min<int64_t>(a0 + 64, q.dim_i);   // i range
        const int64_t b0 = b * 64, b1 = std::min<int64_t>(b0 + 64, q.dim_j);   // j range
        const bool kk = a0 < q.K_i && b0 < q.K_j;
        const bool ls = a0 < q.K_i && b1 > q.K_j, gt = a1 > q.K_i && b0 < q.K_j;
        if (last ? !kk : !(ls || gt)) continue;
        items.push_back((int)pd.size());
        items.push_back((int)(b * ti + c));
        any = true;
      }
    if (any) pd.push_back(d);
  }
  return run_spectral_step(ctx, st, gd, pd, items, [&](const FdmPairDev *pdev, const int2 *idev) {
    CfsArgs ca;
    ca.pairs = pdev;
    ca.T = T; ca.Z = Z; ca.msign = -(double)sign;
    ca.log10emin = grid->log10emin; ca.bpd = grid->bins_per_decade; ca.discard = grid->discard_immediately;
    ca.nbins = grid->nbins;
    ca.last = last == 2 ? 2 : last ? 1 : 0;
    ca.priv = ctx->ws.as<double>();
    replicas_zero(ctx, st, ca.priv, (size_t)FDM_COPIES * 4 * grid->nbins);
    NRG_TIMED(ctx, st, "k_cfs_bin", hipLaunchKernelGGL(k_cfs_bin, dim3((unsigned)(items.size() / 2)), dim3(256), 0, st, ca,
                                                       idev));
    // replica layout (gt_pos, gt_neg, ls_pos, ls_neg) per copy; FT reduces the first two only
    replicas_reduce(ctx, st, ca.priv, grid->nbins, 4, last == 2 ? 2 : 4, gt_pos, gt_neg, ls_pos, ls_neg);
  });
}

}  // namespace nrg
