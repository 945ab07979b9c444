// syev.hip - batched real symmetric eigensolver for gfx950.
//
// Replaces the per-block LAPACK dsyev loop of NRG Ljubljana
// (diagonalisations_OpenMP, c++/core.hpp:90-110 -> diagonalise_dsyev, c++/diag.hpp:61-82).
// All blocks of one NRG step go through the same short sequence of batched launches:
//
//   1. k_symm       mirror the stored triangle, giving a full symmetric column-major work matrix
//   2. k_panel      blocked Householder reduction to tridiagonal form (LAPACK dsytrd/dlatrd
//                   algorithm, lower variant), one workgroup per block and panel of NB columns;
//      k_syr2k     trailing rank-2k update A22 -= V W^T + W V^T over 64x64 tiles of every block
//   3. k_dc_prep    divide and conquer on the tridiagonal matrix (Cuppen's method with the
//      k_dc_merge   Gu-Eisenstat eigenvector formula, as LAPACK dstedc/dlaed0-3): leaves of order
//      k_gemm_nn    1-2 at a common tree depth, then one merge launch per depth for all blocks
//   4. k_bt_tfactor back-transformation V = Q_H Z with compact-WY panels (dlarft + dlarfb)
//      k_bt_apply
//   5. k_finish     eigenvalues (rescaled) and eigenvectors back into the caller's arenas
//
// Storage: a row-major block with its upper triangle filled is, byte for byte, a column-major
// matrix with its lower triangle filled, so every kernel works column-major: A(r,c) = A[c*n + r].
// A column-major eigenvector matrix with eigenvector i in column i is, byte for byte, the
// row-major matrix with eigenvector i in row i that copy_vec() builds (c++/diag.hpp:41-48).
#include "common.hpp"
#include <algorithm>
#include <cstring>
#include <cmath>

namespace nrg {

constexpr int NB = 32;        // panel width of the Householder reduction
constexpr int NT_PANEL = 512; // threads of the panel kernel
constexpr int RMAX = (NMAX + NT_PANEL - 1) / NT_PANEL;
constexpr int NT_DC = 256;    // threads of the merge kernel
constexpr int SMALLM = 64;    // merges up to this order do their own block-diagonal product
constexpr int CW = 32;        // column-strip width of the back-transformation

struct BlkDesc {
  int64_t n, offA, offZ, offV, offWp, offT, offH, offw;
  int L, P;
};

template <typename T>
struct SyevArgs {
  const BlkDesc *bd;
  T *A;            // work matrices (column-major, full symmetric, then reflectors)
  T *Wp;           // dlatrd W panels, n x NB per block
  T *tau;          // Householder scalars, n per block
  T *Tf;           // compact-WY T factors, P x NB x NB per block
  T *Zt;           // eigenvector matrix used by the back-transformation (== Z0 for FP64)
  double *Z0, *Z1, *S;
  double *d, *e, *lam, *scale;
  double *dscr;    // 8 arrays of Ntot doubles
  int *iscr;       // 8 arrays of Ntot ints
  int64_t Ntot;
  int *flags;
  double *H, *w;
};

struct GemmNN {  // C = A*B, column-major, all with leading dimension ld
  const double *A, *B;
  double *C;
  int M, N, K, ld;
};

// ------------------------------------------------------------------------------------------
// 1. symmetrisation: lower triangle of the (double) input -> full (T) work matrix
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_symm(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n;
  const int t = it.y;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
  while (ti * (ti + 1) / 2 > t) ti--;
  const int tj = t - ti * (ti + 1) / 2;
  const double *Hs = a.H + bd.offH;
  T *Ad = a.A + bd.offA;
  __shared__ T tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int cl = ty + 8 * q;
    const int r = ti * 32 + tx, c = tj * 32 + cl;
    T v = 0;
    if (r < n && c < n && r >= c) {
      v = (T)Hs[(size_t)c * n + r];
      Ad[(size_t)c * n + r] = v;
    }
    tile[cl][tx] = v;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int rl = ty + 8 * q;           // local row in the lower tile = column of the mirror
    const int row = tj * 32 + tx, col = ti * 32 + rl;  // mirror element (row, col) = lower (col, row)
    if (row < n && col < n && col > row) Ad[(size_t)col * n + row] = tile[tx][rl];
  }
}

// ------------------------------------------------------------------------------------------
// 2a. panel reduction (dlatrd, lower): one workgroup per block
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT_PANEL) void k_panel(SyevArgs<T> a, const int *blist, int p) {
  const BlkDesc bd = a.bd[blist[blockIdx.x]];
  const int n = (int)bd.n;
  const int i0 = p * NB;
  const int m = n - i0;
  T *A = a.A + bd.offA + (size_t)i0 * n + i0;
  T *W = a.Wp + bd.offWp;
  double *d = a.d + bd.offV + i0;
  double *e = a.e + bd.offV + i0;
  T *tau = a.tau + bd.offV + i0;
  const bool final = (m - 1 <= NB);
  const int w = final ? m - 1 : NB;
  const int tid = threadIdx.x;
  constexpr int NW = NT_PANEL / WAVE;
  const int lane = tid & (WAVE - 1), wid = tid / WAVE;

  __shared__ T vs[NMAX];
  __shared__ T vrow[NB], wrow[NB], cw[NB], cv[NB];
  __shared__ T red[NW];

  for (int j = 0; j < w; j++) {
    if (j > 0) {
      if (tid < j) {
        vrow[tid] = A[(size_t)tid * n + j];
        wrow[tid] = W[(size_t)tid * n + j];
      }
      __syncthreads();
      for (int r = j + tid; r < m; r += NT_PANEL) {
        T acc = 0;
        for (int k = 0; k < j; k++) acc += A[(size_t)k * n + r] * wrow[k] + W[(size_t)k * n + r] * vrow[k];
        A[(size_t)j * n + r] -= acc;
      }
      __syncthreads();
    }
    if (j < m - 1) {
      T *colj = A + (size_t)j * n;
      const T alpha = colj[j + 1];
      T part = 0;
      for (int r = j + 2 + tid; r < m; r += NT_PANEL) part += colj[r] * colj[r];
      const T xn2 = block_sum<NT_PANEL>(part, red);
      T beta, tv, scal = 1;
      if (xn2 == (T)0) {
        tv = 0;
        beta = alpha;
      } else {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tv = (beta - alpha) / beta;
        scal = (T)1 / (alpha - beta);
      }
      for (int r = j + 1 + tid; r < m; r += NT_PANEL) {
        if (r == j + 1) {
          vs[r] = 1;
        } else {
          const T x = colj[r] * scal;
          colj[r] = x;
          vs[r] = x;
        }
      }
      __syncthreads();
      if (tid == 0) {
        e[j] = (double)beta;
        tau[j] = tv;
        colj[j + 1] = 1;
      }
      T *Wj = W + (size_t)j * n;
      if (tv != (T)0) {
        const int r0 = j + 1 + tid;
        const int nq = (m - j - 1 + NT_PANEL - 1) / NT_PANEL;
        T acc[RMAX];
#pragma unroll
        for (int q = 0; q < RMAX; q++) acc[q] = 0;
        // y = A22 * v on the (not yet updated) trailing matrix
        for (int c = j + 1; c < m; c++) {
          const T vc = vs[c];
          const T *col = A + (size_t)c * n;
#pragma unroll
          for (int q = 0; q < RMAX; q++) {
            const int r = r0 + q * NT_PANEL;
            if (q < nq && r < m) acc[q] += col[r] * vc;
          }
        }
        if (j > 0) {
          for (int k = wid; k < j; k += NW) {
            T sw = 0, sv = 0;
            const T *Wk = W + (size_t)k * n, *Ak = A + (size_t)k * n;
            for (int r = j + 1 + lane; r < m; r += WAVE) {
              sw += Wk[r] * vs[r];
              sv += Ak[r] * vs[r];
            }
            sw = wave_sum(sw);
            sv = wave_sum(sv);
            if (lane == 0) {
              cw[k] = sw;
              cv[k] = sv;
            }
          }
          __syncthreads();
#pragma unroll
          for (int q = 0; q < RMAX; q++) {
            const int r = r0 + q * NT_PANEL;
            if (q < nq && r < m) {
              T s = 0;
              for (int k = 0; k < j; k++) s += A[(size_t)k * n + r] * cw[k] + W[(size_t)k * n + r] * cv[k];
              acc[q] -= s;
            }
          }
        }
        T dp = 0;
#pragma unroll
        for (int q = 0; q < RMAX; q++) {
          const int r = r0 + q * NT_PANEL;
          if (q < nq && r < m) {
            acc[q] *= tv;
            dp += acc[q] * vs[r];
          }
        }
        const T a2 = (T)(-0.5) * tv * block_sum<NT_PANEL>(dp, red);
#pragma unroll
        for (int q = 0; q < RMAX; q++) {
          const int r = r0 + q * NT_PANEL;
          if (q < nq && r < m) Wj[r] = acc[q] + a2 * vs[r];
        }
      } else {
        for (int r = j + 1 + tid; r < m; r += NT_PANEL) Wj[r] = 0;
      }
      __syncthreads();
    }
    if (tid == 0) d[j] = (double)A[(size_t)j * n + j];
  }
  if (final && tid == 0) {
    T dd = A[(size_t)(m - 1) * n + (m - 1)];
    for (int k = 0; k < w; k++) dd -= (T)2 * A[(size_t)k * n + (m - 1)] * W[(size_t)k * n + (m - 1)];
    A[(size_t)(m - 1) * n + (m - 1)] = dd;
    d[m - 1] = (double)dd;
  }
}

// ------------------------------------------------------------------------------------------
// 2b. trailing update A22 -= V W^T + W V^T (full square), 64x64 tiles
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_syr2k(SyevArgs<T> a, const int2 *items, int p) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n;
  const int i0 = p * NB;
  const int mt = n - i0 - NB;                 // trailing order
  const int td = (mt + 63) / 64;
  const int tr = it.y % td, tc = it.y / td;
  const T *Vp = a.A + bd.offA + (size_t)i0 * n + (i0 + NB);  // V(r,k) = Vp[k*n + r]
  const T *Wq = a.Wp + bd.offWp + NB;                        // W(r,k) = Wq[k*n + r]
  T *C = a.A + bd.offA + (size_t)(i0 + NB) * n + (i0 + NB);
  __shared__ T Vr[NB][64], Wr[NB][64], Vc[NB][64], Wc[NB][64];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < NB * 64; idx += 256) {
    const int k = idx / 64, rr = idx % 64;
    const int row = tr * 64 + rr, col = tc * 64 + rr;
    Vr[k][rr] = row < mt ? Vp[(size_t)k * n + row] : (T)0;
    Wr[k][rr] = row < mt ? Wq[(size_t)k * n + row] : (T)0;
    Vc[k][rr] = col < mt ? Vp[(size_t)k * n + col] : (T)0;
    Wc[k][rr] = col < mt ? Wq[(size_t)k * n + col] : (T)0;
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0;
#pragma unroll 4
  for (int k = 0; k < NB; k++) {
    T av[4], aw[4], bv[4], bw[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      av[i] = Vr[k][tx + 16 * i];
      aw[i] = Wr[k][tx + 16 * i];
      bv[i] = Vc[k][ty + 16 * i];
      bw[i] = Wc[k][ty + 16 * i];
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bw[j] + aw[i] * bv[j];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int col = tc * 64 + ty + 16 * j;
    if (col >= mt) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int row = tr * 64 + tx + 16 * i;
      if (row < mt) C[(size_t)col * n + row] -= acc[i][j];
    }
  }
}

// ------------------------------------------------------------------------------------------
// 3a. divide and conquer: scaling, Cuppen tears and the 1x1 / 2x2 leaves
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool is_leaf_start(int i, int n, int L) {
  if (i <= 0 || i >= n) return false;
  const int q = (int)(((int64_t)i << L) + n - 1) / n;
  return q < (1 << L) && (int)(((int64_t)q * n) >> L) == i;
}

__global__ __launch_bounds__(256) void k_dc_prep(SyevArgs<double> a, const int *blist) {
  const int b = blist[blockIdx.x];
  const BlkDesc bd = a.bd[b];
  const int n = (int)bd.n, L = bd.L;
  double *d = a.d + bd.offV, *e = a.e + bd.offV, *lam = a.lam + bd.offV;
  double *dp = a.dscr + bd.offV;  // scratch array 0
  __shared__ double red[4];
  double mx = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    mx = fmax(mx, fabs(d[i]));
    if (i < n - 1) mx = fmax(mx, fabs(e[i]));
    if (!isfinite(d[i]) || (i < n - 1 && !isfinite(e[i]))) mx = INFINITY;
  }
  mx = block_max<256>(mx, red);
  const bool bad = !isfinite(mx);
  if (bad) {
    // non-finite input: report the block and carry on with a zero tridiagonal matrix, so that no
    // NaN reaches the data-dependent orderings of the merges
    if (threadIdx.x == 0) a.flags[b] = 1;
    mx = 1;
  }
  const double s = mx > 0 ? mx : 1.0;
  if (threadIdx.x == 0) a.scale[b] = s;
  for (int i = threadIdx.x; i < n; i += 256) {
    double di = bad ? 0.0 : d[i] / s;
    if (!bad && is_leaf_start(i, n, L)) di -= fabs(e[i - 1] / s);
    if (!bad && is_leaf_start(i + 1, n, L)) di -= fabs(e[i] / s);
    dp[i] = di;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n - 1; i += 256) e[i] = bad ? 0.0 : e[i] / s;
  __syncthreads();
  double *Z = (L & 1) ? a.Z1 : a.Z0;
  Z += bd.offZ;
  for (int q = threadIdx.x; q < (1 << L); q += 256) {
    const int lo = (int)(((int64_t)q * n) >> L), hi = (int)(((int64_t)(q + 1) * n) >> L);
    if (hi - lo == 1) {
      lam[lo] = dp[lo];
      Z[(size_t)lo * n + lo] = 1.0;
    } else {
      const double x = dp[lo], y = dp[lo + 1], bb = e[lo];
      double l1, l2, v1a, v1b, v2a, v2b;
      if (bb == 0.0) {
        l1 = x; l2 = y; v1a = 1; v1b = 0; v2a = 0; v2b = 1;
      } else {
        const double th = (y - x) / (2.0 * bb);
        const double t = copysign(1.0, th) / (fabs(th) + sqrt(1.0 + th * th));
        const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
        l1 = x - t * bb; l2 = y + t * bb;
        v1a = cs; v1b = -sn; v2a = sn; v2b = cs;
      }
      if (l2 < l1) {
        double tmp = l1; l1 = l2; l2 = tmp;
        tmp = v1a; v1a = v2a; v2a = tmp;
        tmp = v1b; v1b = v2b; v2b = tmp;
      }
      lam[lo] = l1; lam[lo + 1] = l2;
      Z[(size_t)lo * n + lo] = v1a; Z[(size_t)lo * n + lo + 1] = v1b;
      Z[(size_t)(lo + 1) * n + lo] = v2a; Z[(size_t)(lo + 1) * n + lo + 1] = v2b;
    }
  }
}

// ------------------------------------------------------------------------------------------
// 3b. divide and conquer: one merge (secular equation + Loewner eigenvectors), one workgroup
// ------------------------------------------------------------------------------------------
// Root j of 1 + rho sum_i z_i^2 / (d_i - lambda) = 0, d ascending, z nonzero, rho > 0.
// The origin is the nearer pole o; lambda = d_o + tau. Newton iteration on the function with
// both adjacent poles multiplied out, safeguarded by bisection on the bracket.
__device__ void secular_root(int j, int k, const double *d, const double *z, double rho, double zz, int &org,
                             double &tau_out) {
  const double EPS = 2.220446049250313e-16;
  int o, o2;
  double lo, hi;
  if (j < k - 1) {
    const double g = d[j + 1] - d[j], mid = 0.5 * g;
    double f = 1.0;
    for (int i = 0; i < k; i++) f += rho * z[i] * z[i] / ((d[i] - d[j]) - mid);
    if (f >= 0) { o = j; o2 = j + 1; lo = 0; hi = mid; }
    else { o = j + 1; o2 = j; lo = -mid; hi = 0; }
  } else {
    o = k - 1; o2 = -1; lo = 0; hi = rho * zz;
  }
  const double dorg = d[o];
  const double d2 = o2 >= 0 ? d[o2] - dorg : 0.0;
  const double za = rho * z[o] * z[o];
  const double zb = o2 >= 0 ? rho * z[o2] * z[o2] : 0.0;
  double t = 0.5 * (lo + hi);
  for (int it = 0; it < 100; it++) {
    double R = 0, Rp = 0, Ra = 0;
    for (int i = 0; i < k; i++) {
      if (i == o || i == o2) continue;
      const double q = 1.0 / ((d[i] - dorg) - t);
      const double tt = rho * z[i] * z[i] * q;
      R += tt; Rp += tt * q; Ra += fabs(tt);
    }
    double f, err, gv, gp;
    const double pa = -za / t;
    if (o2 >= 0) {
      const double pb = zb / (d2 - t);
      f = 1.0 + R + pa + pb;
      err = 8.0 * (Ra + fabs(pa) + fabs(pb)) + 2.0;
      gv = (-t) * (d2 - t) * (1 + R) + za * (d2 - t) - zb * t;
      gp = (2 * t - d2) * (1 + R) + (-t) * (d2 - t) * Rp - za - zb;
    } else {
      f = 1.0 + R + pa;
      err = 8.0 * (Ra + fabs(pa)) + 2.0;
      gv = t * (1 + R) - za;
      gp = (1 + R) + t * Rp;
    }
    if (fabs(f) <= EPS * err) break;
    if (f < 0) lo = t; else hi = t;
    double tn = t - gv / gp;
    // the pole (tau = 0) is the only open end of the bracket
    if (!(tn >= lo && tn <= hi && tn != 0.0)) tn = 0.5 * (lo + hi);
    if (fabs(tn - t) <= 2 * EPS * fabs(tn)) { t = tn; break; }
    t = tn;
  }
  org = o;
  tau_out = t;
}

// Orderings of eigenvalues use a strict total order: NaN sorts as +inf, ties break by index, so
// merge positions and ranks are permutations whatever the values.
__device__ __forceinline__ double okey(double x) { return isnan(x) ? INFINITY : x; }
__device__ __forceinline__ int lower_bound_d(const double *v, int len, double x) {
  x = okey(x);
  int lo = 0, hi = len;
  while (lo < hi) { const int mid = (lo + hi) >> 1; if (okey(v[mid]) < x) lo = mid + 1; else hi = mid; }
  return lo;
}
__device__ __forceinline__ int upper_bound_d(const double *v, int len, double x) {
  x = okey(x);
  int lo = 0, hi = len;
  while (lo < hi) { const int mid = (lo + hi) >> 1; if (okey(v[mid]) <= x) lo = mid + 1; else hi = mid; }
  return lo;
}
__device__ __forceinline__ int clampi(int v, int hi) { return v < 0 ? 0 : (v >= hi ? hi - 1 : v); }

__global__ __launch_bounds__(NT_DC) void k_dc_merge(SyevArgs<double> a, const int2 *items, int depth, int mcap) {
  extern __shared__ double dyn[];
  double *dnd = dyn, *znd = dyn + mcap, *rt = dyn + 2 * mcap, *vals = dyn + 3 * mcap;
  int *ro = (int *)(dyn + 4 * mcap);
  __shared__ double red[NT_DC / WAVE];
  __shared__ int sh_k, sh_nrot;
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n, q = it.y, tid = threadIdx.x;
  const int a0 = (int)(((int64_t)q * n) >> depth);
  const int mid = (int)(((int64_t)(2 * q + 1) * n) >> (depth + 1));
  const int a1 = (int)(((int64_t)(q + 1) * n) >> depth);
  const int m1 = mid - a0, m2 = a1 - mid, m = m1 + m2;
  const size_t ld = n;
  const double *Qin = ((depth + 1) & 1 ? a.Z1 : a.Z0) + bd.offZ;
  double *Qout = (depth & 1 ? a.Z1 : a.Z0) + bd.offZ;
  double *S = a.S + bd.offZ;
  double *lam = a.lam + bd.offV;
  const double *e = a.e + bd.offV;
  const int64_t N = a.Ntot;
  double *ds = a.dscr + 0 * N + bd.offV + a0, *zs = a.dscr + 1 * N + bd.offV + a0;
  double *dcat = a.dscr + 2 * N + bd.offV + a0, *zcat = a.dscr + 3 * N + bd.offV + a0;
  double *invn = a.dscr + 4 * N + bd.offV + a0, *rc = a.dscr + 5 * N + bd.offV + a0;
  double *rs = a.dscr + 6 * N + bd.offV + a0;
  int *perm = a.iscr + 0 * N + bd.offV + a0, *pinv = a.iscr + 1 * N + bd.offV + a0;
  int *cmp = a.iscr + 2 * N + bd.offV + a0, *cpos = a.iscr + 3 * N + bd.offV + a0;
  int *colsrc = a.iscr + 4 * N + bd.offV + a0, *rpj = a.iscr + 5 * N + bd.offV + a0;
  int *rnj = a.iscr + 6 * N + bd.offV + a0;

  const double beta = e[mid - 1];
  const double sgn = beta >= 0 ? 1.0 : -1.0;
  const double rho = 2.0 * fabs(beta);
  const double r2 = 0.70710678118654752440;
  // 1. the two sorted spectra and z = (last row of Q1, +-first row of Q2) / sqrt(2)
  for (int i = tid; i < m; i += NT_DC) {
    dcat[i] = lam[a0 + i];
    zcat[i] = (i < m1 ? Qin[(size_t)(a0 + i) * ld + (mid - 1)] : sgn * Qin[(size_t)(a0 + i) * ld + mid]) * r2;
  }
  __syncthreads();
  // 2. merge the two sorted lists
  double mx = 0;
  for (int i = tid; i < m; i += NT_DC) {
    int pos;
    if (i < m1) pos = i + lower_bound_d(dcat + m1, m2, dcat[i]);
    else pos = (i - m1) + upper_bound_d(dcat, m1, dcat[i]);
    pos = clampi(pos, m);
    perm[pos] = i;
    pinv[i] = pos;
    ds[pos] = dcat[i];
    zs[pos] = zcat[i];
    mx = fmax(mx, fmax(fabs(dcat[i]), fabs(zcat[i])));
  }
  const double dzmax = block_max<NT_DC>(mx, red);  // includes a barrier
  // 3. deflation (dlaed2), sequential
  if (tid == 0) {
    const double EPS = 2.220446049250313e-16;
    const double tol = 8.0 * EPS * dzmax;
    double zmax = 0;
    for (int p = 0; p < m; p++) zmax = fmax(zmax, fabs(zs[p]));
    int k = 0, nd = 0, nrot = 0;
    if (rho * zmax <= tol) {
      for (int p = 0; p < m; p++) { cmp[p] = -1; cpos[m - 1 - nd] = p; nd++; }
    } else {
      int pj = -1;
      for (int nj = 0; nj < m; nj++) {
        if (rho * fabs(zs[nj]) <= tol) { cmp[nj] = -1; cpos[m - 1 - nd] = nj; nd++; continue; }
        if (pj < 0) { pj = nj; continue; }
        double s = zs[pj], c = zs[nj];
        const double tau = hypot(c, s), t = ds[nj] - ds[pj];
        c /= tau; s = -s / tau;
        if (fabs(t * c * s) <= tol) {
          zs[nj] = tau; zs[pj] = 0.0;
          rpj[nrot] = pj; rnj[nrot] = nj; rc[nrot] = c; rs[nrot] = s; nrot++;
          const double t2 = ds[pj] * c * c + ds[nj] * s * s;
          ds[nj] = ds[pj] * s * s + ds[nj] * c * c;
          ds[pj] = t2;
          cmp[pj] = -1; cpos[m - 1 - nd] = pj; nd++;
          pj = nj;
        } else {
          cmp[pj] = k; cpos[k] = pj; dnd[k] = ds[pj]; znd[k] = zs[pj]; k++;
          pj = nj;
        }
      }
      if (pj >= 0) { cmp[pj] = k; cpos[k] = pj; dnd[k] = ds[pj]; znd[k] = zs[pj]; k++; }
    }
    sh_k = k;
    sh_nrot = nrot;
  }
  __syncthreads();
  const int k = sh_k, nrot = sh_nrot;
  // deflated positions were stored from the back: cpos[k..m) holds them in reverse scan order
  // 4. secular roots
  double zz = 0;
  for (int i = 0; i < k; i++) zz += znd[i] * znd[i];
  for (int j = tid; j < k; j += NT_DC) {
    int o; double t;
    secular_root(j, k, dnd, znd, rho, zz, o, t);
    ro[j] = o;
    rt[j] = t;
  }
  __syncthreads();
  // 5. Loewner: zhat_i^2 = prod_j (lambda_j - d_i) / prod_{j != i} (d_j - d_i)
  for (int i = tid; i < k; i += NT_DC) {
    const double di = dnd[i];
    double W = (di - dnd[ro[i]]) - rt[i];
    for (int j = 0; j < k; j++) {
      if (j == i) continue;
      W *= ((di - dnd[ro[j]]) - rt[j]) / (di - dnd[j]);
    }
    zcat[i] = copysign(sqrt(fabs(W)), znd[i]);
  }
  __syncthreads();
  for (int i = tid; i < k; i += NT_DC) znd[i] = zcat[i];
  __syncthreads();
  // 6. norms of the eigenvectors u_ij = zhat_i / (d_i - lambda_j)
  for (int j = tid; j < k; j += NT_DC) {
    const double dj = dnd[ro[j]], tj = rt[j];
    double s = 0;
    for (int i = 0; i < k; i++) {
      const double u = znd[i] / ((dnd[i] - dj) - tj);
      s += u * u;
    }
    invn[j] = 1.0 / sqrt(s);
  }
  // 7. new eigenvalues and their order
  for (int u = tid; u < m; u += NT_DC) vals[u] = u < k ? dnd[clampi(ro[u], k)] + rt[u] : ds[clampi(cpos[u], m)];
  __syncthreads();
  for (int u = tid; u < m; u += NT_DC) {
    const double x = vals[u], kx = okey(x);
    int r = 0;
    for (int v = 0; v < m; v++) {
      const double kv = okey(vals[v]);
      r += (kv < kx) || (kv == kx && v < u);
    }
    colsrc[r] = u;
    lam[a0 + r] = x;
  }
  __syncthreads();
  // 8. S in the concatenated row order: column c is the eigenvector of the c-th eigenvalue
  for (int idx = tid; idx < m * m; idx += NT_DC) {
    const int c = idx / m, r = idx - c * m;
    const int u = clampi(colsrc[c], m);
    const int p = clampi(pinv[r], m);
    double val;
    if (u < k) {
      const int ci = cmp[p];
      val = ci < 0 ? 0.0 : znd[clampi(ci, k)] / ((dnd[clampi(ci, k)] - dnd[clampi(ro[u], k)]) - rt[u]) * invn[u];
    } else {
      val = (p == clampi(cpos[u], m)) ? 1.0 : 0.0;
    }
    S[(size_t)(a0 + c) * ld + (a0 + r)] = val;
  }
  __syncthreads();
  // 9. undo the deflation rotations (last generated first) on the rows of S
  if (nrot > 0) {
    for (int c = tid; c < m; c += NT_DC) {
      double *Sc = S + (size_t)(a0 + c) * ld + a0;
      for (int t = nrot - 1; t >= 0; t--) {
        const int r1 = clampi(perm[clampi(rpj[t], m)], m), r2i = clampi(perm[clampi(rnj[t], m)], m);
        const double cc = rc[t], ss = rs[t];
        const double x = Sc[r1], y = Sc[r2i];
        Sc[r1] = cc * x - ss * y;
        Sc[r2i] = ss * x + cc * y;
      }
    }
    __syncthreads();
  }
  // 10. small merges form Qout = diag(Q1, Q2) * S here
  if (m <= SMALLM) {
    for (int idx = tid; idx < m * m; idx += NT_DC) {
      const int c = idx / m, r = idx - c * m;
      const double *Sc = S + (size_t)(a0 + c) * ld + a0;
      double acc = 0;
      if (r < m1) {
        for (int t = 0; t < m1; t++) acc += Qin[(size_t)(a0 + t) * ld + (a0 + r)] * Sc[t];
      } else {
        for (int t = m1; t < m; t++) acc += Qin[(size_t)(a0 + t) * ld + (a0 + r)] * Sc[t];
      }
      Qout[(size_t)(a0 + c) * ld + (a0 + r)] = acc;
    }
  }
}

// ------------------------------------------------------------------------------------------
// 3c. C = A * B (column-major FP64), 64x64 tiles, for the large merges
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm_nn(const GemmNN *probs, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const GemmNN P = probs[it.x];
  const int tm = (P.M + 63) / 64;
  const int ti = it.y % tm, tj = it.y / tm;
  __shared__ double As[16][64], Bs[16][65];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0;
  for (int k0 = 0; k0 < P.K; k0 += 16) {
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int idx = tid + 256 * t;
      const int r = idx & 63, kk = idx >> 6;
      const int gr = ti * 64 + r, gk = k0 + kk;
      As[kk][r] = (gr < P.M && gk < P.K) ? P.A[(size_t)gk * P.ld + gr] : 0.0;
      const int kb = idx & 15, c = idx >> 4;
      const int gc = tj * 64 + c, gkb = k0 + kb;
      Bs[kb][c] = (gc < P.N && gkb < P.K) ? P.B[(size_t)gc * P.ld + gkb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      double av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) { av[i] = As[kk][tx + 16 * i]; bv[i] = Bs[kk][ty + 16 * i]; }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int gc = tj * 64 + ty + 16 * j;
    if (gc >= P.N) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int gr = ti * 64 + tx + 16 * i;
      if (gr < P.M) P.C[(size_t)gc * P.ld + gr] = acc[i][j];
    }
  }
}

// ------------------------------------------------------------------------------------------
// 4a. compact-WY T factor of panel p (dlarft, forward, columnwise)
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_bt_tfactor(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n, p = it.y, i0 = p * NB;
  const int m = n - i0;
  const int w = (m - 1 <= NB) ? m - 1 : NB;
  const T *A = a.A + bd.offA;  // reflector jj of this panel: column i0+jj, unit at row i0+jj+1
  const T *tau = a.tau + bd.offV + i0;
  T *Tp = a.Tf + bd.offT + (size_t)p * NB * NB;  // Tp(r,c) = Tp[c*NB + r]
  __shared__ T G[NB][NB + 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // G(x,y) = y_x^T y_y for x < y
  for (int pr = wid; pr < w * w; pr += 4) {
    const int x = pr % w, y = pr / w;
    if (x >= y) continue;
    const T *cx = A + (size_t)(i0 + x) * n, *cy = A + (size_t)(i0 + y) * n;
    T s = 0;
    for (int r = i0 + y + 2 + lane; r < n; r += 64) s += cx[r] * cy[r];
    s = wave_sum(s);
    if (lane == 0) G[x][y] = s + cx[i0 + y + 1];  // row i0+y+1: y_y = 1, y_x stored
  }
  __syncthreads();
  // recurrence, one column at a time; threads x < i compute T(x, i)
  __shared__ T Ts[NB][NB + 1];
  for (int idx = threadIdx.x; idx < NB * NB; idx += 256) Ts[idx % NB][idx / NB] = 0;
  __syncthreads();
  for (int i = 0; i < w; i++) {
    const T ti = tau[i];
    if (threadIdx.x < i) {
      const int x = threadIdx.x;
      T s = 0;
      for (int c = x; c < i; c++) s += Ts[x][c] * G[c][i];
      Ts[x][i] = -ti * s;
    }
    if (threadIdx.x == 0) Ts[i][i] = ti;
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < NB * NB; idx += 256) Tp[idx] = Ts[idx % NB][idx / NB];
}

// ------------------------------------------------------------------------------------------
// 4b. V = Q_H Z on a strip of CW columns: Z <- (I - Y T Y^T) Z, last panel first
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_bt_apply(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n;
  const int c0 = it.y * CW;
  const int ncol = min(CW, n - c0);
  const T *A = a.A + bd.offA;
  T *Z = a.Zt + bd.offZ;
  __shared__ T Ys[32][NB + 1];   // [row][reflector]
  __shared__ T Zs[32][CW + 1];   // [row][column]
  __shared__ T W1[NB][CW + 1], W2[NB][CW + 1];
  __shared__ T Ts[NB][NB + 1];
  const int tid = threadIdx.x;
  const int ta = tid & 31, tc = tid >> 5;  // (reflector or row, column group of 4)
  for (int p = bd.P - 1; p >= 0; p--) {
    const int i0 = p * NB, m = n - i0;
    const int w = (m - 1 <= NB) ? m - 1 : NB;
    const T *Tp = a.Tf + bd.offT + (size_t)p * NB * NB;
    for (int idx = tid; idx < NB * NB; idx += 256) Ts[idx % NB][idx / NB] = Tp[idx];
    // W1 = Y^T Z over rows i0+1 .. n-1
    T acc[4] = {0, 0, 0, 0};
    for (int r0 = i0 + 1; r0 < n; r0 += 32) {
      __syncthreads();
      for (int idx = tid; idx < 32 * NB; idx += 256) {
        const int rr = idx & 31, x = idx >> 5;
        const int r = r0 + rr;
        T y = 0;
        if (r < n && x < w) {
          const int unit = i0 + x + 1;
          y = r == unit ? (T)1 : (r > unit ? A[(size_t)(i0 + x) * n + r] : (T)0);
        }
        Ys[rr][x] = y;
      }
      for (int idx = tid; idx < 32 * CW; idx += 256) {
        const int rr = idx & 31, c = idx >> 5;
        const int r = r0 + rr;
        Zs[rr][c] = (r < n && c < ncol) ? Z[(size_t)(c0 + c) * n + r] : (T)0;
      }
      __syncthreads();
#pragma unroll 8
      for (int rr = 0; rr < 32; rr++) {
        const T y = Ys[rr][ta];
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] += y * Zs[rr][tc * 4 + q];
      }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) W1[ta][tc * 4 + q] = acc[q];
    __syncthreads();
    // W2 = T W1 (T upper triangular)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int c = tc * 4 + q;
      T s = 0;
      for (int y = ta; y < NB; y++) s += Ts[ta][y] * W1[y][c];
      W2[ta][c] = s;
    }
    // Z -= Y W2
    for (int r0 = i0 + 1; r0 < n; r0 += 32) {
      __syncthreads();
      for (int idx = tid; idx < 32 * NB; idx += 256) {
        const int rr = idx & 31, x = idx >> 5;
        const int r = r0 + rr;
        T y = 0;
        if (r < n && x < w) {
          const int unit = i0 + x + 1;
          y = r == unit ? (T)1 : (r > unit ? A[(size_t)(i0 + x) * n + r] : (T)0);
        }
        Ys[rr][x] = y;
      }
      __syncthreads();
      const int r = r0 + ta;
      if (r < n) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int c = tc * 4 + q;
          if (c < ncol) {
            T s = 0;
#pragma unroll 8
            for (int x = 0; x < NB; x++) s += Ys[ta][x] * W2[x][c];
            Z[(size_t)(c0 + c) * n + r] -= s;
          }
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// 5. results back into the caller's arenas
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_finish(SyevArgs<T> a, const int *blist) {
  const int b = blist[blockIdx.x];
  const BlkDesc bd = a.bd[b];
  const int n = (int)bd.n;
  const double s = a.scale[b];
  const double *lam = a.lam + bd.offV;
  double *w = a.w + bd.offw;
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double v = lam[i] * s;
    w[i] = v;
    bad |= !isfinite(v);
  }
  const T *Z = a.Zt + bd.offZ;
  double *H = a.H + bd.offH;
  const size_t nn = (size_t)n * n;
  for (size_t i = threadIdx.x; i < nn; i += 256) H[i] = (double)Z[i];
  if (bad) a.flags[b] = 2;
}

template <typename T>
__global__ void k_convert(const double *src, T *dst, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (T)src[i];
}

// ==========================================================================================
// host driver
// ==========================================================================================
int syev_run(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H, const int64_t *off_w,
             double *w, int mode, int *info, hipStream_t st) {
  if (nblk < 0 || (nblk > 0 && (!n || !off_H || !off_w || !H || !w))) {
    set_error("nrg_syev_batched: invalid arguments");
    return NRG_E_ARG;
  }
  if (mode != NRG_SYEV_FP64 && mode != NRG_SYEV_MIXED) {
    set_error("nrg_syev_batched: unknown mode");
    return NRG_E_ARG;
  }
  if (nblk == 0) return NRG_OK;
  for (int b = 0; b < nblk; b++)
    if (n[b] < 1 || n[b] > NMAX) {
      set_error("nrg_syev_batched: block order out of range [1, 4096]");
      return NRG_E_ARG;
    }
  if (mode == NRG_SYEV_MIXED) {
    extern int syev_mixed_run(nrg_ctx *, int, const int64_t *, const int64_t *, double *, const int64_t *,
                              double *, int *, hipStream_t);
    return syev_mixed_run(ctx, nblk, n, off_H, H, off_w, w, info, st);
  }
  extern int syev_fp64_run(nrg_ctx *, int, const int64_t *, const int64_t *, double *, const int64_t *, double *,
                           int *, hipStream_t);
  return syev_fp64_run(ctx, nblk, n, off_H, H, off_w, w, info, st);
}

// ---------------------------------------------------------------------------------------
// Host-side planning shared by both forms.
// ---------------------------------------------------------------------------------------
struct SyevPlan {
  std::vector<BlkDesc> bd;
  int64_t sumn = 0, sumn2 = 0, sumWp = 0, sumT = 0;
  int Pmax = 0, Lmax = 0, nmax = 0;
  // work lists, all packed into one int array; offsets in ints
  std::vector<int> ibuf;
  size_t symm_off = 0, symm_cnt = 0;
  std::vector<size_t> pan_off, pan_cnt, upd_off, upd_cnt;
  size_t all_off = 0;  // list 0..nblk-1
  std::vector<size_t> mrg_off, mrg_cnt, gemm_off, gemm_cnt, prob_first, prob_cnt;
  size_t tf_off = 0, tf_cnt = 0, bt_off = 0, bt_cnt = 0;
  std::vector<int> mcap;
  std::vector<GemmNN> probs;  // filled once the arena pointers are known
};

static int panels_of(int64_t n) {
  if (n <= 1) return n == 1 ? 1 : 0;
  int P = 1;
  while (n - (int64_t)NB * (P - 1) - 1 > NB) P++;
  return P;
}

static void build_plan(SyevPlan &pl, int nblk, const int64_t *n, const int64_t *off_H, const int64_t *off_w,
                       bool fp64_inplace) {
  pl.bd.resize(nblk);
  int64_t zo = 0, vo = 0, wo = 0, to = 0, ao = 0;
  for (int b = 0; b < nblk; b++) {
    BlkDesc &d = pl.bd[b];
    d.n = n[b];
    d.offZ = zo; zo += n[b] * n[b];
    d.offA = fp64_inplace ? off_H[b] : ao; ao += n[b] * n[b];
    d.offV = vo; vo += n[b];
    d.offWp = wo; wo += n[b] * NB;
    d.P = panels_of(n[b]);
    d.offT = to; to += (int64_t)d.P * NB * NB;
    d.offH = off_H[b];
    d.offw = off_w[b];
    d.L = 31 - __builtin_clz((unsigned)n[b]);
    pl.Pmax = std::max(pl.Pmax, d.P);
    pl.Lmax = std::max(pl.Lmax, d.L);
    pl.nmax = std::max(pl.nmax, (int)n[b]);
  }
  pl.sumn2 = zo; pl.sumn = vo; pl.sumWp = wo; pl.sumT = to;
  std::vector<int> &I = pl.ibuf;
  I.clear();
  auto push2 = [&](int x, int y) { I.push_back(x); I.push_back(y); };
  // all blocks, largest first (as TaskList, c++/subspaces.hpp:114-121)
  std::vector<int> order(nblk);
  for (int b = 0; b < nblk; b++) order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return n[x] > n[y]; });
  pl.all_off = I.size();
  for (int b : order) I.push_back(b);
  // symmetrisation tiles
  pl.symm_off = I.size();
  for (int b : order) {
    const int64_t t = (n[b] + 31) / 32;
    for (int64_t q = 0; q < t * (t + 1) / 2; q++) push2(b, (int)q);
  }
  pl.symm_cnt = (I.size() - pl.symm_off) / 2;
  // panels and updates
  pl.pan_off.assign(pl.Pmax, 0); pl.pan_cnt.assign(pl.Pmax, 0);
  pl.upd_off.assign(pl.Pmax, 0); pl.upd_cnt.assign(pl.Pmax, 0);
  for (int p = 0; p < pl.Pmax; p++) {
    pl.pan_off[p] = I.size();
    for (int b : order) if (pl.bd[b].P > p) I.push_back(b);
    pl.pan_cnt[p] = I.size() - pl.pan_off[p];
    pl.upd_off[p] = I.size();
    for (int b : order)
      if (pl.bd[b].P > p + 1) {
        const int64_t mt = n[b] - (int64_t)NB * (p + 1);
        const int64_t td = (mt + 63) / 64;
        for (int64_t t = 0; t < td * td; t++) push2(b, (int)t);
      }
    pl.upd_cnt[p] = (I.size() - pl.upd_off[p]) / 2;
  }
  // divide and conquer merges, deepest first
  pl.mrg_off.assign(pl.Lmax, 0); pl.mrg_cnt.assign(pl.Lmax, 0);
  pl.gemm_off.assign(pl.Lmax, 0); pl.gemm_cnt.assign(pl.Lmax, 0);
  pl.prob_first.assign(pl.Lmax, 0); pl.prob_cnt.assign(pl.Lmax, 0);
  pl.mcap.assign(pl.Lmax, 0);
  for (int depth = pl.Lmax - 1; depth >= 0; depth--) {
    pl.mrg_off[depth] = I.size();
    int mc = 0;
    for (int b : order)
      if (pl.bd[b].L > depth)
        for (int q = 0; q < (1 << depth); q++) {
          push2(b, q);
          const int a0 = (int)(((int64_t)q * n[b]) >> depth), a1 = (int)(((int64_t)(q + 1) * n[b]) >> depth);
          mc = std::max(mc, a1 - a0);
        }
    pl.mrg_cnt[depth] = (I.size() - pl.mrg_off[depth]) / 2;
    pl.mcap[depth] = mc;
  }
  // compact-WY factors and back-transformation strips
  pl.tf_off = I.size();
  for (int b : order)
    if (n[b] >= 2)
      for (int p = 0; p < pl.bd[b].P; p++) push2(b, p);
  pl.tf_cnt = (I.size() - pl.tf_off) / 2;
  pl.bt_off = I.size();
  for (int b : order)
    if (n[b] >= 3)
      for (int s = 0; s < (n[b] + CW - 1) / CW; s++) push2(b, s);
  pl.bt_cnt = (I.size() - pl.bt_off) / 2;
}

// GEMM problems of the large merges at each depth (pointers into the arenas)
static void build_gemms(SyevPlan &pl, double *Z0, double *Z1, double *S) {
  pl.probs.clear();
  for (int depth = pl.Lmax - 1; depth >= 0; depth--) {
    pl.prob_first[depth] = pl.probs.size();
    const int *it = pl.ibuf.data() + pl.mrg_off[depth];
    for (size_t i = 0; i < pl.mrg_cnt[depth]; i++) {
      const int b = it[2 * i], q = it[2 * i + 1];
      const BlkDesc &d = pl.bd[b];
      const int nb = (int)d.n;
      const int a0 = (int)(((int64_t)q * nb) >> depth);
      const int mid = (int)(((int64_t)(2 * q + 1) * nb) >> (depth + 1));
      const int a1 = (int)(((int64_t)(q + 1) * nb) >> depth);
      const int m = a1 - a0;
      if (m <= SMALLM) continue;
      const double *Qin = ((depth + 1) & 1 ? Z1 : Z0) + d.offZ;
      double *Qout = (depth & 1 ? Z1 : Z0) + d.offZ;
      const double *Sb = S + d.offZ;
      for (int half = 0; half < 2; half++) {
        const int ro = half ? mid - a0 : 0;
        const int mh = half ? a1 - mid : mid - a0;
        GemmNN g;
        g.A = Qin + (size_t)(a0 + ro) * nb + (a0 + ro);
        g.B = Sb + (size_t)a0 * nb + (a0 + ro);
        g.C = Qout + (size_t)a0 * nb + (a0 + ro);
        g.M = mh; g.N = m; g.K = mh; g.ld = nb;
        pl.probs.push_back(g);
      }
    }
    pl.prob_cnt[depth] = pl.probs.size() - pl.prob_first[depth];
  }
  // tile items for every depth appended to ibuf
  for (int depth = pl.Lmax - 1; depth >= 0; depth--) {
    pl.gemm_off[depth] = pl.ibuf.size();
    for (size_t pi = pl.prob_first[depth]; pi < pl.prob_first[depth] + pl.prob_cnt[depth]; pi++) {
      const GemmNN &g = pl.probs[pi];
      const int tiles = ((g.M + 63) / 64) * ((g.N + 63) / 64);
      for (int t = 0; t < tiles; t++) { pl.ibuf.push_back((int)pi); pl.ibuf.push_back(t); }
    }
    pl.gemm_cnt[depth] = (pl.ibuf.size() - pl.gemm_off[depth]) / 2;
  }
}

struct SyevLaunch {
  const int *dI;      // device copy of ibuf
  const GemmNN *dP;   // device copy of probs
};

// Launch sequence of the reduction + D&C + back-transformation for a typed work matrix.
template <typename T>
static int launch_core(nrg_ctx *ctx, SyevPlan &pl, SyevArgs<T> &aT, SyevArgs<double> &aD, const SyevLaunch &L,
                       hipStream_t st, bool zt_is_z0, int phase0) {
  const int *I = L.dI;
  auto grid = [](size_t c) { return dim3((unsigned)c); };
  // 1. symmetrisation
  if (pl.symm_cnt)
    NRG_TIMED(ctx, st, "k_symm", hipLaunchKernelGGL(k_symm<T>, grid(pl.symm_cnt), dim3(256), 0, st, aT, (const int2 *)(I + pl.symm_off)));
  // 2. reduction
  for (int p = 0; p < pl.Pmax; p++) {
    if (pl.pan_cnt[p])
      NRG_TIMED(ctx, st, "k_panel", hipLaunchKernelGGL(k_panel<T>, grid(pl.pan_cnt[p]), dim3(NT_PANEL), 0, st, aT, I + pl.pan_off[p], p));
    if (pl.upd_cnt[p])
      NRG_TIMED(ctx, st, "k_syr2k", hipLaunchKernelGGL(k_syr2k<T>, grid(pl.upd_cnt[p]), dim3(256), 0, st, aT, (const int2 *)(I + pl.upd_off[p]), p));
  }
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[phase0 + 1], st));
  // 3. divide and conquer
  {
    int mc = 0;
    for (int depth = 0; depth < pl.Lmax; depth++) mc = std::max(mc, pl.mcap[depth]);
    const size_t lds = (size_t)mc * (4 * sizeof(double) + sizeof(int)) + 16;
    (void)hipFuncSetAttribute((const void *)k_dc_merge, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  }
  NRG_TIMED(ctx, st, "k_dc_prep", hipLaunchKernelGGL(k_dc_prep, grid(pl.bd.size()), dim3(256), 0, st, aD, I + pl.all_off));
  for (int depth = pl.Lmax - 1; depth >= 0; depth--) {
    if (!pl.mrg_cnt[depth]) continue;
    const int mc = pl.mcap[depth];
    const size_t lds = (size_t)mc * (4 * sizeof(double) + sizeof(int)) + 16;
    NRG_TIMED(ctx, st, "k_dc_merge", hipLaunchKernelGGL(k_dc_merge, grid(pl.mrg_cnt[depth]), dim3(NT_DC), lds, st, aD,
                       (const int2 *)(I + pl.mrg_off[depth]), depth, mc));
    if (pl.gemm_cnt[depth])
      NRG_TIMED(ctx, st, "k_gemm_nn", hipLaunchKernelGGL(k_gemm_nn, grid(pl.gemm_cnt[depth]), dim3(256), 0, st, L.dP,
                         (const int2 *)(I + pl.gemm_off[depth])));
  }
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[phase0 + 2], st));
  // 4. back-transformation
  if (!zt_is_z0) {
    const size_t cnt = (size_t)pl.sumn2;
    NRG_TIMED(ctx, st, "k_convert", hipLaunchKernelGGL(k_convert<T>, dim3(2048), dim3(256), 0, st, (const double *)aD.Z0, aT.Zt, cnt));
  }
  if (pl.tf_cnt)
    NRG_TIMED(ctx, st, "k_bt_tfactor", hipLaunchKernelGGL(k_bt_tfactor<T>, grid(pl.tf_cnt), dim3(256), 0, st, aT, (const int2 *)(I + pl.tf_off)));
  if (pl.bt_cnt)
    NRG_TIMED(ctx, st, "k_bt_apply", hipLaunchKernelGGL(k_bt_apply<T>, grid(pl.bt_cnt), dim3(256), 0, st, aT, (const int2 *)(I + pl.bt_off)));
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[phase0 + 3], st));
  return NRG_OK;
}

// Workspace layout helper
struct Carver {
  size_t off = 0;
  template <typename T> size_t take(size_t count) {
    const size_t o = off;
    off += (count * sizeof(T) + 255) & ~(size_t)255;
    return o;
  }
};

static int finish_flags(nrg_ctx *ctx, int nblk, int *info, hipStream_t st) {
  NRG_HIP(hipMemcpyAsync(ctx->hflags.p, ctx->flags.p, sizeof(int) * nblk, hipMemcpyDeviceToHost, st));
  NRG_HIP(hipStreamSynchronize(st));
  const int *f = (const int *)ctx->hflags.p;
  int bad = 0;
  for (int b = 0; b < nblk; b++) {
    if (info) info[b] = f[b];
    bad |= f[b] != 0;
  }
  if (bad) {
    set_error("nrg_syev_batched: at least one block failed (non-finite input or result)");
    return NRG_E_NOCONV;
  }
  return NRG_OK;
}

int syev_fp64_run(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H, const int64_t *off_w,
                  double *w, int *info, hipStream_t st) {
  SyevPlan pl;
  build_plan(pl, nblk, n, off_H, off_w, true);
  Carver cv;
  const size_t oZ0 = cv.take<double>(pl.sumn2), oZ1 = cv.take<double>(pl.sumn2), oS = cv.take<double>(pl.sumn2);
  const size_t oWp = cv.take<double>(pl.sumWp), oTau = cv.take<double>(pl.sumn), oTf = cv.take<double>(pl.sumT);
  const size_t od = cv.take<double>(pl.sumn), oe = cv.take<double>(pl.sumn), ol = cv.take<double>(pl.sumn);
  const size_t osc = cv.take<double>(nblk), odscr = cv.take<double>(8 * pl.sumn), oiscr = cv.take<int>(8 * pl.sumn);
  if (int r = ctx->ws.ensure(cv.off)) return r;
  double *Z0 = ctx->ws.as<double>(oZ0), *Z1 = ctx->ws.as<double>(oZ1), *S = ctx->ws.as<double>(oS);
  build_gemms(pl, Z0, Z1, S);
  // descriptor upload: [BlkDesc][GemmNN][ints]
  const size_t bBD = sizeof(BlkDesc) * pl.bd.size(), bP = sizeof(GemmNN) * pl.probs.size(), bI = sizeof(int) * pl.ibuf.size();
  const size_t oBD = 0, oP = (bBD + 255) & ~(size_t)255, oI = oP + ((bP + 255) & ~(size_t)255);
  std::vector<char> blob(oI + bI);
  memcpy(blob.data() + oBD, pl.bd.data(), bBD);
  if (bP) memcpy(blob.data() + oP, pl.probs.data(), bP);
  memcpy(blob.data() + oI, pl.ibuf.data(), bI);
  if (int r = ctx->flags.ensure(sizeof(int) * nblk)) return r;
  if (int r = ctx->hflags.ensure(sizeof(int) * nblk)) return r;
  NRG_HIP(hipEventRecord(ctx->ev[0], st));
  if (int r = upload_desc(ctx, blob.data(), blob.size(), st)) return r;
  NRG_HIP(hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * nblk, st));
  SyevArgs<double> a;
  a.bd = (const BlkDesc *)(ctx->desc.as<char>(oBD));
  a.A = H;
  a.Wp = ctx->ws.as<double>(oWp); a.tau = ctx->ws.as<double>(oTau); a.Tf = ctx->ws.as<double>(oTf);
  a.Zt = Z0; a.Z0 = Z0; a.Z1 = Z1; a.S = S;
  a.d = ctx->ws.as<double>(od); a.e = ctx->ws.as<double>(oe); a.lam = ctx->ws.as<double>(ol);
  a.scale = ctx->ws.as<double>(osc);
  a.dscr = ctx->ws.as<double>(odscr); a.iscr = ctx->ws.as<int>(oiscr); a.Ntot = pl.sumn;
  a.flags = (int *)ctx->flags.p;
  a.H = H; a.w = w;
  SyevLaunch L;
  L.dI = (const int *)(ctx->desc.as<char>(oI));
  L.dP = (const GemmNN *)(ctx->desc.as<char>(oP));
  if (int r = launch_core<double>(ctx, pl, a, a, L, st, true, 0)) return r;
  NRG_TIMED(ctx, st, "k_finish", hipLaunchKernelGGL(k_finish<double>, dim3((unsigned)nblk), dim3(256), 0, st, a, L.dI + pl.all_off));
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[4], st));
  const int r = finish_flags(ctx, nblk, info, st);
  if (ctx->kt.on) ctx->kt.collect();
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]);
  ctx->total_ms = ms;
  for (int ph = 0; ph < 4; ph++) {
    ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[ph], ctx->ev[ph + 1]);
    ctx->phase_ms[ph] = ms;
  }
  return r;
}

}  // namespace nrg
