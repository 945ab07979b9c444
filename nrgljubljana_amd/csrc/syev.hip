// syev.hip - batched real symmetric eigensolver for gfx950.
//
// Replaces the per-block LAPACK dsyev loop of NRG Ljubljana
// (diagonalisations_OpenMP, c++/core.hpp:90-110 -> diagonalise_dsyev, c++/diag.hpp:61-82).
// All blocks of one NRG step go through the same short sequence of batched launches:
//
//   1. k_symm       mirror the stored triangle, giving a full symmetric column-major work matrix
//   2. k_col_prep   blocked Householder reduction to tridiagonal form (LAPACK dsytrd/dlatrd
//      k_col_symv   algorithm, lower variant): per reduced column, one workgroup per block prepares
//                   the reflector and many workgroups per block form the symmetric product over
//                   64x64 tiles of the lower triangle;
//      k_gemm_f64   after each panel of NB columns, the trailing rank-2k update A22 -= V W^T + W V^T
//                   (tiles on or below the first tile super-diagonal only)
//      k_tail       the last trailing matrix of order <= TAIL (whole blocks up to TAIL), reduced
//                   column by column in one workgroup's LDS, one launch for all blocks
//   3. k_dc_prep    divide and conquer on the tridiagonal matrix (Cuppen's method with the
//      k_dc_*       Gu-Eisenstat eigenvector formula, as LAPACK dstedc/dlaed0-3): leaves of order
//      k_gemm_f64   1-2 at a common tree depth, then five merge launches per depth for all blocks
//   4. k_bt_*       back-transformation V = Q_H Z with compact-WY super-panels (dlarft + dlarfb) as
//      k_gemm_f64   FP64 matrix-core products
//   5. k_finish     eigenvalues (rescaled) and eigenvectors back into the caller's arenas
//
// Storage: a row-major block with its upper triangle filled is, byte for byte, a column-major
// matrix with its lower triangle filled, so every kernel works column-major: A(r,c) = A[c*n + r].
// A column-major eigenvector matrix with eigenvector i in column i is, byte for byte, the
// row-major matrix with eigenvector i in row i that copy_vec() builds (c++/diag.hpp:41-48).
#include "common.hpp"
#include "gemm.hpp"
#include <algorithm>
#include <array>
#include <map>
#include <cstring>
#include <cstdlib>
#include <cmath>

namespace nrg {

constexpr int NB = 32;        // panel width of the Householder reduction
constexpr int NT_PANEL = 1024; // threads of the column-preparation kernel
constexpr int TAIL = 128;      // trailing order reduced inside one workgroup's LDS (k_tail)
constexpr int RED_GROUPS = 2;  // block groups whose column chains run concurrently, one stream each
constexpr int SMALLM = 64;    // merges up to this order do their own block-diagonal product
// work-list entry of the column kernels: block index in the low bits, its panel above them
constexpr int COL_PANEL_SHIFT = 20;
constexpr int COL_BLK_MASK = (1 << COL_PANEL_SHIFT) - 1;
inline int col_entry(int b, int p) { return b | (p << COL_PANEL_SHIFT); }
// Panel of a block with P panels in panel slot q (-1: none). Every block starts in slot 0; a
// right-aligned schedule (all blocks ending together) measured no faster, since the bandwidth of
// the first columns belongs to the largest blocks either way.
inline int panel_in_slot(int q, int P) { return q < P ? q : -1; }

struct BlkDesc {
  int64_t n, offA, offZ, offV, offWp, offH, offw, offYp, offDt;
  int64_t offAB, offV2, offT2;   // two-stage reduction: band, bulge-chasing reflectors and scalars
  int L, P;   // P: panels reduced by the column kernels (before k_wg_red and the tail)
  int two;    // 1: reduced in two stages (dense -> band -> tridiagonal)
  int roff;   // reflector j of the reduction starts at row j + roff (1, or SB in two stages)
};

template <typename T>
struct SyevArgs {
  const BlkDesc *bd;
  T *A;            // work matrices (column-major, full symmetric, then reflectors)
  T *Wp;           // dlatrd W panels, n x NB per block
  T *tau;          // Householder scalars, n per block
  T *vb;           // current reflector of every block, n per block
  T *Yp;           // symmetric-product tile partials, n x ceil(n/64) per block
  T *Ym, *Yh;      // back-transformation: explicit reflectors Y and Yhat = Y T^T (n x n per block)
  T *Gb, *Tb;      // back-transformation: G_s and T_s of every super-panel
  T *Zt;           // eigenvector matrix used by the back-transformation (== Z0 for FP64)
  double *Z0, *Z1, *S;
  double *d, *e, *lam, *scale;
  double *dscr;    // 8 arrays of Ntot doubles
  int *iscr;       // 8 arrays of Ntot ints
  int64_t Ntot;
  int *flags;
  double *amax;    // per-block max-norm of the input (bit pattern updated with atomicMax)
  double *H, *w;
  // column preparation of blocks with more than NT_PANEL trailing rows (k_col_yred)
  T *cs;                // the reduced column after the panel terms, n per block
  T *dpp;               // per 256-row chunk of the trailing rows: its part of (w^T v)
  // two-stage reduction (syev_sbr.hpp)
  T *sbV, *sbY, *sbW;   // explicit panel reflectors V, Y = A22 V and W, n x SB per block (ld n)
  T *sbT;               // T factor of each block's current panel, SB x SB
  T *sbAB;              // band storage, 2 SB x n per block
  T *sbV2, *sbT2;       // bulge-chasing reflectors and their scalars
};


// ------------------------------------------------------------------------------------------
// 1. symmetrisation: lower triangle of the (double) input -> full (T) work matrix
// ------------------------------------------------------------------------------------------
// Max-norm of every block's stored triangle (atomicMax on the bit pattern of a non-negative
// double, whose integer order is the numeric order); non-finite entries give +inf.
template <typename T>
__global__ __launch_bounds__(256) void k_absmax(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n;
  const int t = it.y;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
  while (ti * (ti + 1) / 2 > t) ti--;
  const int tj = t - ti * (ti + 1) / 2;
  const double *Hs = a.H + bd.offH;
  __shared__ double red[4];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  double mx = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int r = ti * 32 + tx, c = tj * 32 + ty + 8 * q;
    if (r < n && c < n && r >= c) {
      const double v = Hs[(size_t)c * n + r];
      mx = isfinite(v) ? fmax(mx, fabs(v)) : INFINITY;
    }
  }
  mx = block_max<256>(mx, red);
  if (threadIdx.x == 0 && mx > 0)
    atomicMax((unsigned long long *)(a.amax + it.x), (unsigned long long)__double_as_longlong(mx));
}

// Power-of-two factor bringing the block's max-norm into [1, 2): exact, and it keeps squares of
// the entries away from the subnormal range in the reflector norms (dsyev scales for the same
// reason when the norm leaves [sqrt(safmin/eps), sqrt(bignum)]).
__device__ __forceinline__ double block_sigma(double amax) {
  return (amax > 0 && isfinite(amax)) ? ldexp(1.0, -ilogb(amax)) : 1.0;
}

// dst_h: write the full matrix in place into the caller's arena (FP64 only) instead of the work matrix
template <typename T>
__global__ __launch_bounds__(256) void k_symm(SyevArgs<T> a, const int2 *items, int dst_h) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n;
  const int t = it.y;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ti++;
  while (ti * (ti + 1) / 2 > t) ti--;
  const int tj = t - ti * (ti + 1) / 2;
  const double *Hs = a.H + bd.offH;
  T *Ad = dst_h ? (T *)(a.H + bd.offH) : a.A + bd.offA;
  const double sigma = block_sigma(a.amax[it.x]);
  __shared__ T tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int cl = ty + 8 * q;
    const int r = ti * 32 + tx, c = tj * 32 + cl;
    T v = 0;
    if (r < n && c < n && r >= c) {
      v = (T)(Hs[(size_t)c * n + r] * sigma);
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
// 2a. column preparation (dlatrd, lower): one workgroup per block and reduced column.
//     Finishes the W column of the previous reflector from the partials of k_col_symv, applies
//     the panel's earlier reflectors to column jl and generates reflector jl. A thread owns the
//     rows tid + q*NT_PANEL of the trailing part and issues every load of a row together (with one
//     row per thread, the first sixteen panel columns of the update too), so a column costs a few
//     memory round trips instead of one per dot product.
// ------------------------------------------------------------------------------------------
constexpr int KCH = 16;   // panel columns loaded per batch
constexpr int NT_YR = 256;   // rows per workgroup of k_col_yred

// RPT: rows per thread, 1 when every block of the launch has at most NT rows [jl, m); NT threads
// (fewer than NT_PANEL when the launch's blocks have few rows: fewer waves per barrier and reduction)
template <typename T, int RPT, int NT = NT_PANEL>
__global__ __launch_bounds__(NT) void k_col_prep(SyevArgs<T> a, const int *blist, int jl) {
  const int ent = blist[blockIdx.x], p = ent >> COL_PANEL_SHIFT;
  const BlkDesc bd = a.bd[ent & COL_BLK_MASK];
  const int n = (int)bd.n;
  const int i0 = p * NB;
  const int m = n - i0;
  T *A = a.A + bd.offA + (size_t)i0 * n + i0;   // A(r,c) = A[c*n + r], relative to the panel
  T *W = a.Wp + bd.offWp;                        // W(r,k) = W[k*n + r]
  double *d = a.d + bd.offV + i0;
  double *e = a.e + bd.offV + i0;
  T *tau = a.tau + bd.offV + i0;
  T *vb = a.vb + bd.offV + i0;                   // reflector handed to k_col_symv
  const T *yp = a.Yp + bd.offYp + i0;            // symmetric-product partials: yp[slot*n + r]
  const T *dt = a.Yp + bd.offDt;                 // per 64-row segment: W^T v (0..31), V^T v (32..63)
  const bool final = (m - 1 <= NB);
  const int w = final ? m - 1 : NB;
  const bool do_col = jl < w || final;           // column jl is updated here (and reduced if jl < w)
  const int jp = jl - 1;                         // reflector whose W column is finished here
  const int tid = threadIdx.x;
  constexpr int NW = NT / WAVE;
  constexpr int KC = KCH;

  __shared__ T part[NW][64];
  __shared__ T cwv[2 * NB];          // cw[k] = W(:,k)^T v and cv[k] = V(:,k)^T v over the trailing rows
  __shared__ T vrow[NB], wrow[NB];   // A(jl,k) and W(jl,k)
  __shared__ T red[NW];
  __shared__ T sh[2];                // W(jl,jp); A(jl+1,jl) after the update

  const T tv = jl > 0 ? tau[jp] : (T)0;
  const int Tt = (m - jl + 63) / 64;             // 64-row segments of the rows [jl, m)
  T cf[RPT];   // column jl after its update, rows jl + tid + q NT
  if (m - jl > NT) {
    T v[RPT], y[RPT], aj[RPT];
    T *c = cf;
    // more rows than threads. The first column of a panel is only loaded. Otherwise k_col_yred has
    // reduced the symmetric-product partials and applied the panel terms row by row in many
    // workgroups: W(:,jp) holds tau (A22 v - V cw - W cv) and cs the column after the terms k < jp.
    // Finish W(:,jp) with a2 v and the column with the term k = jp.
    if (jl == 0) {
#pragma unroll
      for (int q = 0; q < RPT; q++) {
        const int r = tid + q * NT;
        c[q] = r < m ? A[r] : (T)0;
      }
    } else {
    const T *cs = a.cs + bd.offV + i0;
    const T *dpp = a.dpp + bd.offV;
    const int nch = (m - jl + NT_YR - 1) / NT_YR;   // at most NMAX / NT_YR <= NT chunks
    // every load first, then the block sum that needs the chunks' partials
    const T vjp = (tid == 0 && do_col) ? A[(size_t)jp * n + jl] : (T)0;   // V(jl, jp)
    const T dv = tid < nch ? dpp[tid] : (T)0;
    T wo[RPT];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int r = jl + tid + q * NT;
      v[q] = y[q] = c[q] = aj[q] = wo[q] = 0;
      if (r < m) {
        v[q] = vb[r];
        wo[q] = W[(size_t)jp * n + r];
        if (do_col) {
          c[q] = cs[r];
          aj[q] = A[(size_t)jp * n + r];
        }
      }
    }
    if (tid == 0) sh[1] = vjp;
    // no closing barrier: the one below, before the reflector, comes first
    const T a2 = (T)(-0.5) * tv * block_sum<NT, false>(dv, red);
#pragma unroll
    for (int q = 0; q < RPT; q++) {
      const int r = jl + tid + q * NT;
      if (r < m) {
        y[q] = wo[q] + a2 * v[q];
        W[(size_t)jp * n + r] = y[q];
        if (r == jl) sh[0] = y[q];
      }
    }
    __syncthreads();
    if (do_col) {
      const T wj = sh[0], vj = sh[1];
#pragma unroll
      for (int q = 0; q < RPT; q++) c[q] -= aj[q] * wj + y[q] * vj;
    }
    __syncthreads();   // sh[1] is reused by the reflector below
    }
  } else {
  // at most NT rows [jl, m): one row per thread. Every load of the first phase is issued before any
  // of them is used (a loop over a run-time count of segments, or an LDS store of a loaded value,
  // between them would wait for each load in turn)
  constexpr int R1 = 1;
  constexpr int TS = NT / 64;   // Tt <= TS segments, and at most one segment per wave
  T v[R1], y[R1], c[R1], aj[R1], ys[R1][TS];
  // 1. loads that depend on nothing computed here
  T pl = 0, vr = 0, wr = 0;   // pl: this wave's segment partial of column tid & 63
  if (jl > 0 && ((tid & 63) & 31) < jp && (tid >> 6) < Tt) pl = dt[(tid >> 6) * 64 + (tid & 63)];
  if (tid < NB) {
    vr = (do_col && tid < jl) ? A[(size_t)tid * n + jl] : (T)0;
    wr = (do_col && tid < jp) ? W[(size_t)tid * n + jl] : (T)0;
  }
#pragma unroll
  for (int q = 0; q < R1; q++) {
    const int r = jl + tid + q * NT;
    v[q] = y[q] = c[q] = aj[q] = 0;
    if (r < m) {
      if (jl > 0) {
        v[q] = vb[r];
        aj[q] = A[(size_t)jp * n + r];
#pragma unroll
        for (int sl = 0; sl < TS; sl++) ys[q][sl] = sl < Tt ? yp[(size_t)sl * n + r] : (T)0;
      }
      if (do_col) c[q] = A[(size_t)jl * n + r];
    }
  }
  // the first KP panel columns of the update, loaded with the loads above (none when a thread owns
  // several rows: registers). KP and NB are multiples of KC, so the batches below index the panel
  // arrays cwv, wrow, vrow below NB (columns past jp carry zeros)
  constexpr int KP = KC;
  static_assert(KP % KC == 0 && NB % KC == 0, "panel batches must tile the panel");
  T ap[R1][KP], wp[R1][KP];
#pragma unroll
  for (int q = 0; q < R1; q++) {
    const int r = jl + tid + q * NT;
#pragma unroll
    for (int k = 0; k < KP; k++) {
      const bool in = r < m && k < jp;
      ap[q][k] = in ? A[(size_t)k * n + r] : (T)0;
      wp[q][k] = in ? W[(size_t)k * n + r] : (T)0;
    }
  }
  // the loaded values into the LDS, and the product's segment partials summed in segment order
  if (jl > 0) part[tid >> 6][tid & 63] = pl;   // summed from +0 below, as if added to a +0 here
  if (tid < NB) {
    vrow[tid] = vr;
    wrow[tid] = wr;
  }
#pragma unroll
  for (int q = 0; q < R1; q++) {
    const int r = jl + tid + q * NT;
    if (r < m && jl > 0) {
      T s = 0;
#pragma unroll
      for (int sl = 0; sl < TS; sl++)
        if (sl < Tt) s += ys[q][sl];
      y[q] = s;
    }
  }
  if (jl > 0) {
    __syncthreads();
    if (tid < 2 * NB) {
      T s = 0;
#pragma unroll
      for (int q = 0; q < NW; q++) s += part[q][tid];
      cwv[tid] = s;
    }
    __syncthreads();
    // 2. W(:,jp) = tau (A22 v - V cw - W cv) - tau/2 (w^T v) v, and the terms k < jp of the update
    //    of column jl: A(:,jl) -= V W(jl,:)^T + W V(jl,:)^T
    T dp = 0;
#pragma unroll
    for (int q = 0; q < R1; q++) {
      const int r = jl + tid + q * NT;
      if (r < m) {
        T ay = 0, ac = 0;
#pragma unroll
        for (int k = 0; k < KP; k++) {
          ay += ap[q][k] * cwv[k] + wp[q][k] * cwv[NB + k];
          ac += ap[q][k] * wrow[k] + wp[q][k] * vrow[k];
        }
        for (int k0 = KP; k0 < jp; k0 += KC) {
          T ak[KC], wk[KC];
#pragma unroll
          for (int k = 0; k < KC; k++) {
            const bool in = k0 + k < jp;
            ak[k] = in ? A[(size_t)(k0 + k) * n + r] : (T)0;
            wk[k] = in ? W[(size_t)(k0 + k) * n + r] : (T)0;
          }
#pragma unroll
          for (int k = 0; k < KC; k++) {
            ay += ak[k] * cwv[k0 + k] + wk[k] * cwv[NB + k0 + k];
            ac += ak[k] * wrow[k0 + k] + wk[k] * vrow[k0 + k];
          }
        }
        const T yy = tv == (T)0 ? (T)0 : tv * (y[q] - ay);
        y[q] = yy;
        dp += yy * v[q];
        c[q] -= ac;
      }
    }
    const T a2 = (T)(-0.5) * tv * block_sum<NT, false>(dp, red);   // the barrier after W(:,jp) follows
#pragma unroll
    for (int q = 0; q < R1; q++) {
      const int r = jl + tid + q * NT;
      if (r < m) {
        y[q] += a2 * v[q];
        W[(size_t)jp * n + r] = y[q];
        if (r == jl) sh[0] = y[q];
      }
    }
    __syncthreads();
    if (do_col) {
      const T wj = sh[0], vj = vrow[jp];
#pragma unroll
      for (int q = 0; q < R1; q++) c[q] -= aj[q] * wj + y[q] * vj;
    }
  }
  cf[0] = c[0];
#pragma unroll
  for (int q = 1; q < RPT; q++) cf[q] = 0;
  }
  if (!do_col) return;
  if (jl == w) {   // last diagonal element of the block (row jl is owned by thread 0)
    if (tid == 0) {
      A[(size_t)jl * n + jl] = cf[0];
      d[jl] = (double)cf[0];
    }
    return;
  }
  // 3. reflector annihilating A(jl+2:m, jl) (dlarfg)
  T nrm = 0;
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int r = jl + tid + q * NT;
    if (r >= jl + 2 && r < m) nrm += cf[q] * cf[q];
    if (r == jl + 1) sh[1] = cf[q];
  }
  const T xn2 = block_sum<NT, false>(nrm, red);   // red is not written again
  const T alpha = sh[1];
  T beta, tn, scal = 1;
  if (xn2 == (T)0) {
    tn = 0;
    beta = alpha;
  } else {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tn = (beta - alpha) / beta;
    scal = (T)1 / (alpha - beta);
  }
  T *colj = A + (size_t)jl * n;
#pragma unroll
  for (int q = 0; q < RPT; q++) {
    const int r = jl + tid + q * NT;
    if (r < m) {
      if (r == jl) {
        colj[r] = cf[q];
        d[jl] = (double)cf[q];
      } else if (r == jl + 1) {
        colj[r] = 1;
        vb[r] = 1;
      } else {
        const T x = cf[q] * scal;
        colj[r] = x;
        vb[r] = x;
      }
    }
  }
  if (tid == 0) {
    e[jl] = (double)beta;
    tau[jl] = tn;
  }
}

// Sum over the 64 lanes of x[c(lane)], c(lane) = 8 b5 + 4 b4 + 2 b3 + b2 (bits of the lane id):
// a butterfly that halves the number of live values per lane at every exchange, without the LDS
// (permlane swaps across the wave halves and the rows, DPP mirrors within rows and quads).
template <typename T>
__device__ __forceinline__ T butterfly16(T (&x)[16]) {
  const int l = threadIdx.x & (WAVE - 1);
  // bit 5: lanes 0-31 keep x[j] + x[j] of lane + 32, lanes 32-63 keep x[j + 8] + x[j + 8] of lane - 32
#pragma unroll
  for (int j = 0; j < 8; j++) {
    T a = x[j], b = x[j + 8];
    swap_rows<true>(a, b);
    x[j] = a + b;
  }
  // bit 4, rows 0/2 against rows 1/3
#pragma unroll
  for (int j = 0; j < 4; j++) {
    T a = x[j], b = x[j + 4];
    swap_rows<false>(a, b);
    x[j] = a + b;
  }
  {   // bit 3: the partner lane 15 - i of the row has the other bit 3
    const bool up = l & 8;
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const T send = up ? x[j] : x[j + 2];
      const T keep = up ? x[j + 2] : x[j];
      x[j] = keep + dpp<DPP_MIRROR>(send);
    }
  }
  {   // bit 2: the partner lane 7 - i of the eight has the other bit 2
    const bool up = l & 4;
    const T send = up ? x[0] : x[1];
    const T keep = up ? x[1] : x[0];
    x[0] = keep + dpp<DPP_HALF_MIRROR>(send);
  }
  T s = x[0];
  s += dpp<DPP_QXOR2>(s);
  s += dpp<DPP_QXOR1>(s);
  return s;
}

// ------------------------------------------------------------------------------------------
// 2a'. the row-parallel part of column jl's preparation for blocks with more than NT_PANEL rows
//     [jl, m): item (block entry, 256-row chunk). Reduces the symmetric-product partials of the
//     previous reflector's row, applies the panel terms (V (W^T v) + W (V^T v) to the product, and
//     V W(jl,:)^T + W V(jl,:)^T to column jl, k < jp), and leaves W(:,jp) = tau (A22 v - ...), the
//     updated column in cs and the chunk's part of (W(:,jp)^T v) in dpp for k_col_prep, so that
//     its one workgroup only touches a few vectors of the block's rows.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT_YR) void k_col_yred(SyevArgs<T> a, const int2 *items, int jl) {
  const int2 it = items[blockIdx.x];
  const int ent = it.x, p = ent >> COL_PANEL_SHIFT;
  const BlkDesc bd = a.bd[ent & COL_BLK_MASK];
  const int n = (int)bd.n;
  const int i0 = p * NB;
  const int m = n - i0;
  const T *A = a.A + bd.offA + (size_t)i0 * n + i0;
  T *W = a.Wp + bd.offWp;
  const T *yp = a.Yp + bd.offYp + i0;
  const T *dt = a.Yp + bd.offDt;
  const T *vb = a.vb + bd.offV + i0;
  const bool final = (m - 1 <= NB);
  const int w = final ? m - 1 : NB;
  const bool do_col = jl < w || final;
  const int jp = jl - 1;
  const T tv = a.tau[bd.offV + i0 + jp];
  const int Tt = (m - jl + 63) / 64;
  constexpr int NW = NT_YR / WAVE;
  __shared__ T part[NW][64];
  __shared__ T cwv[2 * NB], vrow[NB], wrow[NB];
  __shared__ T red[NW];
  const int tid = threadIdx.x;
  // The loads of each batch below are issued before any of them is used (loads whose values are
  // summed or stored to the LDS as they arrive would each wait for the one before); the sums keep
  // their order: segments sg = wave, wave + NW, ... into ps; slots of full 32-slot chunks alternately
  // into y and y2, those of the last partial chunk into y, then y += y2.
  constexpr int DB = 8, YB = 32;
  const int cl = tid & 63;
  T ps = 0, vr = 0, wr = 0;
  if (tid < NB) {
    vr = (do_col && tid < jl) ? A[(size_t)tid * n + jl] : (T)0;
    wr = (do_col && tid < jp) ? W[(size_t)tid * n + jl] : (T)0;
  }
  // the first batch of segment partials (segments sgw + u NW) before the product's loads
  const int sgw = tid >> 6;
  const bool dcol = (cl & 31) < jp;
  T d0[DB];
#pragma unroll
  for (int u = 0; u < DB; u++) d0[u] = dcol && sgw + u * NW < Tt ? dt[(sgw + u * NW) * 64 + cl] : (T)0;
  const int r = jl + it.y * NT_YR + tid;
  const bool rin = r < m;
  T y = 0, cc = 0, v = 0;
  if (rin) {
    v = vb[r];
    if (do_col) cc = A[(size_t)jl * n + r];
    T y2 = 0;
    for (int sl = 0; sl < Tt; sl += YB) {
      T ys[YB];
#pragma unroll
      for (int u = 0; u < YB; u++) ys[u] = sl + u < Tt ? yp[(size_t)(sl + u) * n + r] : (T)0;
      if (sl + YB <= Tt) {
#pragma unroll
        for (int u = 0; u < YB; u += 2) {
          y += ys[u];
          y2 += ys[u + 1];
        }
      } else {
#pragma unroll
        for (int u = 0; u < YB; u++)
          if (sl + u < Tt) y += ys[u];
      }
    }
    y += y2;
  }
  ps = d0[0];   // as 0 + d0[0] up to the sign of a zero, which the sum from +0 over the waves drops
  if (dcol) {
#pragma unroll
    for (int u = 1; u < DB; u++)
      if (sgw + u * NW < Tt) ps += d0[u];
    for (int sg0 = sgw + DB * NW; sg0 < Tt; sg0 += DB * NW) {
      T ds[DB];
#pragma unroll
      for (int u = 0; u < DB; u++) ds[u] = sg0 + u * NW < Tt ? dt[(sg0 + u * NW) * 64 + cl] : (T)0;
#pragma unroll
      for (int u = 0; u < DB; u++)
        if (sg0 + u * NW < Tt) ps += ds[u];
    }
  }
  part[tid >> 6][cl] = ps;
  if (tid < NB) {
    vrow[tid] = vr;
    wrow[tid] = wr;
  }
  __syncthreads();
  if (tid < 2 * NB) {
    T s = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) s += part[q][tid];
    cwv[tid] = s;
  }
  __syncthreads();
  T ay = 0, ac = 0;
  if (rin)
    for (int k0 = 0; k0 < jp; k0 += KCH) {
      T ak[KCH], wk[KCH];
#pragma unroll
      for (int k = 0; k < KCH; k++) {
        const bool in = k0 + k < jp;
        ak[k] = in ? A[(size_t)(k0 + k) * n + r] : (T)0;
        wk[k] = in ? W[(size_t)(k0 + k) * n + r] : (T)0;
      }
#pragma unroll
      for (int k = 0; k < KCH; k++) {
        const int kk = k0 + k < NB ? k0 + k : NB - 1;   // ak = wk = 0 past jp
        ay += ak[k] * cwv[kk] + wk[k] * cwv[NB + kk];
        ac += ak[k] * wrow[kk] + wk[k] * vrow[kk];
      }
    }
  const T yy = (rin && tv != (T)0) ? tv * (y - ay) : (T)0;
  if (rin) {
    W[(size_t)jp * n + r] = yy;
    if (do_col) a.cs[bd.offV + i0 + r] = cc - ac;
  }
  const T dps = block_sum<NT_YR>(yy * v, red);
  if (tid == 0) a.dpp[bd.offV + it.y] = dps;
}

// ------------------------------------------------------------------------------------------
// 2b. symmetric matrix-vector product y = A22 v of reduced column jl over 64x64 tiles of the
//     lower triangle of every block's trailing matrix. Tile (I,J), I >= J, writes A_IJ v_J into
//     slot J of row segment I and A_IJ^T v_I into slot I of row segment J: every slot of every
//     segment is written exactly once, so the sum k_col_prep forms is deterministic.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_col_symv(SyevArgs<T> a, const int *blist, const int *tpre, int nact, int jl) {
  // block of this workgroup: the largest i with tpre[i] <= blockIdx.x, i.e. the count of such entries
  // less one (tpre ascends from 0). Every wave counts them with ballots over 256 entries per round
  // trip (a binary search would chain log2(nact) dependent loads before the tile's own)
  const int g = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int cnt = 0;
  for (int base = 0; base < nact; base += 4 * WAVE) {
    int tv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = base + u * WAVE + lane;
      tv[u] = i < nact ? tpre[i] : 0x7fffffff;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) cnt += __popcll(__ballot(tv[u] <= g));
  }
  const int lo = cnt - 1;
  const int ent = blist[lo], p = ent >> COL_PANEL_SHIFT;
  const BlkDesc bd = a.bd[ent & COL_BLK_MASK];
  const int tl = g - tpre[lo];
  const int n = (int)bd.n;
  const int i0 = p * NB;
  const int m = n - i0;
  const int r0 = jl + 1;                          // trailing rows/cols [r0, m) of the panel
  const int t = m - r0;
  const int nseg = (t + 63) / 64, ntri = nseg * (nseg + 1) / 2;
  if (tl >= ntri) {
    // segment partials of W^T v and V^T v for the next column's k_col_prep (only columns k < jl
    // exist): wave 0/1 take W(:, 0..15 / 16..31), wave 2/3 take V(:, 0..15 / 16..31)
    const int I = tl - ntri;
    const int kb = (wv & 1) * KCH;
    if (kb >= jl) return;
    const int r = I * 64 + lane;
    const bool rin = r < t;
    const T *P = wv < 2 ? a.Wp + bd.offWp + r0 : a.A + bd.offA + (size_t)i0 * n + i0 + r0;
    const T vr = rin ? a.vb[bd.offV + i0 + r0 + r] : (T)0;
    T z[16];
#pragma unroll
    for (int q = 0; q < 16; q++) z[q] = (rin && kb + q < jl) ? P[(size_t)(kb + q) * n + r] : (T)0;
#pragma unroll
    for (int q = 0; q < 16; q++) z[q] *= vr;
    const T s = butterfly16(z);
    const int cl = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
    if ((lane & 3) == 0) a.Yp[bd.offDt + I * 64 + (wv >> 1) * 32 + kb + cl] = s;
    return;
  }
  int I = (int)((sqrt(8.0 * tl + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= tl) I++;
  while (I * (I + 1) / 2 > tl) I--;
  const int J = tl - I * (I + 1) / 2;
  const T *S = a.A + bd.offA + (size_t)(i0 + r0) * n + (i0 + r0);   // S(r,c) = S[c*n + r]
  const T *v = a.vb + bd.offV + i0 + r0;
  T *yp = a.Yp + bd.offYp + i0 + r0;             // yp[slot*n + r]
  __shared__ T vI[64], vJ[64];
  __shared__ T red[4][64];
  // the tile's loads first: they do not depend on the reflector, so their round trip overlaps the
  // staging of v below
  const int r = I * 64 + lane;
  const bool rin = r < t;
  T x[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int c = J * 64 + wv * 16 + q;
    x[q] = (rin && c < t) ? S[(size_t)c * n + r] : (T)0;
  }
  if (tid < 64) {
    const int rr = I * 64 + tid;
    vI[tid] = rr < t ? v[rr] : (T)0;
  } else if (tid < 128) {
    const int c = J * 64 + tid - 64;
    vJ[tid - 64] = c < t ? v[c] : (T)0;
  }
  __syncthreads();
  T yi = 0;
#pragma unroll
  for (int q = 0; q < 16; q++) yi += x[q] * vJ[wv * 16 + q];
  red[wv][lane] = yi;
  if (I != J) {
    const T vr = vI[lane];
#pragma unroll
    for (int q = 0; q < 16; q++) x[q] *= vr;
    const T yj = butterfly16(x);
    const int cl = ((lane >> 5) & 1) * 8 + ((lane >> 4) & 1) * 4 + ((lane >> 3) & 1) * 2 + ((lane >> 2) & 1);
    const int c = J * 64 + wv * 16 + cl;
    if ((lane & 3) == 0 && c < t) yp[(size_t)I * n + c] = yj;
  }
  __syncthreads();
  if (tid < 64 && rin) yp[(size_t)J * n + r] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
}

// ------------------------------------------------------------------------------------------
// 2d. one workgroup per block for trailing orders up to WGR_MAX: the blocked Householder reduction
//     (dlatrd, then the rank-2k update) of every panel that starts with more than TAIL rows, with
//     the matrix in global memory (a few MB: L2-resident). A thread owns a row of the trailing
//     matrix and keeps the row's V and W in registers; panels here are KW columns wide (the
//     column kernels' NB-column panels, as two halves: W is internal, the reflectors are stored as
//     the column kernels store them). Each column's symmetric product runs over 64x64 lower tiles
//     of the trailing matrix, one wave per tile, into per-wave partial sums in LDS (deterministic);
//     each panel's trailing update runs over the lower triangle, a column tile at a time with the
//     tile's V and W rows in LDS. It replaces, for these orders, the chain of about 33 launches per
//     panel of the column kernels, and blocks of order up to WGR_MAX run beside the larger blocks'
//     chains instead of in them. Opt-in (NRG_WG_RED): slower than the chains on MI355X.
// ------------------------------------------------------------------------------------------
// Raw buffer loads: one 32-bit per-lane offset register shared by many loads whose uniform parts
// sit in scalar registers. Callers keep every offset inside the buffer.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t rs, int voff, int soff) {
  if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}

constexpr int WGR_MAX = 512;          // largest trailing order of k_wg_red: one row per thread
constexpr int NW_WGR = WGR_MAX / WAVE;
constexpr int KW = 16;                // panel width of k_wg_red
static_assert(NB % KW == 0, "k_wg_red panels must tile the column kernels' panels");

template <typename T>
__global__ __launch_bounds__(WGR_MAX) void k_wg_red(SyevArgs<T> a, const int *blist) {
  const int ent = blist[blockIdx.x], p0 = ent >> COL_PANEL_SHIFT;
  const BlkDesc bd = a.bd[ent & COL_BLK_MASK];
  const int n = (int)bd.n;
  const int iend = n > TAIL ? (n - TAIL + NB - 1) / NB * NB : 0;   // k_tail starts at row iend
  // the wave index, in a scalar register: tile indices and load offsets derived from it are uniform
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = __builtin_amdgcn_readfirstlane(tid / WAVE);
  __shared__ T ypart[NW_WGR][WGR_MAX];   // per-wave partial sums of A22 v
  __shared__ T vsh[WGR_MAX + 64];        // the current reflector (panel-relative rows), zero padded
  __shared__ T rowj[2 * KW];             // V(jl, :) and W(jl, :)
  __shared__ T dots[NW_WGR][2 * KW];     // per-wave partials of W^T v and V^T v
  __shared__ T cwv[2 * KW];              // W^T v (0..KW) and V^T v (KW..2KW)
  __shared__ T red[NW_WGR];
  __shared__ T sh[2];
  __shared__ T tV[64 * KW], tW[64 * KW]; // V and W rows of a column tile of the trailing update
  const int r = tid;                     // this thread's row of the panel
  if (tid < 64) vsh[WGR_MAX + tid] = 0;
  for (int i0 = p0 * NB; i0 < iend; i0 += KW) {
    const int m = n - i0;                // TAIL < m <= WGR_MAX
    T *A = a.A + bd.offA + (size_t)i0 * n + i0;   // A(r,c) = A[c*n + r], relative to the panel
    T *W = a.Wp + bd.offWp;                        // W(r,k) = W[k*n + r]
    double *d = a.d + bd.offV + i0, *e = a.e + bd.offV + i0;
    T *tau = a.tau + bd.offV + i0;
    const bool rin = r < m;
    T vrow[KW], wrow[KW];   // V(r, k), W(r, k) of the panel's columns done so far (zero beyond)
#pragma unroll
    for (int k = 0; k < KW; k++) vrow[k] = wrow[k] = 0;
#pragma unroll 1
    for (int jl = 0; jl < KW; jl++) {
      // 1. column jl minus the panel's earlier reflectors: A(:,jl) -= V W(jl,:)^T + W V(jl,:)^T
      if (r == jl) {
#pragma unroll
        for (int k = 0; k < KW; k++) {
          rowj[k] = vrow[k];
          rowj[KW + k] = wrow[k];
        }
      }
      __syncthreads();
      T c = 0;
      if (rin && r >= jl) {
        c = A[(size_t)jl * n + r];
#pragma unroll
        for (int k = 0; k < KW; k++) c -= vrow[k] * rowj[KW + k] + wrow[k] * rowj[k];
      }
      // 2. reflector annihilating A(jl+2:m, jl) (dlarfg)
      if (r == jl + 1) sh[0] = c;
      const T xn2 = block_sum<WGR_MAX>((rin && r >= jl + 2) ? c * c : (T)0, red);
      const T alpha = sh[0];
      T beta, tv, scal = 1;
      if (xn2 == (T)0) {
        tv = 0;
        beta = alpha;
      } else {
        beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
        tv = (beta - alpha) / beta;
        scal = (T)1 / (alpha - beta);
      }
      T v = 0;
      if (rin) {
        if (r == jl) {
          A[(size_t)jl * n + r] = c;
          d[jl] = (double)c;
          e[jl] = (double)beta;
          tau[jl] = tv;
        } else if (r == jl + 1) {
          v = 1;
          A[(size_t)jl * n + r] = v;
        } else if (r > jl + 1) {
          v = c * scal;
          A[(size_t)jl * n + r] = v;
        }
      }
      vsh[r] = v;
#pragma unroll
      for (int k = 0; k < KW; k++) vrow[k] = k == jl ? v : vrow[k];
      for (int idx = tid; idx < NW_WGR * WGR_MAX; idx += WGR_MAX) (&ypart[0][0])[idx] = 0;
      __syncthreads();
      // 3. y = A22 v over the trailing rows [r0, m): lower 64x64 tiles, one wave per tile
      const int r0 = jl + 1, t = m - r0;
      const int nseg = (t + 63) / 64, ntri = nseg * (nseg + 1) / 2;
      const T *S = A + (size_t)r0 * n + r0;   // S(r,c) = S[c*n + r]
      // every load below reads an element of S(0:t, 0:t) (indices clamped), so stays in the block
      const __amdgpu_buffer_rsrc_t rs = buf_rsrc(S, (t - 1) * n * (int)sizeof(T) + t * (int)sizeof(T));
      for (int tl = wv; tl < ntri; tl += NW_WGR) {
        int I = (int)((sqrtf(8.0f * tl + 1.0f) - 1.0f) * 0.5f);
        while ((I + 1) * (I + 2) / 2 <= tl) I++;
        while (I * (I + 1) / 2 > tl) I--;
        const int J = tl - I * (I + 1) / 2;
        const int rr = I * 64 + lane;
        const bool rok = rr < t;
        const int voff = min(rr, t - 1) * (int)sizeof(T);
        const T vI = rok ? vsh[r0 + rr] : (T)0;
        T yrow = 0;
#pragma unroll 1
        for (int qb = 0; qb < 64; qb += 16) {
          T x[16];
#pragma unroll
          for (int q = 0; q < 16; q++) {
            const int cc = min(J * 64 + qb + q, t - 1);
            x[q] = buf_load<T>(rs, voff, cc * n * (int)sizeof(T));
          }
#pragma unroll
          for (int q = 0; q < 16; q++) {
            const int cc = J * 64 + qb + q;
            const bool in = rok && cc < t && (I > J || rr >= cc);
            x[q] = in ? x[q] : (T)0;
            yrow += x[q] * vsh[r0 + cc];   // zero past the trailing rows
            x[q] = (I == J && rr == cc) ? (T)0 : x[q] * vI;   // the transpose, without the diagonal
          }
          const T sz = butterfly16(x);   // lane l: column qb + ((l >> 2) & 15)
          const int cc = J * 64 + qb + ((lane >> 2) & 15);
          if ((lane & 3) == 0 && cc < t) ypart[wv][cc] += sz;
        }
        if (rok) ypart[wv][rr] += yrow;
      }
      __syncthreads();
      T y = 0;
      const bool rt = rin && r >= r0;
      if (rt) {
#pragma unroll
        for (int w = 0; w < NW_WGR; w++) y += ypart[w][r - r0];
      }
      // 4. W(:,k)^T v and V(:,k)^T v, k < jl, over the trailing rows
#pragma unroll 1
      for (int g = 0; g < 2; g++) {
        T z[KW];
#pragma unroll
        for (int k = 0; k < KW; k++) z[k] = (rt && k < jl) ? (g == 0 ? wrow[k] : vrow[k]) * v : (T)0;
        const T sz = butterfly16(z);
        if ((lane & 3) == 0) dots[wv][g * KW + ((lane >> 2) & 15)] = sz;
      }
      __syncthreads();
      if (tid < 2 * KW) {
        T sdot = 0;
#pragma unroll
        for (int w = 0; w < NW_WGR; w++) sdot += dots[w][tid];
        cwv[tid] = sdot;
      }
      __syncthreads();
      // 5. W(:,jl) = tau (y - V (W^T v) - W (V^T v)) - tau/2 (w^T v) v
      T yy = 0;
      if (rt && tv != (T)0) {
        T ay = 0;
#pragma unroll
        for (int k = 0; k < KW; k++) ay += vrow[k] * cwv[k] + wrow[k] * cwv[KW + k];
        yy = tv * (y - ay);
      }
      const T a2 = (T)(-0.5) * tv * block_sum<WGR_MAX>(yy * v, red);
      const T wv_r = rt ? yy + a2 * v : (T)0;
      if (rt) W[(size_t)jl * n + r] = wv_r;
#pragma unroll
      for (int k = 0; k < KW; k++) wrow[k] = k == jl ? wv_r : wrow[k];
    }
    // 6. trailing update of the lower triangle: A(KW:m, KW:m) -= V W^T + W V^T, a 64-column tile at a
    //    time; thread r updates A(r, c) for the tile's columns c <= r
    const int t = m - KW, nct = (t + 63) / 64;
    const __amdgpu_buffer_rsrc_t rsA = buf_rsrc(A, (m - 1) * n * (int)sizeof(T) + m * (int)sizeof(T));
    for (int J = 0; J < nct; J++) {
      __syncthreads();   // the previous tile's rows are no longer read
      for (int idx = tid; idx < 64 * KW; idx += WGR_MAX) {
        const int cl = idx % 64, k = idx / 64, cr = KW + J * 64 + cl;
        tV[cl * KW + k] = cr < m ? A[(size_t)k * n + cr] : (T)0;   // V(cr, k), cr >= KW > k
        tW[cl * KW + k] = cr < m ? W[(size_t)k * n + cr] : (T)0;
      }
      __syncthreads();
      if (rin && r >= KW + J * 64) {
        // columns [0, cmax) of the tile, 16 loads at a time; column indices are clamped to the tile
        // (uniform), so a lane whose row ends earlier reads elements above the diagonal it leaves alone
        const int tw = min(64, t - J * 64), cmax = min(tw, r - KW - J * 64 + 1);
        const int cb = KW + J * 64;
        for (int c0 = 0; c0 < cmax; c0 += 16) {
          T x[16];
#pragma unroll
          for (int q = 0; q < 16; q++) x[q] = buf_load<T>(rsA, r * (int)sizeof(T), (cb + min(c0 + q, tw - 1)) * n * (int)sizeof(T));
#pragma unroll
          for (int q = 0; q < 16; q++) {
            const int cl = c0 + q;
#pragma unroll
            for (int k = 0; k < KW; k++) x[q] -= vrow[k] * tW[cl * KW + k] + wrow[k] * tV[cl * KW + k];
            if (cl < cmax) A[(size_t)(cb + cl) * n + r] = x[q];
          }
        }
      }
    }
    __syncthreads();   // the next panel reads the updated matrix
  }
}

#include "syev_sbr.hpp"

// ------------------------------------------------------------------------------------------
// 2c. tail of the reduction: the trailing matrix of order m <= TAIL left after the last panel
//     (the whole block when n <= TAIL) is reduced column by column (dsytd2, lower) in LDS by one
//     workgroup, with the reflectors, tau, d and e stored as the column kernels store them.
// ------------------------------------------------------------------------------------------
constexpr int NT_TAIL = NT_PANEL;   // threads of k_tail (256 took 0.68 ms per call against 0.49)
template <typename T>
__global__ __launch_bounds__(NT_TAIL) void k_tail(SyevArgs<T> a, const int *blist) {
  const int ent = blist[blockIdx.x], p = ent >> COL_PANEL_SHIFT;
  const BlkDesc bd = a.bd[ent & COL_BLK_MASK];
  const int n = (int)bd.n;
  const int i0 = p * NB;
  const int m = n - i0;
  T *A = a.A + bd.offA + (size_t)i0 * n + i0;   // A(r,c) = A[c*n + r]
  double *d = a.d + bd.offV + i0;
  double *e = a.e + bd.offV + i0;
  T *tau = a.tau + bd.offV + i0;
  constexpr int LD = TAIL + 1;
  constexpr int NW = NT_TAIL / WAVE;
  __shared__ T S[TAIL * LD];        // S(r,c) at S[c*LD + r], full symmetric
  __shared__ T vv[TAIL], yy[TAIL];
  __shared__ T red[NW];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < m * m; idx += NT_TAIL) {
    const int r = idx % m, c = idx / m;
    if (r >= c) {
      const T x = A[(size_t)c * n + r];
      S[c * LD + r] = x;
      S[r * LD + c] = x;
    }
  }
  __syncthreads();
  for (int j = 0; j + 1 < m; j++) {
    // reflector annihilating S(j+2:m, j) (dlarfg)
    const int r1 = j + 1;
    T part = 0;
    for (int r = r1 + 1 + tid; r < m; r += NT_TAIL) part += S[j * LD + r] * S[j * LD + r];
    const T xn2 = block_sum<NT_TAIL, false>(part, red);   // the barrier after the reflector follows
    const T alpha = S[j * LD + r1];
    T beta, tv, scal = 1;
    if (xn2 == (T)0) {
      tv = 0;
      beta = alpha;
    } else {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      tv = (beta - alpha) / beta;
      scal = (T)1 / (alpha - beta);
    }
    for (int r = r1 + tid; r < m; r += NT_TAIL) {
      const T x = r == r1 ? (T)1 : S[j * LD + r] * scal;
      vv[r] = x;
      A[(size_t)j * n + r] = x;
    }
    if (tid == 0) {
      d[j] = (double)S[j * LD + j];
      e[j] = (double)beta;
      tau[j] = tv;
    }
    __syncthreads();
    if (tv != (T)0) {
      // y = tau S22 v over the rows [r1, m): 16 threads per row
      const int mt = m - r1;
      T dp = 0;
      for (int base = 0; base < mt; base += NT_TAIL / 16) {
        const int rr = base + tid / 16, sub = tid % 16;
        T y = 0;
        if (rr < mt)
          for (int c = sub; c < mt; c += 16) y += S[(r1 + c) * LD + r1 + rr] * vv[r1 + c];
        y = row16_sum(y);   // the 16 threads of a row are a row of 16 lanes
        if (rr < mt && sub == 0) {
          y *= tv;
          yy[r1 + rr] = y;
          dp += y * vv[r1 + rr];
        }
      }
      const T a2 = (T)(-0.5) * tv * block_sum<NT_TAIL, false>(dp, red);   // the barrier after the update follows
      // w = y + a2 v; S22 -= v w^T + w v^T
      for (int idx = tid; idx < mt * mt; idx += NT_TAIL) {
        const int rr = r1 + idx % mt, cc = r1 + idx / mt;
        const T wr = yy[rr] + a2 * vv[rr], wc = yy[cc] + a2 * vv[cc];
        S[cc * LD + rr] -= vv[rr] * wc + wr * vv[cc];
      }
      __syncthreads();
    }
  }
  if (tid == 0 && m >= 1) d[m - 1] = (double)S[(m - 1) * LD + (m - 1)];
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
// 3b. divide and conquer merges. Each merge of two sorted spectra with the rank-one tear
//     rho z z^T (dlaed1) runs as five launches over all merges of one tree depth:
//       k_dc_deflate   merge order, deflation (dlaed2), one workgroup per merge
//       k_dc_roots     secular equation roots (dlaed4), 16 lanes per root, 16 roots per workgroup
//       k_dc_zhat      Loewner recomputation of z (Gu-Eisenstat, dlaed3)
//       k_dc_normrank  eigenvector norms and the ascending order of all eigenvalues
//       k_dc_fill      eigenvector matrix S of the merge (+ undoing the deflation rotations),
//                      and Qout = diag(Q1, Q2) S for small merges
// ------------------------------------------------------------------------------------------
struct MergeState {
  int k, nrot;
  double rho;
};
enum { DS_DS = 0, DS_ZS, DS_DCAT, DS_ZCAT, DS_DND, DS_ZND, DS_ZHAT, DS_RT, DS_DORG, DS_INVN, DS_RC, DS_RS, DS_COUNT };
enum { IS_PERM = 0, IS_PINV, IS_CMP, IS_CPOS, IS_COLSRC, IS_RPJ, IS_RNJ, IS_RO, IS_COUNT };

struct NodeGeo {
  int a0, mid, a1, m1, m2, m;
};
__device__ __forceinline__ NodeGeo node_geo(int n, int depth, int q) {
  NodeGeo g;
  g.a0 = (int)(((int64_t)q * n) >> depth);
  g.mid = (int)(((int64_t)(2 * q + 1) * n) >> (depth + 1));
  g.a1 = (int)(((int64_t)(q + 1) * n) >> depth);
  g.m1 = g.mid - g.a0;
  g.m2 = g.a1 - g.mid;
  g.m = g.m1 + g.m2;
  return g;
}
__device__ __forceinline__ double *dsa(const SyevArgs<double> &a, const BlkDesc &bd, int which, int a0) {
  return a.dscr + (size_t)which * a.Ntot + bd.offV + a0;
}
__device__ __forceinline__ int *isa(const SyevArgs<double> &a, const BlkDesc &bd, int which, int a0) {
  return a.iscr + (size_t)which * a.Ntot + bd.offV + a0;
}

// Root j of 1 + rho sum_i z_i^2 / (d_i - lambda) = 0, d ascending, z nonzero, rho > 0.
// The origin is the nearer pole o (lambda = d_o + tau); the first iterate solves the model with
// the two adjacent poles exact and the rest frozen; then Newton on the function with both
// adjacent poles multiplied out, safeguarded by bisection.
// A group of SL consecutive lanes solves one root: every sum over the poles is split across the
// group (lane `sub` takes i = sub mod SL) and combined pairwise within the row (row16_sum), which
// gives every lane of the group the same bits (the pairwise additions commute), so the group's lanes
// take the same iterations and branches.
constexpr int SL = 16;   // lanes per secular root: a row of 16 lanes
__device__ __forceinline__ double group_sum(double v) { return row16_sum(v); }

__device__ void secular_root(int j, int k, const double *d, const double *z, double rho, double zz, int sub, int &org,
                             double &tau_out) {
  const double EPS = 2.220446049250313e-16;
  int o, o2;
  double lo, hi, t, za, zb = 0.0, d2 = 0.0;
  if (j < k - 1) {
    const double g = d[j + 1] - d[j], mid = 0.5 * g;
    double R0 = 0.0;
    for (int i = sub; i < k; i += SL) {
      if (i == j || i == j + 1) continue;
      R0 += rho * z[i] * z[i] * rcp_nr((d[i] - d[j]) - mid);
    }
    R0 = group_sum(R0);
    const double f = 1.0 + R0 - rho * z[j] * z[j] / mid + rho * z[j + 1] * z[j + 1] / (g - mid);
    if (f >= 0) { o = j; o2 = j + 1; lo = 0; hi = mid; }
    else { o = j + 1; o2 = j; lo = -mid; hi = 0; }
    d2 = d[o2] - d[o];
    za = rho * z[o] * z[o];
    zb = rho * z[o2] * z[o2];
    t = 0.5 * (lo + hi);
    const double c = 1.0 + R0, B = c * d2 + za + zb, disc = B * B - 4.0 * c * za * d2;
    if (disc >= 0) {
      const double q = 0.5 * (B + copysign(sqrt(disc), B));
      const double t1 = c != 0 ? q / c : 0.0, t2 = q != 0 ? za * d2 / q : 0.0;
      if (t1 >= lo && t1 <= hi && t1 != 0.0) t = t1;
      else if (t2 >= lo && t2 <= hi && t2 != 0.0) t = t2;
    }
  } else {
    o = k - 1; o2 = -1; lo = 0; hi = rho * zz;
    za = rho * z[o] * z[o];
    const double t0 = 0.5 * hi;
    double R0 = 0.0;
    for (int i = sub; i < k - 1; i += SL) R0 += rho * z[i] * z[i] * rcp_nr((d[i] - d[o]) - t0);
    R0 = group_sum(R0);
    t = t0;
    if (1.0 + R0 > 0) {
      const double tc = za / (1.0 + R0);
      if (tc >= lo && tc <= hi && tc != 0.0) t = tc;
    }
  }
  const double dorg = d[o];
  for (int it = 0; it < 100; it++) {
    // the two adjacent poles are masked out (their denominators -t and d2 - t are nonzero);
    // independent divisions in flight hide their latency
    double R = 0, Rp = 0, Ra = 0, R2 = 0, Rp2 = 0, Ra2 = 0;
    int i = sub;
    for (; i + SL < k; i += 2 * SL) {
      const double q0 = rcp_nr((d[i] - dorg) - t), q1 = rcp_nr((d[i + SL] - dorg) - t);
      const double t0 = (i == o || i == o2) ? 0.0 : rho * z[i] * z[i] * q0;
      const double t1 = (i + SL == o || i + SL == o2) ? 0.0 : rho * z[i + SL] * z[i + SL] * q1;
      R += t0; Rp += t0 * q0; Ra += fabs(t0);
      R2 += t1; Rp2 += t1 * q1; Ra2 += fabs(t1);
    }
    if (i < k) {
      const double q0 = rcp_nr((d[i] - dorg) - t);
      const double t0 = (i == o || i == o2) ? 0.0 : rho * z[i] * z[i] * q0;
      R += t0; Rp += t0 * q0; Ra += fabs(t0);
    }
    R = group_sum(R + R2);
    Rp = group_sum(Rp + Rp2);
    Ra = group_sum(Ra + Ra2);
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

constexpr int NT_DEF = 512;
// dynamic LDS of k_dc_deflate for merges of order <= mc: d and z (doubles), predecessor (int) and
// flags (byte) per index
inline int dc_deflate_lds(int mc) { return (int)(2 * sizeof(double) * (size_t)mc + 5 * (size_t)mc + 64); }
// dynamic LDS a merge kernel may ask for: the 160 KiB of a workgroup less room for static arrays
constexpr int DYN_LDS_MAX = 160 * 1024 - 8 * 1024;

// G: merges too large for LDS keep d, z, predecessors and flags at gscr + merge * gstride (global
// memory, one region per workgroup) instead of in dynamic LDS, through generic pointers. With the
// scratch in LDS (G = false) the compiler sees LDS loads and global stores (dcg pointers), so the
// serial scan's LDS loads do not wait for its stores of the merge's lists.
template <bool G, typename T>
using dcg = std::conditional_t<G, T *, __attribute__((address_space(1))) T *>;
template <bool G>
__global__ __launch_bounds__(NT_DEF) void k_dc_deflate(SyevArgs<double> a, const int2 *items, int depth,
                                                       MergeState *mst, double *gscr, int64_t gstride) {
  extern __shared__ double dyn_lds[];
  double *dyn = G ? (gscr ? gscr + (size_t)blockIdx.x * gstride : dyn_lds) : dyn_lds;
  __shared__ double red[NT_DEF / WAVE];
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n, tid = threadIdx.x;
  const NodeGeo g = node_geo(n, depth, it.y);
  const int m = g.m, m1 = g.m1, m2 = g.m2, a0 = g.a0, mid = g.mid;
  double *dsl = dyn, *zsl = dyn + m;
  const size_t ld = n;
  const double *Qin = ((depth + 1) & 1 ? a.Z1 : a.Z0) + bd.offZ;
  const double *lam = a.lam + bd.offV;
  const double beta = (a.e + bd.offV)[mid - 1];
  const double sgn = beta >= 0 ? 1.0 : -1.0;
  const double rho = 2.0 * fabs(beta);
  const double r2 = 0.70710678118654752440;
  using GD = dcg<G, double>;
  using GI = dcg<G, int>;
  GD dcat = (GD)dsa(a, bd, DS_DCAT, a0), zcat = (GD)dsa(a, bd, DS_ZCAT, a0);
  GD ds = (GD)dsa(a, bd, DS_DS, a0);
  GD dnd = (GD)dsa(a, bd, DS_DND, a0), znd = (GD)dsa(a, bd, DS_ZND, a0);
  GD rc = (GD)dsa(a, bd, DS_RC, a0), rs = (GD)dsa(a, bd, DS_RS, a0);
  GI perm = (GI)isa(a, bd, IS_PERM, a0), pinv = (GI)isa(a, bd, IS_PINV, a0);
  GI cmp = (GI)isa(a, bd, IS_CMP, a0), cpos = (GI)isa(a, bd, IS_CPOS, a0);
  GI rpj = (GI)isa(a, bd, IS_RPJ, a0), rnj = (GI)isa(a, bd, IS_RNJ, a0);
  // the two sorted spectra and z = (last row of Q1, +-first row of Q2) / sqrt(2)
  for (int i = tid; i < m; i += NT_DEF) {
    dcat[i] = lam[a0 + i];
    zcat[i] = (i < m1 ? Qin[(size_t)(a0 + i) * ld + (mid - 1)] : sgn * Qin[(size_t)(a0 + i) * ld + mid]) * r2;
  }
  __syncthreads();
  double mx = 0;
  for (int i = tid; i < m; i += NT_DEF) {
    int pos;
    if (i < m1) pos = i + lower_bound_d((const double *)(dcat + m1), m2, dcat[i]);
    else pos = (i - m1) + upper_bound_d((const double *)dcat, m1, dcat[i]);
    pos = clampi(pos, m);
    perm[pos] = i;
    pinv[i] = pos;
    dsl[pos] = dcat[i];
    zsl[pos] = zcat[i];
    mx = fmax(mx, fmax(fabs(dcat[i]), fabs(zcat[i])));
  }
  const double dzmax = block_max<NT_DEF>(mx, red);  // includes a barrier
  const double EPS = 2.220446049250313e-16;
  const double tol = 8.0 * EPS * dzmax;
  // Parallel part of the deflation scan (dlaed2): which z are negligible, the previous
  // non-negligible index of every index, and the Givens deflation test of every index against that
  // predecessor with the unrotated values. The serial scan below then only repeats the test after a
  // rotation has changed its predecessor, and computes the rotation itself when the test passes.
  int *pv = (int *)(dyn + 2 * m);
  unsigned char *flg = (unsigned char *)(pv + m);   // bit 0: negligible z, bit 1: the test deflates
  __shared__ int carry[NT_DEF];
  __shared__ int wlast[NT_DEF / WAVE];
  double zm = 0;
  for (int p = tid; p < m; p += NT_DEF) {
    zm = fmax(zm, fabs(zsl[p]));
    flg[p] = rho * fabs(zsl[p]) <= tol ? 1 : 0;
  }
  const double zmax = block_max<NT_DEF>(zm, red);
  const int ch = (m + NT_DEF - 1) / NT_DEF, c0 = tid * ch, c1 = min(m, c0 + ch);
  int last = -1;
  for (int p = c0; p < c1; p++)
    if (!(flg[p] & 1)) last = p;
  {   // inclusive max-scan of `last` over the threads
    const int lane = tid & (WAVE - 1), w = tid / WAVE;
    int v = last;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int y = __shfl_up(v, o, WAVE);
      if (lane >= o) v = max(v, y);
    }
    if (lane == WAVE - 1) wlast[w] = v;
    __syncthreads();
    for (int q = 0; q < w; q++) v = max(v, wlast[q]);
    carry[tid] = v;
    __syncthreads();
  }
  int cur = tid > 0 ? carry[tid - 1] : -1;
  for (int p = c0; p < c1; p++) {
    pv[p] = cur;
    if (!(flg[p] & 1)) cur = p;
  }
  for (int p = c0; p < c1; p++) {
    const int q = pv[p];
    if ((flg[p] & 1) || q < 0) continue;
    double s = zsl[q], c = zsl[p];
    const double tau = hypot(c, s), t = dsl[p] - dsl[q];
    c /= tau; s = -s / tau;
    if (fabs(t * c * s) <= tol) flg[p] |= 2;
  }
  __syncthreads();
  if (tid == 0) {
    int k = 0, nd = 0, nrot = 0;
    if (rho * zmax <= tol) {
      for (int p = 0; p < m; p++) { cmp[p] = -1; cpos[m - 1 - nd] = p; nd++; }
    } else {
      int pj = -1;
      bool rotated = false;   // pj's values were changed by the last rotation
      for (int nj = 0; nj < m; nj++) {
        if (flg[nj] & 1) { cmp[nj] = -1; cpos[m - 1 - nd] = nj; nd++; continue; }
        if (pj < 0) { pj = nj; continue; }
        bool defl = (flg[nj] & 2) != 0;
        double c = 0, s = 0, tau = 0;
        if (rotated || defl) {
          s = zsl[pj]; c = zsl[nj];
          tau = hypot(c, s);
          const double t = dsl[nj] - dsl[pj];
          c /= tau; s = -s / tau;
          defl = fabs(t * c * s) <= tol;
        }
        if (defl) {
          zsl[nj] = tau; zsl[pj] = 0.0;
          rpj[nrot] = pj; rnj[nrot] = nj; rc[nrot] = c; rs[nrot] = s; nrot++;
          const double dp = dsl[pj], dn = dsl[nj];
          dsl[pj] = dp * c * c + dn * s * s;
          dsl[nj] = dp * s * s + dn * c * c;
          cmp[pj] = -1; cpos[m - 1 - nd] = pj; nd++;
          rotated = true;
        } else {
          cmp[pj] = k; cpos[k] = pj; dnd[k] = dsl[pj]; znd[k] = zsl[pj]; k++;
          rotated = false;
        }
        pj = nj;
      }
      if (pj >= 0) { cmp[pj] = k; cpos[k] = pj; dnd[k] = dsl[pj]; znd[k] = zsl[pj]; k++; }
    }
    mst[blockIdx.x].k = k;
    mst[blockIdx.x].nrot = nrot;
    mst[blockIdx.x].rho = rho;
  }
  __syncthreads();
  for (int p = tid; p < m; p += NT_DEF) ds[p] = dsl[p];
}

// chunk items: .x = merge index within the depth's merge list, .y = chunk of RCH roots;
// a 256-thread workgroup solves RCH = 256 / SL roots
constexpr int RCH = 256 / SL;
__global__ __launch_bounds__(256) void k_dc_roots(SyevArgs<double> a, const int2 *mitems, const int2 *citems,
                                                  int depth, const MergeState *mst) {
  extern __shared__ double dyn[];
  __shared__ double red[256 / WAVE];
  const int2 ci = citems[blockIdx.x];
  const int2 it = mitems[ci.x];
  const BlkDesc bd = a.bd[it.x];
  const NodeGeo g = node_geo((int)bd.n, depth, it.y);
  const MergeState ms = mst[ci.x];
  const int k = ms.k;
  if (ci.y * RCH >= k) return;
  double *dl = dyn, *zl = dyn + g.m;
  const double *dnd = dsa(a, bd, DS_DND, g.a0), *znd = dsa(a, bd, DS_ZND, g.a0);
  double zz = 0;
  for (int i = threadIdx.x; i < k; i += 256) {
    dl[i] = dnd[i];
    zl[i] = znd[i];
    zz += znd[i] * znd[i];
  }
  zz = block_sum<256>(zz, red);
  const int j = ci.y * RCH + threadIdx.x / SL, sub = threadIdx.x % SL;
  if (j < k) {   // uniform within each group of SL lanes
    int o; double t;
    secular_root(j, k, dl, zl, ms.rho, zz, sub, o, t);
    if (sub == 0) {
      isa(a, bd, IS_RO, g.a0)[j] = o;
      dsa(a, bd, DS_RT, g.a0)[j] = t;
      dsa(a, bd, DS_DORG, g.a0)[j] = dl[o];
    }
  }
}

// zhat_i^2 = prod_j (lambda_j - d_i) / prod_{j != i} (d_j - d_i), formed from d_i - lambda_j computed
// relative to each root's origin
// LDS: stage the merge's d, origins and offsets in dynamic LDS (merges that fit), else read them from
// the merge's global arrays
template <bool LDS>
__global__ __launch_bounds__(256) void k_dc_zhat(SyevArgs<double> a, const int2 *mitems, const int2 *citems, int depth,
                                                 const MergeState *mst) {
  extern __shared__ double dyn[];
  const int2 ci = citems[blockIdx.x];
  const int2 it = mitems[ci.x];
  const BlkDesc bd = a.bd[it.x];
  const NodeGeo g = node_geo((int)bd.n, depth, it.y);
  const int k = mst[ci.x].k;
  if (ci.y * 256 >= k) return;
  const double *dnd = dsa(a, bd, DS_DND, g.a0), *dorg = dsa(a, bd, DS_DORG, g.a0), *rt = dsa(a, bd, DS_RT, g.a0);
  const double *dl = dnd, *ol = dorg, *tl = rt;
  if constexpr (LDS) {
    double *dls = dyn, *ols = dyn + g.m, *tls = dyn + 2 * g.m;
    for (int i = threadIdx.x; i < k; i += 256) {
      dls[i] = dnd[i];
      ols[i] = dorg[i];
      tls[i] = rt[i];
    }
    __syncthreads();
    dl = dls;
    ol = ols;
    tl = tls;
  }
  const int i = ci.y * 256 + threadIdx.x;
  if (i < k) {
    const double di = dl[i];
    double W = (di - ol[i]) - tl[i];
    for (int j = 0; j < k; j++) {
      if (j == i) continue;
      W *= ((di - ol[j]) - tl[j]) * rcp_nr(di - dl[j]);
    }
    dsa(a, bd, DS_ZHAT, g.a0)[i] = copysign(sqrt(fabs(W)), dsa(a, bd, DS_ZND, g.a0)[i]);
  }
}

// eigenvector norms of the roots, and ranks of all m eigenvalues of the merge
// LDS as for k_dc_zhat; without it the eigenvalue of index v is formed where it is read
template <bool LDS>
__global__ __launch_bounds__(256) void k_dc_normrank(SyevArgs<double> a, const int2 *mitems, const int2 *citems,
                                                     int depth, const MergeState *mst) {
  extern __shared__ double dyn[];
  const int2 ci = citems[blockIdx.x];
  const int2 it = mitems[ci.x];
  const BlkDesc bd = a.bd[it.x];
  const NodeGeo g = node_geo((int)bd.n, depth, it.y);
  const int m = g.m, k = mst[ci.x].k;
  if (ci.y * 256 >= m) return;
  const double *dnd = dsa(a, bd, DS_DND, g.a0), *zhat = dsa(a, bd, DS_ZHAT, g.a0);
  const double *dorg = dsa(a, bd, DS_DORG, g.a0), *rt = dsa(a, bd, DS_RT, g.a0);
  const double *ds = dsa(a, bd, DS_DS, g.a0);
  const int *cpos = isa(a, bd, IS_CPOS, g.a0);
  auto value = [&](int i) { return i < k ? dorg[i] + rt[i] : ds[clampi(cpos[i], m)]; };
  const double *dl = dnd, *zl = zhat;
  double *vl = nullptr;
  if constexpr (LDS) {
    double *dls = dyn, *zls = dyn + m;
    vl = dyn + 2 * m;
    for (int i = threadIdx.x; i < m; i += 256) {
      if (i < k) {
        dls[i] = dnd[i];
        zls[i] = zhat[i];
      }
      vl[i] = value(i);
    }
    __syncthreads();
    dl = dls;
    zl = zls;
  }
  const int u = ci.y * 256 + threadIdx.x;
  if (u < m) {
    const double x = LDS ? vl[u] : value(u), kx = okey(x);
    int r = 0;
    for (int v = 0; v < m; v++) {
      const double kv = okey(LDS ? vl[v] : value(v));
      r += (kv < kx) || (kv == kx && v < u);
    }
    isa(a, bd, IS_COLSRC, g.a0)[r] = u;
    a.lam[bd.offV + g.a0 + r] = x;
    if (u < k) {
      const double dj = dorg[u], tj = rt[u];
      double s = 0;
      for (int i = 0; i < k; i++) {
        const double v = zl[i] * rcp_nr((dl[i] - dj) - tj);
        s += v * v;
      }
      dsa(a, bd, DS_INVN, g.a0)[u] = 1.0 / sqrt(s);
    }
  }
}

constexpr int FILLC = 32;  // columns of S per fill workgroup

// STAGE: the rows' positions pinv[r] and cmp[pinv[r]] and the chunk's column data go to LDS first, so
// the element loop reads no global memory (merges whose arrays fit: fill_lds)
inline int fill_lds(int mc) { return (int)(2 * sizeof(double) * (size_t)mc + 2 * sizeof(int) * (size_t)mc + 16); }
template <bool STAGE>
__global__ __launch_bounds__(256) void k_dc_fill(SyevArgs<double> a, const int2 *mitems, const int2 *citems, int depth,
                                                 const MergeState *mst) {
  extern __shared__ double dyn[];
  const int2 ci = citems[blockIdx.x];
  const int2 it = mitems[ci.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n;
  const NodeGeo g = node_geo(n, depth, it.y);
  const int m = g.m, m1 = g.m1, a0 = g.a0;
  const MergeState ms = mst[ci.x];
  const int k = ms.k, nrot = ms.nrot;
  const int c0 = ci.y * FILLC;
  if (c0 >= m) return;
  const int nc = min(FILLC, m - c0);
  const size_t ld = n;
  double *dl = dyn, *zl = dyn + m;
  const double *dnd = dsa(a, bd, DS_DND, a0), *zhat = dsa(a, bd, DS_ZHAT, a0);
  const double *dorg = dsa(a, bd, DS_DORG, a0), *rt = dsa(a, bd, DS_RT, a0), *invn = dsa(a, bd, DS_INVN, a0);
  const double *rc = dsa(a, bd, DS_RC, a0), *rs = dsa(a, bd, DS_RS, a0);
  const int *perm = isa(a, bd, IS_PERM, a0), *pinv = isa(a, bd, IS_PINV, a0), *cmp = isa(a, bd, IS_CMP, a0);
  const int *cpos = isa(a, bd, IS_CPOS, a0), *colsrc = isa(a, bd, IS_COLSRC, a0);
  const int *rpj = isa(a, bd, IS_RPJ, a0), *rnj = isa(a, bd, IS_RNJ, a0);
  double *S = a.S + bd.offZ;
  for (int i = threadIdx.x; i < k; i += 256) {
    dl[i] = dnd[i];
    zl[i] = zhat[i];
  }
  if (STAGE) {
    int *pr = (int *)(zl + m), *cixr = pr + m;   // pinv[r] and cmp[pinv[r]] of every row
    __shared__ int ucol[FILLC], ccpos[FILLC];
    __shared__ double cdorg[FILLC], crt[FILLC], cinvn[FILLC];
    for (int r = threadIdx.x; r < m; r += 256) {
      const int p = clampi(pinv[r], m);
      pr[r] = p;
      cixr[r] = cmp[p];
    }
    if (threadIdx.x < nc) {
      const int cc = threadIdx.x, u = clampi(colsrc[c0 + cc], m);
      ucol[cc] = u;
      if (u < k) {
        cdorg[cc] = dorg[u];
        crt[cc] = rt[u];
        cinvn[cc] = invn[u];
      } else {
        ccpos[cc] = clampi(cpos[u], m);
      }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nc * m; idx += 256) {
      const int cc = idx / m, r = idx - cc * m;
      const int u = ucol[cc];
      double val;
      if (u < k) {
        const int cix = cixr[r];
        val = cix < 0 ? 0.0 : zl[clampi(cix, k)] * rcp_nr((dl[clampi(cix, k)] - cdorg[cc]) - crt[cc]) * cinvn[cc];
      } else {
        val = (pr[r] == ccpos[cc]) ? 1.0 : 0.0;
      }
      S[(size_t)(a0 + c0 + cc) * ld + (a0 + r)] = val;
    }
  } else {
    __syncthreads();
    for (int idx = threadIdx.x; idx < nc * m; idx += 256) {
      const int cc = idx / m, r = idx - cc * m;
      const int c = c0 + cc;
      const int u = clampi(colsrc[c], m);
      const int p = clampi(pinv[r], m);
      double val;
      if (u < k) {
        const int cix = cmp[p];
        val = cix < 0 ? 0.0 : zl[clampi(cix, k)] * rcp_nr((dl[clampi(cix, k)] - dorg[u]) - rt[u]) * invn[u];
      } else {
        val = (p == clampi(cpos[u], m)) ? 1.0 : 0.0;
      }
      S[(size_t)(a0 + c) * ld + (a0 + r)] = val;
    }
  }
  __syncthreads();
  // undo the deflation rotations (last generated first) on the rows of this chunk's columns
  if (nrot > 0 && threadIdx.x < nc) {
    double *Sc = S + (size_t)(a0 + c0 + threadIdx.x) * ld + a0;
    for (int t = nrot - 1; t >= 0; t--) {
      const int r1 = clampi(perm[clampi(rpj[t], m)], m), r2i = clampi(perm[clampi(rnj[t], m)], m);
      const double cc = rc[t], ss = rs[t];
      const double x = Sc[r1], y = Sc[r2i];
      Sc[r1] = cc * x - ss * y;
      Sc[r2i] = ss * x + cc * y;
    }
  }
  // small merges: Qout = diag(Q1, Q2) S for this chunk's columns
  if (m <= SMALLM) {
    __syncthreads();
    const double *Qin = ((depth + 1) & 1 ? a.Z1 : a.Z0) + bd.offZ;
    double *Qout = (depth & 1 ? a.Z1 : a.Z0) + bd.offZ;
    for (int idx = threadIdx.x; idx < nc * m; idx += 256) {
      const int cc = idx / m, r = idx - cc * m;
      const double *Sc = S + (size_t)(a0 + c0 + cc) * ld + a0;
      double acc = 0;
      if (r < m1) {
        for (int t = 0; t < m1; t++) acc += Qin[(size_t)(a0 + t) * ld + (a0 + r)] * Sc[t];
      } else {
        for (int t = m1; t < m; t++) acc += Qin[(size_t)(a0 + t) * ld + (a0 + r)] * Sc[t];
      }
      Qout[(size_t)(a0 + c0 + cc) * ld + (a0 + r)] = acc;
    }
  }
}

// ------------------------------------------------------------------------------------------
// 4c. FP64 back-transformation with super-panels of KBT reflectors, on the FP64 matrix cores:
//     Y_s (explicit, unit lower trapezoidal), G_s = Y_s^T Y_s, T_s by the dlarft recurrence,
//     Yhat_s = Y_s T_s^T; then for s from last to first: W = Yhat_s^T Z, Z -= Y_s W.
// ------------------------------------------------------------------------------------------
constexpr int KBT = 128;

// Y (column j = reflector j: zeros above row j + roff, 1 there, v below), 64 columns per item
template <typename T>
__global__ __launch_bounds__(256) void k_bt_ymat(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n, ro = bd.roff;
  const int j0 = it.y * 64, j1 = min(n - ro, j0 + 64);
  const T *A = a.A + bd.offA;
  T *Y = a.Ym + bd.offZ;
  for (int j = j0; j < j1; j++)
    for (int r = threadIdx.x; r < n; r += 256)
      Y[(size_t)j * n + r] = r == j + ro ? (T)1 : (r > j + ro ? A[(size_t)j * n + r] : (T)0);
}

// T_s (kb x kb, upper triangular, column-major) of super-panel s from G_s = Y_s^T Y_s and the
// reflector scalars (dlarft, forward, columnwise), blocked: the four diagonal 32x32 blocks follow
// the column recurrence T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i), one wave each; the blocks above
// them follow from T12 = -T11 G12 T22 (the product of two compact-WY transforms).
template <typename T>
__global__ __launch_bounds__(256) void k_bt_tbig(SyevArgs<T> a, const int2 *items, const int64_t *toff) {
  constexpr int BS = 32, NBS = KBT / BS;
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n, sp = it.y;
  const int kb = min(KBT, (n - bd.roff) - sp * KBT);
  const size_t off = (size_t)toff[it.x] + (size_t)sp * KBT * KBT;
  const T *G = a.Gb + off;     // G(c, i) = G[i*kb + c]
  T *Tg = a.Tb + off;
  const T *tau = a.tau + bd.offV + sp * KBT;
  __shared__ T Tp[KBT * (KBT + 1) / 2];   // upper triangle, T(x, i) at i(i+1)/2 + x
  __shared__ T Gs[NBS * BS * (BS + 1)];   // diagonal blocks of G; then G(0:base, block)
  __shared__ T Ys[(KBT - BS) * BS];       // T(0:base, 0:base) G(0:base, block)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  auto P = [](int x, int i) { return i * (i + 1) / 2 + x; };
  for (int idx = tid; idx < NBS * BS * BS; idx += 256) {
    const int b = idx / (BS * BS), c = idx % BS, i = (idx / BS) % BS;
    const int gc = b * BS + c, gi = b * BS + i;
    Gs[(b * BS + c) * (BS + 1) + i] = (gc < kb && gi < kb) ? G[(size_t)gi * kb + gc] : (T)0;
  }
  __syncthreads();
  // 1. diagonal blocks
  {
    const int base = wv * BS, x = lane;
    const T *Gb = Gs + base * (BS + 1);   // G(base + c, base + i) at Gb[c*(BS+1) + i]
    for (int i = 0; i < BS; i++) {
      const int gi = base + i;
      if (gi < kb && x <= i) {
        const T ti = tau[gi];
        if (x < i) {
          T s = 0;
          for (int c = x; c < i; c++) s += Tp[P(base + x, base + c)] * Gb[c * (BS + 1) + i];
          Tp[P(base + x, gi)] = -ti * s;
        } else {
          Tp[P(gi, gi)] = ti;
        }
      }
      __syncthreads();
    }
  }
  // 2. blocks above the diagonal, one column block at a time
  for (int b = 1; b < NBS; b++) {
    const int base = b * BS;
    if (base >= kb) break;
    const int jw = min(BS, kb - base);
    for (int idx = tid; idx < base * BS; idx += 256) {
      const int c = idx / BS, j = idx % BS;
      Gs[idx] = j < jw ? G[(size_t)(base + j) * kb + c] : (T)0;   // X(c, j) = G(c, base + j)
    }
    __syncthreads();
    for (int idx = tid; idx < base * BS; idx += 256) {
      const int x = idx / BS, j = idx % BS;
      T s = 0;
      for (int c = x; c < base; c++) s += Tp[P(x, c)] * Gs[c * BS + j];
      Ys[idx] = s;
    }
    __syncthreads();
    for (int idx = tid; idx < base * BS; idx += 256) {
      const int x = idx / BS, j = idx % BS;
      if (j < jw) {
        T s = 0;
        for (int c = 0; c <= j; c++) s += Ys[x * BS + c] * Tp[P(base + c, base + j)];
        Tp[P(x, base + j)] = -s;
      }
    }
    __syncthreads();
  }
  for (int idx = tid; idx < kb * kb; idx += 256) {
    const int x = idx % kb, i = idx / kb;
    Tg[idx] = x <= i ? Tp[P(x, i)] : (T)0;
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
  const double s = a.scale[b] / block_sigma(a.amax[b]);
  const double *lam = a.lam + bd.offV;
  double *w = a.w + bd.offw;
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double v = lam[i] * s;
    w[i] = v;
    bad |= !isfinite(v);
  }
  if (bad) a.flags[b] = 2;
}

constexpr int COPY_CHUNK = 16384;

// eigenvectors into the caller's arena: item (block, chunk of COPY_CHUNK elements)
template <typename T>
__global__ __launch_bounds__(256) void k_copy_out(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const size_t nn = (size_t)bd.n * bd.n;
  const size_t beg = (size_t)it.y * COPY_CHUNK, end = min(nn, beg + COPY_CHUNK);
  const T *Z = a.Zt + bd.offZ;
  double *H = a.H + bd.offH;
  for (size_t i = beg + threadIdx.x; i < end; i += 256) H[i] = (double)Z[i];
}

template <typename T>
__global__ void k_convert(const double *src, T *dst, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (T)src[i];
}

// ------------------------------------------------------------------------------------------
// 6. Mixed form: Ogita-Aishima refinement in FP64 of the FP32 eigenvectors.
//    For X with A X ~ X L:  R = I - X^T X,  S = X^T A X,  l_i = s_ii / (1 - r_ii),
//    E_ij = (s_ij + l_j r_ij) / (l_j - l_i) for pairs of different clusters, else r_ij / 2,
//    X <- X (I + E + K^2 / 2) with K the skew part of E (the second-order term of exp(K), which
//    keeps the orthogonality error cubic instead of quadratic in K).
//    The close pairs are those of one cluster: a run of neighbouring estimates linked by their own
//    coupling (ref_linked). A cluster's Ritz problem B = S_CC + (R_CC S_CC + S_CC R_CC) / 2 is solved
//    by Jacobi and rotates X_C.
// ------------------------------------------------------------------------------------------
struct RefArgs {
  const BlkDesc *bd;
  const double *A;      // scaled symmetric input, full storage (the H arena)
  double *X, *Xn, *AX, *S, *R, *E;   // n x n per block at offZ
  double *lt;           // eigenvalue estimates, n per block at offV
  double *acc;          // per block: [2] max |l| ([0], [1] unused)
  int *flags;
  int *rid;             // per estimate (n per block at offV): first index of its cluster run
};

// X = (double) V32, R = I
__global__ __launch_bounds__(256) void k_ref_init(RefArgs r, const float *V32, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = r.bd[it.x];
  const int n = (int)bd.n;
  const size_t nn = (size_t)n * n;
  const size_t beg = (size_t)it.y * COPY_CHUNK, end = min(nn, beg + COPY_CHUNK);
  for (size_t i = beg + threadIdx.x; i < end; i += 256) {
    r.X[bd.offZ + i] = (double)V32[bd.offZ + i];
    r.R[bd.offZ + i] = (i % n == i / n) ? 1.0 : 0.0;
  }
}

// R = I again before the next sweep's X^T X product
__global__ __launch_bounds__(256) void k_ref_eye(RefArgs r, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = r.bd[it.x];
  const int n = (int)bd.n;
  const size_t nn = (size_t)n * n;
  const size_t beg = (size_t)it.y * COPY_CHUNK, end = min(nn, beg + COPY_CHUNK);
  for (size_t i = beg + threadIdx.x; i < end; i += 256) r.R[bd.offZ + i] = (i % n == i / n) ? 1.0 : 0.0;
}

// eigenvalue estimates l_i = s_ii / (1 - r_ii), max |l|
__global__ __launch_bounds__(256) void k_ref_lam(RefArgs r, const int *blist) {
  const int b = blist[blockIdx.x];
  const BlkDesc bd = r.bd[b];
  const int n = (int)bd.n;
  __shared__ double red[4];
  double mx = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const size_t di = bd.offZ + (size_t)i * n + i;
    const double l = r.S[di] / (1.0 - r.R[di]);
    r.lt[bd.offV + i] = l;
    mx = fmax(mx, fabs(l));
  }
  mx = block_max<256>(mx, red);
  if (threadIdx.x == 0) r.acc[3 * b + 2] = mx;
}

// Which estimates are clustered. Ogita-Aishima's global threshold 2 (||S - diag l||_2 + ||A|| ||R||_2),
// bounded by Frobenius norms, overstates the coupling by up to sqrt(n) (about 75 at n = 5500); on NRG
// blocks (dense spectra, half the levels in exactly degenerate pairs) it merged a thousand estimates
// into one run, and no global value served both those blocks and the synthetic clustered spectra. So
// neighbours i-1, i are linked by their own coupling: when their gap is at most
// max(CLUSTER_FLOOR ||A||, CLUSTER_KAPPA (|s| + |l| |r|)) of the pair (symmetric parts). Linked runs
// are the clusters; every other pair takes the first-order correction, which is then below
// 1 / CLUSTER_KAPPA for neighbours.
constexpr double CLUSTER_FLOOR = 1e-6;
constexpr double CLUSTER_KAPPA = 8.0;
__device__ __forceinline__ bool ref_linked(const RefArgs &r, const BlkDesc &bd, int n, int i, double anorm) {
  const double *lt = r.lt + bd.offV;
  const double gap = fabs(lt[i] - lt[i - 1]);
  const size_t k = bd.offZ + (size_t)i * n + (i - 1), kt = bd.offZ + (size_t)(i - 1) * n + i;
  const double sij = 0.5 * (r.S[k] + r.S[kt]), rij = 0.5 * (r.R[k] + r.R[kt]);
  const double coup = fabs(sij) + fabs(lt[i]) * fabs(rij);
  return gap <= fmax(CLUSTER_FLOOR * anorm, CLUSTER_KAPPA * coup);
}

// E' = I + E, 16 columns per item
__global__ __launch_bounds__(256) void k_ref_E(RefArgs r, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = r.bd[it.x];
  const int n = (int)bd.n;
  const int c0 = it.y * 16, c1 = min(n, c0 + 16);
  const double *lt = r.lt + bd.offV;
  const int *rid = r.rid + bd.offV;
  // S and R enter through their symmetric parts: the computed X^T (A X) and X^T X are symmetric
  // only up to rounding, and the antisymmetric rest, divided by l_j - l_i, would leave
  // E + E^T != R and stall the orthogonality at (s_ij - s_ji) / gap.
  for (int j = c0; j < c1; j++) {
    const double lj = lt[j];
    for (int i = threadIdx.x; i < n; i += 256) {
      const size_t k = bd.offZ + (size_t)j * n + i, kt = bd.offZ + (size_t)i * n + j;
      const double rij = 0.5 * (r.R[k] + r.R[kt]);
      double e, kij = 0.0;
      if (i == j) {
        e = 1.0 + 0.5 * rij;
      } else {
        const double li = lt[i];
        if (rid[i] != rid[j]) {
          const double sij = 0.5 * (r.S[k] + r.S[kt]);
          const double eij = (sij + lj * rij) / (lj - li), eji = (sij + li * rij) / (li - lj);
          e = eij;
          kij = 0.5 * (eij - eji);
        } else {
          e = 0.5 * rij;
        }
      }
      r.E[k] = e;
      r.AX[k] = kij;   // skew part K of the correction (A X is no longer needed)
    }
  }
}

constexpr int CMAX = 64;   // largest cluster solved as one Ritz problem

// One cluster: estimates [start, start + c) of block b.
struct RefCluster {
  int b, start, c;
};

// Clusters: runs of consecutive estimates linked by ref_linked. A run longer than CMAX is cut into
// pieces of CMAX (and a last piece); pieces of one estimate are not clusters. With `half` set the first
// piece of a run has CMAX / 2 estimates, so that consecutive sweeps cut long runs at different places.
// rid receives, for every estimate, the first index of its run (its own index when unlinked): k_ref_E
// treats two estimates of one run as a cluster pair. The thread at the first index of a run walks it.
__global__ __launch_bounds__(256) void k_ref_find(RefArgs r, const int *blist, RefCluster *cl, int *ncl, int half) {
  const int b = blist[blockIdx.x];
  const BlkDesc bd = r.bd[b];
  const int n = (int)bd.n;
  const double anorm = r.acc[3 * b + 2];
  int *rid = r.rid + bd.offV;
  for (int i = threadIdx.x; i < n; i += 256) {
    if (i > 0 && ref_linked(r, bd, n, i, anorm)) continue;   // inside a run
    rid[i] = i;
    if (i + 1 >= n || !ref_linked(r, bd, n, i + 1, anorm)) continue;   // a run of one
    int end = i + 1;
    while (end < n && ref_linked(r, bd, n, end, anorm)) rid[end++] = i;
    for (int s0 = i; s0 < end;) {
      const int len = (half && s0 == i && end - i > CMAX) ? CMAX / 2 : CMAX;
      const int c = min(len, end - s0);
      if (c >= 2) cl[atomicAdd(ncl, 1)] = RefCluster{b, s0, c};
      s0 += c;
    }
  }
}

// Ritz problem of one cluster C per workgroup: B = S_CC + (R_CC S_CC + S_CC R_CC) / 2, symmetrised,
// diagonalised by parallel cyclic Jacobi in LDS (round-robin ordering: the c/2 rotations of a step act
// on disjoint index pairs, so they are applied together, columns first, then rows); the estimates
// l_C become its eigenvalues in ascending order and X_C <- X_C Q with Q the matching eigenvectors.
// Workgroups past the number of clusters found return at once.
__global__ __launch_bounds__(256) void k_ref_csolve(RefArgs r, const RefCluster *cl, const int *ncl) {
  if ((int)blockIdx.x >= *ncl) return;
  const RefCluster C = cl[blockIdx.x];
  const BlkDesc bd = r.bd[C.b];
  const int n = (int)bd.n, start = C.start, c = C.c;
  const double *S = r.S + bd.offZ, *R = r.R + bd.offZ;
  double *X = r.Xn + bd.offZ;
  double *lt = r.lt + bd.offV;
  __shared__ double B[CMAX][CMAX + 1], Q[CMAX][CMAX + 1];
  __shared__ double rc[CMAX / 2], rs[CMAX / 2], red[2][256 / WAVE];
  __shared__ int rp[CMAX / 2], rq[CMAX / 2], perm[CMAX];
  __shared__ double ev[CMAX];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < c * c; idx += 256) {
    const int i = idx % c, j = idx / c;
    double rsum = 0;
    for (int k = 0; k < c; k++)
      rsum += R[(size_t)(start + k) * n + start + i] * S[(size_t)(start + j) * n + start + k] +
              S[(size_t)(start + k) * n + start + i] * R[(size_t)(start + j) * n + start + k];
    B[i][j] = S[(size_t)(start + j) * n + start + i] + 0.5 * rsum;
  }
  __syncthreads();
  for (int idx = tid; idx < c * c; idx += 256) Q[idx % c][idx / c] = 0.5 * (B[idx % c][idx / c] + B[idx / c][idx % c]);
  __syncthreads();
  for (int idx = tid; idx < c * c; idx += 256) {
    const int i = idx % c, j = idx / c;
    B[i][j] = Q[i][j];
    Q[i][j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  const int m = (c + 1) & ~1, half = m / 2;   // an odd c gets a dummy index c, paired with nothing
  for (int sweep = 0; sweep < 30; sweep++) {
    double off = 0, tot = 0;
    for (int idx = tid; idx < c * c; idx += 256) {
      const double v = B[idx % c][idx / c];
      tot += v * v;
      if (idx % c != idx / c) off += v * v;
    }
    off = wave_sum(off);
    tot = wave_sum(tot);
    if (tid % WAVE == 0) {
      red[0][tid / WAVE] = off;
      red[1][tid / WAVE] = tot;
    }
    __syncthreads();
    off = 0;
    tot = 0;
    for (int w = 0; w < 256 / WAVE; w++) {
      off += red[0][w];
      tot += red[1][w];
    }
    if (off <= 1e-34 * tot || off == 0.0) break;   // uniform: every thread read the same sums
    for (int step = 0; step < m - 1; step++) {
      if (tid < half) {   // circle method: m - 1 stays, the others turn
        int p, q;
        if (tid == 0) {
          p = step;
          q = m - 1;
        } else {
          p = (step + tid) % (m - 1);
          q = (step - tid + (m - 1)) % (m - 1);
        }
        if (p > q) { const int t = p; p = q; q = t; }
        double cc = 1.0, ss = 0.0;
        if (q < c) {
          const double apq = B[p][q];
          if (apq != 0.0) {
            const double th = (B[q][q] - B[p][p]) / (2.0 * apq);
            const double t = copysign(1.0, th) / (fabs(th) + sqrt(1.0 + th * th));
            cc = 1.0 / sqrt(1.0 + t * t);
            ss = t * cc;
          }
        }
        rp[tid] = p;
        rq[tid] = q;
        rc[tid] = cc;
        rs[tid] = ss;
      }
      __syncthreads();
      for (int idx = tid; idx < half * c; idx += 256) {   // columns p, q of B and Q
        const int k = idx / c, row = idx % c;
        const double cc = rc[k], ss = rs[k];
        if (ss == 0.0) continue;
        const int p = rp[k], q = rq[k];
        const double bp = B[row][p], bq = B[row][q];
        B[row][p] = cc * bp - ss * bq;
        B[row][q] = ss * bp + cc * bq;
        const double qp = Q[row][p], qq = Q[row][q];
        Q[row][p] = cc * qp - ss * qq;
        Q[row][q] = ss * qp + cc * qq;
      }
      __syncthreads();
      for (int idx = tid; idx < half * c; idx += 256) {   // rows p, q of B
        const int k = idx / c, col = idx % c;
        const double cc = rc[k], ss = rs[k];
        if (ss == 0.0) continue;
        const int p = rp[k], q = rq[k];
        const double bp = B[p][col], bq = B[q][col];
        B[p][col] = cc * bp - ss * bq;
        B[q][col] = ss * bp + cc * bq;
      }
      __syncthreads();
    }
  }
  // ascending order of the Ritz values
  if (tid == 0) {
    for (int i = 0; i < c; i++) {
      perm[i] = i;
      ev[i] = B[i][i];
    }
    for (int i = 1; i < c; i++) {
      int j = i;
      while (j > 0 && ev[perm[j - 1]] > ev[perm[j]]) {
        const int t = perm[j];
        perm[j] = perm[j - 1];
        perm[j - 1] = t;
        j--;
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < c; i += 256) lt[start + i] = ev[perm[i]];
  // X_C <- X_C Q(:, perm), eigenvector k of the cluster at X[(start + k) n .. + n): 64 components at a
  // time, staged in B (no longer needed)
  for (int r0 = 0; r0 < n; r0 += CMAX) {
    const int w = min(CMAX, n - r0);
    for (int idx = tid; idx < c * CMAX; idx += 256) {
      const int k = idx / CMAX, rr = idx % CMAX;
      if (rr < w) B[k][rr] = X[(size_t)(start + k) * n + r0 + rr];
    }
    __syncthreads();
    for (int idx = tid; idx < c * CMAX; idx += 256) {
      const int j = idx / CMAX, rr = idx % CMAX;
      if (rr >= w) continue;
      const int pj = perm[j];
      double v = 0;
      for (int k = 0; k < c; k++) v += B[k][rr] * Q[k][pj];
      X[(size_t)(start + j) * n + r0 + rr] = v;
    }
    __syncthreads();
  }
}

// results of the mixed form: w = l / sigma, eigenvectors X -> H
__global__ __launch_bounds__(256) void k_ref_finish(RefArgs r, double *H, double *w, const double *amax,
                                                    const int *blist) {
  const int b = blist[blockIdx.x];
  const BlkDesc bd = r.bd[b];
  const int n = (int)bd.n;
  const double sg = block_sigma(amax[b]);
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double v = r.lt[bd.offV + i] / sg;
    w[bd.offw + i] = v;
    bad |= !isfinite(v);
  }
  if (bad) r.flags[b] = 2;
}

__global__ __launch_bounds__(256) void k_ref_copy(RefArgs r, double *H, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = r.bd[it.x];
  const size_t nn = (size_t)bd.n * bd.n;
  const size_t beg = (size_t)it.y * COPY_CHUNK, end = min(nn, beg + COPY_CHUNK);
  for (size_t i = beg + threadIdx.x; i < end; i += 256) H[bd.offH + i] = r.X[bd.offZ + i];
}

// Concurrent column chains of the reduction fixed by NRG_RED_STREAMS (1 .. NAUX: the caller's stream
// and NAUX - 1 extra streams; the last extra stream carries the two-stage chain), or 0 when unset.
static int red_groups_env() {
  const char *e = std::getenv("NRG_RED_STREAMS");
  return e ? std::max(1, std::min(std::atoi(e), nrg_ctx::NAUX)) : 0;
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
      set_error("nrg_syev_batched: block order out of range [1, " + std::to_string(NMAX) + "]");
      return NRG_E_ARG;
    }
  if (int r = ensure_aux_streams(ctx)) return r;
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
  int64_t sumn = 0, sumn2 = 0, sumWp = 0;
  int Pmax = 0, Lmax = 0, nmax = 0;
  // work lists, all packed into one int array; offsets in ints
  std::vector<int> ibuf;
  size_t symm_off = 0, symm_cnt = 0;
  // reduced column jl of panel slot p for the blocks of group g
  struct ColLaunch { int p, jl, g, rows; size_t prep_off, prep_cnt, symv_off, tpre_off, nact, tiles, yred_off, yred_cnt; };
  int ngrp = 1;
  std::vector<int> grp;   // group of every block
  std::vector<ColLaunch> cols;
  int64_t sumYp = 0;
  size_t all_off = 0;  // list 0..nblk-1
  std::vector<size_t> mrg_off, mrg_cnt, gemm_off, gemm_cnt;
  std::vector<size_t> ch_off, ch_cnt, fc_off, fc_cnt;  // 256-wide chunks and FILLC-wide column chunks
  std::vector<size_t> rc_off, rc_cnt;                  // RCH-wide root chunks
  size_t mrg_max = 0;
  size_t cp_off = 0, cp_cnt = 0, tail_off = 0, tail_cnt = 0;
  // single-workgroup reduction (k_wg_red): blocks without column-kernel panels, launched beside the
  // chains, and per group the blocks that continue after their chain
  size_t wg0_off = 0, wg0_cnt = 0;
  std::vector<size_t> wgg_off, wgg_cnt;
  std::vector<int> mcap;
  std::vector<GemmDesc> gd;   // products, filled once the arena pointers are known
  // FP64 back-transformation
  std::vector<int64_t> toff;  // per block: offset of its super-panels in the G/T arenas
  int64_t sumTb = 0, sumW2 = 0;
  int Smax = 0;
  size_t ym_off = 0, ym_cnt = 0, tb_off = 0, tb_cnt = 0;
  size_t g_items = 0, g_cnt = 0, yh_items = 0, yh_cnt = 0;
  std::vector<size_t> w_items, w_cnt, u_items, u_cnt;
  std::vector<size_t> s2_items, s2_cnt;   // trailing rank-2k updates per (panel slot, group): [q * ngrp + g]
  // two-stage reduction of the blocks with two = 1
  int nsb = 0, P1max = 0;                 // blocks reduced in two stages, their most stage-1 panels
  int64_t sumAB = 0, sumV2 = 0, sumT2 = 0;
  std::vector<size_t> sbp_off, sbp_cnt;   // stage-1 panel lists per panel slot (col_entry(b, q))
  std::vector<size_t> sby_items, sby_cnt, sbu_items, sbu_cnt;   // Y = A22 V and the rank-2 SB update
  size_t sbb_off = 0, sbb_cnt = 0, sbc_off = 0, sbc_cnt = 0, sbq_off = 0, sbq_cnt = 0;   // band, chase, Q2
};

// The work lists and product tables of a call, kept in the context for the next call: the lists of
// a 256-block batch take some 4 MB, and fresh memory for them each call costs more host time (page
// faults) than building them.
// A call captured into a HIP graph: its launches read the descriptor tables from a device buffer of
// the graph's own, so other calls on the context (which reuse the shared descriptor buffer) leave it
// intact. `key` identifies the calls it stands for.
struct SyevGraph {
  std::vector<int64_t> key;
  hipGraphExec_t exec = nullptr;
  DevBuf desc;
  int nphase = 0;   // phase events recorded by the graph
  ~SyevGraph() {
    if (exec) hipGraphExecDestroy(exec);
    desc.release();
  }
};

// Autotuner of the number of concurrent column chains. A call is described by a coarse signature of
// its block-size histogram: the form, the octave of the largest order, and the octaves of the numbers
// of blocks above 1024 and in 129..1024. The first call with a signature runs the default and is not
// measured (it pays for first-use costs); the next TUNE_ROUNDS x NTUNE calls run the candidates
// (TUNE_CAND chains) in turn, each measured by the call's own event time per unit of work (Σ n³, so
// that the different steps of an NRG run with one signature compare); then the candidate with the
// smallest time per unit is used for that signature if it beats the default by TUNE_MARGIN.
// NRG_RED_STREAMS fixes the count, NRG_AUTOTUNE=0 keeps RED_GROUPS.
constexpr int TUNE_CAND[] = {2, 3, 4};
constexpr int NTUNE = sizeof(TUNE_CAND) / sizeof(int);
constexpr int TUNE_ROUNDS = 2;
constexpr double TUNE_MARGIN = 0.02;
static_assert(TUNE_CAND[0] == RED_GROUPS, "the first candidate is the default");
struct TuneEntry {
  double unit[NTUNE] = {};   // least ms per 1e9 n^3 measured for each candidate
  float ms[NTUNE] = {};      // call time of that measurement
  int step = -1;             // next measurement (-1: the unmeasured first call; then k = round * NTUNE + cand)
  int best = -1;             // chosen candidate once every measurement has run
  long calls = 0;
};
using TuneSig = std::array<int, 4>;

struct SyevCache {
  SyevPlan pl;
  std::map<TuneSig, TuneEntry> tune;
  std::vector<GemmDesc> g64, gref;
  std::vector<GemmDesc32> g32;
  // graph replay of repeated calls: the key of the previous call, and the graph of the last key
  // that was seen twice in a row
  std::vector<int64_t> last_key;
  std::unique_ptr<SyevGraph> graph;
  bool graph_failed = false;   // capture is not retried on this context after a failure
};

// An empty plan that keeps the memory of the previous call's lists.
static SyevCache &syev_cache(nrg_ctx *ctx) {
  if (!ctx->syev_cache) ctx->syev_cache = std::make_shared<SyevCache>();
  SyevCache &c = *static_cast<SyevCache *>(ctx->syev_cache.get());
  std::vector<int> ibuf = std::move(c.pl.ibuf);
  std::vector<GemmDesc> gd = std::move(c.pl.gd);
  std::vector<SyevPlan::ColLaunch> cols = std::move(c.pl.cols);
  c.pl = SyevPlan();
  ibuf.clear();
  gd.clear();
  cols.clear();
  c.pl.ibuf = std::move(ibuf);
  c.pl.gd = std::move(gd);
  c.pl.cols = std::move(cols);
  c.g64.clear();
  c.gref.clear();
  c.g32.clear();
  return c;
}

static bool autotune_on() {
  const char *e = std::getenv("NRG_AUTOTUNE");
  return !(e && e[0] == '0');
}

static TuneSig tune_signature(int form, int nblk, const int64_t *n) {
  int64_t nmax = 1, big = 0, mid = 0;
  for (int b = 0; b < nblk; b++) {
    nmax = std::max(nmax, n[b]);
    big += n[b] > 1024;
    mid += n[b] > TAIL && n[b] <= 1024;
  }
  auto oct = [](int64_t x) { return x > 0 ? 64 - __builtin_clzll((unsigned long long)x) : 0; };
  return TuneSig{form, oct(nmax), oct(big), oct(mid)};
}

// The chain count of this call: fixed by the environment, the tuner's choice, or what the tuner
// measures next (*cand >= 0: the call's time is recorded as measurement *cand).
static int tune_choose(SyevCache &c, const TuneSig &sig, int *cand) {
  *cand = -1;
  if (const int g = red_groups_env()) return g;
  if (!autotune_on()) return RED_GROUPS;
  TuneEntry &t = c.tune[sig];
  t.calls++;
  if (t.best >= 0) return TUNE_CAND[t.best];
  *cand = t.step;
  return t.step < 0 ? RED_GROUPS : TUNE_CAND[t.step % NTUNE];
}

static void tune_record(SyevCache &c, const TuneSig &sig, int cand, float ms, int nblk, const int64_t *n) {
  if (cand < -1 || !c.tune.count(sig)) return;
  TuneEntry &t = c.tune[sig];
  if (t.best >= 0 || cand != t.step || !(ms > 0)) return;
  if (cand >= 0) {
    double n3 = 0;
    for (int b = 0; b < nblk; b++) n3 += (double)n[b] * n[b] * n[b];
    const double u = ms / (n3 * 1e-9);
    const int k = cand % NTUNE;
    if (t.unit[k] == 0 || u < t.unit[k]) {
      t.unit[k] = u;
      t.ms[k] = ms;
    }
  }
  if (++t.step == TUNE_ROUNDS * NTUNE) {
    // the default (TUNE_CAND[0] = RED_GROUPS) unless another count is at least TUNE_MARGIN faster:
    // smaller differences are within the call-to-call spread, and more chains spread more
    t.best = 0;
    for (int k = 1; k < NTUNE; k++)
      if (t.unit[k] < t.unit[t.best] && t.unit[k] < (1.0 - TUNE_MARGIN) * t.unit[0]) t.best = k;
  }
}

// Back-transformation super-panels of a block: its reflectors (n - roff of them) in groups of KBT;
// none for orders below 3 in one stage (the only reflector has length one).
static int bt_panels(const BlkDesc &d) {
  if (!d.two && d.n < 3) return 0;
  return (int)((d.n - d.roff + KBT - 1) / KBT);
}

// Panels reduced by the column kernels: those that start with more than TAIL trailing rows. The
// trailing matrix of order <= TAIL that remains is reduced by k_tail.
static int main_panels(int64_t n) {
  return n > TAIL ? (int)((n - TAIL + NB - 1) / NB) : 0;
}

// Trailing order from which k_wg_red takes over the panels from the column kernels: NRG_WG_RED in
// the environment (read at every call; at most WGR_MAX). Off (TAIL: every panel to the column
// kernels) by default: on MI355X one CU's symmetric product of an order-500 trailing matrix takes
// about 20 us per column (sequential L2 round trips and cross-lane sums), more than the column
// chain's whole step (about 14 us per column for a lone block, 0.8 MB read by many workgroups), so a
// lone block of order 512 reduces in 12.6 ms this way against 5.5 ms in the chain.
static int wg_red_order() {
  const char *e = std::getenv("NRG_WG_RED");
  if (!e || !*e) return TAIL;
  return std::max(TAIL, std::min(WGR_MAX, atoi(e)));
}

// Panels of a block reduced by the column kernels: those that start with more than `thr` rows.
static int chain_panels(int64_t n, int thr) {
  return n > thr ? (int)((n - thr + NB - 1) / NB) : 0;
}

// Blocks reduced in two stages when NRG_TWO_STAGE=1 in the environment (read at every call): orders
// above TAIL whose stage-1 panels fit one row per thread. Off by default: on MI355X the bulge chase's
// critical path (each sweep three dependent steps behind the previous one, about 3n steps of a few
// microseconds for a block of order n) exceeds the whole one-stage reduction at these orders.
static bool two_stage_on() {
  const char *e = std::getenv("NRG_TWO_STAGE");
  return e && e[0] == '1';
}
static bool use_two_stage(int64_t n) { return two_stage_on() && n > TAIL && n <= SB_NMAX; }

static void build_plan(SyevPlan &pl, int nblk, const int64_t *n, const int64_t *off_H, const int64_t *off_w,
                       bool fp64_inplace, int elem_bytes, int ngrp) {
  pl.ngrp = std::max(1, std::min(ngrp, nblk));
  pl.bd.resize(nblk);
  int64_t zo = 0, vo = 0, wo = 0, ao = 0;
  pl.nsb = 0; pl.P1max = 0; pl.sumAB = 0; pl.sumV2 = 0; pl.sumT2 = 0;
  const int wthr = wg_red_order();
  for (int b = 0; b < nblk; b++) {
    BlkDesc &d = pl.bd[b];
    d.n = n[b];
    d.offZ = zo; zo += n[b] * n[b];
    d.offA = fp64_inplace ? off_H[b] : ao; ao += n[b] * n[b];
    d.offV = vo; vo += n[b];
    d.offWp = wo; wo += n[b] * NB;
    d.two = use_two_stage(n[b]) ? 1 : 0;
    d.roff = d.two ? SB : 1;
    d.offAB = pl.sumAB; d.offV2 = pl.sumV2; d.offT2 = pl.sumT2;
    if (d.two) {
      pl.nsb++;
      pl.P1max = std::max(pl.P1max, sb_panels((int)n[b]));
      pl.sumAB += 2 * SB * n[b];
      pl.sumV2 += (int64_t)SB * n[b] * sb_ks((int)n[b]);
      pl.sumT2 += n[b] * sb_ks((int)n[b]);
    }
    d.P = d.two ? 0 : chain_panels(n[b], wthr);
    d.offYp = pl.sumYp; pl.sumYp += n[b] * ((n[b] + 63) / 64);
    d.offDt = pl.sumYp; pl.sumYp += 64 * ((n[b] + 63) / 64);
    d.offH = off_H[b];
    d.offw = off_w[b];
    d.L = 31 - __builtin_clz((unsigned)n[b]);
    pl.Pmax = std::max(pl.Pmax, d.P);
    pl.Lmax = std::max(pl.Lmax, d.L);
    pl.nmax = std::max(pl.nmax, (int)n[b]);
  }
  pl.sumn2 = zo; pl.sumn = vo; pl.sumWp = wo;
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
  // reduced columns (two launches each) and trailing updates, on a common sequence of Pmax panel
  // slots: block b works on panel panel_in_slot(q, P_b) in slot q. Work-list entries of the column
  // kernels carry the block and its panel (col_entry). The blocks are dealt round-robin, largest
  // first, into ngrp groups whose chains of column launches are independent: each group runs on a
  // stream of its own, so one group's latency-bound column preparation overlaps another group's
  // bandwidth-bound symmetric product. Launches are listed slot by slot, column by column, group by
  // group, the order in which they are issued.
  pl.grp.assign(nblk, 0);
  std::vector<std::vector<int>> members(pl.ngrp);   // the blocks of each group, largest first
  for (size_t k = 0; k < order.size(); k++) {
    pl.grp[order[k]] = (int)(k % pl.ngrp);
    members[k % pl.ngrp].push_back(order[k]);
  }
  pl.cols.clear();
  // Consecutive launches of a group within a slot mostly have the same block list (every active
  // block works on every column of its panel), and often the same tile offsets: a list equal to
  // the group's previous list of that kind is not stored again.
  std::vector<int> tmp;
  struct Last { size_t off = 0, cnt = 0; };
  std::vector<Last> last_blk(pl.ngrp), last_tile(pl.ngrp);
  auto share = [&](Last &last) -> size_t {
    if (last.cnt == tmp.size() && std::equal(tmp.begin(), tmp.end(), I.begin() + last.off)) return last.off;
    last.off = I.size();
    last.cnt = tmp.size();
    I.insert(I.end(), tmp.begin(), tmp.end());
    return last.off;
  };
  for (int q = 0; q < pl.Pmax; q++) {
    auto panel = [&](int b) { return panel_in_slot(q, pl.bd[b].P); };
    auto width = [&](int b) {
      const int64_t m = n[b] - (int64_t)NB * panel(b);
      return (int)(m - 1 <= NB ? m - 1 : NB);
    };
    int wmax = 0;
    for (int b : order) if (panel(b) >= 0) wmax = std::max(wmax, width(b));
    for (int jl = 0; jl <= wmax; jl++)
      for (int g = 0; g < pl.ngrp; g++) {
        // A block works on column jl in slot q while it has a panel there (P_b > q) that is wider
        // than jl. Both bounds grow with the block's order, so the blocks that do are a leading
        // run of the group's list, largest first; the list ends at the first one that does not.
        const std::vector<int> &mb = members[g];
        SyevPlan::ColLaunch c;
        c.p = q; c.jl = jl; c.g = g;
        c.rows = 0;
        size_t np = 0;
        for (; np < mb.size() && panel(mb[np]) >= 0 && jl <= width(mb[np]); np++) {
          const int b = mb[np];
          c.rows = std::max(c.rows, (int)(n[b] - (int64_t)NB * panel(b) - jl));
        }
        if (!np) continue;   // every symv block is a prep block too
        size_t ns = 0;
        while (ns < np && jl < width(mb[ns])) ns++;
        // block lists: the symv blocks are the leading ns of the np prep blocks
        tmp.clear();
        for (size_t k = 0; k < np; k++) tmp.push_back(col_entry(mb[k], panel(mb[k])));
        c.prep_off = share(last_blk[g]);
        c.prep_cnt = np;
        // row-parallel preparation (k_col_yred) of the blocks with more than NT_PANEL rows [jl, m)
        c.yred_off = I.size();
        if (jl > 0)
          for (size_t k = 0; k < np; k++) {
            const int b = mb[k];
            const int64_t rows = n[b] - (int64_t)NB * panel(b) - jl;
            if (rows > NT_PANEL)
              for (int64_t ch = 0; ch < (rows + NT_YR - 1) / NT_YR; ch++) push2(col_entry(b, panel(b)), (int)ch);
          }
        c.yred_cnt = (I.size() - c.yred_off) / 2;
        if (ns == np) {
          c.symv_off = c.prep_off;
        } else {
          tmp.resize(ns);
          c.symv_off = share(last_blk[g]);
        }
        c.nact = ns;
        // tile offsets of the symv blocks
        tmp.clear();
        int64_t tiles = 0;
        for (size_t k = 0; k < ns; k++) {
          const int b = mb[k];
          const int64_t t = n[b] - (int64_t)NB * panel(b) - jl - 1;
          const int64_t T = (t + 63) / 64;
          tmp.push_back((int)tiles);
          tiles += T * (T + 1) / 2 + (jl > 0 ? T : 0);   // + one dot-product workgroup per segment
        }
        c.tpre_off = share(last_tile[g]);
        c.tiles = (size_t)tiles;
        pl.cols.push_back(c);
      }
  }
  // panels left to k_wg_red: from panel P_b to the tail
  auto wg_left = [&](int b) { return !pl.bd[b].two && main_panels(n[b]) > pl.bd[b].P; };
  pl.wg0_off = I.size();
  for (int b : order)
    if (wg_left(b) && pl.bd[b].P == 0) I.push_back(col_entry(b, 0));
  pl.wg0_cnt = I.size() - pl.wg0_off;
  pl.wgg_off.assign(pl.ngrp, 0); pl.wgg_cnt.assign(pl.ngrp, 0);
  for (int g = 0; g < pl.ngrp; g++) {
    pl.wgg_off[g] = I.size();
    for (int b : members[g])
      if (wg_left(b) && pl.bd[b].P > 0) I.push_back(col_entry(b, pl.bd[b].P));
    pl.wgg_cnt[g] = I.size() - pl.wgg_off[g];
  }
  // tail of every block reduced in one stage (one k_tail launch after the column launches)
  pl.tail_off = I.size();
  for (int b : order)
    if (n[b] >= 1 && !pl.bd[b].two) I.push_back(col_entry(b, main_panels(n[b])));
  pl.tail_cnt = I.size() - pl.tail_off;
  // two-stage reduction: stage-1 panels slot by slot (block b has panel q in slot q), the band
  // copies, one bulge-chasing workgroup per block, and the stage-2 back-transformation slabs
  pl.sbp_off.assign(pl.P1max, 0); pl.sbp_cnt.assign(pl.P1max, 0);
  for (int q = 0; q < pl.P1max; q++) {
    pl.sbp_off[q] = I.size();
    for (int b : order)
      if (pl.bd[b].two && sb_panels((int)n[b]) > q) I.push_back(col_entry(b, q));
    pl.sbp_cnt[q] = I.size() - pl.sbp_off[q];
  }
  pl.sbb_off = I.size();
  for (int b : order)
    if (pl.bd[b].two)
      for (int c = 0; c < (int)((n[b] + 63) / 64); c++) push2(b, c);
  pl.sbb_cnt = (I.size() - pl.sbb_off) / 2;
  pl.sbc_off = I.size();
  for (int b : order)
    if (pl.bd[b].two) I.push_back(b);
  pl.sbc_cnt = I.size() - pl.sbc_off;
  pl.sbq_off = I.size();
  for (int b : order)
    if (pl.bd[b].two)
      for (int c = 0; c < (int)((n[b] + WAVE - 1) / WAVE); c++) push2(b, c);
  pl.sbq_cnt = (I.size() - pl.sbq_off) / 2;
  // divide and conquer merges, deepest first
  pl.mrg_off.assign(pl.Lmax, 0); pl.mrg_cnt.assign(pl.Lmax, 0);
  pl.gemm_off.assign(pl.Lmax, 0); pl.gemm_cnt.assign(pl.Lmax, 0);
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
    pl.mrg_max = std::max(pl.mrg_max, pl.mrg_cnt[depth]);
  }
  pl.ch_off.assign(pl.Lmax, 0); pl.ch_cnt.assign(pl.Lmax, 0);
  pl.rc_off.assign(pl.Lmax, 0); pl.rc_cnt.assign(pl.Lmax, 0);
  pl.fc_off.assign(pl.Lmax, 0); pl.fc_cnt.assign(pl.Lmax, 0);
  // chunks of every merge: 256-wide, RCH-wide (roots) and FILLC-wide (columns of S), each list
  // (merge, chunk) pairs merge by merge; sized first and written in place
  std::vector<int> sizes_here;
  for (int depth = pl.Lmax - 1; depth >= 0; depth--) {
    sizes_here.resize(pl.mrg_cnt[depth]);
    size_t nch = 0, nrc = 0, nfc = 0;
    for (size_t mi = 0; mi < pl.mrg_cnt[depth]; mi++) {
      const int b = I[pl.mrg_off[depth] + 2 * mi], q = I[pl.mrg_off[depth] + 2 * mi + 1];
      const int m = (int)((((int64_t)(q + 1) * n[b]) >> depth) - (((int64_t)q * n[b]) >> depth));
      sizes_here[mi] = m;
      nch += (m + 255) / 256;
      nrc += (m + RCH - 1) / RCH;
      nfc += (m + FILLC - 1) / FILLC;
    }
    pl.ch_off[depth] = I.size();
    pl.ch_cnt[depth] = nch;
    pl.rc_off[depth] = pl.ch_off[depth] + 2 * nch;
    pl.rc_cnt[depth] = nrc;
    pl.fc_off[depth] = pl.rc_off[depth] + 2 * nrc;
    pl.fc_cnt[depth] = nfc;
    I.resize(pl.fc_off[depth] + 2 * nfc);
    int *wch = I.data() + pl.ch_off[depth], *wrc = I.data() + pl.rc_off[depth], *wfc = I.data() + pl.fc_off[depth];
    for (size_t mi = 0; mi < sizes_here.size(); mi++) {
      const int m = sizes_here[mi];
      for (int c = 0; c < (m + 255) / 256; c++) { *wch++ = (int)mi; *wch++ = c; }
      for (int c = 0; c < (m + RCH - 1) / RCH; c++) { *wrc++ = (int)mi; *wrc++ = c; }
      for (int c = 0; c < (m + FILLC - 1) / FILLC; c++) { *wfc++ = (int)mi; *wfc++ = c; }
    }
  }
  // FP64 back-transformation: super-panel offsets, Y columns and T factors
  pl.toff.assign(nblk, 0);
  pl.sumTb = 0; pl.sumW2 = 0; pl.Smax = 0;
  for (int b = 0; b < nblk; b++) {
    const int sb = bt_panels(pl.bd[b]);
    pl.toff[b] = pl.sumTb;
    pl.sumTb += (int64_t)sb * KBT * KBT;
    pl.sumW2 += (int64_t)KBT * n[b];
    pl.Smax = std::max(pl.Smax, sb);
  }
  pl.ym_off = I.size();
  for (int b : order)
    if (bt_panels(pl.bd[b]))
      for (int c = 0; c < (int)((n[b] - pl.bd[b].roff + 63) / 64); c++) push2(b, c);
  pl.ym_cnt = (I.size() - pl.ym_off) / 2;
  pl.tb_off = I.size();
  for (int b : order)
    for (int sp = 0; sp < bt_panels(pl.bd[b]); sp++) push2(b, sp);
  pl.tb_cnt = (I.size() - pl.tb_off) / 2;
  pl.cp_off = I.size();
  for (int b : order)
    for (int64_t c = 0; c < (n[b] * n[b] + COPY_CHUNK - 1) / COPY_CHUNK; c++) push2(b, (int)c);
  pl.cp_cnt = (I.size() - pl.cp_off) / 2;
}

// Products of the D&C merges at each depth (FP64, pointers into the arenas); appends items to pl.ibuf.
static void build_dc_gemms(SyevPlan &pl, double *Z0, double *Z1, double *S) {
  pl.gd.clear();
  std::vector<int> &I = pl.ibuf;
  for (int depth = pl.Lmax - 1; depth >= 0; depth--) {
    const size_t d0 = pl.gd.size();
    const int *it = I.data() + pl.mrg_off[depth];
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
        GemmDesc g{};
        g.A = Qin + (size_t)(a0 + ro) * nb + (a0 + ro); g.lda = nb; g.tA = 0;
        g.B = Sb + (size_t)a0 * nb + (a0 + ro); g.ldb = nb; g.tB = 0;
        g.C = Qout + (size_t)a0 * nb + (a0 + ro); g.ldc = nb;
        g.M = mh; g.N = m; g.K = mh; g.K2 = 0; g.alpha = 1.0; g.beta = 0.0;
        pl.gd.push_back(g);
      }
    }
    pl.gemm_off[depth] = I.size();
    gemm_items(pl.gd, d0, pl.gd.size(), I);
    pl.gemm_cnt[depth] = (I.size() - pl.gemm_off[depth]) / 2;
  }
}

// Products of the reduction (trailing rank-2k updates) and of the back-transformation in the
// work precision T; appends items to pl.ibuf.
template <typename T>
static void build_red_bt_gemms(SyevPlan &pl, std::vector<GemmDescT<T>> &gd, T *Abase, T *Wpbase, T *Ym, T *Yh, T *Z,
                               T *Gb, T *Tb, T *W2, T *sbV, T *sbY, T *sbW) {
  gd.clear();
  std::vector<int> &I = pl.ibuf;
  // stage 1 of the two-stage reduction, per panel slot: Y = A22 V, then A22 -= V W^T + W V^T (every
  // tile: the next panel's product reads the whole symmetric A22)
  pl.sby_items.assign(pl.P1max, 0); pl.sby_cnt.assign(pl.P1max, 0);
  pl.sbu_items.assign(pl.P1max, 0); pl.sbu_cnt.assign(pl.P1max, 0);
  for (int q = 0; q < pl.P1max; q++)
    for (int ph = 0; ph < 2; ph++) {
      const size_t d0 = gd.size();
      for (size_t b = 0; b < pl.bd.size(); b++) {
        const BlkDesc &d = pl.bd[b];
        if (!d.two || sb_panels((int)d.n) <= q) continue;
        const int n = (int)d.n, k0 = q * SB, R = n - k0 - SB;
        T *A22 = Abase + d.offA + (size_t)(k0 + SB) * n + (k0 + SB);
        GemmDescT<T> g{};
        if (ph == 0) {
          g.A = A22; g.lda = n; g.B = sbV + d.offWp; g.ldb = n; g.K = R;
          g.C = sbY + d.offWp; g.ldc = n;
          g.M = R; g.N = SB; g.alpha = 1.0; g.beta = 0.0;
        } else {
          g.A = sbV + d.offWp; g.lda = n; g.B = sbW + d.offWp; g.ldb = n; g.tB = 1; g.K = SB;
          g.A2 = sbW + d.offWp; g.lda2 = n; g.B2 = sbV + d.offWp; g.ldb2 = n; g.tB2 = 1; g.K2 = SB;
          g.C = A22; g.ldc = n;
          g.M = R; g.N = R; g.alpha = -1.0; g.beta = 1.0;
        }
        gd.push_back(g);
      }
      size_t &io = ph == 0 ? pl.sby_items[q] : pl.sbu_items[q];
      size_t &ic = ph == 0 ? pl.sby_cnt[q] : pl.sbu_cnt[q];
      io = I.size();
      gemm_items(gd, d0, gd.size(), I);
      ic = (I.size() - io) / 2;
    }
  // trailing updates A22 -= V W^T + W V^T after every panel that is not a block's last
  pl.s2_items.assign((size_t)pl.Pmax * pl.ngrp, 0); pl.s2_cnt.assign((size_t)pl.Pmax * pl.ngrp, 0);
  for (int q = 0; q < pl.Pmax; q++)   // panel slot q, as in build_plan
  for (int gr = 0; gr < pl.ngrp; gr++) {
    const size_t d0 = gd.size();
    for (size_t b = 0; b < pl.bd.size(); b++) {
      const BlkDesc &d = pl.bd[b];
      if (pl.grp[b] != gr) continue;
      const int p = panel_in_slot(q, d.P);
      if (p < 0) continue;
      const int n = (int)d.n, i0 = p * NB, mt = n - i0 - NB;
      GemmDescT<T> g{};
      T *Ab = Abase + d.offA;
      const T *V = Ab + (size_t)i0 * n + (i0 + NB), *Wt = Wpbase + d.offWp + NB;
      g.A = V; g.lda = n; g.tA = 0; g.B = Wt; g.ldb = n; g.tB = 1; g.K = NB;
      g.A2 = Wt; g.lda2 = n; g.tA2 = 0; g.B2 = V; g.ldb2 = n; g.tB2 = 1; g.K2 = NB;
      g.C = Ab + (size_t)(i0 + NB) * n + (i0 + NB); g.ldc = n;
      g.M = mt; g.N = mt; g.alpha = -1.0; g.beta = 1.0;
      gd.push_back(g);
    }
    // only the tiles on or below the first tile super-diagonal: the reduction reads the lower
    // triangle, and the symmetric product's diagonal 64x64 tiles reach at most 63 columns above it
    const size_t qg = (size_t)q * pl.ngrp + gr;
    pl.s2_items[qg] = I.size();
    for (size_t i = d0; i < gd.size(); i++) {
      const int tm = (gd[i].M + GEMM_TILE - 1) / GEMM_TILE;
      const int tiles = tm * ((gd[i].N + GEMM_TILE - 1) / GEMM_TILE);
      for (int k = 0; k < tiles; k++)
        if (k / tm <= k % tm + 1) {
          I.push_back((int)i);
          I.push_back(k);
        }
    }
    pl.s2_cnt[qg] = (I.size() - pl.s2_items[qg]) / 2;
  }
  // back-transformation products
  const int nblk = (int)pl.bd.size();
  auto nsp = [&](int b) { return bt_panels(pl.bd[b]); };
  size_t d0 = gd.size();
  for (int b = 0; b < nblk; b++) {
    const BlkDesc &d = pl.bd[b];
    const int n = (int)d.n;
    for (int sp = 0; sp < nsp(b); sp++) {
      const int kb = std::min(KBT, (n - d.roff) - sp * KBT), r0 = sp * KBT + d.roff;
      GemmDescT<T> g{};
      g.A = Ym + d.offZ + (size_t)sp * KBT * n + r0; g.lda = n; g.tA = 1;
      g.B = g.A; g.ldb = n; g.tB = 0;
      g.C = Gb + pl.toff[b] + (size_t)sp * KBT * KBT; g.ldc = kb;
      g.M = kb; g.N = kb; g.K = n - r0; g.alpha = 1.0; g.beta = 0.0;
      gd.push_back(g);
    }
  }
  pl.g_items = I.size();
  gemm_items(gd, d0, gd.size(), I);
  pl.g_cnt = (I.size() - pl.g_items) / 2;
  d0 = gd.size();
  for (int b = 0; b < nblk; b++) {
    const BlkDesc &d = pl.bd[b];
    const int n = (int)d.n;
    for (int sp = 0; sp < nsp(b); sp++) {
      const int kb = std::min(KBT, (n - d.roff) - sp * KBT), r0 = sp * KBT + d.roff;
      GemmDescT<T> g{};
      g.A = Ym + d.offZ + (size_t)sp * KBT * n + r0; g.lda = n; g.tA = 0;
      g.B = Tb + pl.toff[b] + (size_t)sp * KBT * KBT; g.ldb = kb; g.tB = 1;
      g.C = Yh + d.offZ + (size_t)sp * KBT * n + r0; g.ldc = n;
      g.M = n - r0; g.N = kb; g.K = kb; g.alpha = 1.0; g.beta = 0.0;
      gd.push_back(g);
    }
  }
  pl.yh_items = I.size();
  gemm_items(gd, d0, gd.size(), I);
  pl.yh_cnt = (I.size() - pl.yh_items) / 2;
  pl.w_items.assign(pl.Smax, 0); pl.w_cnt.assign(pl.Smax, 0);
  pl.u_items.assign(pl.Smax, 0); pl.u_cnt.assign(pl.Smax, 0);
  for (int step = 0; step < pl.Smax; step++) {
    for (int phase = 0; phase < 2; phase++) {
      d0 = gd.size();
      int64_t w2o = 0;
      for (int b = 0; b < nblk; b++) {
        const BlkDesc &d = pl.bd[b];
        const int n = (int)d.n;
        const int S_b = nsp(b);
        T *W2b = W2 + w2o;
        w2o += (int64_t)KBT * n;
        if (S_b <= step) continue;
        const int sp = S_b - 1 - step;
        const int kb = std::min(KBT, (n - d.roff) - sp * KBT), r0 = sp * KBT + d.roff;
        GemmDescT<T> g{};
        if (phase == 0) {   // W = Yhat_s^T Z(r0:, :)
          g.A = Yh + d.offZ + (size_t)sp * KBT * n + r0; g.lda = n; g.tA = 1;
          g.B = Z + d.offZ + r0; g.ldb = n; g.tB = 0;
          g.C = W2b; g.ldc = kb;
          g.M = kb; g.N = n; g.K = n - r0; g.alpha = 1.0; g.beta = 0.0;
        } else {            // Z(r0:, :) -= Y_s W
          g.A = Ym + d.offZ + (size_t)sp * KBT * n + r0; g.lda = n; g.tA = 0;
          g.B = W2b; g.ldb = kb; g.tB = 0;
          g.C = Z + d.offZ + r0; g.ldc = n;
          g.M = n - r0; g.N = n; g.K = kb; g.alpha = -1.0; g.beta = 1.0;
        }
        gd.push_back(g);
      }
      size_t &io = phase == 0 ? pl.w_items[step] : pl.u_items[step];
      size_t &ic = phase == 0 ? pl.w_cnt[step] : pl.u_cnt[step];
      io = I.size();
      gemm_items(gd, d0, gd.size(), I);
      ic = (I.size() - io) / 2;
    }
  }
}

struct SyevLaunch {
  const int *dI;      // device copy of ibuf
  const GemmDesc *dG;   // device copy of the FP64 product descriptors (D&C)
  const void *dGT;      // device copy of the reduction/back-transformation descriptors (work precision)
  const int64_t *toff; // device copy of the per-block super-panel offsets
  MergeState *mst;    // per-merge state of the current depth
};

template <typename T>
static void launch_gemm(nrg_ctx *ctx, hipStream_t st, const void *descs, const int *items, const int *items_host,
                        size_t count) {
  if (!count) return;
  if constexpr (sizeof(T) == 8) gemm_launch(ctx, st, (const GemmDesc *)descs, items, items_host, count);
  else gemm_launch(ctx, st, (const GemmDesc32 *)descs, items, items_host, count);
}

// Launch sequence of the reduction + D&C + back-transformation for a typed work matrix.
template <typename T>
static int launch_core(nrg_ctx *ctx, SyevPlan &pl, SyevArgs<T> &aT, SyevArgs<double> &aD, const SyevLaunch &L,
                       hipStream_t st, bool zt_is_z0, int phase0) {
  const int *I = L.dI;
  const int *H = pl.ibuf.data();   // host copy of the work lists (GEMM lists are split by tile shape)
  auto grid = [](size_t c) { return dim3((unsigned)c); };
  // 1. symmetrisation
  if (pl.symm_cnt)
    NRG_TIMED(ctx, st, "k_absmax", hipLaunchKernelGGL(k_absmax<T>, grid(pl.symm_cnt), dim3(256), 0, st, aT, (const int2 *)(I + pl.symm_off)));
  if (pl.symm_cnt)
    NRG_TIMED(ctx, st, "k_symm", hipLaunchKernelGGL(k_symm<T>, grid(pl.symm_cnt), dim3(256), 0, st, aT, (const int2 *)(I + pl.symm_off), 0));
  // 2. reduction: group g's chain on stream S[g], the two-stage chain on Ssb, forked from and joined
  //    back into st
  const int NG = pl.ngrp;
  hipStream_t S[1 + nrg_ctx::NAUX] = {st};
  for (int g = 1; g < NG; g++) S[g] = ctx->aux[g - 1];
  hipStream_t Ssb = ctx->aux[nrg_ctx::NAUX - 1];
  const bool side = pl.nsb || pl.wg0_cnt;   // work on Ssb beside the chains
  if (NG > 1 || side) {
    NRG_HIP(hipEventRecord(ctx->fork_ev, st));
    for (int g = 1; g < NG; g++) NRG_HIP(hipStreamWaitEvent(S[g], ctx->fork_ev, 0));
    if (side) NRG_HIP(hipStreamWaitEvent(Ssb, ctx->fork_ev, 0));
  }
  if (pl.wg0_cnt)
    NRG_TIMED(ctx, Ssb, "k_wg_red", hipLaunchKernelGGL(k_wg_red<T>, grid(pl.wg0_cnt), dim3(WGR_MAX), 0, Ssb, aT,
                                                       I + pl.wg0_off));
  if (pl.nsb) {
    for (int q = 0; q < pl.P1max; q++) {
      const int *lst = I + pl.sbp_off[q];
      NRG_TIMED(ctx, Ssb, "k_sb_panel", hipLaunchKernelGGL(k_sb_panel<T>, grid(pl.sbp_cnt[q]), dim3(NT_PANEL), 0, Ssb,
                                                           aT, lst));
      launch_gemm<T>(ctx, Ssb, L.dGT, I + pl.sby_items[q], H + pl.sby_items[q], pl.sby_cnt[q]);
      NRG_TIMED(ctx, Ssb, "k_sb_w", hipLaunchKernelGGL(k_sb_w<T>, grid(pl.sbp_cnt[q]), dim3(NT_SBW), 0, Ssb, aT, lst));
      launch_gemm<T>(ctx, Ssb, L.dGT, I + pl.sbu_items[q], H + pl.sbu_items[q], pl.sbu_cnt[q]);
    }
    NRG_TIMED(ctx, Ssb, "k_sb_toband", hipLaunchKernelGGL(k_sb_toband<T>, grid(pl.sbb_cnt), dim3(256), 0, Ssb, aT,
                                                          (const int2 *)(I + pl.sbb_off)));
    NRG_TIMED(ctx, Ssb, "k_sb_chase", hipLaunchKernelGGL(k_sb_chase<T>, grid(pl.sbc_cnt), dim3(SB_WAVES * WAVE), 0, Ssb,
                                                         aT, I + pl.sbc_off));
  }
  size_t ci = 0;
  for (int q = 0; q < pl.Pmax; q++) {
    for (; ci < pl.cols.size() && pl.cols[ci].p == q; ci++) {
      const SyevPlan::ColLaunch &c = pl.cols[ci];
      hipStream_t sg = S[c.g];
      if (c.yred_cnt)
        NRG_TIMED(ctx, sg, "k_col_yred", hipLaunchKernelGGL(k_col_yred<T>, grid(c.yred_cnt), dim3(NT_YR), 0, sg, aT,
                                                            (const int2 *)(I + c.yred_off), c.jl));
      if (c.rows <= 256)
        NRG_TIMED(ctx, sg, "k_col_prep", hipLaunchKernelGGL((k_col_prep<T, 1, 256>), grid(c.prep_cnt), dim3(256), 0, sg, aT,
                                                            I + c.prep_off, c.jl));
      else if (c.rows <= 512)
        NRG_TIMED(ctx, sg, "k_col_prep", hipLaunchKernelGGL((k_col_prep<T, 1, 512>), grid(c.prep_cnt), dim3(512), 0, sg, aT,
                                                            I + c.prep_off, c.jl));
      else if (c.rows <= NT_PANEL)
        NRG_TIMED(ctx, sg, "k_col_prep", hipLaunchKernelGGL((k_col_prep<T, 1>), grid(c.prep_cnt), dim3(NT_PANEL), 0, sg, aT,
                                                            I + c.prep_off, c.jl));
      else if (c.rows <= 4 * NT_PANEL)
        NRG_TIMED(ctx, sg, "k_col_prep", hipLaunchKernelGGL((k_col_prep<T, 4>), grid(c.prep_cnt), dim3(NT_PANEL), 0, sg, aT,
                                                            I + c.prep_off, c.jl));
      else
        NRG_TIMED(ctx, sg, "k_col_prep", hipLaunchKernelGGL((k_col_prep<T, NMAX / NT_PANEL>), grid(c.prep_cnt), dim3(NT_PANEL),
                                                            0, sg, aT, I + c.prep_off, c.jl));
      if (c.tiles)
        NRG_TIMED(ctx, sg, "k_col_symv", hipLaunchKernelGGL(k_col_symv<T>, grid(c.tiles), dim3(256), 0, sg, aT,
                                                            I + c.symv_off, I + c.tpre_off, (int)c.nact, c.jl));
    }
    for (int g = 0; g < NG; g++) {
      const size_t qg = (size_t)q * NG + g;
      launch_gemm<T>(ctx, S[g], L.dGT, I + pl.s2_items[qg], H + pl.s2_items[qg], pl.s2_cnt[qg]);
    }
  }
  for (int g = 0; g < NG; g++)
    if (pl.wgg_cnt[g])
      NRG_TIMED(ctx, S[g], "k_wg_red", hipLaunchKernelGGL(k_wg_red<T>, grid(pl.wgg_cnt[g]), dim3(WGR_MAX), 0, S[g], aT,
                                                          I + pl.wgg_off[g]));
  if (NG > 1)
    for (int g = 1; g < NG; g++) {
      NRG_HIP(hipEventRecord(ctx->join_ev[g - 1], S[g]));
      NRG_HIP(hipStreamWaitEvent(st, ctx->join_ev[g - 1], 0));
    }
  if (side) {
    NRG_HIP(hipEventRecord(ctx->join_ev[nrg_ctx::NAUX - 1], Ssb));
    NRG_HIP(hipStreamWaitEvent(st, ctx->join_ev[nrg_ctx::NAUX - 1], 0));
  }
  if (pl.tail_cnt)
    NRG_TIMED(ctx, st, "k_tail", hipLaunchKernelGGL(k_tail<T>, grid(pl.tail_cnt), dim3(NT_TAIL), 0, st, aT, I + pl.tail_off));
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[phase0 + 1], st));
  // 3. divide and conquer
  {
    int mc = 0;
    for (int depth = 0; depth < pl.Lmax; depth++) mc = std::max(mc, pl.mcap[depth]);
    const int lds3 = (int)(3 * sizeof(double) * (size_t)mc + 16), lds2 = (int)(2 * sizeof(double) * (size_t)mc + 16);
    (void)hipFuncSetAttribute((const void *)k_dc_deflate<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              std::min(dc_deflate_lds(mc), DYN_LDS_MAX));
    (void)hipFuncSetAttribute((const void *)k_dc_roots, hipFuncAttributeMaxDynamicSharedMemorySize, lds2);
    (void)hipFuncSetAttribute((const void *)k_dc_zhat<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              std::min(lds3, DYN_LDS_MAX));
    (void)hipFuncSetAttribute((const void *)k_dc_normrank<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              std::min(lds3, DYN_LDS_MAX));
    (void)hipFuncSetAttribute((const void *)k_dc_fill<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lds2);
    (void)hipFuncSetAttribute((const void *)k_dc_fill<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              std::min(fill_lds(mc), DYN_LDS_MAX));
  }
  // merges whose deflation does not fit the LDS keep its arrays in global memory, one region each
  double *dc_gscr = nullptr;
  int64_t dc_gstride = 0;
  {
    size_t need = 0;
    for (int depth = 0; depth < pl.Lmax; depth++)
      if (pl.mrg_cnt[depth] && dc_deflate_lds(pl.mcap[depth]) > DYN_LDS_MAX) {
        dc_gstride = std::max<int64_t>(dc_gstride, (dc_deflate_lds(pl.mcap[depth]) + 7) / 8);
        need = std::max(need, (size_t)pl.mrg_cnt[depth]);
      }
    if (need) {
      if (int r = ctx->dcbuf.ensure(sizeof(double) * need * (size_t)dc_gstride)) return r;
      dc_gscr = ctx->dcbuf.as<double>();
    }
  }
  NRG_TIMED(ctx, st, "k_dc_prep", hipLaunchKernelGGL(k_dc_prep, grid(pl.bd.size()), dim3(256), 0, st, aD, I + pl.all_off));
  for (int depth = pl.Lmax - 1; depth >= 0; depth--) {
    if (!pl.mrg_cnt[depth]) continue;
    const int mc = pl.mcap[depth];
    const size_t lds2 = 2 * sizeof(double) * (size_t)mc + 16, lds3 = 3 * sizeof(double) * (size_t)mc + 16;
    const int2 *mi = (const int2 *)(I + pl.mrg_off[depth]);
    const int2 *ch = (const int2 *)(I + pl.ch_off[depth]);
    const int2 *fc = (const int2 *)(I + pl.fc_off[depth]);
    const bool dg = dc_deflate_lds(mc) > DYN_LDS_MAX, lg = (int)lds3 > DYN_LDS_MAX;
    if (dg)
      NRG_TIMED(ctx, st, "k_dc_deflate", hipLaunchKernelGGL(k_dc_deflate<true>, grid(pl.mrg_cnt[depth]), dim3(NT_DEF), 0,
                                                            st, aD, mi, depth, L.mst, dc_gscr, dc_gstride));
    else
      NRG_TIMED(ctx, st, "k_dc_deflate", hipLaunchKernelGGL(k_dc_deflate<false>, grid(pl.mrg_cnt[depth]), dim3(NT_DEF),
                                                            dc_deflate_lds(mc), st, aD, mi, depth, L.mst, nullptr,
                                                            dc_gstride));
    NRG_TIMED(ctx, st, "k_dc_roots", hipLaunchKernelGGL(k_dc_roots, grid(pl.rc_cnt[depth]), dim3(256), lds2, st, aD, mi,
                                                        (const int2 *)(I + pl.rc_off[depth]), depth, (const MergeState *)L.mst));
    if (lg) {
      NRG_TIMED(ctx, st, "k_dc_zhat", hipLaunchKernelGGL(k_dc_zhat<false>, grid(pl.ch_cnt[depth]), dim3(256), 0, st, aD,
                                                         mi, ch, depth, (const MergeState *)L.mst));
      NRG_TIMED(ctx, st, "k_dc_normrank", hipLaunchKernelGGL(k_dc_normrank<false>, grid(pl.ch_cnt[depth]), dim3(256), 0,
                                                             st, aD, mi, ch, depth, (const MergeState *)L.mst));
    } else {
      NRG_TIMED(ctx, st, "k_dc_zhat", hipLaunchKernelGGL(k_dc_zhat<true>, grid(pl.ch_cnt[depth]), dim3(256), lds3, st, aD,
                                                         mi, ch, depth, (const MergeState *)L.mst));
      NRG_TIMED(ctx, st, "k_dc_normrank", hipLaunchKernelGGL(k_dc_normrank<true>, grid(pl.ch_cnt[depth]), dim3(256), lds3,
                                                             st, aD, mi, ch, depth, (const MergeState *)L.mst));
    }
    if (fill_lds(mc) <= DYN_LDS_MAX)
      NRG_TIMED(ctx, st, "k_dc_fill", hipLaunchKernelGGL(k_dc_fill<true>, grid(pl.fc_cnt[depth]), dim3(256), fill_lds(mc), st,
                                                         aD, mi, fc, depth, (const MergeState *)L.mst));
    else
      NRG_TIMED(ctx, st, "k_dc_fill", hipLaunchKernelGGL(k_dc_fill<false>, grid(pl.fc_cnt[depth]), dim3(256), lds2, st, aD,
                                                         mi, fc, depth, (const MergeState *)L.mst));
    if (pl.gemm_cnt[depth]) gemm_launch(ctx, st, L.dG, I + pl.gemm_off[depth], H + pl.gemm_off[depth], pl.gemm_cnt[depth]);
  }
  NRG_HIP(hipGetLastError());
  NRG_HIP(hipEventRecord(ctx->ev[phase0 + 2], st));
  // 4. back-transformation
  if (!zt_is_z0) {
    const size_t cnt = (size_t)pl.sumn2;
    NRG_TIMED(ctx, st, "k_convert", hipLaunchKernelGGL(k_convert<T>, dim3(2048), dim3(256), 0, st, (const double *)aD.Z0, aT.Zt, cnt));
  }
  if (pl.ym_cnt)
    NRG_TIMED(ctx, st, "k_bt_ymat", hipLaunchKernelGGL(k_bt_ymat<T>, grid(pl.ym_cnt), dim3(256), 0, st, aT, (const int2 *)(I + pl.ym_off)));
  launch_gemm<T>(ctx, st, L.dGT, I + pl.g_items, H + pl.g_items, pl.g_cnt);
  if (pl.tb_cnt)
    NRG_TIMED(ctx, st, "k_bt_tbig", hipLaunchKernelGGL(k_bt_tbig<T>, grid(pl.tb_cnt), dim3(256), 0, st, aT, (const int2 *)(I + pl.tb_off), L.toff));
  launch_gemm<T>(ctx, st, L.dGT, I + pl.yh_items, H + pl.yh_items, pl.yh_cnt);
  // two-stage blocks: the bulge-chasing reflectors first (V = Q1 (Q2 Z))
  if (pl.sbq_cnt)
    NRG_TIMED(ctx, st, "k_sb_q2", hipLaunchKernelGGL(k_sb_q2<T>, grid(pl.sbq_cnt), dim3(WAVE), 0, st, aT,
                                                    (const int2 *)(I + pl.sbq_off)));
  for (int step = 0; step < pl.Smax; step++) {
    launch_gemm<T>(ctx, st, L.dGT, I + pl.w_items[step], H + pl.w_items[step], pl.w_cnt[step]);
    launch_gemm<T>(ctx, st, L.dGT, I + pl.u_items[step], H + pl.u_items[step], pl.u_cnt[step]);
  }
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

// ---------------------------------------------------------------------------------------
// Graph replay, with NRG_SYEV_GRAPH=1 in the environment (read at every call). A call whose key
// (form, block orders and offsets, the caller's arenas and stream, the context's workspaces, the
// launch-shaping settings) equals the previous call's is captured into a HIP graph; later calls with
// that key launch the graph and skip the host work (planning, upload, ~2000 launches). The launches
// clear memory with fill kernels, not hipMemsetAsync. Off by default: on ROCm 7.2 the replay of the
// 256-block call takes the same GPU time as the eager launches (30.95 against 30.97 ms) and ~4 ms
// more host time before its first kernel starts.
// ---------------------------------------------------------------------------------------
static bool graphs_on() {
  const char *e = std::getenv("NRG_SYEV_GRAPH");
  return e && e[0] == '1';
}

static std::vector<int64_t> syev_key(nrg_ctx *ctx, int form, int nblk, const int64_t *n, const int64_t *off_H,
                                     const int64_t *off_w, const double *H, const double *w, hipStream_t st,
                                     int setting, int groups) {
  std::vector<int64_t> k;
  k.reserve(3 * (size_t)nblk + 16);
  k.push_back(form); k.push_back(nblk); k.push_back(groups); k.push_back(two_stage_on()); k.push_back(setting);
  k.push_back(wg_red_order());
  for (const void *q : {(const void *)H, (const void *)w, (const void *)st, (const void *)ctx->ws.p,
                        (const void *)ctx->ws2.p, (const void *)ctx->dcbuf.p, (const void *)ctx->flags.p})
    k.push_back((int64_t)(intptr_t)q);
  k.push_back((int64_t)ctx->ws.bytes); k.push_back((int64_t)ctx->ws2.bytes); k.push_back((int64_t)ctx->dcbuf.bytes);
  k.insert(k.end(), n, n + nblk);
  k.insert(k.end(), off_H, off_H + nblk);
  k.insert(k.end(), off_w, off_w + nblk);
  return k;
}

template <typename T>
__global__ __launch_bounds__(256) void k_zero(T *p, size_t count) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) p[i] = (T)0;
}
template <typename T>
static void zero_fill(hipStream_t st, T *p, size_t count) {
  hipLaunchKernelGGL(k_zero<T>, dim3((unsigned)std::min<size_t>((count + 255) / 256, 1024)), dim3(256), 0, st, p, count);
}

// What a call does with the graph cache: launch a cached graph, capture this call, or run eagerly.
enum class GraphUse { eager, capture };

// Before planning: launches the cached graph if the key matches (*done = true, *rc = its status);
// otherwise decides whether this call is captured (the previous call had the same key).
static GraphUse graph_begin(nrg_ctx *ctx, SyevCache &c, const std::vector<int64_t> &key, bool on, bool *done) {
  *done = false;
  if (!on) {
    c.last_key.clear();
    return GraphUse::eager;
  }
  if (c.graph && c.graph->key == key) {
    *done = true;
    return GraphUse::eager;
  }
  const bool again = key == c.last_key;
  c.last_key = key;
  return again ? GraphUse::capture : GraphUse::eager;
}

static int finish_flags(nrg_ctx *ctx, int nblk, int *info, hipStream_t st) {
  // the kernels complete before the copy is issued (a DMA copy was seen to overtake kernels queued
  // before it under rocprofv3's kernel tracing)
  NRG_HIP(hipStreamSynchronize(st));
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

// Phase times of the last call from its events ev[0..nph] (phase k between ev[k] and ev[k+1]).
static void record_phase_times(nrg_ctx *ctx, int nph) {
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[nph]);
  ctx->total_ms = ms;
  for (int ph = 0; ph < nph; ph++) {
    ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[ph], ctx->ev[ph + 1]);
    ctx->phase_ms[ph] = ms;
  }
}

// Launches the cached graph of this key and collects the call's status words and phase times.
static int graph_replay(nrg_ctx *ctx, SyevCache &c, int nblk, int *info, hipStream_t st) {
  NRG_HIP(hipEventRecord(ctx->gfork, st));
  NRG_HIP(hipStreamWaitEvent(ctx->gstream, ctx->gfork, 0));
  NRG_HIP(hipGraphLaunch(c.graph->exec, ctx->gstream));
  NRG_HIP(hipEventRecord(ctx->gjoin, ctx->gstream));
  NRG_HIP(hipStreamWaitEvent(st, ctx->gjoin, 0));
  const int r = finish_flags(ctx, nblk, info, st);
  record_phase_times(ctx, c.graph->nphase);
  return r;
}

// Copies the descriptor blob staged in ctx->hdesc into a new graph's own buffer and starts capturing
// the context's graph stream, after the caller's stream st; *dbase is where the launches read the
// tables from. The call's launches then go to ctx->gstream.
static int graph_capture_start(nrg_ctx *ctx, SyevCache &c, const std::vector<int64_t> &key, size_t bytes,
                               hipStream_t st, char **dbase) {
  auto g = std::make_unique<SyevGraph>();
  g->key = key;
  if (int r = g->desc.ensure(bytes)) return r;
  NRG_HIP(hipMemcpyAsync(g->desc.p, ctx->hdesc.p, bytes, hipMemcpyHostToDevice, st));
  NRG_HIP(hipStreamSynchronize(st));
  NRG_HIP(hipStreamBeginCapture(ctx->gstream, hipStreamCaptureModeRelaxed));
  *dbase = (char *)g->desc.p;
  c.graph = std::move(g);
  return NRG_OK;
}

// Ends the capture started by graph_capture_start and launches the graph. The graph is kept for
// later calls when this call's key still describes it (no workspace moved during the call). A
// failure to capture or instantiate drops the graph and turns capture off for this context; the
// call is then repeated eagerly by the caller (*retry).
static int graph_capture_end(nrg_ctx *ctx, SyevCache &c, const std::vector<int64_t> &key_now, int launch_rc, int nph,
                             hipStream_t st, bool *retry) {
  *retry = false;
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->gstream, &graph);
  hipError_t ei = hipErrorUnknown;
  if (e == hipSuccess && graph && launch_rc == NRG_OK) {
    ei = hipGraphInstantiate(&c.graph->exec, graph, nullptr, nullptr, 0);
  }
  if (graph) hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    (void)hipGetLastError();
    c.graph.reset();
    c.graph_failed = true;
    if (launch_rc != NRG_OK) return launch_rc;
    *retry = true;
    return NRG_OK;
  }
  c.graph->nphase = nph;
  NRG_HIP(hipGraphLaunch(c.graph->exec, ctx->gstream));
  NRG_HIP(hipEventRecord(ctx->gjoin, ctx->gstream));
  NRG_HIP(hipStreamWaitEvent(st, ctx->gjoin, 0));
  if (c.graph->key != key_now) {   // a workspace moved during the call: do not keep the graph
    NRG_HIP(hipStreamSynchronize(st));
    c.graph.reset();
  }
  return NRG_OK;
}

int syev_fp64_run(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H, const int64_t *off_w,
                  double *w, int *info, hipStream_t st) {
  SyevCache &cache = syev_cache(ctx);
  const TuneSig sig = tune_signature(NRG_SYEV_FP64, nblk, n);
  int cand = -1;
  const int groups = tune_choose(cache, sig, &cand);
  const bool gon = graphs_on() && !ctx->kt.on && !cache.graph_failed;
  std::vector<int64_t> key;
  if (gon) key = syev_key(ctx, NRG_SYEV_FP64, nblk, n, off_H, off_w, H, w, st, 0, groups);
  bool replay = false;
  const GraphUse gu = graph_begin(ctx, cache, key, gon, &replay);
  if (replay) {
    const int r = graph_replay(ctx, cache, nblk, info, st);
    tune_record(cache, sig, cand, ctx->total_ms, nblk, n);
    return r;
  }
  for (int attempt = 0; attempt < 2; attempt++) {
    const bool capture = gu == GraphUse::capture && attempt == 0;
    SyevPlan &pl = cache.pl;
    build_plan(pl, nblk, n, off_H, off_w, true, 8, groups);
    Carver cv;
    const size_t oZ0 = cv.take<double>(pl.sumn2), oZ1 = cv.take<double>(pl.sumn2), oS = cv.take<double>(pl.sumn2);
    const size_t oWp = cv.take<double>(pl.sumWp), oTau = cv.take<double>(pl.sumn);
    const size_t od = cv.take<double>(pl.sumn), oe = cv.take<double>(pl.sumn), ol = cv.take<double>(pl.sumn);
    const size_t osc = cv.take<double>(nblk), odscr = cv.take<double>(DS_COUNT * pl.sumn), oiscr = cv.take<int>(IS_COUNT * pl.sumn);
    const size_t ovb = cv.take<double>(pl.sumn), oYp = cv.take<double>(pl.sumYp);
    const size_t omst = cv.take<MergeState>(pl.mrg_max + 1), oamax = cv.take<double>(nblk);
    const size_t oGb = cv.take<double>(pl.sumTb + 1), oTb = cv.take<double>(pl.sumTb + 1), oW2 = cv.take<double>(pl.sumW2);
    const size_t sbw = pl.nsb ? (size_t)pl.sumWp : 1;
    const size_t osbV = cv.take<double>(sbw), osbY = cv.take<double>(sbw), osbW = cv.take<double>(sbw);
    const size_t osbT = cv.take<double>(pl.nsb ? (size_t)nblk * SB * SB : 1), osbAB = cv.take<double>(pl.sumAB + 1);
    const size_t osbV2 = cv.take<double>(pl.sumV2 + 1), osbT2 = cv.take<double>(pl.sumT2 + 1);
    const size_t ocs = cv.take<double>(pl.sumn), odpp = cv.take<double>(pl.sumn);
    if (int r = ctx->ws.ensure(cv.off)) return r;
    double *Z0 = ctx->ws.as<double>(oZ0), *Z1 = ctx->ws.as<double>(oZ1), *S = ctx->ws.as<double>(oS);
    build_dc_gemms(pl, Z0, Z1, S);
    std::vector<GemmDesc> &gdT = cache.g64;
    build_red_bt_gemms<double>(pl, gdT, H, ctx->ws.as<double>(oWp), S, Z1, Z0, ctx->ws.as<double>(oGb),
                               ctx->ws.as<double>(oTb), ctx->ws.as<double>(oW2), ctx->ws.as<double>(osbV),
                               ctx->ws.as<double>(osbY), ctx->ws.as<double>(osbW));
    // descriptor upload: [BlkDesc][GemmDesc D&C][GemmDesc work][toff][ints]
    const size_t bBD = sizeof(BlkDesc) * pl.bd.size(), bP = sizeof(GemmDesc) * pl.gd.size(), bI = sizeof(int) * pl.ibuf.size();
    const size_t bQ = sizeof(GemmDesc) * gdT.size(), bT = sizeof(int64_t) * pl.toff.size();
    const size_t oBD = 0, oP = (bBD + 255) & ~(size_t)255, oQ = oP + ((bP + 255) & ~(size_t)255);
    const size_t oT = oQ + ((bQ + 255) & ~(size_t)255);
    const size_t oI = oT + ((bT + 255) & ~(size_t)255);
    if (int r = ctx->flags.ensure(sizeof(int) * nblk)) return r;
    if (int r = ctx->hflags.ensure(sizeof(int) * nblk)) return r;
    char *blob = nullptr;
    if (int r = stage_desc(ctx, oI + bI, st, &blob)) return r;
    memcpy(blob + oBD, pl.bd.data(), bBD);
    if (bP) memcpy(blob + oP, pl.gd.data(), bP);
    if (bQ) memcpy(blob + oQ, gdT.data(), bQ);
    if (bT) memcpy(blob + oT, pl.toff.data(), bT);
    memcpy(blob + oI, pl.ibuf.data(), bI);
    char *dbase = ctx->desc.as<char>(0);
    if (capture) {
      if (int r = graph_capture_start(ctx, cache, key, oI + bI, st, &dbase)) return r;
    } else if (int r = upload_staged(ctx, oI + bI, st)) {
      return r;
    }
    const hipStream_t rs = capture ? ctx->gstream : st;   // the stream the launches go to
    int lrc = NRG_OK;
    do {
      if (hipEventRecord(ctx->ev[0], rs) != hipSuccess) { lrc = NRG_E_HIP; break; }
      zero_fill(rs, (int *)ctx->flags.p, (size_t)nblk);
      zero_fill(rs, ctx->ws.as<double>(oamax), (size_t)nblk);
      SyevArgs<double> a;
      a.bd = (const BlkDesc *)(dbase + oBD);
      a.A = H;
      a.Wp = ctx->ws.as<double>(oWp); a.tau = ctx->ws.as<double>(oTau);
      a.vb = ctx->ws.as<double>(ovb); a.Yp = ctx->ws.as<double>(oYp);
      a.Zt = Z0; a.Z0 = Z0; a.Z1 = Z1; a.S = S;
      a.d = ctx->ws.as<double>(od); a.e = ctx->ws.as<double>(oe); a.lam = ctx->ws.as<double>(ol);
      a.scale = ctx->ws.as<double>(osc);
      a.dscr = ctx->ws.as<double>(odscr); a.iscr = ctx->ws.as<int>(oiscr); a.Ntot = pl.sumn;
      a.flags = (int *)ctx->flags.p;
      a.amax = ctx->ws.as<double>(oamax);
      a.H = H; a.w = w;
      SyevLaunch L;
      L.dI = (const int *)(dbase + oI);
      L.dG = (const GemmDesc *)(dbase + oP);
      L.dGT = (const void *)(dbase + oQ);
      L.toff = (const int64_t *)(dbase + oT);
      a.Gb = ctx->ws.as<double>(oGb); a.Tb = ctx->ws.as<double>(oTb);
      a.Ym = S; a.Yh = Z1;
      a.sbV = ctx->ws.as<double>(osbV); a.sbY = ctx->ws.as<double>(osbY); a.sbW = ctx->ws.as<double>(osbW);
      a.sbT = ctx->ws.as<double>(osbT); a.sbAB = ctx->ws.as<double>(osbAB);
      a.sbV2 = ctx->ws.as<double>(osbV2); a.sbT2 = ctx->ws.as<double>(osbT2);
      a.cs = ctx->ws.as<double>(ocs); a.dpp = ctx->ws.as<double>(odpp);
      L.mst = ctx->ws.as<MergeState>(omst);
      if ((lrc = launch_core<double>(ctx, pl, a, a, L, rs, true, 0))) break;
      NRG_TIMED(ctx, rs, "k_finish", hipLaunchKernelGGL(k_finish<double>, dim3((unsigned)nblk), dim3(256), 0, rs, a, L.dI + pl.all_off));
      NRG_TIMED(ctx, rs, "k_copy_out", hipLaunchKernelGGL(k_copy_out<double>, dim3((unsigned)pl.cp_cnt), dim3(256), 0, rs, a,
                                                          (const int2 *)(L.dI + pl.cp_off)));
      if (hipGetLastError() != hipSuccess || hipEventRecord(ctx->ev[4], rs) != hipSuccess) lrc = NRG_E_HIP;
    } while (false);
    if (capture) {
      bool retry = false;
      const std::vector<int64_t> key_now = syev_key(ctx, NRG_SYEV_FP64, nblk, n, off_H, off_w, H, w, st, 0, groups);
      if (int r = graph_capture_end(ctx, cache, key_now, lrc, 4, st, &retry)) return r;
      if (retry) continue;
    } else if (lrc) {
      return lrc;
    }
    const int r = finish_flags(ctx, nblk, info, st);
    if (ctx->kt.on) ctx->kt.collect();
    record_phase_times(ctx, 4);
    if (!ctx->kt.on) tune_record(cache, sig, cand, ctx->total_ms, nblk, n);
    return r;
  }
  set_error("nrg_syev_batched: graph capture failed");
  return NRG_E_HIP;
}


// ------------------------------------------------------------------------------------------
// Mixed form: FP32 reduction and back-transformation on the FP32 matrix cores, FP64 divide and
// conquer on the tridiagonal matrix, then NREF FP64 refinement sweeps.
// ------------------------------------------------------------------------------------------
// Sweeps of the mixed form. Two leave ~1e-11 eigenvalue errors on clustered spectra; on two-channel
// NRG blocks of order ~5500 (half the spectrum in exactly degenerate pairs, a run of 1145 estimates
// closer than the first sweep's delta) three leave eigenvalue errors of 1e-9 and normalisation errors of
// 3e-11, four reach 1e-13 and 1e-14.
constexpr int NREF_DEFAULT = 4;

// number of refinement sweeps: NRG_MIXED_SWEEPS in the environment (1..6), default NREF_DEFAULT
static int refine_sweeps() {
  const char *e = getenv("NRG_MIXED_SWEEPS");
  const int v = e ? atoi(e) : NREF_DEFAULT;
  return v < 1 ? 1 : (v > 6 ? 6 : v);
}

int syev_mixed_run(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H, const int64_t *off_w,
                   double *w, int *info, hipStream_t st) {
  SyevCache &cache = syev_cache(ctx);
  const TuneSig sig = tune_signature(NRG_SYEV_MIXED, nblk, n);
  int cand = -1;
  const int groups = tune_choose(cache, sig, &cand);
  const bool gon = graphs_on() && !ctx->kt.on && !cache.graph_failed;
  std::vector<int64_t> key;
  if (gon) key = syev_key(ctx, NRG_SYEV_MIXED, nblk, n, off_H, off_w, H, w, st, refine_sweeps(), groups);
  bool replay = false;
  const GraphUse gu = graph_begin(ctx, cache, key, gon, &replay);
  if (replay) {
    const int r = graph_replay(ctx, cache, nblk, info, st);
    tune_record(cache, sig, cand, ctx->total_ms, nblk, n);
    return r;
  }
  for (int attempt = 0; attempt < 2; attempt++) {
  const bool capture = gu == GraphUse::capture && attempt == 0;
  SyevPlan &pl = cache.pl;
  build_plan(pl, nblk, n, off_H, off_w, false, 4, groups);
  // per-block chunk items for the elementwise refinement kernels
  const size_t rc_off = pl.ibuf.size();
  for (int b = 0; b < nblk; b++)
    for (int64_t c = 0; c < (n[b] * n[b] + COPY_CHUNK - 1) / COPY_CHUNK; c++) { pl.ibuf.push_back(b); pl.ibuf.push_back((int)c); }
  const size_t rc_cnt = (pl.ibuf.size() - rc_off) / 2;
  const size_t col_off = pl.ibuf.size();
  for (int b = 0; b < nblk; b++)
    for (int64_t c = 0; c < (n[b] + 15) / 16; c++) { pl.ibuf.push_back(b); pl.ibuf.push_back((int)c); }
  const size_t col_cnt = (pl.ibuf.size() - col_off) / 2;
  // FP64 workspace
  Carver cv;
  const size_t oZ0 = cv.take<double>(pl.sumn2), oZ1 = cv.take<double>(pl.sumn2), oS = cv.take<double>(pl.sumn2);
  const size_t oAX = cv.take<double>(pl.sumn2), oR = cv.take<double>(pl.sumn2), oE = cv.take<double>(pl.sumn2);
  const size_t od = cv.take<double>(pl.sumn), oe = cv.take<double>(pl.sumn), ol = cv.take<double>(pl.sumn);
  const size_t olt = cv.take<double>(pl.sumn);
  const size_t osc = cv.take<double>(nblk), odscr = cv.take<double>(DS_COUNT * pl.sumn), oiscr = cv.take<int>(IS_COUNT * pl.sumn);
  const size_t omst = cv.take<MergeState>(pl.mrg_max + 1), oamax = cv.take<double>(nblk), oacc = cv.take<double>(3 * nblk);
  const size_t ccap = (size_t)pl.sumn / 2 + 1;   // clusters have two estimates or more and do not overlap
  const size_t ocl = cv.take<RefCluster>(ccap), oncl = cv.take<int>(1);
  const size_t orid = cv.take<int>(pl.sumn);
  if (int r = ctx->ws.ensure(cv.off)) return r;
  // FP32 workspace
  Carver cf;
  const size_t fA = cf.take<float>(pl.sumn2), fWp = cf.take<float>(pl.sumWp), fTau = cf.take<float>(pl.sumn);
  const size_t fvb = cf.take<float>(pl.sumn), fYp = cf.take<float>(pl.sumYp);
  const size_t fYm = cf.take<float>(pl.sumn2), fYh = cf.take<float>(pl.sumn2), fZt = cf.take<float>(pl.sumn2);
  const size_t fGb = cf.take<float>(pl.sumTb + 1), fTb = cf.take<float>(pl.sumTb + 1), fW2 = cf.take<float>(pl.sumW2);
  const size_t sbw = pl.nsb ? (size_t)pl.sumWp : 1;
  const size_t fsbV = cf.take<float>(sbw), fsbY = cf.take<float>(sbw), fsbW = cf.take<float>(sbw);
  const size_t fsbT = cf.take<float>(pl.nsb ? (size_t)nblk * SB * SB : 1), fsbAB = cf.take<float>(pl.sumAB + 1);
  const size_t fsbV2 = cf.take<float>(pl.sumV2 + 1), fsbT2 = cf.take<float>(pl.sumT2 + 1);
  const size_t fcs = cf.take<float>(pl.sumn), fdpp = cf.take<float>(pl.sumn);
  if (int r = ctx->ws2.ensure(cf.off)) return r;
  double *Z0 = ctx->ws.as<double>(oZ0), *Z1 = ctx->ws.as<double>(oZ1), *S = ctx->ws.as<double>(oS);
  build_dc_gemms(pl, Z0, Z1, S);
  std::vector<GemmDesc32> &gdT = cache.g32;
  float *A32 = ctx->ws2.as<float>(fA);
  build_red_bt_gemms<float>(pl, gdT, A32, ctx->ws2.as<float>(fWp), ctx->ws2.as<float>(fYm), ctx->ws2.as<float>(fYh),
                            ctx->ws2.as<float>(fZt), ctx->ws2.as<float>(fGb), ctx->ws2.as<float>(fTb),
                            ctx->ws2.as<float>(fW2), ctx->ws2.as<float>(fsbV), ctx->ws2.as<float>(fsbY),
                            ctx->ws2.as<float>(fsbW));
  // refinement products: AX = A X; S = X^T AX and R = I - X^T X; Xn = X (I + E). Both sweep
  // directions (X in Z0 -> Xn in Z1, then back) are prepared.
  std::vector<GemmDesc> &gref = cache.gref;
  double *Xbuf[2] = {Z0, Z1};
  double *AX = ctx->ws.as<double>(oAX), *Rm = ctx->ws.as<double>(oR), *Em = ctx->ws.as<double>(oE);
  size_t rg_off[2][4], rg_cnt[2][4];
  for (int dir = 0; dir < 2; dir++) {
    double *X = Xbuf[dir], *Xn = Xbuf[1 - dir];
    for (int ph = 0; ph < 4; ph++) {
      const size_t d0 = gref.size();
      for (int b = 0; b < nblk; b++) {
        const BlkDesc &d = pl.bd[b];
        const int nb = (int)d.n;
        GemmDesc g{};
        g.M = nb; g.N = nb; g.K = nb; g.ldc = nb; g.lda = nb; g.ldb = nb;
        if (ph == 0) {
          g.A = H + d.offH; g.B = X + d.offZ; g.C = AX + d.offZ; g.alpha = 1.0; g.beta = 0.0;
          gref.push_back(g);
        } else if (ph == 1) {
          g.A = X + d.offZ; g.tA = 1; g.B = AX + d.offZ; g.C = S + d.offZ; g.alpha = 1.0; g.beta = 0.0;
          gref.push_back(g);
          GemmDesc h{};
          h.M = nb; h.N = nb; h.K = nb; h.ldc = nb; h.lda = nb; h.ldb = nb;
          h.A = X + d.offZ; h.tA = 1; h.B = X + d.offZ; h.C = Rm + d.offZ; h.alpha = -1.0; h.beta = 1.0;
          gref.push_back(h);
        } else if (ph == 2) {
          g.A = X + d.offZ; g.B = Em + d.offZ; g.C = Xn + d.offZ; g.alpha = 1.0; g.beta = 0.0;
          gref.push_back(g);
        } else {   // E += K K / 2
          g.A = AX + d.offZ; g.B = AX + d.offZ; g.C = Em + d.offZ; g.alpha = 0.5; g.beta = 1.0;
          gref.push_back(g);
        }
      }
      rg_off[dir][ph] = pl.ibuf.size();
      gemm_items(gref, d0, gref.size(), pl.ibuf);
      rg_cnt[dir][ph] = (pl.ibuf.size() - rg_off[dir][ph]) / 2;
    }
  }
  // descriptor blob: [BlkDesc][D&C GemmDesc][FP32 GemmDesc][refinement GemmDesc][toff][ints]
  const size_t bBD = sizeof(BlkDesc) * pl.bd.size(), bP = sizeof(GemmDesc) * pl.gd.size();
  const size_t bQ = sizeof(GemmDesc32) * gdT.size(), bR = sizeof(GemmDesc) * gref.size();
  const size_t bT = sizeof(int64_t) * pl.toff.size(), bI = sizeof(int) * pl.ibuf.size();
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t oBD = 0, oP = al(bBD), oQ = oP + al(bP), oRr = oQ + al(bQ), oT = oRr + al(bR), oI = oT + al(bT);
  if (int r = ctx->flags.ensure(sizeof(int) * nblk)) return r;
  if (int r = ctx->hflags.ensure(sizeof(int) * nblk)) return r;
  char *blob = nullptr;
  if (int r = stage_desc(ctx, oI + bI, st, &blob)) return r;
  memcpy(blob + oBD, pl.bd.data(), bBD);
  if (bP) memcpy(blob + oP, pl.gd.data(), bP);
  if (bQ) memcpy(blob + oQ, gdT.data(), bQ);
  if (bR) memcpy(blob + oRr, gref.data(), bR);
  if (bT) memcpy(blob + oT, pl.toff.data(), bT);
  memcpy(blob + oI, pl.ibuf.data(), bI);
  char *dbase = ctx->desc.as<char>(0);
  if (capture) {
    if (int r = graph_capture_start(ctx, cache, key, oI + bI, st, &dbase)) return r;
  } else if (int r = upload_staged(ctx, oI + bI, st)) {
    return r;
  }
  const hipStream_t rs = capture ? ctx->gstream : st;   // the stream the launches go to
  int lrc = NRG_OK;
  do {
  if (hipEventRecord(ctx->ev[0], rs) != hipSuccess) { lrc = NRG_E_HIP; break; }
  zero_fill(rs, (int *)ctx->flags.p, (size_t)nblk);
  zero_fill(rs, ctx->ws.as<double>(oamax), (size_t)nblk);
  const BlkDesc *dbd = (const BlkDesc *)(dbase + oBD);
  const int *I = (const int *)(dbase + oI);
  // FP64 arguments (D&C, scaling), FP32 arguments (reduction, back-transformation)
  SyevArgs<double> aD{};
  aD.bd = dbd; aD.A = H; aD.Z0 = Z0; aD.Z1 = Z1; aD.S = S; aD.Zt = Z0;
  aD.d = ctx->ws.as<double>(od); aD.e = ctx->ws.as<double>(oe); aD.lam = ctx->ws.as<double>(ol);
  aD.scale = ctx->ws.as<double>(osc); aD.dscr = ctx->ws.as<double>(odscr); aD.iscr = ctx->ws.as<int>(oiscr);
  aD.Ntot = pl.sumn; aD.flags = (int *)ctx->flags.p; aD.amax = ctx->ws.as<double>(oamax); aD.H = H; aD.w = w;
  SyevArgs<float> aF{};
  aF.bd = dbd; aF.A = A32; aF.Wp = ctx->ws2.as<float>(fWp); aF.tau = ctx->ws2.as<float>(fTau);
  aF.vb = ctx->ws2.as<float>(fvb); aF.Yp = ctx->ws2.as<float>(fYp);
  aF.Ym = ctx->ws2.as<float>(fYm); aF.Yh = ctx->ws2.as<float>(fYh); aF.Zt = ctx->ws2.as<float>(fZt);
  aF.Gb = ctx->ws2.as<float>(fGb); aF.Tb = ctx->ws2.as<float>(fTb);
  aF.sbV = ctx->ws2.as<float>(fsbV); aF.sbY = ctx->ws2.as<float>(fsbY); aF.sbW = ctx->ws2.as<float>(fsbW);
  aF.sbT = ctx->ws2.as<float>(fsbT); aF.sbAB = ctx->ws2.as<float>(fsbAB);
  aF.sbV2 = ctx->ws2.as<float>(fsbV2); aF.sbT2 = ctx->ws2.as<float>(fsbT2);
  aF.cs = ctx->ws2.as<float>(fcs); aF.dpp = ctx->ws2.as<float>(fdpp);
  aF.Z0 = Z0; aF.Z1 = Z1; aF.S = S; aF.d = aD.d; aF.e = aD.e; aF.lam = aD.lam; aF.scale = aD.scale;
  aF.dscr = aD.dscr; aF.iscr = aD.iscr; aF.Ntot = pl.sumn; aF.flags = aD.flags; aF.amax = aD.amax; aF.H = H; aF.w = w;
  SyevLaunch L;
  L.dI = I;
  L.dG = (const GemmDesc *)(dbase + oP);
  L.dGT = (const void *)(dbase + oQ);
  L.toff = (const int64_t *)(dbase + oT);
  L.mst = ctx->ws.as<MergeState>(omst);
  if ((lrc = launch_core<float>(ctx, pl, aF, aD, L, rs, false, 0))) break;
  // FP64 copy of the scaled matrix for the residuals (in place in the caller's arena)
  NRG_TIMED(ctx, rs, "k_symm", hipLaunchKernelGGL(k_symm<double>, dim3((unsigned)pl.symm_cnt), dim3(256), 0, rs, aD,
                                                  (const int2 *)(I + pl.symm_off), 1));
  RefArgs ra;
  ra.bd = dbd; ra.A = H; ra.AX = AX; ra.S = S; ra.R = Rm; ra.E = Em;
  ra.lt = ctx->ws.as<double>(olt); ra.acc = ctx->ws.as<double>(oacc); ra.flags = aD.flags;
  RefCluster *clist = ctx->ws.as<RefCluster>(ocl);
  int *ncl = ctx->ws.as<int>(oncl);
  ra.rid = ctx->ws.as<int>(orid);
  ra.X = Z0; ra.Xn = Z1;
  const GemmDesc *dR = (const GemmDesc *)(dbase + oRr);
  const int2 *rci = (const int2 *)(I + rc_off), *cci = (const int2 *)(I + col_off);
  NRG_TIMED(ctx, rs, "k_ref_init", hipLaunchKernelGGL(k_ref_init, dim3((unsigned)rc_cnt), dim3(256), 0, rs, ra,
                                                      (const float *)aF.Zt, rci));
  const int nref = refine_sweeps();
  for (int sweep = 0; sweep < nref; sweep++) {
    const int dir = sweep & 1;
    ra.X = Xbuf[dir];
    ra.Xn = Xbuf[1 - dir];
    if (sweep > 0)
      NRG_TIMED(ctx, rs, "k_ref_eye", hipLaunchKernelGGL(k_ref_eye, dim3((unsigned)rc_cnt), dim3(256), 0, rs, ra, rci));
    zero_fill(rs, ra.acc, (size_t)3 * nblk);
    zero_fill(rs, ncl, (size_t)1);
    for (int ph = 0; ph < 2; ph++)
      gemm_launch(ctx, rs, dR, I + rg_off[dir][ph], pl.ibuf.data() + rg_off[dir][ph], rg_cnt[dir][ph]);
    NRG_TIMED(ctx, rs, "k_ref_lam", hipLaunchKernelGGL(k_ref_lam, dim3((unsigned)nblk), dim3(256), 0, rs, ra, I + pl.all_off));
    NRG_TIMED(ctx, rs, "k_ref_find", hipLaunchKernelGGL(k_ref_find, dim3((unsigned)nblk), dim3(256), 0, rs, ra,
                                                        I + pl.all_off, clist, ncl, sweep & 1));
    NRG_TIMED(ctx, rs, "k_ref_E", hipLaunchKernelGGL(k_ref_E, dim3((unsigned)col_cnt), dim3(256), 0, rs, ra, cci));
    gemm_launch(ctx, rs, dR, I + rg_off[dir][3], pl.ibuf.data() + rg_off[dir][3], rg_cnt[dir][3]);
    gemm_launch(ctx, rs, dR, I + rg_off[dir][2], pl.ibuf.data() + rg_off[dir][2], rg_cnt[dir][2]);
    NRG_TIMED(ctx, rs, "k_ref_csolve", hipLaunchKernelGGL(k_ref_csolve, dim3((unsigned)ccap), dim3(256), 0, rs, ra,
                                                          (const RefCluster *)clist, (const int *)ncl));
  }
  if (hipEventRecord(ctx->ev[4], rs) != hipSuccess) { lrc = NRG_E_HIP; break; }
  ra.X = Xbuf[nref & 1];
  NRG_TIMED(ctx, rs, "k_ref_finish", hipLaunchKernelGGL(k_ref_finish, dim3((unsigned)nblk), dim3(256), 0, rs, ra, H, w,
                                                        (const double *)aD.amax, I + pl.all_off));
  NRG_TIMED(ctx, rs, "k_ref_copy", hipLaunchKernelGGL(k_ref_copy, dim3((unsigned)rc_cnt), dim3(256), 0, rs, ra, H, rci));
  if (hipGetLastError() != hipSuccess || hipEventRecord(ctx->ev[5], rs) != hipSuccess) lrc = NRG_E_HIP;
  } while (false);
  if (capture) {
    bool retry = false;
    const std::vector<int64_t> key_now = syev_key(ctx, NRG_SYEV_MIXED, nblk, n, off_H, off_w, H, w, st, refine_sweeps(), groups);
    if (int r = graph_capture_end(ctx, cache, key_now, lrc, 5, st, &retry)) return r;
    if (retry) continue;
  } else if (lrc) {
    return lrc;
  }
  const int r = finish_flags(ctx, nblk, info, st);
  if (ctx->kt.on) ctx->kt.collect();
  record_phase_times(ctx, 5);
  if (!ctx->kt.on) tune_record(cache, sig, cand, ctx->total_ms, nblk, n);
  return r;
  }
  set_error("nrg_syev_batched: graph capture failed");
  return NRG_E_HIP;
}

// The tuner's table as JSON text: one entry per signature with its candidates' times and choice.
std::string syev_tuning_json(nrg_ctx *ctx) {
  std::string out = "{";
  if (ctx->syev_cache) {
    const SyevCache &c = *static_cast<SyevCache *>(ctx->syev_cache.get());
    bool first = true;
    for (const auto &[sig, t] : c.tune) {
      char buf[1024];
      int len = snprintf(buf, sizeof buf, "%s\"%s maxn<2^%d big<2^%d mid<2^%d\": {\"calls\": %ld, \"chains\": %d, \"ms\": {",
                         first ? "" : ", ", sig[0] == NRG_SYEV_MIXED ? "mixed" : "fp64", sig[1], sig[2], sig[3], t.calls,
                         t.best >= 0 ? TUNE_CAND[t.best] : -1);
      bool f2 = true;
      for (int k = 0; k < NTUNE && len < (int)sizeof buf - 64; k++)
        if (t.ms[k] > 0) {
          len += snprintf(buf + len, sizeof buf - len, "%s\"%d\": %.3f", f2 ? "" : ", ", TUNE_CAND[k], t.ms[k]);
          f2 = false;
        }
      len += snprintf(buf + len, sizeof buf - len, "}, \"ms_per_1e9_n3\": {");
      f2 = true;
      for (int k = 0; k < NTUNE && len < (int)sizeof buf - 64; k++)
        if (t.unit[k] > 0) {
          len += snprintf(buf + len, sizeof buf - len, "%s\"%d\": %.4f", f2 ? "" : ", ", TUNE_CAND[k], t.unit[k]);
          f2 = false;
        }
      snprintf(buf + len, sizeof buf - len, "}}");
      out += buf;
      first = false;
    }
  }
  return out + "}";
}

}  // namespace nrg
