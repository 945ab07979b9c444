// syev_sbr.hpp - two-stage reduction to tridiagonal form (successive band reduction) for the blocks
// of order TAIL < n <= SB_NMAX, included by syev.hip.
//
// Stage 1 (dense -> band of half-width SB), panel by panel for all blocks of a step at once:
//   k_sb_panel   QR of the panel below the band, P = A(k+SB:n, k:k+SB), one workgroup per block with
//                one panel row per thread in registers: Householder vectors V (kept in A below the
//                band, as dgeqrf leaves them, and explicitly in a workspace), tau and the compact-WY
//                factor T (dlarft, forward, columnwise: Q = I - V T V^T)
//   k_gemm       Y = A22 V                                   (matrix cores)
//   k_sb_w       W = (Y - V T^T (V^T Y) / 2) T
//   k_gemm       A22 -= V W^T + W V^T                        (matrix cores, rank 2 SB)
// Stage 2 (band -> tridiagonal), bulge chasing in one workgroup per block (k_sb_chase): sweep i
// annihilates column i of the band with a reflector on rows i+1 .. i+SB, and chases the bulge it
// creates below the band by SB rows per step. The waves of the workgroup take sweeps round-robin;
// sweep i may run step k once sweep i-1 has finished step k+2 (the regions two sweeps touch are
// then disjoint, so the result is that of the sequential order).
// Back-transformation: V = Q1 Q2 Z. k_sb_q2 applies the bulge-chasing reflectors to the divide and
// conquer eigenvectors (one column per lane, groups of SB consecutive sweeps from the last to the
// first, each group step by step down the matrix); Q1 is the compact-WY back-transformation of the
// one-stage path with its reflectors starting SB rows (instead of one row) below their column.

constexpr int SB = 32;                    // band half-width (lower bandwidth) after stage 1
constexpr int SB_NMAX = SB + NT_PANEL;    // stage-1 panels of at most NT_PANEL rows (a row per thread)
constexpr int SB_WAVES = 8;               // waves of the bulge-chasing workgroup
constexpr int SB_SPIN = 1 << 24;          // s_sleep rounds a bulge-chasing wave waits before it gives up

// steps of bulge-chasing sweep i in a block of order n: windows [i+1+k SB, i+(k+1) SB] within [0, n)
__host__ __device__ inline int sb_steps(int i, int n) { return (n - 1 - i + SB - 1) / SB; }
// the most steps of any sweep (sweep 0): the stride of the reflector store
__host__ __device__ inline int sb_ks(int n) { return n >= 2 ? (n - 2 + SB) / SB : 0; }
// stage-1 panels: panel p (columns p SB .. p SB + SB - 1) while at least two rows lie below the band
__host__ __device__ inline int sb_panels(int n) { return n >= SB + 2 ? (n - SB - 2) / SB + 1 : 0; }

// A reflector is generated only for vectors with sum of squares at least this; below it the vector
// is taken as zero (tau = 0, its entries cleared). Such entries are far below the rounding error of
// the scaled matrix (max-norm in [1, 2)), but their squares reach the subnormal range, where the sum
// of squares is too inaccurate for tau and v to form an orthogonal reflector (rank-deficient panels
// decay that far within one panel).
template <typename T> __device__ __forceinline__ T sb_sig_min();
template <> __device__ __forceinline__ double sb_sig_min<double>() { return 0x1p-960; }
template <> __device__ __forceinline__ float sb_sig_min<float>() { return 0x1p-100f; }

// LDS written by some lanes of a wave and read by others of the same wave
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------------------------
// Stage 1, panel factorisation (dgeqr2 on the panel below the band + dlarft), item = (block, panel)
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(NT_PANEL) void k_sb_panel(SyevArgs<T> a, const int *blist) {
  const int ent = blist[blockIdx.x], p = ent >> COL_PANEL_SHIFT, b = ent & COL_BLK_MASK;
  const BlkDesc bd = a.bd[b];
  const int n = (int)bd.n, k0 = p * SB, R = n - k0 - SB;   // R >= 2 rows below the band
  const int nc = R < SB ? R : SB;                          // reflectors of this panel
  T *P = a.A + bd.offA + (size_t)k0 * n + (k0 + SB);      // P(r, c) = P[c n + r], r < R
  T *Vw = a.sbV + bd.offWp;                                // explicit V, R x SB, ld n
  T *tau = a.tau + bd.offV + k0;
  T *Tw = a.sbT + (size_t)b * SB * SB;                     // T(i, c) at Tw[c SB + i]
  constexpr int NW = NT_PANEL / WAVE;
  __shared__ T red[NW * 16];
  __shared__ T wred[SB];
  __shared__ T Tm[SB * SB];
  __shared__ T s_alpha;
  const int tid = threadIdx.x, lane = tid & (WAVE - 1), wv = tid / WAVE;
  T x[SB];
#pragma unroll
  for (int c = 0; c < SB; c++) x[c] = tid < R ? P[(size_t)c * n + tid] : (T)0;
  // one copy of the column step (a loop unrolled over the columns would not fit the instruction
  // cache); the register columns are addressed by compile-time indices under run-time selects
#pragma unroll 1
  for (int j = 0; j < SB; j++) {
    if (j < nc) {
      T xj = 0;
#pragma unroll
      for (int c = 0; c < SB; c++) xj = c == j ? x[c] : xj;
      // reflector of column j (dlarfg): rows j .. R-1
      T s = tid > j ? xj * xj : (T)0;
      s = wave_sum(s);
      if (lane == 0) red[wv] = s;
      if (tid == j) s_alpha = xj;
      __syncthreads();
      T sig = 0;
#pragma unroll
      for (int w = 0; w < NW; w++) sig += red[w];
      const T alpha = s_alpha;
      T beta = alpha, tv = 0, scal = 0;
      if (sig >= sb_sig_min<T>()) {
        beta = -copysign(sqrt(alpha * alpha + sig), alpha);
        tv = (beta - alpha) / beta;
        scal = (T)1 / (alpha - beta);
      }
      const T v = tid == j ? (T)1 : (tid > j ? xj * scal : (T)0);
      // v . x_c: for c > j the update of column c, for c < j the Gram column V^T v of the T factor
      // (column c < j holds reflector c below row c, and R above)
#pragma unroll
      for (int half = 0; half < 2; half++) {
        T pv[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const int c = half * 16 + q;
          pv[q] = c > j ? v * x[c] : (c < j ? v * (tid > c ? x[c] : (tid == c ? (T)1 : (T)0)) : (T)0);
        }
        const T ps = butterfly16(pv);   // lane l: the sum of value (l >> 2) & 15
        __syncthreads();   // red is free again (sig read, or the first half reduced)
        if ((lane & 3) == 0) red[wv * 16 + (lane >> 2)] = ps;
        __syncthreads();
        if (tid < 16) {
          T t = 0;
#pragma unroll
          for (int w = 0; w < NW; w++) t += red[w * 16 + tid];
          wred[half * 16 + tid] = t;
        }
      }
      __syncthreads();
      const T xn = tid == j ? beta : (tid > j ? v : xj);
#pragma unroll
      for (int c = 0; c < SB; c++) {
        const T upd = x[c] - tv * v * wred[c];
        x[c] = c > j ? upd : (c == j ? xn : x[c]);
      }
      // T(0:j, j) = -tau_j T(0:j, 0:j) V(:, 0:j)^T v_j, T(j, j) = tau_j
      if (tid < j) {
        T acc = 0;
        for (int l = tid; l < j; l++) acc += Tm[l * SB + tid] * wred[l];
        Tm[j * SB + tid] = -tv * acc;
      } else if (tid == j) {
        Tm[j * SB + j] = tv;
        tau[j] = tv;
      } else if (tid < SB) {
        Tm[j * SB + tid] = 0;
      }
    } else if (tid < SB) {
      Tm[j * SB + tid] = 0;
    }
  }
  __syncthreads();
  // V below the diagonal and R on and above it, in place (as dgeqrf); the explicit, zero-padded V
#pragma unroll
  for (int c = 0; c < SB; c++)
    if (tid < R) {
      P[(size_t)c * n + tid] = x[c];
      Vw[(size_t)c * n + tid] = c < nc ? (tid == c ? (T)1 : (tid > c ? x[c] : (T)0)) : (T)0;
    }
  Tw[tid] = Tm[tid];   // NT_PANEL == SB * SB
  // the block's last reflector has length one (tau = 0) when this panel leaves SB + 1 rows
  if (tid == 0 && R == SB + 1) tau[SB] = 0;
}

// W = Y T - V N / 2 with N = T^T (V^T Y) T (that is, W = (Y - V T^T (V^T Y) / 2) T), Y = A22 V,
// item = (block, panel); 512 threads, rows tid and tid + 512
constexpr int NT_SBW = 512;
template <typename T>
__global__ __launch_bounds__(NT_SBW) void k_sb_w(SyevArgs<T> a, const int *blist) {
  const int ent = blist[blockIdx.x], p = ent >> COL_PANEL_SHIFT, b = ent & COL_BLK_MASK;
  const BlkDesc bd = a.bd[b];
  const int n = (int)bd.n, R = n - p * SB - SB;
  const T *V = a.sbV + bd.offWp, *Y = a.sbY + bd.offWp;
  T *W = a.sbW + bd.offWp;
  const T *Tw = a.sbT + (size_t)b * SB * SB;
  constexpr int CH = 64;
  __shared__ T Vs[CH * (SB + 1)], Ys[CH * (SB + 1)];
  __shared__ T Pm[SB * SB], Tm[SB * SB], Mm[SB * SB];
  const int tid = threadIdx.x, c1 = tid >> 5, c2 = tid & 31;   // outputs (c1, c2) and (c1 + 16, c2)
  T acc0 = 0, acc1 = 0;
  for (int r0 = 0; r0 < R; r0 += CH) {
    for (int idx = tid; idx < CH * SB; idx += NT_SBW) {
      const int r = idx % CH, c = idx / CH;
      const bool in = r0 + r < R;
      Vs[r * (SB + 1) + c] = in ? V[(size_t)c * n + r0 + r] : (T)0;
      Ys[r * (SB + 1) + c] = in ? Y[(size_t)c * n + r0 + r] : (T)0;
    }
    __syncthreads();
#pragma unroll 8
    for (int r = 0; r < CH; r++) {
      const T y = Ys[r * (SB + 1) + c2];
      acc0 += Vs[r * (SB + 1) + c1] * y;
      acc1 += Vs[r * (SB + 1) + c1 + 16] * y;
    }
    __syncthreads();
  }
  Pm[c1 * SB + c2] = acc0;   // P(l, c) = (V^T Y)(l, c) at Pm[l SB + c]
  Pm[(c1 + 16) * SB + c2] = acc1;
  Tm[tid] = Tw[tid];         // T(i, c) at Tm[c SB + i]
  Tm[tid + NT_SBW] = Tw[tid + NT_SBW];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; q++) {   // M(i, c) = sum_{l <= i} T(l, i) P(l, c), at Mm[i SB + c]
    const int i = c1 + 16 * q, c = c2;
    T s = 0;
    for (int l = 0; l <= i; l++) s += Tm[i * SB + l] * Pm[l * SB + c];
    Mm[i * SB + c] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; q++) {   // N(i, c) = sum_{l <= c} M(i, l) T(l, c), at Pm[c SB + i]
    const int i = c2, c = c1 + 16 * q;
    T s = 0;
    for (int l = 0; l <= c; l++) s += Mm[i * SB + l] * Tm[c * SB + l];
    Pm[c * SB + i] = s;
  }
  __syncthreads();
  for (int r = tid; r < R; r += NT_SBW) {
    T v[SB], y[SB];
#pragma unroll
    for (int c = 0; c < SB; c++) {
      v[c] = V[(size_t)c * n + r];
      y[c] = Y[(size_t)c * n + r];
    }
#pragma unroll 1
    for (int c = 0; c < SB; c++) {
      T s = 0, t = 0;
#pragma unroll
      for (int i = 0; i < SB; i++) {
        s += (i <= c ? y[i] : (T)0) * Tm[c * SB + i];
        t += v[i] * Pm[c * SB + i];
      }
      W[(size_t)c * n + r] = s - (T)0.5 * t;
    }
  }
}

// the band after stage 1, lower band storage with room for the bulge: AB(r, c) = AB[c 2SB + r - c]
// for 0 <= r - c < 2 SB; item = (block, chunk of 64 columns)
template <typename T>
__global__ __launch_bounds__(256) void k_sb_toband(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n;
  const T *A = a.A + bd.offA;
  T *AB = a.sbAB + bd.offAB;
  for (int idx = threadIdx.x; idx < 64 * 2 * SB; idx += 256) {
    const int c = it.y * 64 + idx / (2 * SB), d = idx % (2 * SB);
    if (c < n) AB[(size_t)c * 2 * SB + d] = (d <= SB && c + d < n) ? A[(size_t)c * n + c + d] : (T)0;
  }
}

// ------------------------------------------------------------------------------------------
// Stage 2: bulge chasing, one workgroup per block. A wave holds 32 x 32 blocks as half columns:
// lane l has column l & 31, rows (l >> 5) * 16 + t, t < 16.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(SB_WAVES * WAVE) void k_sb_chase(SyevArgs<T> a, const int *blist) {
  const int b = blist[blockIdx.x];
  const BlkDesc bd = a.bd[b];
  const int n = (int)bd.n, KS = sb_ks(n), nsw = n - 2;
  T *AB = a.sbAB + bd.offAB;
  T *V2 = a.sbV2 + bd.offV2;   // reflector (i, k): V2[(i KS + k) SB + t]
  T *t2 = a.sbT2 + bd.offT2;   // its scalar: t2[i KS + k]
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  const int c = lane & 31, h = lane >> 5;
  __shared__ int prog[SB_WAVES];   // sweep * 64 + steps done of the wave's current sweep
  __shared__ int abort_flag;
  __shared__ T xs[SB_WAVES][SB], vs[SB_WAVES][SB], ws[SB_WAVES][SB];
  __shared__ T tile[SB_WAVES][SB * (SB + 1)];
  if (threadIdx.x < SB_WAVES) prog[threadIdx.x] = -1;
  if (threadIdx.x == 0) abort_flag = 0;
  __syncthreads();
  auto ab = [&](int r, int cc) -> T & { return AB[(size_t)cc * (2 * SB) + (r - cc)]; };
  const int pw = (wv + SB_WAVES - 1) % SB_WAVES;
  T *tl = tile[wv];
  for (int i = wv; i < nsw; i += SB_WAVES) {
    const int ns = sb_steps(i, n), nsp = i >= 1 ? sb_steps(i - 1, n) : 0;
    T Bl[16];
#pragma unroll
    for (int t = 0; t < 16; t++) Bl[t] = 0;
    for (int k = 0; k < ns; k++) {
      if (i >= 1) {
        const int target = (i - 1) * 64 + min(k + 3, nsp);
        int spins = 0;
        while (__hip_atomic_load(&prog[pw], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target) {
          if (__hip_atomic_load(&abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
          if (++spins > SB_SPIN) {
            __hip_atomic_store(&abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (lane == 0) a.flags[b] = 3;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (__hip_atomic_load(&abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
      }
      const int s0 = i + 1 + k * SB, s1 = min(s0 + SB - 1, n - 1), L = s1 - s0 + 1;
      const int t0 = s1 + 1, Ln = min(SB, n - t0);   // rows of the block below the window
      // every load of the step at once: column i (step 0), the diagonal block of the window and the
      // block below it (transposed: lane row c, columns h*16+t); the stores of this step go elsewhere
      const T x0 = (k == 0 && lane < L) ? ab(s0 + lane, i) : (T)0;
      T D[16], Bn[16];
#pragma unroll
      for (int t = 0; t < 16; t++) {
        const int r = h * 16 + t;
        D[t] = (r < L && c < L) ? (r >= c ? ab(s0 + r, s0 + c) : ab(s0 + c, s0 + r)) : (T)0;
        Bn[t] = (c < Ln && r < L) ? ab(t0 + c, s0 + r) : (T)0;
      }
      // the vector to annihilate: column i (step 0) or the first column of the bulge block
      if (k == 0) {
        if (lane < SB) xs[wv][lane] = x0;
      } else if (c == 0) {
#pragma unroll
        for (int t = 0; t < 16; t++) xs[wv][h * 16 + t] = Bl[t];
      }
      wave_lds_sync();
      const T xl = lane < SB ? xs[wv][lane] : (T)0;
      const T alpha = xs[wv][0];
      T sg = (lane >= 1 && lane < SB) ? xl * xl : (T)0;
      sg = wave_sum(sg);
      T beta = alpha, tv = 0, scal = 0;
      if (sg >= sb_sig_min<T>()) {
        beta = -copysign(sqrt(alpha * alpha + sg), alpha);
        tv = (beta - alpha) / beta;
        scal = (T)1 / (alpha - beta);
      }
      if (lane < SB) {
        const T v = lane == 0 ? (T)1 : xl * scal;
        vs[wv][lane] = v;
        V2[(size_t)(i * KS + k) * SB + lane] = v;
      }
      if (lane == 0) t2[(size_t)i * KS + k] = tv;
      wave_lds_sync();
      T vr[16];
#pragma unroll
      for (int t = 0; t < 16; t++) vr[t] = vs[wv][h * 16 + t];
      const T vc = vs[wv][c];
      if (k == 0) {
        if (lane < L) ab(s0 + lane, i) = lane == 0 ? beta : (T)0;
      } else {
        // from the left on the bulge block (rows of the window, the previous window's columns)
        T wc = 0;
#pragma unroll
        for (int t = 0; t < 16; t++) wc += vr[t] * Bl[t];
        wc += __shfl_xor(wc, 32, WAVE);
#pragma unroll
        for (int t = 0; t < 16; t++) Bl[t] -= tv * vr[t] * wc;
        if (c == 0) {
#pragma unroll
          for (int t = 0; t < 16; t++) Bl[t] = (h * 16 + t == 0) ? beta : (T)0;
        }
        const int p0 = s0 - SB;
#pragma unroll
        for (int t = 0; t < 16; t++)
          if (h * 16 + t < L) ab(s0 + h * 16 + t, p0 + c) = Bl[t];
      }
      // both sides on the diagonal block of the window
      T pc = 0;
#pragma unroll
      for (int t = 0; t < 16; t++) pc += D[t] * vr[t];
      pc += __shfl_xor(pc, 32, WAVE);
      pc *= tv;                                  // p = tau D v
      T kk = lane < SB ? pc * vc : (T)0;
      kk = (T)0.5 * tv * wave_sum(kk);           // tau (p . v) / 2
      const T wcol = pc - kk * vc;               // w = p - K v
      if (lane < SB) ws[wv][lane] = wcol;
      wave_lds_sync();
#pragma unroll
      for (int t = 0; t < 16; t++) {
        const int r = h * 16 + t;
        D[t] -= vr[t] * wcol + ws[wv][r] * vc;
        if (r < L && c < L && r >= c) ab(s0 + r, s0 + c) = D[t];
      }
      // from the right on the block below the window (held transposed: lane row c, columns h*16+t)
      if (Ln > 0) {
        T u = 0;
#pragma unroll
        for (int t = 0; t < 16; t++) u += Bn[t] * vr[t];
        u += __shfl_xor(u, 32, WAVE);
#pragma unroll
        for (int t = 0; t < 16; t++) Bn[t] -= tv * u * vr[t];
        // to the layout of a window block: lane column c, rows h*16+t
#pragma unroll
        for (int t = 0; t < 16; t++) tl[(h * 16 + t) * (SB + 1) + c] = Bn[t];
        wave_lds_sync();
#pragma unroll
        for (int t = 0; t < 16; t++) Bl[t] = tl[c * (SB + 1) + h * 16 + t];
        wave_lds_sync();
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      if (lane == 0) __hip_atomic_store(&prog[wv], i * 64 + k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (__hip_atomic_load(&abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
  }
  __syncthreads();
  double *d = a.d + bd.offV, *e = a.e + bd.offV;
  for (int cc = threadIdx.x; cc < n; cc += SB_WAVES * WAVE) {
    d[cc] = (double)ab(cc, cc);
    if (cc + 1 < n) e[cc] = (double)ab(cc + 1, cc);
  }
}

// ------------------------------------------------------------------------------------------
// Stage-2 back-transformation Z <- Q2 Z, item = (block, slab of 64 columns), one wave: lane =
// column. The sweeps are taken in groups of SB from the last group to the first; a group walks
// down the matrix step by step, applying its sweeps' reflectors of that step from the last sweep
// to the first on a window of 2 SB rows held in registers.
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(WAVE) void k_sb_q2(SyevArgs<T> a, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const BlkDesc bd = a.bd[it.x];
  const int n = (int)bd.n, KS = sb_ks(n), nsw = n - 2;
  const int lane = threadIdx.x, col = it.y * WAVE + lane;
  const bool on = col < n;
  T *Z = a.Zt + bd.offZ + (size_t)(on ? col : 0) * n;
  const T *V2 = a.sbV2 + bd.offV2, *t2 = a.sbT2 + bd.offT2;
  constexpr int G = SB;
  __shared__ T vsh[G * SB];
  __shared__ T tsh[G];
  const int ng = (nsw + G - 1) / G;
  for (int g = ng - 1; g >= 0; g--) {
    const int sA = g * G, sB = min(sA + G, nsw);
    const int kmax = sb_steps(sA, n);
    int base = sA + 1;   // first row of the window
    T zw[2 * SB];
#pragma unroll
    for (int q = 0; q < 2 * SB; q++) zw[q] = (on && base + q < n) ? Z[base + q] : (T)0;
    for (int k = 0; k < kmax; k++) {
      for (int idx = lane; idx < G * SB; idx += WAVE) {
        const int s = sA + idx / SB, t = idx % SB;
        vsh[idx] = (s < sB && k < sb_steps(s, n)) ? V2[(size_t)(s * KS + k) * SB + t] : (T)0;
      }
      if (lane < G) {
        const int s = sA + lane;
        tsh[lane] = (s < sB && k < sb_steps(s, n)) ? t2[(size_t)s * KS + k] : (T)0;
      }
      wave_lds_sync();
      // the rows the next step brings in (this group has not touched them yet), loaded early
      T zn[SB];
#pragma unroll
      for (int q = 0; q < SB; q++) zn[q] = (on && k + 1 < kmax && base + 2 * SB + q < n) ? Z[base + 2 * SB + q] : (T)0;
#pragma unroll
      for (int o = G - 1; o >= 0; o--) {
        T v[SB];
#pragma unroll
        for (int t = 0; t < SB; t++) v[t] = vsh[o * SB + t];
        T d0 = 0, d1 = 0, d2 = 0, d3 = 0;
#pragma unroll
        for (int t = 0; t < SB; t += 4) {
          d0 += v[t] * zw[o + t];
          d1 += v[t + 1] * zw[o + t + 1];
          d2 += v[t + 2] * zw[o + t + 2];
          d3 += v[t + 3] * zw[o + t + 3];
        }
        const T dot = ((d0 + d1) + (d2 + d3)) * tsh[o];
#pragma unroll
        for (int t = 0; t < SB; t++) zw[o + t] -= v[t] * dot;
      }
      wave_lds_sync();
      // rows base .. base+SB-1 are final for this group
#pragma unroll
      for (int q = 0; q < SB; q++)
        if (on && base + q < n) Z[base + q] = zw[q];
      base += SB;
#pragma unroll
      for (int q = 0; q < SB; q++) {
        zw[q] = zw[q + SB];
        zw[q + SB] = zn[q];
      }
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      if (on && base + q < n) Z[base + q] = zw[q];
  }
}
