// ozaki.hip - split-precision form of nrg_gemm3_grouped (NRG_GEMM_OZAKI).
//
// Same table semantics and evaluation order as the FP64 form (recalc_f / recalc_general / cdmI,
// c++/recalc.hpp:52-108, c++/dmnrg.hpp:54-80, c++/numerics_ublas.hpp:108-140); each FP64 product is
// evaluated with the Ozaki error-free splitting on the INT8 matrix cores (v_mfma_i32_16x16x64_i8):
//
//   1. k_oz_split: every operand row x (a row of op(A), or a column of op(B)) gets an exponent E with
//      |x_k| < 2^E, and OZ_S signed integer slices m_p(k) with |m_p| < 2^OZ_BITS (int8),
//          x_k = sum_p m_p(k) 2^(E - OZ_BITS (p + 1)) + remainder,  |remainder| < 2^(E - OZ_S OZ_BITS).
//   2. k_oz_gemm: for every pair of slices with p + q < OZ_S (36 pairs) the slice products are summed
//      on the matrix cores in INT32, exactly: a product of two slices is at most 127^2 < 2^14 in
//      magnitude and at most OZ_S pairs share a diagonal d = p + q, so a diagonal's sum over OZ_KFLUSH
//      K values stays below 2^31. Each diagonal is converted once per OZ_KFLUSH of K (once for the
//      shapes of a recalc table), scaled by 2^(-7 d) and by the row/column exponents, and added into
//      FP64 accumulators.
// Dropped terms (p + q >= OZ_S, remainders) are below K 2^(E_row + F_col) 10 2^-56, i.e. a normwise
// relative error of order 1e-15 for the shapes of a recalc table.
#include <cstring>
#include <map>
#include <tuple>

#include "gemm.hpp"

namespace nrg {

constexpr int OZ_S = 8;        // slices per operand
constexpr int OZ_BITS = 7;     // bits per slice (magnitude; int8)
constexpr int OZ_KC = 64;      // K of one INT8 matrix-core step
constexpr int OZ_LDK = 80;     // LDS row stride (bytes): 64 + 16 to spread the banks
constexpr int OZ_KFLUSH = 16384;   // K values summed in INT32 before a conversion: 8 * 2^14 * 16384 = 2^31
static_assert(OZ_KFLUSH % OZ_KC == 0, "conversions fall on K steps");

typedef int i32x4 __attribute__((ext_vector_type(4)));   // 16 int8 (an MFMA operand) or 4 int32

// Operand rows to split: row r of an R x K matrix Y whose element (r, k) is
// t ? X[r * ld + k] : X[k * ld + r] (column-major storage conventions, as GemmPart).
struct OzSplit {
  const double *X;
  int ld, t, R, K, Kp;   // Kp = K rounded up to OZ_KC; slices are zero beyond K
  int8_t *S;             // slice p of row r at S + ((size_t)p * R + r) * Kp
  int *E;                // per-row exponent
};

struct OzPart {          // scale * op(A) op(B): A rows = slices of op(A) rows, B rows = slices of op(B) columns
  const int8_t *SA, *SB;
  const int *EA, *EB;
  int MA, NB, Kp;
  double scale;
};

struct OzDesc {          // column-major C (M x N) = alpha * sum of parts + beta * C
  double *C;
  int M, N, ldc;
  double alpha, beta;
  int p0, np;
};

// One workgroup per 64 rows of an operand. Reads run along the operand's storage (lanes along r when
// rows are strided, along k when they are contiguous); the slices of a 64 x 64 (row, k) block are
// formed in LDS and written out along k, 16 bytes per thread.
constexpr int OZ_SR = 64;   // rows per split workgroup
__global__ __launch_bounds__(256) void k_oz_split(const OzSplit *sp, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const OzSplit S = sp[it.x];
  const int r0 = it.y * OZ_SR, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ double red[4][OZ_SR];
  __shared__ int Er[OZ_SR];
  __shared__ __attribute__((aligned(16))) int8_t T[OZ_S][OZ_SR][64 + 16];
  // element (r0 + a, k) with a < 64: strided rows are read with lanes along a, contiguous rows along k
  auto elem = [&](int a, int k) -> double {
    const int r = r0 + a;
    if (r >= S.R || k >= S.K) return 0.0;
    return S.t ? S.X[(size_t)r * S.ld + k] : S.X[(size_t)k * S.ld + r];
  };
  // 1. row exponents: |x| < 2^E over the row
  {
    if (!S.t) {   // lane = row; wave w reads k = 16 w + 64 i + [0, 16): 16 loads in flight per thread
      double mx = 0.0;
      for (int kb = 16 * wv; kb < S.K; kb += 64) {
        double x[16];
#pragma unroll
        for (int q = 0; q < 16; q++) x[q] = elem(lane, kb + q);
#pragma unroll
        for (int q = 0; q < 16; q++) mx = fmax(mx, fabs(x[q]));
      }
      red[wv][lane] = mx;
    } else {      // wave w: rows w, w + 4, ...; lanes along k
      for (int a = wv; a < OZ_SR; a += 4) {
        double m = 0.0;
        for (int kb = 0; kb < S.K; kb += 64 * 8) {
          double x[8];
#pragma unroll
          for (int q = 0; q < 8; q++) x[q] = elem(a, kb + 64 * q + lane);
#pragma unroll
          for (int q = 0; q < 8; q++) m = fmax(m, fabs(x[q]));
        }
        m = wave_max(m);
        if (lane == 0) red[0][a] = m;
      }
      if (lane < OZ_SR) red[1][lane] = red[2][lane] = red[3][lane] = 0.0;   // only red[0] carries values
    }
    __syncthreads();
    if (tid < OZ_SR) {
      const double m = fmax(fmax(red[0][tid], red[1][tid]), fmax(red[2][tid], red[3][tid]));
      int E = 0;
      if (m > 0.0) frexp(m, &E);   // m < 2^E
      Er[tid] = E;
      if (r0 + tid < S.R) S.E[r0 + tid] = E;
    }
    __syncthreads();
  }
  // 2. slices, one 64-wide K block at a time (Kp is a multiple of 64; zero beyond K)
  for (int k0 = 0; k0 < S.Kp; k0 += 64) {
    // element (a, kl) of the block: strided rows a = lane, kl = 16 wv + q; contiguous rows a = 16 wv + q,
    // kl = lane; the 16 loads go out together
    double xs[16];
#pragma unroll
    for (int q = 0; q < 16; q++) xs[q] = elem(S.t ? 16 * wv + q : lane, k0 + (S.t ? lane : 16 * wv + q));
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int a = S.t ? 16 * wv + q : lane, kl = S.t ? lane : 16 * wv + q;
      double x = xs[q];
      const int E = Er[a];
#pragma unroll
      for (int p = 0; p < OZ_S; p++) {
        const int sh = E - OZ_BITS * (p + 1);
        const double m = trunc(ldexp(x, -sh));   // |m| < 2^OZ_BITS, exact
        x -= ldexp(m, sh);                       // exact remainder
        T[p][a][kl] = (int8_t)m;
      }
    }
    __syncthreads();
    // write out: slice p, row a, 16 bytes at 16 c; 8 x 64 x 4 chunks over 256 threads
    for (int v = tid; v < OZ_S * OZ_SR * 4; v += 256) {
      const int p = v / (OZ_SR * 4), a = (v / 4) % OZ_SR, c = v % 4;
      if (r0 + a < S.R)
        *(i32x4 *)(S.S + ((size_t)p * S.R + r0 + a) * S.Kp + k0 + 16 * c) = *(const i32x4 *)&T[p][a][16 * c];
    }
    __syncthreads();
  }
}

// one wave per SIMD: the INT32 accumulators (128 registers) and the next K step's slices in flight
// take more than the 256 registers per lane of two waves
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_oz_gemm(const OzDesc *descs, const OzPart *parts, const int2 *items) {
  const int2 it = items[blockIdx.x];
  const OzDesc G = descs[it.x];
  const int tm = (G.M + 63) / 64;
  const int m0 = (it.y % tm) * 64, n0 = (it.y / tm) * 64;
  __shared__ __attribute__((aligned(16))) int8_t As[OZ_S][64][OZ_LDK];   // slices of op(A) rows m0..m0+63
  __shared__ __attribute__((aligned(16))) int8_t Bs[OZ_S][64][OZ_LDK];   // slices of op(B) columns n0..n0+63
  __shared__ int Ea[64], Eb[64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
  double accC[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) accC[i][j][r] = 0.0;
  for (int pi = 0; pi < G.np; pi++) {
    const OzPart P = parts[G.p0 + pi];
    __syncthreads();
    if (tid < 64) Ea[tid] = (m0 + tid < P.MA) ? P.EA[m0 + tid] : 0;
    else if (tid < 128) Eb[tid - 64] = (n0 + tid - 64 < P.NB) ? P.EB[n0 + tid - 64] : 0;
    __syncthreads();
    // this lane's outputs: D = C^T tile, lane l register r holds (n = 4 (l / 16) + r, m = l % 16)
    int sh[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          sh[i][j][r] = Ea[wm + 16 * i + (lane & 15)] + Eb[wn + 16 * j + 4 * (lane >> 4) + r] - 2 * OZ_BITS;
    i32x4 acc[OZ_S][2][2];   // exact INT32 sums per diagonal p + q
    // the slices of one K step, 16 bytes of each of OZ_S rows of op(A) and of op(B) per thread: the
    // next step's are loaded while the matrix cores work on this one
    i32x4 ra[OZ_S], rb[OZ_S];
    const int lrow = (tid >> 2) & 63, lkv = tid & 3;
    auto load_step = [&](int k0) {
#pragma unroll
      for (int p = 0; p < OZ_S; p++) {
        ra[p] = (i32x4){0, 0, 0, 0};
        rb[p] = (i32x4){0, 0, 0, 0};
        if (m0 + lrow < P.MA) ra[p] = *(const i32x4 *)(P.SA + ((size_t)p * P.MA + m0 + lrow) * P.Kp + k0 + 16 * lkv);
        if (n0 + lrow < P.NB) rb[p] = *(const i32x4 *)(P.SB + ((size_t)p * P.NB + n0 + lrow) * P.Kp + k0 + 16 * lkv);
      }
    };
    load_step(0);
    for (int k0 = 0; k0 < P.Kp; k0 += OZ_KC) {
      if (k0 % OZ_KFLUSH == 0) {
#pragma unroll
        for (int d = 0; d < OZ_S; d++)
#pragma unroll
          for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) acc[d][i][j] = (i32x4){0, 0, 0, 0};
      }
      // 256 threads x 16 bytes cover the 64 rows x 64 bytes of a slice: thread t, row (t / 4) % 64,
      // bytes 16 (t % 4), slice p
#pragma unroll
      for (int p = 0; p < OZ_S; p++) {
        *(i32x4 *)&As[p][lrow][16 * lkv] = ra[p];
        *(i32x4 *)&Bs[p][lrow][16 * lkv] = rb[p];
      }
      __syncthreads();
      if (k0 + OZ_KC < P.Kp) load_step(k0 + OZ_KC);
      // MFMA operands: lane l holds 16 K values (16 (l / 16) .. + 16) of row l % 16; A operand: op(B)
      // columns (rows of D), B operand: op(A) rows (columns of D)
#pragma unroll
      for (int p = 0; p < OZ_S; p++) {
        i32x4 fa[2];
#pragma unroll
        for (int i = 0; i < 2; i++) fa[i] = *(const i32x4 *)&As[p][wm + 16 * i + (lane & 15)][16 * (lane >> 4)];
#pragma unroll
        for (int q = 0; q < OZ_S - p; q++) {
          i32x4 fb[2];
#pragma unroll
          for (int j = 0; j < 2; j++) fb[j] = *(const i32x4 *)&Bs[q][wn + 16 * j + (lane & 15)][16 * (lane >> 4)];
#pragma unroll
          for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
              acc[p + q][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fb[j], fa[i], acc[p + q][i][j], 0, 0, 0);
        }
      }
      __syncthreads();
      if ((k0 + OZ_KC) % OZ_KFLUSH == 0 || k0 + OZ_KC >= P.Kp) {
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
          for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              double t = 0.0;
#pragma unroll
              for (int d = OZ_S - 1; d >= 0; d--) t += ldexp((double)acc[d][i][j][r], -OZ_BITS * d);
              accC[i][j][r] = fma(P.scale, ldexp(t, sh[i][j][r]), accC[i][j][r]);
            }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + 16 * i + (lane & 15);
        const int n = n0 + wn + 16 * j + 4 * (lane >> 4) + r;
        if (m < G.M && n < G.N) {
          double *c = G.C + (size_t)n * G.ldc + m;
          const double v = G.alpha * accC[i][j][r];
          *c = G.beta == 0.0 ? v : v + G.beta * *c;
        }
      }
}

namespace {

struct OzBuild {
  std::vector<OzSplit> splits;
  std::vector<OzPart> parts;
  std::vector<OzDesc> descs;
  std::map<std::tuple<const double *, int, int, int, int>, int> seen;
  size_t ws_bytes = 0;
  // index of the split of an R x K operand (element (r, k) = t ? X[r ld + k] : X[k ld + r])
  int split(const double *X, int ld, int t, int R, int K) {
    const auto key = std::make_tuple(X, ld, t, R, K);
    if (auto f = seen.find(key); f != seen.end()) return f->second;
    OzSplit s{};
    s.X = X; s.ld = ld; s.t = t; s.R = R; s.K = K;
    s.Kp = (K + OZ_KC - 1) / OZ_KC * OZ_KC;
    s.S = (int8_t *)(uintptr_t)ws_bytes;   // offset for now, rebased later
    ws_bytes += ((size_t)OZ_S * R * s.Kp + 255) & ~(size_t)255;
    s.E = (int *)(uintptr_t)ws_bytes;
    ws_bytes += ((size_t)R * sizeof(int) + 255) & ~(size_t)255;
    splits.push_back(s);
    return seen[key] = (int)splits.size() - 1;
  }
  // one part: scale * op(A) op(B) with op(A) M x K (flags of GemmPart) and op(B) K x N
  void part(const double *A, int lda, int tA, const double *B, int ldb, int tB, int M, int N, int K, double scale) {
    OzPart p{};
    p.SA = (const int8_t *)(uintptr_t)split(A, lda, tA, M, K);   // split indices for now
    p.SB = (const int8_t *)(uintptr_t)split(B, ldb, !tB, N, K);
    p.MA = M; p.NB = N; p.Kp = (K + OZ_KC - 1) / OZ_KC * OZ_KC; p.scale = scale;
    parts.push_back(p);
  }
};

}  // namespace

int gemm3_ozaki(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, hipStream_t st) {
  Gemm3Plan pl;
  if (int r = gemm3_plan(ctx, terms, nterms, pl)) return r;
  if (pl.gd.empty()) {
    ctx->total_ms = 0;
    return NRG_OK;
  }
  OzBuild b;
  size_t split_end[2], desc_end[2];
  for (int stage = 0; stage < 2; stage++) {
    const size_t g0 = stage ? pl.nd1 : 0, g1 = stage ? pl.gd.size() : pl.nd1;
    for (size_t g = g0; g < g1; g++) {
      const GemmDesc &d = pl.gd[g];
      OzDesc o{};
      o.C = d.C; o.M = d.M; o.N = d.N; o.ldc = d.ldc; o.alpha = d.alpha; o.beta = d.beta;
      o.p0 = (int)b.parts.size();
      if (stage == 0) {
        b.part(d.A, d.lda, d.tA, d.B, d.ldb, d.tB, d.M, d.N, d.K, 1.0);
      } else {
        for (size_t q = pl.prange[g - pl.nd1].first; q < pl.prange[g - pl.nd1].second; q++) {
          const GemmPart &P = pl.parts[q];
          b.part(P.A, P.lda, P.tA, P.B, P.ldb, P.tB, d.M, d.N, P.K, P.scale);
        }
      }
      o.np = (int)b.parts.size() - o.p0;
      b.descs.push_back(o);
    }
    split_end[stage] = b.splits.size();
    desc_end[stage] = b.descs.size();
  }
  if (int r = ctx->ws.ensure(b.ws_bytes + 256)) return r;
  char *wsb = ctx->ws.as<char>();
  for (OzSplit &s : b.splits) {
    s.S = (int8_t *)(wsb + (uintptr_t)s.S);
    s.E = (int *)(wsb + (uintptr_t)s.E);
  }
  for (OzPart &p : b.parts) {
    const OzSplit &sa = b.splits[(size_t)(uintptr_t)p.SA], &sb = b.splits[(size_t)(uintptr_t)p.SB];
    p.SA = sa.S; p.EA = sa.E;
    p.SB = sb.S; p.EB = sb.E;
  }
  // work items: split (split, group of OZ_SR rows); product (descriptor, 64 x 64 tile)
  std::vector<int> si[2], gi[2];
  for (int stage = 0; stage < 2; stage++) {
    for (size_t s = stage ? split_end[0] : 0; s < split_end[stage]; s++)
      for (int g = 0; g < (b.splits[s].R + OZ_SR - 1) / OZ_SR; g++) { si[stage].push_back((int)s); si[stage].push_back(g); }
    for (size_t d = stage ? desc_end[0] : 0; d < desc_end[stage]; d++) {
      const int t = ((b.descs[d].M + 63) / 64) * ((b.descs[d].N + 63) / 64);
      for (int k = 0; k < t; k++) { gi[stage].push_back((int)d); gi[stage].push_back(k); }
    }
  }
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t oS = 0, oP = al(sizeof(OzSplit) * b.splits.size());
  const size_t oD = oP + al(sizeof(OzPart) * b.parts.size());
  size_t off = oD + al(sizeof(OzDesc) * b.descs.size());
  size_t oSI[2], oGI[2];
  for (int stage = 0; stage < 2; stage++) { oSI[stage] = off; off += al(sizeof(int) * si[stage].size()); }
  for (int stage = 0; stage < 2; stage++) { oGI[stage] = off; off += al(sizeof(int) * gi[stage].size()); }
  std::vector<char> blob(off + 8);
  memcpy(blob.data() + oS, b.splits.data(), sizeof(OzSplit) * b.splits.size());
  memcpy(blob.data() + oP, b.parts.data(), sizeof(OzPart) * b.parts.size());
  memcpy(blob.data() + oD, b.descs.data(), sizeof(OzDesc) * b.descs.size());
  for (int stage = 0; stage < 2; stage++) {
    if (!si[stage].empty()) memcpy(blob.data() + oSI[stage], si[stage].data(), sizeof(int) * si[stage].size());
    if (!gi[stage].empty()) memcpy(blob.data() + oGI[stage], gi[stage].data(), sizeof(int) * gi[stage].size());
  }
  NRG_HIP(hipEventRecord(ctx->ev[0], st));
  if (int r = upload_desc(ctx, blob.data(), blob.size(), st)) return r;
  const OzSplit *dS = (const OzSplit *)ctx->desc.as<char>(oS);
  const OzPart *dP = (const OzPart *)ctx->desc.as<char>(oP);
  const OzDesc *dD = (const OzDesc *)ctx->desc.as<char>(oD);
  for (int stage = 0; stage < 2; stage++) {
    if (!si[stage].empty())
      NRG_TIMED(ctx, st, "k_oz_split", hipLaunchKernelGGL(k_oz_split, dim3((unsigned)(si[stage].size() / 2)), dim3(256), 0, st,
                                                          dS, (const int2 *)ctx->desc.as<char>(oSI[stage])));
    if (!gi[stage].empty())
      NRG_TIMED(ctx, st, "k_oz_gemm", hipLaunchKernelGGL(k_oz_gemm, dim3((unsigned)(gi[stage].size() / 2)), dim3(256), 0, st,
                                                         dD, dP, (const int2 *)ctx->desc.as<char>(oGI[stage])));
    if (stage == 0) NRG_HIP(hipEventRecord(ctx->ev[1], st));
  }
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
