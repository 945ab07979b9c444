// gemm3.hip - grouped triple products C += f opA O opB (nrg_gemm3_grouped).
//
// Replaces the table-row loops of recalc_f / recalc_general (c++/recalc.hpp:52-108) and the
// combs loop of cdmI (c++/dmnrg.hpp:54-80), whose arithmetic is product / transform / rotate
// (c++/numerics_ublas.hpp:108-140). Two launches for the whole table:
//   1. every term with a middle factor forms its intermediate in the reference's order:
//        transform (opA = A):    T = O opB          (k x n)   then  C += f A T
//        rotate    (opA = A^T):  T = opA O          (m x l)   then  C += f T opB
//   2. every distinct C gets one descriptor whose K-parts are all its terms, so C is read and
//      written once and the sum over table rows happens in the matrix-core accumulators.
// Matrices are row-major as in the reference; a row-major X is the column-major X^T in the
// same memory, so row-major C = op(X) op(Y) is column-major C^T = op(Y)^T op(X)^T.
#include "gemm.hpp"
#include <cstring>
#include <map>

namespace nrg {

namespace {

struct RmOperand {   // a row-major operand and whether the product uses its transpose
  const double *p;
  int64_t ld;
  int trans;
};

// column-major descriptor of the row-major product C(MxN) = alpha op(X) op(Y) + beta C
GemmDesc rm_desc(double *C, int64_t ldc, RmOperand X, RmOperand Y, int64_t M, int64_t N, int64_t K, double alpha,
                 double beta) {
  GemmDesc g{};
  g.A = Y.p; g.lda = (int)Y.ld; g.tA = Y.trans;
  g.B = X.p; g.ldb = (int)X.ld; g.tB = X.trans;
  g.C = C; g.ldc = (int)ldc;
  g.M = (int)N; g.N = (int)M; g.K = (int)K; g.K2 = 0;
  g.alpha = alpha; g.beta = beta;
  g.parts = nullptr; g.nparts = 0;
  return g;
}

GemmPart rm_part(RmOperand X, RmOperand Y, int64_t K, double scale) {
  GemmPart q;
  q.A = Y.p; q.lda = (int)Y.ld; q.tA = Y.trans;
  q.B = X.p; q.ldb = (int)X.ld; q.tB = X.trans;
  q.K = (int)K; q.scale = scale;
  return q;
}

}  // namespace

// Builds the two stages of a table: gd[0, nd1) form the intermediates (stored in ws2), gd[nd1, ...)
// accumulate into each distinct C, whose K-parts are parts[prange[g].first, prange[g].second).
int gemm3_plan(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, Gemm3Plan &pl) {
  std::vector<int64_t> toff(nterms, -1);
  int64_t tsum = 0;
  for (int i = 0; i < nterms; i++) {
    const nrg_gemm3_term &t = terms[i];
    if (t.m == 0 || t.n == 0 || t.k == 0 || t.l == 0 || !t.O) continue;
    toff[i] = tsum;
    tsum += t.transA ? t.m * t.l : t.k * t.n;
  }
  if (int r = ctx->ws2.ensure(sizeof(double) * (size_t)(tsum + 1))) return r;
  double *Tbase = ctx->ws2.as<double>();
  pl.gd.clear();
  pl.parts.clear();
  pl.prange.clear();
  for (int i = 0; i < nterms; i++) {
    if (toff[i] < 0) continue;
    const nrg_gemm3_term &t = terms[i];
    double *T = Tbase + toff[i];
    if (t.transA)   // T = A^T O  (m x l)
      pl.gd.push_back(rm_desc(T, t.l, {t.A, t.lda, 1}, {t.O, t.ldo, 0}, t.m, t.l, t.k, 1.0, 0.0));
    else            // T = O opB  (k x n)
      pl.gd.push_back(rm_desc(T, t.n, {t.O, t.ldo, 0}, {t.B, t.ldb, !t.transB}, t.k, t.n, t.l, 1.0, 0.0));
  }
  pl.nd1 = pl.gd.size();
  // grouped accumulation, one descriptor per distinct C
  std::map<double *, std::vector<int>> groups;
  std::vector<double *> order;
  for (int i = 0; i < nterms; i++) {
    const nrg_gemm3_term &t = terms[i];
    if (t.m == 0 || t.n == 0 || t.k == 0 || t.l == 0) continue;
    if (!groups.count(t.C)) order.push_back(t.C);
    groups[t.C].push_back(i);
  }
  for (double *C : order) {
    const std::vector<int> &g = groups[C];
    const size_t p0 = pl.parts.size();
    for (int i : g) {
      const nrg_gemm3_term &t = terms[i];
      if (t.m != terms[g[0]].m || t.n != terms[g[0]].n || t.ldc != terms[g[0]].ldc) {
        set_error("nrg_gemm3_grouped: terms sharing C disagree on its shape");
        return NRG_E_ARG;
      }
      const RmOperand opB{t.B, t.ldb, !t.transB};
      if (!t.O) pl.parts.push_back(rm_part({t.A, t.lda, t.transA}, opB, t.k, t.f));
      else if (t.transA) pl.parts.push_back(rm_part({Tbase + toff[i], t.l, 0}, opB, t.l, t.f));
      else pl.parts.push_back(rm_part({t.A, t.lda, 0}, {Tbase + toff[i], t.n, 0}, t.k, t.f));
    }
    const nrg_gemm3_term &t0 = terms[g[0]];
    GemmDesc d{};
    d.C = C; d.ldc = (int)t0.ldc; d.M = (int)t0.n; d.N = (int)t0.m;
    d.alpha = 1.0; d.beta = 1.0; d.nparts = (int)(pl.parts.size() - p0);
    for (size_t q = p0; q < pl.parts.size(); q++) d.K += pl.parts[q].K;   // total K: a tile-shape hint only
    pl.gd.push_back(d);
    pl.prange.push_back({p0, pl.parts.size()});
  }
  return NRG_OK;
}

int gemm3_fp64(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, hipStream_t st) {
  Gemm3Plan pl;
  if (int r = gemm3_plan(ctx, terms, nterms, pl)) return r;
  std::vector<GemmDesc> &gd = pl.gd;
  const size_t nd1 = pl.nd1;
  std::vector<GemmPart> &parts = pl.parts;
  if (gd.empty()) {
    ctx->total_ms = 0;
    return NRG_OK;
  }
  // descriptor blob: [GemmDesc][GemmPart][items]
  std::vector<int> items1, items2;
  gemm_items(gd, 0, nd1, items1);
  gemm_items(gd, nd1, gd.size(), items2);
  const size_t bD = sizeof(GemmDesc) * gd.size(), bP = sizeof(GemmPart) * parts.size();
  const size_t oD = 0, oP = (bD + 255) & ~(size_t)255, oI1 = oP + ((bP + 255) & ~(size_t)255);
  const size_t oI2 = oI1 + sizeof(int) * items1.size();
  const size_t total = oI2 + sizeof(int) * items2.size();
  if (int r = ctx->desc.ensure(total)) return r;
  const GemmPart *dparts = (const GemmPart *)(ctx->desc.as<char>(oP));
  for (size_t gi = nd1; gi < gd.size(); gi++) gd[gi].parts = dparts + pl.prange[gi - nd1].first;
  std::vector<char> blob(total);
  memcpy(blob.data() + oD, gd.data(), bD);
  if (bP) memcpy(blob.data() + oP, parts.data(), bP);
  if (!items1.empty()) memcpy(blob.data() + oI1, items1.data(), sizeof(int) * items1.size());
  if (!items2.empty()) memcpy(blob.data() + oI2, items2.data(), sizeof(int) * items2.size());
  NRG_HIP(hipEventRecord(ctx->ev[0], st));
  if (int r = upload_desc(ctx, blob.data(), blob.size(), st)) return r;
  const GemmDesc *dd = (const GemmDesc *)ctx->desc.as<char>(oD);
  if (!items1.empty()) gemm_launch(ctx, st, dd, ctx->desc.as<int>(oI1), items1.data(), items1.size() / 2);
  NRG_HIP(hipEventRecord(ctx->ev[1], st));
  if (!items2.empty()) gemm_launch(ctx, st, dd, ctx->desc.as<int>(oI2), items2.data(), items2.size() / 2);
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

int gemm3_ozaki(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, hipStream_t st);

int gemm3_run(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, int precision, hipStream_t st) {
  if (nterms < 0 || (nterms > 0 && !terms)) {
    set_error("nrg_gemm3_grouped: invalid arguments");
    return NRG_E_ARG;
  }
  for (int i = 0; i < nterms; i++) {
    const nrg_gemm3_term &t = terms[i];
    if (t.m < 0 || t.n < 0 || t.k < 0 || t.l < 0 || (!t.O && t.k != t.l)) {
      set_error("nrg_gemm3_grouped: term " + std::to_string(i) + " has inconsistent sizes");
      return NRG_E_ARG;
    }
    const bool nz = t.m && t.n && t.k && t.l;
    if (nz && (!t.A || !t.B || !t.C)) {
      set_error("nrg_gemm3_grouped: term " + std::to_string(i) + " has a NULL matrix");
      return NRG_E_ARG;
    }
    if (nz && (t.m > INT32_MAX || t.n > INT32_MAX || t.k > INT32_MAX || t.l > INT32_MAX)) {
      set_error("nrg_gemm3_grouped: dimension too large");
      return NRG_E_ARG;
    }
    if (nz && (t.ldc < t.n || t.lda < (t.transA ? t.m : t.k) || t.ldb < (t.transB ? t.n : t.l) ||
               (t.O && t.ldo < t.l))) {
      set_error("nrg_gemm3_grouped: term " + std::to_string(i) + " has a leading dimension below its row length");
      return NRG_E_ARG;
    }
  }
  if (precision == NRG_GEMM_FP64) return gemm3_fp64(ctx, terms, nterms, st);
  if (precision == NRG_GEMM_OZAKI) return gemm3_ozaki(ctx, terms, nterms, st);
  set_error("nrg_gemm3_grouped: unknown precision");
  return NRG_E_ARG;
}

}  // namespace nrg
