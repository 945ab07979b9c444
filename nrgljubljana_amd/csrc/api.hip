// api.cpp - C ABI entry points, context, workspace and error reporting.
#include "common.hpp"
#include <cstring>
#include <exception>
#include <string>
#include <algorithm>

#ifndef NRGCORE_VERSION
#define NRGCORE_VERSION "dev"
#endif

namespace nrg {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int DevBuf::ensure(size_t need) {
  if (need <= bytes) return NRG_OK;
  if (p) {
    hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  size_t want = need + need / 8 + 4096;
  if (hipMalloc(&p, want) != hipSuccess) {
    p = nullptr;
    set_error("device workspace allocation of " + std::to_string(want) + " bytes failed");
    return NRG_E_NOMEM;
  }
  bytes = want;
  return NRG_OK;
}
void DevBuf::release() {
  if (p) hipFree(p);
  p = nullptr;
  bytes = 0;
}
int PinnedBuf::ensure(size_t need) {
  if (need <= bytes) return NRG_OK;
  if (p) {
    hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  size_t want = need + need / 8 + 4096;
  if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
    p = nullptr;
    set_error("pinned host allocation failed");
    return NRG_E_NOMEM;
  }
  bytes = want;
  return NRG_OK;
}
void PinnedBuf::release() {
  if (p) hipHostFree(p);
  p = nullptr;
  bytes = 0;
}

int stage_desc(nrg_ctx *ctx, size_t bytes, hipStream_t st, char **host) {
  // both buffers are reused: make sure the previous upload and its readers have left them
  NRG_HIP(hipStreamSynchronize(st));
  if (int r = ctx->desc.ensure(bytes)) return r;
  if (int r = ctx->hdesc.ensure(bytes)) return r;
  *host = (char *)ctx->hdesc.p;
  return NRG_OK;
}

int upload_staged(nrg_ctx *ctx, size_t bytes, hipStream_t st) {
  if (bytes > ctx->hdesc.bytes || bytes > ctx->desc.bytes) {
    set_error("upload_staged: more bytes than were staged");
    return NRG_E_ARG;
  }
  NRG_HIP(hipMemcpyAsync(ctx->desc.p, ctx->hdesc.p, bytes, hipMemcpyHostToDevice, st));
  return NRG_OK;
}

int upload_desc(nrg_ctx *ctx, const void *src, size_t bytes, hipStream_t st) {
  char *h = nullptr;
  if (int r = stage_desc(ctx, bytes, st, &h)) return r;
  memcpy(h, src, bytes);
  return upload_staged(ctx, bytes, st);
}

size_t KTimer::begin(hipStream_t st) {
  if (used + 2 > pool.size()) {
    const size_t grow = std::max<size_t>(256, pool.size());
    for (size_t i = 0; i < grow; i++) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) break;
      pool.push_back(e);
    }
  }
  if (used + 2 > pool.size()) return (size_t)-1;
  const size_t b = used;
  used += 2;
  hipEventRecord(pool[b], st);
  return b;
}
void KTimer::end(hipStream_t st, size_t b, const char *name) {
  if (b == (size_t)-1) return;
  hipEventRecord(pool[b + 1], st);
  marks.push_back({name, b, b + 1});
}
void KTimer::collect() {
  for (const Mark &m : marks) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pool[m.b], pool[m.e]) == hipSuccess) {
      auto &a = acc[m.name];
      a.first += ms;
      a.second += 1;
    }
  }
  marks.clear();
  used = 0;
}
void KTimer::reset() {
  marks.clear();
  used = 0;
  acc.clear();
}
void KTimer::release() {
  for (hipEvent_t e : pool) hipEventDestroy(e);
  pool.clear();
  reset();
}

int ensure_aux_streams(nrg_ctx *ctx) {
  if (ctx->fork_ev) return NRG_OK;
  for (int i = 0; i < nrg_ctx::NAUX; i++) {
    if (!ctx->aux[i]) NRG_HIP(hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking));
    if (!ctx->join_ev[i]) NRG_HIP(hipEventCreateWithFlags(&ctx->join_ev[i], hipEventDisableTiming));
  }
  NRG_HIP(hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
  NRG_HIP(hipEventCreateWithFlags(&ctx->gfork, hipEventDisableTiming));
  NRG_HIP(hipEventCreateWithFlags(&ctx->gjoin, hipEventDisableTiming));
  NRG_HIP(hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
  return NRG_OK;
}

}  // namespace nrg

using namespace nrg;

extern "C" {

const char *nrg_last_error(void) { return g_err.c_str(); }

const char *nrg_version(void) { return "nrgcore " NRGCORE_VERSION " gfx950"; }

int nrg_ctx_create(nrg_ctx **out, int device) {
  if (!out) {
    set_error("nrg_ctx_create: NULL output pointer");
    return NRG_E_ARG;
  }
  *out = nullptr;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
      set_error("nrg_ctx_create: no such HIP device");
      return NRG_E_NODEV;
    }
    hipDeviceProp_t prop;
    NRG_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
      set_error(std::string("nrg_ctx_create: device is ") + prop.gcnArchName + ", this library is built for gfx950");
      return NRG_E_NODEV;
    }
    NRG_HIP(hipSetDevice(device));
    if (int r = scratch_setup(device)) return r;
    nrg_ctx *c = new nrg_ctx();
    c->device = device;
    for (int i = 0; i <= NPHASE; i++) {
      if (hipEventCreate(&c->ev[i]) != hipSuccess) {
        delete c;
        set_error("nrg_ctx_create: hipEventCreate failed");
        return NRG_E_HIP;
      }
    }
    for (int i = 0; i < NPHASE; i++) c->phase_ms[i] = 0;
    *out = c;
    return NRG_OK;
  } catch (const std::exception &e) {
    set_error(std::string("nrg_ctx_create: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_ctx_destroy(nrg_ctx *c) {
  if (!c) return NRG_OK;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  c->syev_cache.reset();   // the eigensolver's graphs and their descriptor buffers
  c->ws.release();
  c->dcbuf.release();
  c->ws2.release();
  c->desc.release();
  c->flags.release();
  c->hdesc.release();
  c->hflags.release();
  c->kt.release();
  for (int i = 0; i <= NPHASE; i++) hipEventDestroy(c->ev[i]);
  for (int i = 0; i < nrg_ctx::NAUX; i++) {
    if (c->aux[i]) hipStreamDestroy(c->aux[i]);
    if (c->join_ev[i]) hipEventDestroy(c->join_ev[i]);
  }
  if (c->fork_ev) hipEventDestroy(c->fork_ev);
  if (c->gstream) hipStreamDestroy(c->gstream);
  if (c->gfork) hipEventDestroy(c->gfork);
  if (c->gjoin) hipEventDestroy(c->gjoin);
  delete c;
  return NRG_OK;
}

int nrg_syev_batched(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H_arena,
                     const int64_t *off_w, double *w_arena, int mode, int *info, void *stream) {
  if (!ctx) {
    set_error("nrg_syev_batched: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return syev_run(ctx, nblk, n, off_H, H_arena, off_w, w_arena, mode, info, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_syev_batched: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_gemm3_grouped(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, int precision, void *stream) {
  if (!ctx) {
    set_error("nrg_gemm3_grouped: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return gemm3_run(ctx, terms, nterms, precision, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_gemm3_grouped: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_fdm_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double wnf,
                int sign, double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg, void *stream) {
  if (!ctx) {
    set_error("nrg_fdm_bin: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return fdm_run(ctx, pairs, npairs, grid, T, wnf, sign, gt_pos, gt_neg, ls_pos, ls_neg, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_fdm_bin: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_dmnrg_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double emin,
                  double emax, int sign, double *out_pos, double *out_neg, void *stream) {
  if (!ctx) {
    set_error("nrg_dmnrg_bin: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return dmnrg_run(ctx, pairs, npairs, grid, emin, emax, sign, out_pos, out_neg, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_dmnrg_bin: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_hamiltonian(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H_arena,
                    const nrg_ham_term *terms, int nterms, void *stream) {
  if (!ctx) {
    set_error("nrg_hamiltonian: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return hamiltonian_run(ctx, nblk, n, off_H, H_arena, terms, nterms, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_hamiltonian: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_cfs_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z,
                int last_step, int sign, double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg, void *stream) {
  if (!ctx) {
    set_error("nrg_cfs_bin: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return cfs_run(ctx, pairs, npairs, grid, T, Z, last_step, sign, gt_pos, gt_neg, ls_pos, ls_neg, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_cfs_bin: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_ft_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z,
               int sign, double *out_pos, double *out_neg, void *stream) {
  if (!ctx) {
    set_error("nrg_ft_bin: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return cfs_run(ctx, pairs, npairs, grid, T, Z, 2, sign, out_pos, out_neg, nullptr, nullptr, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_ft_bin: ") + e.what());
    return NRG_E_HIP;
  }
}

static int mats_entry(const char *who, nrg_ctx *ctx, int mode, const nrg_fdm_pair *pairs, int npairs, int64_t mats,
                      double T, double wnf, double Z, int sign, double *G, void *stream) {
  if (!ctx) {
    set_error(std::string(who) + ": NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return mats_run(ctx, mode, pairs, npairs, mats, T, wnf, Z, sign, G, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string(who) + ": " + e.what());
    return NRG_E_HIP;
  }
}

int nrg_fdm_mats(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, double wnf, int sign,
                 double *G, void *stream) {
  return mats_entry("nrg_fdm_mats", ctx, 0, pairs, npairs, mats, T, wnf, 1.0, sign, G, stream);
}

int nrg_dmnrg_mats(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, int sign, double *G,
                   void *stream) {
  return mats_entry("nrg_dmnrg_mats", ctx, 1, pairs, npairs, mats, T, 1.0, 1.0, sign, G, stream);
}

int nrg_ft_mats(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, double Z, int sign,
                double *G, void *stream) {
  return mats_entry("nrg_ft_mats", ctx, 2, pairs, npairs, mats, T, 1.0, Z, sign, G, stream);
}

int nrg_temp_sum(nrg_ctx *ctx, int kind, const nrg_fdm_pair *pairs, int npairs, double temperature, double Z,
                 double *out, void *stream) {
  if (!ctx) {
    set_error("nrg_temp_sum: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return temp_run(ctx, kind, pairs, npairs, temperature, Z, out, (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_temp_sum: ") + e.what());
    return NRG_E_HIP;
  }
}

int nrg_broaden(nrg_ctx *ctx, const double *pos_e, const double *pos_w, int64_t npos, const double *neg_e,
                const double *neg_w, int64_t nneg, const double *mesh, int64_t nmesh, double alpha, double omega0,
                double *out_pos, double *out_neg, void *stream) {
  if (!ctx) {
    set_error("nrg_broaden: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return broaden_run(ctx, pos_e, pos_w, npos, neg_e, neg_w, nneg, mesh, nmesh, alpha, omega0, out_pos, out_neg,
                       (hipStream_t)stream);
  } catch (const std::exception &e) {
    set_error(std::string("nrg_broaden: ") + e.what());
    return NRG_E_HIP;
  }
}

double nrg_last_time_ms(nrg_ctx *ctx, int phase) {
  if (!ctx) return -1.0;
  if (phase == 0) return ctx->total_ms;
  if (phase < 1 || phase > NPHASE) return -1.0;
  return ctx->phase_ms[phase - 1];
}

int nrg_kernel_timing(nrg_ctx *ctx, int enable) {
  if (!ctx) {
    set_error("nrg_kernel_timing: NULL context");
    return NRG_E_ARG;
  }
  ctx->kt.reset();
  ctx->kt.on = enable != 0;
  return NRG_OK;
}

int nrg_kernel_time(nrg_ctx *ctx, const char *name, double *total_ms, long *launches) {
  if (!ctx || !name) {
    set_error("nrg_kernel_time: NULL argument");
    return NRG_E_ARG;
  }
  auto it = ctx->kt.acc.find(name);
  if (total_ms) *total_ms = it == ctx->kt.acc.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == ctx->kt.acc.end() ? 0 : it->second.second;
  return NRG_OK;
}

int nrg_kernel_names(nrg_ctx *ctx, char *buf, int buflen) {
  if (!ctx || !buf || buflen <= 0) {
    set_error("nrg_kernel_names: invalid argument");
    return NRG_E_ARG;
  }
  std::string s;
  for (const auto &kv : ctx->kt.acc) {
    if (!s.empty()) s += ",";
    s += kv.first;
  }
  if ((int)s.size() >= buflen) {
    set_error("nrg_kernel_names: buffer too small");
    return NRG_E_ARG;
  }
  memcpy(buf, s.c_str(), s.size() + 1);
  return NRG_OK;
}

int nrg_syev_tuning(nrg_ctx *ctx, char *buf, int buflen) {
  if (!ctx || !buf || buflen <= 0) {
    set_error("nrg_syev_tuning: invalid argument");
    return NRG_E_ARG;
  }
  const std::string s = syev_tuning_json(ctx);
  if ((int)s.size() >= buflen) {
    set_error("nrg_syev_tuning: buffer too small");
    return NRG_E_ARG;
  }
  memcpy(buf, s.c_str(), s.size() + 1);
  return NRG_OK;
}

const char *nrg_phase_name(int phase) {
  static const char *names[] = {"total",          "reduce (tridiagonal / dense products)",
                                "tridiagonal eigensolver (divide and conquer)",
                                "back-transformation",
                                "refinement",     "finish",
                                "phase6",         "phase7",
                                "phase8"};
  if (phase < 0 || phase > NPHASE) return "";
  return names[phase];
}

}  // extern "C"
