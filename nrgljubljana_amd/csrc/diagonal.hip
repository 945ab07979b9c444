// diagonal.hip - diagonals of many device matrices in one call (nrg_diagonal_batched).
//
// The density-matrix sweeps set the diagonal of every shell matrix (rho(i,i) = exp(-E_i/T) w_N / Z_N,
// c++/dmnrg.hpp:145-162) and take the trace of every matrix of a step (check_trace_rho,
// c++/dmnrg.hpp:29-33). Element-wise host transfers of a diagonal are one DMA row per element; here
// the values of all matrices of a call travel in one contiguous copy and one launch gathers or
// scatters them.
#include <cstring>

#include "common.hpp"

namespace nrg {

namespace {

constexpr int DIAG_CHUNK = 1024;   // diagonal entries per workgroup

// work item (k, c): entries [c DIAG_CHUNK, (c + 1) DIAG_CHUNK) of item k
__global__ __launch_bounds__(256) void k_diag_get(const nrg_diag_item *items, const int2 *work, double *vals) {
  const int2 w = work[blockIdx.x];
  const nrg_diag_item it = items[w.x];
  const int64_t i0 = (int64_t)w.y * DIAG_CHUNK, i1 = min(it.count, i0 + DIAG_CHUNK);
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) vals[it.off + i] = it.A[(it.first + i) * (it.ld + 1)];
}

__global__ __launch_bounds__(256) void k_diag_set(const nrg_diag_item *items, const int2 *work, const double *vals) {
  const int2 w = work[blockIdx.x];
  const nrg_diag_item it = items[w.x];
  const int64_t i0 = (int64_t)w.y * DIAG_CHUNK, i1 = min(it.count, i0 + DIAG_CHUNK);
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) it.A[(it.first + i) * (it.ld + 1)] = vals[it.off + i];
}

}  // namespace

}  // namespace nrg

using namespace nrg;

extern "C" int nrg_diagonal_batched(nrg_ctx *ctx, int op, const nrg_diag_item *items, int nitems, double *values,
                                    void *stream) {
  try {
    if (!ctx || nitems < 0 || (nitems > 0 && (!items || !values)) || (op != NRG_DIAG_GET && op != NRG_DIAG_SET)) {
      set_error("nrg_diagonal_batched: invalid arguments");
      return NRG_E_ARG;
    }
    if (nitems == 0) return NRG_OK;
    int64_t nval = 0;
    std::vector<int> work;
    for (int k = 0; k < nitems; k++) {
      const nrg_diag_item &it = items[k];
      if (it.count == 0) continue;
      if (it.count < 0 || it.first < 0 || it.off < 0 || it.ld < 1 || !it.A || it.first + it.count > it.ld) {
        set_error("nrg_diagonal_batched: item " + std::to_string(k) + " has an invalid range");
        return NRG_E_ARG;
      }
      nval = std::max(nval, it.off + it.count);
      for (int64_t c = 0; c * DIAG_CHUNK < it.count; c++) {
        work.push_back(k);
        work.push_back((int)c);
      }
    }
    if (work.empty()) return NRG_OK;
    NRG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    // blob: [items][work][values (SET)]
    const size_t bI = sizeof(nrg_diag_item) * nitems, bW = sizeof(int) * work.size(), bV = sizeof(double) * nval;
    const size_t oW = (bI + 255) & ~(size_t)255, oV = oW + ((bW + 255) & ~(size_t)255);
    const size_t total = op == NRG_DIAG_SET ? oV + bV : oW + bW;
    char *h = nullptr;
    if (int r = stage_desc(ctx, total, st, &h)) return r;
    std::memcpy(h, items, bI);
    std::memcpy(h + oW, work.data(), bW);
    if (op == NRG_DIAG_SET) std::memcpy(h + oV, values, bV);
    if (int r = upload_staged(ctx, total, st)) return r;
    const nrg_diag_item *di = (const nrg_diag_item *)ctx->desc.as<char>(0);
    const int2 *dw = (const int2 *)ctx->desc.as<char>(oW);
    const dim3 grid((unsigned)(work.size() / 2));
    if (op == NRG_DIAG_SET) {
      hipLaunchKernelGGL(k_diag_set, grid, dim3(256), 0, st, di, dw, (const double *)ctx->desc.as<char>(oV));
      NRG_HIP(hipGetLastError());
      NRG_HIP(hipStreamSynchronize(st));   // the staging buffer is reused by the next call
      return NRG_OK;
    }
    if (int r = ctx->ws2.ensure(bV)) return r;
    double *dv = ctx->ws2.as<double>();
    hipLaunchKernelGGL(k_diag_get, grid, dim3(256), 0, st, di, dw, dv);
    NRG_HIP(hipGetLastError());
    if (int r = ctx->hflags.ensure(bV)) return r;
    NRG_HIP(hipStreamSynchronize(st));
    NRG_HIP(hipMemcpyAsync(ctx->hflags.p, dv, bV, hipMemcpyDeviceToHost, st));
    NRG_HIP(hipStreamSynchronize(st));
    std::memcpy(values, ctx->hflags.p, bV);
    return NRG_OK;
  } catch (const std::exception &e) {
    set_error(std::string("nrg_diagonal_batched: ") + e.what());
    return NRG_E_HIP;
  }
}
