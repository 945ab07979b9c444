// noalloc.hip - nrg_temp_sum_noalloc: the temperature-dependent sums of nrg_temp_sum with no device
// allocation in the call.
//
// nrg_temp_sum stages its pair and tile descriptors and its FDM_COPIES replicated partial sums in the
// context's grow-only buffers, which are allocated (hipMalloc) the first time a call needs more. This
// variant keeps both in `scratch`, one file-scope array of SCRATCH_BYTES per GPU, whose address on
// each device is looked up once, when the first context on that device is created. The call then
// copies the descriptors in, clears the partial sums with a fill kernel, runs the same k_temp_sum
// kernel and adds the sums into *out with the same reduction kernel, and allocates nothing.
// All calls on one GPU share its array: calls that overlap on one GPU race.
#include <mutex>

#include "spectral.hpp"

namespace nrg {

constexpr size_t SCRATCH_BYTES = 4u << 20;   // 4 MiB per GPU
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
std::mutex scratch_mu;
void *scratch_ptr[MAX_DEVICES] = {};   // device address of `scratch` on each device, set by scratch_setup
}  // namespace

// Called by nrg_ctx_create with `device` current: looks up this device's copy of the array once.
int scratch_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) {
    set_error("nrg_ctx_create: device index beyond the scratch table");
    return NRG_E_ARG;
  }
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_ptr[device]) return NRG_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("nrg_ctx_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return NRG_E_HIP;
  }
  scratch_ptr[device] = p;
  return NRG_OK;
}

static void *scratch_of(int device) {
  std::lock_guard<std::mutex> lock(scratch_mu);
  return device >= 0 && device < MAX_DEVICES ? scratch_ptr[device] : nullptr;
}

static int temp_noalloc_run(nrg_ctx *ctx, int kind, const nrg_fdm_pair *pairs, int npairs, double temperature, double Z,
                            double *out, hipStream_t st) {
  if (kind < NRG_TEMP_GT || kind > NRG_TEMP_CHIT || npairs < 0 || (npairs > 0 && !pairs) || !out ||
      !(temperature > 0) || !(Z > 0)) {
    set_error("nrg_temp_sum_noalloc: invalid arguments");
    return NRG_E_ARG;
  }
  if (int r = check_pairs("nrg_temp_sum_noalloc", pairs, npairs, RhoUse::none)) return r;
  unsigned char *base = (unsigned char *)scratch_of(ctx->device);
  if (!base) {
    set_error("nrg_temp_sum_noalloc: no scratch address for this device");
    return NRG_E_HIP;
  }
  // descriptors as nrg_temp_sum builds them: kept x kept pairs, one item per 64 x 64 tile
  std::vector<FdmPairDev> pd;
  std::vector<int> items;
  for (int p = 0; p < npairs; p++) {
    const nrg_fdm_pair &q = pairs[p];
    if (!(q.K_i > 0 && q.K_j > 0)) continue;
    FdmPairDev d{};
    d.dim_i = q.dim_i;
    d.dim_j = q.dim_j;
    d.K_i = q.K_i;
    d.K_j = q.K_j;
    d.ld_op = q.ld_op;
    d.E_i = q.E_i;
    d.E_j = q.E_j;
    d.op1 = q.op1;
    d.op2 = q.op2;
    d.factor = q.factor;
    const int64_t ti = (q.K_i + 63) / 64, tj = (q.K_j + 63) / 64;
    for (int64_t t = 0; t < ti * tj; t++) {
      items.push_back((int)pd.size());
      items.push_back((int)t);
    }
    pd.push_back(d);
  }
  // layout of the scratch: [FDM_COPIES partial sums | pair descriptors | items]
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t oPriv = 0, oPd = al(sizeof(double) * FDM_COPIES), oI = oPd + al(sizeof(FdmPairDev) * pd.size());
  const size_t bytes = oI + sizeof(int) * items.size();
  if (bytes > SCRATCH_BYTES) {
    set_error("nrg_temp_sum_noalloc: the step needs " + std::to_string(bytes) + " bytes of descriptors, beyond the " +
              std::to_string(SCRATCH_BYTES) + "-byte scratch; use nrg_temp_sum");
    return NRG_E_ARG;
  }
  double *priv = (double *)(base + oPriv);
  if (!pd.empty()) NRG_HIP(hipMemcpyAsync(base + oPd, pd.data(), sizeof(FdmPairDev) * pd.size(), hipMemcpyHostToDevice, st));
  if (!items.empty()) NRG_HIP(hipMemcpyAsync(base + oI, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice, st));
  replicas_zero_untimed(st, priv, FDM_COPIES);
  temp_launch(st, kind, (const FdmPairDev *)(base + oPd), (const int2 *)(base + oI), (unsigned)(items.size() / 2),
              1.0 / temperature, Z, priv);
  replicas_reduce_untimed(st, priv, 1, 1, 1, out, out, out, out);
  NRG_HIP(hipGetLastError());
  // the descriptors were copied from this call's host vectors: finish before they go
  NRG_HIP(hipStreamSynchronize(st));
  return NRG_OK;
}

}  // namespace nrg

extern "C" int nrg_temp_sum_noalloc(nrg_ctx *ctx, int kind, const nrg_fdm_pair *pairs, int npairs, double temperature,
                                    double Z, double *out, void *stream) {
  if (!ctx) {
    nrg::set_error("nrg_temp_sum_noalloc: NULL context");
    return NRG_E_ARG;
  }
  try {
    NRG_HIP(hipSetDevice(ctx->device));
    return nrg::temp_noalloc_run(ctx, kind, pairs, npairs, temperature, Z, out, (hipStream_t)stream);
  } catch (const std::exception &e) {
    nrg::set_error(std::string("nrg_temp_sum_noalloc: ") + e.what());
    return NRG_E_HIP;
  }
}
