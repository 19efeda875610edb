// noalloc.hip - the allocation-free entry point smcm_estimate_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc) get the accuracy-mode estimate c_[0][L-1] of smcm_estimate
// with every scratch buffer of the estimator (row densities R_i, the hub-row factor buffer, the
// work-item lists, partial sums and the cost table) carved from one file-scope device array,
// `scratch`.  The array's device address is looked up once per device with hipGetSymbolAddress
// (at context creation) and only that address is handed to kernels and fills.  The call queues
// its work on the context's stream and returns: nothing is allocated, freed, copied to the host
// or waited for.  Every call on a device shares that device's array, so calls that overlap on one
// GPU (other contexts, other threads) race; calls on one context are ordered by its stream.
#include <mutex>

#include "api_internal.h"
#include "kernels.h"
#include "smcm.h"

using namespace smcm_host;

namespace {

// the scratch of smcm_estimate_noalloc on each device
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

__global__ void copy_f64_kernel(const double* __restrict__ src, double* __restrict__ dst) { *dst = *src; }

}  // namespace

namespace smcm_host {

// Device address of `scratch` on `device` (the current device), looked up once per device.
void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) fail(SMCM_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      fail(SMCM_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

}  // namespace smcm_host

extern "C" {

int smcm_estimate_noalloc(smcm_ctx* ctx, smcm_graph* g, const int32_t* vtypes, const int32_t* etypes, int32_t L,
                          double* d_est_nnz) {
  return guarded([&] {
    check_ctx(ctx);
    if (!g || !vtypes || !etypes || !d_est_nnz) fail(SMCM_EINVAL, "null argument");
    if (g->ctx != ctx) fail(SMCM_EINVAL, "the graph belongs to another context");
    if (L < 1 || L > 16) fail(SMCM_EINVAL, "the meta-path is incorrect: L must be in [1, 16]");
    for (int32_t q = 0; q <= L; ++q)
      if (vtypes[q] < 0 || vtypes[q] >= g->ntypes) fail(SMCM_EINVAL, "the meta-path is incorrect: vertex type not in graph");
    for (int32_t q = 0; q < L; ++q)
      if (etypes[q] < 0) fail(SMCM_EINVAL, "the meta-path is incorrect: negative edge type");
    // the typed matrices must already be on the device (building one allocates)
    std::vector<smcm::Csr> mats(L);
    {
      std::lock_guard<std::recursive_mutex> lk(g->cache_mu);
      for (int32_t i = 0; i < L; ++i) {
        auto it = g->cache.find(std::make_tuple(vtypes[i], etypes[i], vtypes[i + 1]));
        if (it == g->cache.end())
          fail(SMCM_EINVAL, "typed matrix " + std::to_string(i) + " of the meta-path is not on the device yet: "
                            "run smcm_estimate or smcm_chain_run on this meta-path once first");
        mats[i] = it->second;
      }
    }
    // the row densities alone must fit; the pass temporaries are checked as they are carved
    size_t need = (size_t)L * L * sizeof(double);
    for (int32_t i = 0; i < L; ++i) need += (size_t)mats[i].nrows * (size_t)(L - i) * sizeof(double);
    if (need > kScratchBytes)
      fail(SMCM_EINVAL, "the estimate needs more than the " + std::to_string(kScratchBytes >> 20) +
                        " MiB of scratch of smcm_estimate_noalloc (" + std::to_string(need) +
                        " bytes of row densities alone); use smcm_estimate");
    smcm::DevAlloc arena;
    arena.stream = ctx->stream;
    arena.arena = (unsigned char*)noalloc_scratch(ctx->device);
    arena.arena_cap = kScratchBytes;
    hipStream_t st = ctx->stream;
    double* c = arena.alloc_n<double>((int64_t)L * L);
    smcm::fill_i64((int64_t*)c, (int64_t)L * L, 0, st);  // +0.0; a fill kernel, no memset node
    std::vector<double*> R(L, nullptr);
    for (int32_t i = L - 1; i >= 0; --i) {
      const int W = L - 1 - i;
      R[i] = arena.alloc_n<double>((int64_t)mats[i].nrows * (W + 1));
      std::vector<double> dimj(W > 0 ? W : 1);
      for (int q = 0; q < W; ++q) dimj[q] = (double)g->dims[vtypes[i + 2 + q]];
      smcm::EstimatorPass p;
      p.M = &mats[i];
      p.leaf_dim = (double)g->dims[vtypes[i + 1]];
      p.Rn = W > 0 ? R[i + 1] : nullptr;
      p.W = W;
      p.dimj = dimj.data();
      p.Ri = R[i];
      p.c_dev = c + (size_t)i * L + i + 1;
      p.exact_sum = ctx->exact_cost_sum;
      p.tm = nullptr;
      smcm::estimator_pass(p, arena);
    }
    copy_f64_kernel<<<1, 1, 0, st>>>(c + (L - 1), d_est_nnz);
    SMCM_LAUNCH_CHECK();
  });
}

}  // extern "C"
