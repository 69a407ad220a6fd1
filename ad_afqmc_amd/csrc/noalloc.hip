// Comb resampling with no allocation: afq_walkers_sr_noalloc, the variant of
// afq_walkers_sr whose workspace (cumulative weights, indices, gather staging)
// is a file-scope device array of this library instead of a caller buffer.
//
// The array's host-side name is only a stand-in for the symbol: host code uses
// the device address hipGetSymbolAddress returns, fetched once per device with
// that device current (on the first call made on it) and kept per device.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (on different streams, or from different threads) race.
#include <mutex>

#include "kernels.h"

namespace {

constexpr size_t kScratchBytes = 16u << 20;  // 16 MiB per device
constexpr int kMaxWalkers = 65535;  // grid y-dimension limit of the gather launch
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices] = {};

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// dst (W x ncols) = columns [c0, c0 + ncols) of walker idx[w] of src (W x per).
__global__ void gather_cols_kernel(const double2* __restrict__ src, long per, long c0, long ncols,
                                   const int32_t* __restrict__ idx, double2* __restrict__ dst) {
  const long w = blockIdx.y;
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= ncols) return;
  dst[w * ncols + i] = src[(long)idx[w] * per + c0 + i];
}

int device_scratch(void** out) {
  int dev = 0;
  AFQ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return afq::fail(AFQ_E_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    AFQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return AFQ_OK;
}

}  // namespace

extern "C" {

int64_t afq_walkers_sr_noalloc_max_walkers(void) { return kMaxWalkers; }

int afq_walkers_sr_noalloc(void* up, int64_t elems_up, void* dn, int64_t elems_dn, int32_t W, double* weights,
                           const double* zeta, int32_t* idx_out, void* stream) {
  if (!up || !weights || !zeta || W < 1 || elems_up < 1 || (dn && elems_dn < 1))
    return afq::fail(AFQ_E_INVALID, "bad argument");
  if (W > kMaxWalkers)
    return afq::fail(AFQ_E_INVALID, "afq_walkers_sr_noalloc: n_walkers " + std::to_string(W) +
                                        " exceeds the static scratch limit of " + std::to_string(kMaxWalkers));
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  void* base = nullptr;
  AFQ_TRY(device_scratch(&base));
  char* ws = static_cast<char*>(base);
  double* cum = reinterpret_cast<double*>(ws);
  ws += align256(sizeof(double) * (size_t)W);
  int32_t* idx = reinterpret_cast<int32_t*>(ws);
  ws += align256(sizeof(int32_t) * (size_t)W);
  double2* stage = reinterpret_cast<double2*>(ws);
  // walker elements staged per chunk: the gather of columns [c0, c0 + chunk)
  // reads and then overwrites only those columns, so chunks are independent
  const long chunk = (long)((kScratchBytes - (size_t)(ws - static_cast<char*>(base))) / (sizeof(double2) * (size_t)W));
  AFQ_TRY(afq::k_sr_index(weights, zeta, W, cum, idx, st));
  void* bufs[2] = {up, dn};
  const int64_t elems[2] = {elems_up, elems_dn};
  for (int s = 0; s < 2; ++s) {
    if (!bufs[s]) continue;
    double2* Wp = static_cast<double2*>(bufs[s]);
    const long per = elems[s];
    for (long c0 = 0; c0 < per; c0 += chunk) {
      const long nc = std::min(chunk, per - c0);
      hipLaunchKernelGGL(gather_cols_kernel, dim3((unsigned)((nc + 255) / 256), (unsigned)W), dim3(256), 0, st, Wp,
                         per, c0, nc, idx, stage);
      AFQ_TRY(afq::launch_status("gather_cols_kernel"));
      AFQ_HIP(hipMemcpy2DAsync(Wp + c0, sizeof(double2) * per, stage, sizeof(double2) * nc, sizeof(double2) * nc,
                               (size_t)W, hipMemcpyDeviceToDevice, st));
    }
  }
  if (idx_out) AFQ_HIP(hipMemcpyAsync(idx_out, idx, sizeof(int32_t) * W, hipMemcpyDeviceToDevice, st));
  return AFQ_OK;
}

}  // extern "C"
