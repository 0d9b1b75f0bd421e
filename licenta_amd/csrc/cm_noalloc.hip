// cm_tri_update_noalloc: the swept leaf AABBs and Morton keys of cm_tri_update
// without any device allocation, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc). The
// normalised 30-bit path needs the triangle centroids and per-block centroid boxes
// between its three kernels; here they live in a file-scope device array.
#include "cm_internal.hpp"

#include <mutex>

namespace cm {
namespace {

// Largest triangle count the static scratch serves.
constexpr int kNoAllocTriMax = 1 << 18;  // 262,144 triangles
constexpr int kNoAllocBlocks = (kNoAllocTriMax + kBlock - 1) / kBlock;
// centroids [0, kNoAllocTriMax), then the block boxes and the reduced box
// (2 * blocks + 2 entries): 4,227,104 bytes per GPU, held until the process ends
constexpr size_t kScratchF4 = (size_t)kNoAllocTriMax + 2 * (size_t)kNoAllocBlocks + 2;

// One copy per GPU. In host code this name is only a stand-in for the symbol:
// kernels get the device address that hipGetSymbolAddress returns.
static __device__ __align__(256) float4 scratch[kScratchF4];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float4* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch, looked up once per device with
// that device current.
int scratch_address(float4** out) {
  int dev = 0;
  CM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("cm_tri_update_noalloc: device ordinal out of range");
    return CM_ERR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("cm_tri_update_noalloc: hipGetSymbolAddress: ") +
                hipGetErrorString(e));
      return CM_ERR_HIP;
    }
    g_scratch_ptr[dev] = (float4*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return CM_OK;
}

}  // namespace
}  // namespace cm

using namespace cm;

extern "C" int cm_tri_update_noalloc(int32_t n, int32_t t, const void* d_x, const void* d_p,
                                     const void* d_tri, float thickness, int32_t morton_bits,
                                     int32_t morton_normalize, void* d_aabb_min, void* d_aabb_max,
                                     uint64_t* d_keys, void* stream) {
  if (t < 1 || n < 1 || !d_x || !d_p || !d_tri || !d_aabb_min || !d_aabb_max || !d_keys ||
      (morton_bits != 30 && morton_bits != 63)) {
    set_error("cm_tri_update_noalloc: invalid argument");
    return CM_ERR_ARG;
  }
  if (t > kNoAllocTriMax) {
    set_error("cm_tri_update_noalloc: " + std::to_string(t) +
              " triangles exceed the static scratch (at most " + std::to_string(kNoAllocTriMax) +
              "); use cm_tri_update");
    return CM_ERR_ARG;
  }
  const bool norm = morton_bits == 30 && morton_normalize;
  float4* cen = nullptr;
  float4* box = nullptr;
  if (norm) {
    float4* base = nullptr;
    const int rc = scratch_address(&base);
    if (rc) return rc;
    cen = base;                    // t entries, each written before it is read
    box = base + kNoAllocTriMax;   // 2 * nblocks(t) + 2 entries, likewise
  }
  return launch_tri_update(n, t, (const float4*)d_x, (const float4*)d_p, (const int4*)d_tri,
                           thickness, morton_bits, morton_normalize, (float4*)d_aabb_min,
                           (float4*)d_aabb_max, d_keys, nullptr, cen, box, (hipStream_t)stream);
}
