// Device-wide primitives of libclothmi, written for CDNA4 wave64:
//  * exclusive int32 scan over a device-side element count (reduce / scan
//    partials / rescan, fixed grid so the sequence is graph-capturable);
//  * stable LSD radix sort of (uint64 key, uint32 value) with 8-bit digits.
// They replace the CUB primitives the reference calls through
// LicentaV2/LicentaV2/Core/CubWrappers.cuh:24-183 (DeviceScan::ExclusiveSum,
// DeviceSelect::Flagged, DeviceRadixSort::SortPairs). Integer results are exact
// and independent of the launch shape.
#include "cm_internal.hpp"

namespace cm {

namespace {

constexpr int kScanGrid = 256;  // blocks of the reduce / rescan kernels

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ int wave_reduce(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exclusive scan across a 256-thread block; lds needs 5 ints.
__device__ __forceinline__ int block_excl_scan(int v, int* lds, int& total) {
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  int inc = wave_incl_scan(v);
  if (lane == 63) lds[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
      int t = lds[w];
      lds[w] = run;
      run += t;
    }
    lds[kBlock / 64] = run;
  }
  __syncthreads();
  int res = inc - v + lds[wid];
  total = lds[kBlock / 64];
  __syncthreads();
  return res;
}

__device__ __forceinline__ void scan_range(int n_max, const int* d_n, int& n, int& begin,
                                           int& end) {
  n = d_n ? min(*d_n, n_max) : n_max;
  if (n < 0) n = 0;
  int chunk = (n + kScanGrid - 1) / kScanGrid;
  chunk = (chunk + kBlock - 1) / kBlock * kBlock;
  begin = min((int)blockIdx.x * chunk, n);
  end = min(begin + chunk, n);
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const int* __restrict__ in, int n_max,
                                                        const int* __restrict__ d_n,
                                                        int* __restrict__ partials) {
  __shared__ int lds[kBlock / 64];
  int n, begin, end;
  scan_range(n_max, d_n, n, begin, end);
  int s = 0;
  for (int i = begin + threadIdx.x; i < end; i += kBlock) s += in[i];
  s = wave_reduce(s);
  if (lane_id() == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
    partials[blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_partials(int* __restrict__ partials,
                                                          int* __restrict__ d_total) {
  __shared__ int lds[kBlock / 64 + 1];
  // kScanGrid == kBlock: one partial per thread
  int v = partials[threadIdx.x];
  int total;
  int ex = block_excl_scan(v, lds, total);
  partials[threadIdx.x] = ex;
  if (threadIdx.x == 0 && d_total) *d_total = total;
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(const int* __restrict__ in,
                                                       int* __restrict__ out, int n_max,
                                                       const int* __restrict__ d_n,
                                                       const int* __restrict__ partials) {
  __shared__ int lds[kBlock / 64 + 1];
  int n, begin, end;
  scan_range(n_max, d_n, n, begin, end);
  int base = partials[blockIdx.x];
  for (int tile = begin; tile < end; tile += kBlock) {
    int i = tile + threadIdx.x;
    int v = i < end ? in[i] : 0;
    int total;
    int ex = block_excl_scan(v, lds, total);
    if (i < end) out[i] = base + ex;
    base += total;
  }
}

__global__ void k_fill_i32(int* __restrict__ p, int v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ---- radix sort ------------------------------------------------------------
constexpr int kSortItems = 8;                    // keys per thread per tile
constexpr int kSortTile = kBlock * kSortItems;   // 2048 keys per block
constexpr int kWaves = kBlock / 64;

__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint64_t* __restrict__ keys, int n,
                                                       int shift, int nblk,
                                                       int* __restrict__ hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    int i = base + k * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
  }
  __syncthreads();
  hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: wave w ranks the contiguous slice [w*64*ITEMS, (w+1)*64*ITEMS)
// of the tile in order, using 8 ballots per 64-key chunk to find equal digits.
__global__ void __launch_bounds__(kBlock) k_radix_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, int n, int shift, int nblk,
    const int* __restrict__ offsets) {
  __shared__ int cnt[kWaves][256];
  __shared__ int wbase[kWaves][256];
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  for (int d = threadIdx.x; d < kWaves * 256; d += kBlock) (&cnt[0][0])[d] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile + wid * 64 * kSortItems;
  uint64_t key[kSortItems];
  uint32_t val[kSortItems];
  int rank[kSortItems];
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int c = 0; c < kSortItems; ++c) {
    const int i = base + c * 64 + lane;
    const bool valid = i < n;
    key[c] = valid ? kin[i] : 0ull;
    val[c] = valid ? vin[i] : 0u;
    const uint32_t d = (uint32_t)(key[c] >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    int r = 0;
    if (valid) {
      const int before = __popcll(m & lt);
      const int old = cnt[wid][d];
      r = old + before;
      if (before == 0) cnt[wid][d] = old + __popcll(m);
    }
    rank[c] = r;
  }
  __syncthreads();
  {
    const int d = threadIdx.x;  // kBlock == 256 digits
    int run = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      wbase[w][d] = run;
      run += cnt[w][d];
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < kSortItems; ++c) {
    const int i = base + c * 64 + lane;
    if (i < n) {
      const uint32_t d = (uint32_t)(key[c] >> shift) & 255u;
      const int pos = offsets[d * nblk + blockIdx.x] + wbase[wid][d] + rank[c];
      kout[pos] = key[c];
      vout[pos] = val[c];
    }
  }
}

}  // namespace

size_t scan_temp_bytes(int /*n_max*/) { return align_up(sizeof(int) * kScanGrid); }

int scan_exclusive(const int* in, int* out, int n_max, const int* d_n, int* d_total, void* temp,
                   hipStream_t s) {
  int* partials = (int*)temp;
  hipLaunchKernelGGL(k_scan_reduce, dim3(kScanGrid), dim3(kBlock), 0, s, in, n_max, d_n, partials);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kBlock), 0, s, partials, d_total);
  hipLaunchKernelGGL(k_scan_apply, dim3(kScanGrid), dim3(kBlock), 0, s, in, out, n_max, d_n,
                     partials);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int fill_i32(int* p, int v, int64_t n, hipStream_t s) {
  if (n <= 0) return CM_OK;
  int grid = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 4096);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid), dim3(kBlock), 0, s, p, v, n);
  CM_LAUNCH_CHECK();
  return CM_OK;
}
int fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
  return fill_i32((int*)p, (int)v, n, s);
}

static int sort_blocks(int n) { return (n + kSortTile - 1) / kSortTile; }

size_t radix_sort_temp_bytes(int n) {
  int nb = sort_blocks(n);
  return align_up(sizeof(uint64_t) * (size_t)n) + align_up(sizeof(uint32_t) * (size_t)n) +
         align_up(sizeof(int) * 256 * (size_t)nb) + scan_temp_bytes(256 * nb);
}

int radix_sort(uint64_t* keys, uint32_t* vals, int n, int key_bits, void* temp, hipStream_t s) {
  if (n <= 1) return CM_OK;
  const int nb = sort_blocks(n);
  char* tp = (char*)temp;
  uint64_t* kalt = (uint64_t*)tp;
  tp += align_up(sizeof(uint64_t) * (size_t)n);
  uint32_t* valt = (uint32_t*)tp;
  tp += align_up(sizeof(uint32_t) * (size_t)n);
  int* hist = (int*)tp;
  tp += align_up(sizeof(int) * 256 * (size_t)nb);
  void* stemp = tp;
  int passes = (key_bits + 7) / 8;
  if (passes & 1) ++passes;  // even number of passes: result ends in the input buffers
  uint64_t* ka = keys;
  uint32_t* va = vals;
  uint64_t* kb = kalt;
  uint32_t* vb = valt;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = 8 * pass;
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(kBlock), 0, s, ka, n, shift, nb, hist);
    int rc = scan_exclusive(hist, hist, 256 * nb, nullptr, nullptr, stemp, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(kBlock), 0, s, ka, va, kb, vb, n, shift,
                       nb, hist);
    CM_LAUNCH_CHECK();
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  return CM_OK;
}

}  // namespace cm
