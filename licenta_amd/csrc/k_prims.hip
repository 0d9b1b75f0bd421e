// Device-wide primitives of libclothmi, written for CDNA4 wave64:
//  * single-pass exclusive int32 scan over a device-side element count
//    (decoupled look-back, one launch, graph-capturable);
//  * stable LSD radix sort of (uint64 key, uint32 value) with 8-bit digits.
// They replace the CUB primitives the reference calls through
// LicentaV2/LicentaV2/Core/CubWrappers.cuh:24-183 (DeviceScan::ExclusiveSum,
// DeviceSelect::Flagged, DeviceRadixSort::SortPairs). Integer results are exact
// and independent of the launch shape.
#include "cm_internal.hpp"

namespace cm {

namespace {

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

// ---- single-pass exclusive scan -------------------------------------------------
// Tiles of kScanTile elements are taken in order from a ticket counter; each tile
// publishes its sum, looks back over its predecessors (one wave reads 64 tile
// states at a time) until it meets an inclusive prefix, then publishes its own
// inclusive prefix and writes its outputs. A tile state is one 64-bit word (flag
// in bits 62-63: 0 empty, 1 tile sum, 2 inclusive prefix; value in bits 0-31)
// written and read with device-scope relaxed atomics, so no cache-wide fences are
// needed. The last workgroup to finish clears the states and counters, leaving the
// temp ready for the next scan on the stream.
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;  // 4096
constexpr int kScanMaxGrid = 256;
constexpr unsigned long long kFlagAgg = 1ull << 62, kFlagInc = 2ull << 62;

__device__ __forceinline__ unsigned long long ld_state(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_state(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(kBlock) k_scan(const int* in, int* out, int n_max,
                                                 const int* __restrict__ d_n,
                                                 int* __restrict__ d_total, int* __restrict__ ctr,
                                                 unsigned long long* __restrict__ state) {
  __shared__ int lds[kBlock / 64 + 1];
  __shared__ int s_tile, s_prefix, s_last;
  const int n = max(0, d_n ? min(*d_n, n_max) : n_max);
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  const int lane = lane_id();
  if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0 && d_total) *d_total = 0;
  const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(&ctr[0], 1);
    __syncthreads();
    const int tile = s_tile;
    if (tile >= ntiles) break;
    const int base = tile * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    if (vec && base + kScanItems <= n) {
#pragma unroll
      for (int q = 0; q < kScanItems / 4; ++q) {
        const int4 t4 = reinterpret_cast<const int4*>(in + base)[q];
        v[4 * q] = t4.x; v[4 * q + 1] = t4.y; v[4 * q + 2] = t4.z; v[4 * q + 3] = t4.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? in[base + k] : 0;
    }
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    int total;
    const int ex = block_excl_scan(sum, lds, total);
    if (threadIdx.x < 64) {
      int prefix = 0;
      if (tile > 0) {
        if (lane == 0) st_state(&state[tile], kFlagAgg | (unsigned)total);
        int j = tile - 1;
        for (;;) {
          const int idx = j - lane;  // lane 0 is the nearest predecessor
          const unsigned long long st = idx >= 0 ? ld_state(&state[idx]) : kFlagInc;
          const unsigned flag = (unsigned)(st >> 62);
          const unsigned long long pm = __ballot(flag == 2u), zm = __ballot(flag == 0u);
          const int first = pm ? __ffsll((long long)pm) - 1 : 64;
          const unsigned long long need = first == 64 ? ~0ull : ((2ull << first) - 1ull);
          if (zm & need) continue;  // a predecessor up to the first prefix is still empty
          prefix += wave_reduce(lane <= first ? (int)(unsigned)st : 0);
          if (first < 64) break;
          j -= 64;
        }
      }
      if (lane == 0) {
        st_state(&state[tile], kFlagInc | (unsigned)(prefix + total));
        s_prefix = prefix;
      }
    }
    __syncthreads();
    int run = s_prefix + ex;
    if (vec && base + kScanItems <= n) {
#pragma unroll
      for (int q = 0; q < kScanItems / 4; ++q) {
        int4 o;
        o.x = run; run += v[4 * q];
        o.y = run; run += v[4 * q + 1];
        o.z = run; run += v[4 * q + 2];
        o.w = run; run += v[4 * q + 3];
        reinterpret_cast<int4*>(out + base)[q] = o;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) {
          out[base + k] = run;
          run += v[k];
        }
    }
    if (tile == ntiles - 1 && threadIdx.x == 0 && d_total) *d_total = s_prefix + total;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    s_last = atomicAdd(&ctr[1], 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (s_last) {
    for (int i = threadIdx.x; i < ntiles; i += kBlock) st_state(&state[i], 0ull);
    if (threadIdx.x == 0) {
      ctr[0] = 0;
      ctr[1] = 0;
    }
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

// ---- radix sort of 32-bit keys with a device-side count --------------------------
// A fixed grid of kDevSortBlocks workgroups; workgroup b takes the contiguous run of
// tiles [b * tpb, (b + 1) * tpb) of the n = min(*d_n, n_max) keys (tpb = ceil(tiles /
// grid)), so one histogram per workgroup keeps the sort stable and the launch shape
// does not depend on n.
constexpr int kDevSortBlocks = 1024;

__device__ __forceinline__ void dev_sort_range(const int* d_n, int n_max, int& n, int& t0,
                                               int& t1) {
  n = min(*d_n, n_max);
  const int ntiles = (n + kSortTile - 1) / kSortTile;
  const int tpb = (ntiles + gridDim.x - 1) / gridDim.x;
  t0 = min(ntiles, blockIdx.x * tpb);
  t1 = min(ntiles, t0 + tpb);
}

__global__ void __launch_bounds__(kBlock) k_radix_hist_dev(const uint32_t* __restrict__ keys,
                                                           const int* __restrict__ d_n, int n_max,
                                                           int shift, int* __restrict__ hist) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int n, t0, t1;
  dev_sort_range(d_n, n_max, n, t0, t1);
  for (int tile = t0; tile < t1; ++tile) {
#pragma unroll
    for (int k = 0; k < kSortItems; ++k) {
      const int i = tile * kSortTile + k * kBlock + threadIdx.x;
      if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
    }
  }
  __syncthreads();
  hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// k_radix_scatter over the workgroup's run of tiles, each digit's running base carried
// from tile to tile (thread d owns digit d of carry, wbase and the count sums)
__global__ void __launch_bounds__(kBlock) k_radix_scatter_dev(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, const int* __restrict__ d_n,
    int n_max, int shift, const int* __restrict__ offsets) {
  __shared__ int cnt[kWaves][256];
  __shared__ int wbase[kWaves][256];
  __shared__ int carry[256];
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  int n, t0, t1;
  dev_sort_range(d_n, n_max, n, t0, t1);
  carry[threadIdx.x] = 0;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int tile = t0; tile < t1; ++tile) {
    for (int d = threadIdx.x; d < kWaves * 256; d += kBlock) (&cnt[0][0])[d] = 0;
    __syncthreads();
    const int base = tile * kSortTile + wid * 64 * kSortItems;
    uint32_t key[kSortItems];
    uint32_t val[kSortItems];
    int rank[kSortItems];
#pragma unroll
    for (int c = 0; c < kSortItems; ++c) {
      const int i = base + c * 64 + lane;
      const bool valid = i < n;
      key[c] = valid ? kin[i] : 0u;
      val[c] = valid ? vin[i] : 0u;
      const uint32_t d = (key[c] >> shift) & 255u;
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
      int run = carry[d];
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        wbase[w][d] = run;
        run += cnt[w][d];
      }
      carry[d] = run;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kSortItems; ++c) {
      const int i = base + c * 64 + lane;
      if (i < n) {
        const uint32_t d = (key[c] >> shift) & 255u;
        const int pos = offsets[d * gridDim.x + blockIdx.x] + wbase[wid][d] + rank[c];
        kout[pos] = key[c];
        vout[pos] = val[c];
      }
    }
    __syncthreads();
  }
}

// ---- scan of per-tile sums ----------------------------------------------------------
// Kernels that produce counts in tiles of consecutive items also write each tile's sum;
// this scan turns the sums into tile prefixes, and the consuming kernel adds the
// exclusive scan within its tile. The counts are then read once instead of three
// times. A packed sum holds two counts below 2^16 (low and high half), scanned
// separately; the totals are clamped to the buffers they index. It is the
// single-pass look-back scan above carried over to a pair of counts: a thread adds
// both as one 64-bit sum (first count in the low word), and a tile state holds the
// two running counts in 31 bits each (every count and total is below 2^31).
__device__ __forceinline__ unsigned long long wave_incl_scan64(unsigned long long v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_reduce64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long pair_of(int v, int packed) {
  return packed ? (unsigned long long)(v & 0xffff) | ((unsigned long long)((unsigned)v >> 16) << 32)
                : (unsigned long long)(unsigned)v;
}
// 64-bit pair sum <-> the 62 value bits of a tile state
__device__ __forceinline__ unsigned long long pack31(unsigned long long x) {
  return (x & 0x7fffffffull) | (((x >> 32) & 0x7fffffffull) << 31);
}
__device__ __forceinline__ unsigned long long unpack31(unsigned long long st) {
  return (st & 0x7fffffffull) | (((st >> 31) & 0x7fffffffull) << 32);
}

__global__ void __launch_bounds__(kBlock) k_scan_tiles(
    const int* __restrict__ d_items, int per_tile, int ntile_max, const int* sums, int* pre,
    int packed, int64_t cap0, int64_t cap1, int* __restrict__ d_tot0, int* __restrict__ d_tot1,
    int* __restrict__ ctr, unsigned long long* __restrict__ state) {
  __shared__ unsigned long long lds[kBlock / 64 + 1];
  __shared__ unsigned long long s_prefix;
  __shared__ int s_tile, s_last;
  const int items = d_items ? max(0, *d_items) : ntile_max * per_tile;
  const int n = min(ntile_max, (items + per_tile - 1) / per_tile);
  const int ntiles = (n + kScanTile - 1) / kScanTile;
  const int lane = lane_id(), wid = threadIdx.x >> 6;
  if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    if (d_tot0) *d_tot0 = 0;
    if (packed && d_tot1) *d_tot1 = 0;
  }
  const bool vin = (reinterpret_cast<uintptr_t>(sums) & 15) == 0;
  const bool vout = (reinterpret_cast<uintptr_t>(pre) & 15) == 0;
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(&ctr[0], 1);
    __syncthreads();
    const int tile = s_tile;
    if (tile >= ntiles) break;
    const int base = tile * kScanTile + threadIdx.x * kScanItems;
    const bool full = base + kScanItems <= n;
    int v[kScanItems];
    if (vin && full) {
#pragma unroll
      for (int q = 0; q < kScanItems / 4; ++q) {
        const int4 t4 = reinterpret_cast<const int4*>(sums + base)[q];
        v[4 * q] = t4.x; v[4 * q + 1] = t4.y; v[4 * q + 2] = t4.z; v[4 * q + 3] = t4.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? sums[base + k] : 0;
    }
    unsigned long long sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += pair_of(v[k], packed);
    // exclusive scan of the thread sums over the workgroup
    const unsigned long long inc = wave_incl_scan64(sum);
    if (lane == 63) lds[wid] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long run = 0;
      for (int w = 0; w < kBlock / 64; ++w) {
        const unsigned long long t = lds[w];
        lds[w] = run;
        run += t;
      }
      lds[kBlock / 64] = run;
    }
    __syncthreads();
    const unsigned long long ex = inc - sum + lds[wid];
    const unsigned long long total = lds[kBlock / 64];
    if (threadIdx.x < 64) {
      unsigned long long prefix = 0;
      if (tile > 0) {
        if (lane == 0) st_state(&state[tile], kFlagAgg | pack31(total));
        int j = tile - 1;
        for (;;) {
          const int idx = j - lane;  // lane 0 is the nearest predecessor
          const unsigned long long st = idx >= 0 ? ld_state(&state[idx]) : kFlagInc;
          const unsigned flag = (unsigned)(st >> 62);
          const unsigned long long pm = __ballot(flag == 2u), zm = __ballot(flag == 0u);
          const int first = pm ? __ffsll((long long)pm) - 1 : 64;
          const unsigned long long need = first == 64 ? ~0ull : ((2ull << first) - 1ull);
          if (zm & need) continue;  // a predecessor up to the first prefix is still empty
          prefix += wave_reduce64(lane <= first ? unpack31(st) : 0ull);
          if (first < 64) break;
          j -= 64;
        }
      }
      if (lane == 0) {
        st_state(&state[tile], kFlagInc | pack31(prefix + total));
        s_prefix = prefix;
      }
    }
    __syncthreads();
    unsigned long long run = s_prefix + ex;
    if (vout && full) {
      if (packed) {
#pragma unroll
        for (int q = 0; q < kScanItems / 2; ++q) {
          int4 o;
          o.x = (int)(unsigned)run;
          o.y = (int)(run >> 32);
          run += pair_of(v[2 * q], 1);
          o.z = (int)(unsigned)run;
          o.w = (int)(run >> 32);
          run += pair_of(v[2 * q + 1], 1);
          reinterpret_cast<int4*>(pre + 2 * base)[q] = o;
        }
      } else {
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
          int4 o;
          o.x = (int)run; run += (unsigned)v[4 * q];
          o.y = (int)run; run += (unsigned)v[4 * q + 1];
          o.z = (int)run; run += (unsigned)v[4 * q + 2];
          o.w = (int)run; run += (unsigned)v[4 * q + 3];
          reinterpret_cast<int4*>(pre + base)[q] = o;
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) {
          if (packed) {
            pre[2 * (base + k)] = (int)(unsigned)run;
            pre[2 * (base + k) + 1] = (int)(run >> 32);
          } else {
            pre[base + k] = (int)(unsigned)run;
          }
          run += pair_of(v[k], packed);
        }
    }
    if (tile == ntiles - 1 && threadIdx.x == 0) {
      const unsigned long long all = s_prefix + total;
      if (d_tot0) *d_tot0 = (int)min<int64_t>((int64_t)(all & 0xffffffffull), cap0);
      if (packed && d_tot1) *d_tot1 = (int)min<int64_t>((int64_t)(all >> 32), cap1);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    __threadfence();
    s_last = atomicAdd(&ctr[1], 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (s_last) {
    for (int i = threadIdx.x; i < ntiles; i += kBlock) st_state(&state[i], 0ull);
    if (threadIdx.x == 0) {
      ctr[0] = 0;
      ctr[1] = 0;
    }
  }
}

}  // namespace

static int lookback_tiles(int64_t n_max) {
  return (int)std::max<int64_t>(1, (n_max + kScanTile - 1) / kScanTile);
}

size_t scan_temp_bytes(int n_max) {
  return 256 + align_up(sizeof(unsigned long long) * (size_t)lookback_tiles(n_max));
}

int scan_temp_clear(void* temp, int n_max, hipStream_t s) {
  return fill_i32((int*)temp, 0, (int64_t)(scan_temp_bytes(n_max) / sizeof(int)), s);
}

int scan_exclusive(const int* in, int* out, int n_max, const int* d_n, int* d_total, void* temp,
                   hipStream_t s) {
  int* ctr = (int*)temp;
  unsigned long long* state = (unsigned long long*)((char*)temp + 256);
  const int grid = std::min(lookback_tiles(n_max), kScanMaxGrid);
  hipLaunchKernelGGL(k_scan, dim3(grid), dim3(kBlock), 0, s, in, out, n_max, d_n, d_total, ctr,
                     state);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int scan_tiles(const int* d_items, int per_tile, int ntile_max, const int* sums, int* pre,
               int packed, int64_t cap0, int64_t cap1, int* d_tot0, int* d_tot1, void* temp,
               hipStream_t s) {
  int* ctr = (int*)temp;
  unsigned long long* state = (unsigned long long*)((char*)temp + 256);
  const int grid = std::min(lookback_tiles(ntile_max), kScanMaxGrid);
  hipLaunchKernelGGL(k_scan_tiles, dim3(grid), dim3(kBlock), 0, s, d_items, per_tile, ntile_max,
                     sums, pre, packed, cap0, cap1, d_tot0, d_tot1, ctr, state);
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

// The scan part of a sort temp must start zeroed (each scan leaves it zeroed).
int radix_sort_clear(void* temp, int n, hipStream_t s) {
  const int nb = sort_blocks(n);
  char* tp = (char*)temp + align_up(sizeof(uint64_t) * (size_t)n) +
             align_up(sizeof(uint32_t) * (size_t)n) + align_up(sizeof(int) * 256 * (size_t)nb);
  return scan_temp_clear(tp, 256 * nb, s);
}

size_t radix_sort_dev_temp_bytes() {
  return align_up(sizeof(int) * 256 * kDevSortBlocks) + scan_temp_bytes(256 * kDevSortBlocks);
}

int radix_sort_dev_clear(void* temp, hipStream_t s) {
  return scan_temp_clear((char*)temp + align_up(sizeof(int) * 256 * kDevSortBlocks),
                         256 * kDevSortBlocks, s);
}

int radix_sort_dev(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                   int n_max, const int* d_n, int key_bits, void* temp, int* in_alt,
                   hipStream_t s) {
  int* hist = (int*)temp;
  void* stemp = (char*)temp + align_up(sizeof(int) * 256 * kDevSortBlocks);
  const int passes = std::min(4, (key_bits + 7) / 8);
  uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = 8 * pass;
    hipLaunchKernelGGL(k_radix_hist_dev, dim3(kDevSortBlocks), dim3(kBlock), 0, s, ka, d_n, n_max,
                       shift, hist);
    int rc = scan_exclusive(hist, hist, 256 * kDevSortBlocks, nullptr, nullptr, stemp, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_radix_scatter_dev, dim3(kDevSortBlocks), dim3(kBlock), 0, s, ka, va, kb,
                       vb, d_n, n_max, shift, hist);
    CM_LAUNCH_CHECK();
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  *in_alt = passes & 1;
  return CM_OK;
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
