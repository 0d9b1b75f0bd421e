// Morton/LBVH stages of the substep as gfx950 kernels.
//
//  k_tri_update   <- DeformableUtils::_UpdateTrianglesContinuous
//                    (LicentaV2/LicentaV2/Physics/CudaDeformableUtils.cu:120-151)
//  k_build        <- CudaBVH::_GenerateBVH2 / FindRange / FindSplit
//                    (Collision/CudaBVH.cu:163-242), Delta/sgn (Core/CudaUtils.cu:338-362)
//  k_refit        <- CudaBVH::ComputeTreeAABBs (Collision/CudaBVH.cu:128-149)
//  k_traverse     <- CudaBVH::_GetAABBCollisions (Collision/CudaBVH.cu:23-113)
//
// Layout: node AABBs are float4 min / float4 max arrays of 2T-1 entries, internal
// nodes [0, T-1) then leaves [T-1, 2T-1), as in the reference. Parent writes are
// made deterministic with claim records (count, highest claimant and the first
// four claimant ids): parents[c] = highest claimant, which is what the
// reference's last writer in thread order produces; the refit climbs through
// every claimant so a node shared by several parents (midpoint split on tied
// codes) is still the union of its two children.
#include "cm_internal.hpp"

namespace cm {

namespace {

__device__ __forceinline__ int atomic_add_acq_rel(int* p, int v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- per-triangle swept AABB and Morton key --------------------------------
__device__ __forceinline__ void swept_aabb(const float4* __restrict__ x,
                                           const float4* __restrict__ p, int4 tr, float th,
                                           f3& smn, f3& smx) {
  const f3 p1 = xyz(p[tr.x]), p2 = xyz(p[tr.y]), p3 = xyz(p[tr.z]);
  const f3 q1 = xyz(x[tr.x]), q2 = xyz(x[tr.y]), q3 = xyz(x[tr.z]);
  const f3 mn = subs(min3(p1, p2, p3), th), mx = adds(max3(p1, p2, p3), th);
  const f3 pmn = subs(min3(q1, q2, q3), th), pmx = adds(max3(q1, q2, q3), th);
  smn = vmin(mn, pmn);
  smx = vmax(mx, pmx);
}

__global__ void __launch_bounds__(kBlock) k_tri_update(
    int t, const float4* __restrict__ x, const float4* __restrict__ p,
    const int4* __restrict__ tri, float th, int bits, float4* __restrict__ amin,
    float4* __restrict__ amax, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
    float4* __restrict__ centroid, float4* __restrict__ box_partials) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  f3 c = mk3(0, 0, 0);
  bool valid = i < t;
  if (valid) {
    f3 smn, smx;
    swept_aabb(x, p, tri[i], th, smn, smx);
    amin[t - 1 + i] = f4(smn, 0.f);
    amax[t - 1 + i] = f4(smx, 0.f);
    c = add(smn, divs(sub(smx, smn), 2.f));
    vals[i] = (uint32_t)i;
    if (centroid) {
      centroid[i] = f4(c, 0.f);
    } else if (bits == 30) {
      keys[i] = morton30(c, mk3(0, 0, 0), mk3(1, 1, 1));
    } else {
      keys[i] = morton63(c);
    }
  }
  if (box_partials) {  // block min/max of the centroids (exact)
    __shared__ float smin[3][kBlock], smax[3][kBlock];
    const float inf = __builtin_huge_valf();
    smin[0][threadIdx.x] = valid ? c.x : inf;
    smin[1][threadIdx.x] = valid ? c.y : inf;
    smin[2][threadIdx.x] = valid ? c.z : inf;
    smax[0][threadIdx.x] = valid ? c.x : -inf;
    smax[1][threadIdx.x] = valid ? c.y : -inf;
    smax[2][threadIdx.x] = valid ? c.z : -inf;
    __syncthreads();
    for (int s = kBlock / 2; s >= 1; s >>= 1) {
      if (threadIdx.x < s)
        for (int a = 0; a < 3; ++a) {
          smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
          smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
        }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      box_partials[2 * blockIdx.x] = make_float4(smin[0][0], smin[1][0], smin[2][0], 0.f);
      box_partials[2 * blockIdx.x + 1] = make_float4(smax[0][0], smax[1][0], smax[2][0], 0.f);
    }
  }
}

// single block: reduce the per-block boxes; result in box[2*nb], box[2*nb+1]
__global__ void __launch_bounds__(kBlock) k_box_reduce(float4* __restrict__ box, int nb) {
  __shared__ float smin[3][kBlock], smax[3][kBlock];
  const float inf = __builtin_huge_valf();
  float mn[3] = {inf, inf, inf}, mx[3] = {-inf, -inf, -inf};
  for (int b = threadIdx.x; b < nb; b += kBlock) {
    float4 a = box[2 * b], c = box[2 * b + 1];
    mn[0] = fminf(mn[0], a.x); mn[1] = fminf(mn[1], a.y); mn[2] = fminf(mn[2], a.z);
    mx[0] = fmaxf(mx[0], c.x); mx[1] = fmaxf(mx[1], c.y); mx[2] = fmaxf(mx[2], c.z);
  }
  for (int a = 0; a < 3; ++a) {
    smin[a][threadIdx.x] = mn[a];
    smax[a][threadIdx.x] = mx[a];
  }
  __syncthreads();
  for (int s = kBlock / 2; s >= 1; s >>= 1) {
    if (threadIdx.x < s)
      for (int a = 0; a < 3; ++a) {
        smin[a][threadIdx.x] = fminf(smin[a][threadIdx.x], smin[a][threadIdx.x + s]);
        smax[a][threadIdx.x] = fmaxf(smax[a][threadIdx.x], smax[a][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    box[2 * nb] = make_float4(smin[0][0], smin[1][0], smin[2][0], 0.f);
    box[2 * nb + 1] = make_float4(smax[0][0], smax[1][0], smax[2][0], 0.f);
  }
}

__global__ void __launch_bounds__(kBlock) k_keys30(int t, const float4* __restrict__ centroid,
                                                   const float4* __restrict__ box, int nb,
                                                   uint64_t* __restrict__ keys) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= t) return;
  const f3 lo = xyz(box[2 * nb]), hi = xyz(box[2 * nb + 1]);
  keys[i] = morton30(xyz(centroid[i]), lo, hi);
}

// ---- Karras build -----------------------------------------------------------
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int i, int j, int maxj) {
  if (j < 0 || j > maxj) return -1;
  const uint64_t x = k[i] ^ k[j];
  return x ? clz64(x) : clz64((uint64_t)i ^ (uint64_t)j) + 64;
}
__device__ __forceinline__ int sgn(int a) { return signbit((float)a) ? -1 : 1; }

__global__ void __launch_bounds__(kBlock) k_build_init(int t, int* __restrict__ cnt,
                                                       int* __restrict__ hi,
                                                       int* __restrict__ visited,
                                                       uint32_t* __restrict__ warn) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < 2 * t - 1) {
    cnt[i] = 0;
    hi[i] = -1;
  }
  if (i < t - 1) visited[i] = 0;
  // the refit-fallback request is per build; the CM_WARN_* bits stay sticky
  if (i == 0) atomicAnd(warn, ~kRefitOverflow);
}

__global__ void __launch_bounds__(kBlock) k_build(int t, const uint64_t* __restrict__ k,
                                                  int midpoint, int* __restrict__ lefts,
                                                  int* __restrict__ rights, int* __restrict__ cnt,
                                                  int* __restrict__ hi, int* __restrict__ list,
                                                  int2* __restrict__ ranges,
                                                  uint32_t* __restrict__ warn) {
  const int id = blockIdx.x * kBlock + threadIdx.x;
  const int nin = t - 1;
  if (id >= nin) return;
  // range (Collision/CudaBVH.cu:190-212)
  const int d = sgn(delta(k, id, id + 1, nin) - delta(k, id, id - 1, nin));
  const int dmin = delta(k, id, id - d, nin);
  int lmax = 2;
  while (delta(k, id, id + lmax * d, nin) > dmin) lmax <<= 1;
  int l = 0;
  for (int s = lmax / 2; s >= 1; s /= 2)
    if (delta(k, id, id + (l + s) * d, nin) > dmin) l += s;
  const int j = id + l * d;
  const int first = min(id, j), last = max(id, j);
  ranges[id] = make_int2(first, last);
  // split (Collision/CudaBVH.cu:214-242)
  const uint64_t fc = k[first], lc = k[last];
  int split;
  if (midpoint && fc == lc) {
    split = (first + last) >> 1;
  } else {
    const int com = midpoint ? clz64(fc ^ lc) : delta(k, first, last, nin);
    split = first;
    int step = last - first;
    do {
      step = (step + 1) >> 1;
      const int ns = split + step;
      if (ns < last) {
        const int pref = midpoint ? clz64(fc ^ k[ns]) : delta(k, first, ns, nin);
        if (pref > com) split = ns;
      }
    } while (step > 1);
  }
  const int left = split == first ? nin + split : split;
  const int right = split + 1 == last ? nin + split + 1 : split + 1;
  lefts[id] = left;
  rights[id] = right;
  const int o1 = atomicAdd(&cnt[left], 1);
  if (o1 < kMaxClaims) list[kMaxClaims * (int64_t)left + o1] = id;
  atomicMax(&hi[left], id);
  const int o2 = atomicAdd(&cnt[right], 1);
  if (o2 < kMaxClaims) list[kMaxClaims * (int64_t)right + o2] = id;
  atomicMax(&hi[right], id);
  if (o1 > 0 || o2 > 0) atomicOr(warn, (uint32_t)CM_WARN_BROKEN_TREE);
}

__global__ void __launch_bounds__(kBlock) k_parents(int t, const int* __restrict__ cnt,
                                                    const int* __restrict__ hi,
                                                    int* __restrict__ parents,
                                                    uint32_t* __restrict__ warn) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= 2 * t - 1) return;
  const int k = cnt[c];
  if (k > 0) parents[c] = hi[c];
  if (k > kMaxClaims) atomicOr(warn, kRefitOverflow);
}

// ---- bottom-up refit ------------------------------------------------------------
// Packed node record read by the traversal, 4 float4 per internal node:
//   [0] left, right, rmll, rmlr (int bits)   [1] left min xyz, left max x
//   [2] left max yz, right min xy             [3] right min z, right max xyz
__device__ __forceinline__ void write_node(float4* __restrict__ nodes, int q, int l, int r,
                                           int a, int b, f3 lmn, f3 lmx, f3 rmn, f3 rmx) {
  float4* o = nodes + 4 * (int64_t)q;
  o[0] = make_float4(__int_as_float(l), __int_as_float(r), __int_as_float(a), __int_as_float(b));
  o[1] = make_float4(lmn.x, lmn.y, lmn.z, lmx.x);
  o[2] = make_float4(lmx.y, lmx.z, rmn.x, rmn.y);
  o[3] = make_float4(rmn.z, rmx.x, rmx.y, rmx.z);
}

// node q = union of its two children (Collision/CudaBVH.cu:140-144), all in global memory
__device__ __forceinline__ void union_global(int q, int nin, const int* __restrict__ lefts,
                                             const int* __restrict__ rights,
                                             float4* __restrict__ amin, float4* __restrict__ amax,
                                             int* __restrict__ rmll, int* __restrict__ rmlr,
                                             float4* __restrict__ nodes) {
  const int l = lefts[q], r = rights[q];
  const f3 lmn = xyz(amin[l]), lmx = xyz(amax[l]), rmn = xyz(amin[r]), rmx = xyz(amax[r]);
  amin[q] = f4(vmin(lmn, rmn), 0.f);
  amax[q] = f4(vmax(lmx, rmx), 0.f);
  const int a = l >= nin ? l : max(rmll[l], rmlr[l]);
  const int b = r >= nin ? r : max(rmll[r], rmlr[r]);
  rmll[q] = a;
  rmlr[q] = b;
  write_node(nodes, q, l, r, a, b, lmn, lmx, rmn, rmx);
}

constexpr int kRefitCH = 1024;  // leaves per refit block
constexpr int kClimbStack = 8;

// One block per chunk of 1024 consecutive leaves. A node whose key range lies in
// the chunk ("local") is finished in LDS with workgroup-scope arrival counters;
// arrivals at nodes whose range crosses the chunk ("escapes", a few per chunk per
// tree level) then climb with device-scope counters. Every claimant of a node is
// notified, so a node shared by several parents (midpoint split on tied codes) is
// still the union of its children; more than kMaxClaims claimants sets
// kRefitOverflow and the single-block fallback recomputes the tree.
__global__ void __launch_bounds__(kRefitCH) k_refit_chunk(
    int t, const int* __restrict__ lefts, const int* __restrict__ rights,
    const int2* __restrict__ ranges, const int* __restrict__ cnt, const int* __restrict__ list,
    int* __restrict__ visited, float4* __restrict__ amin,
    float4* __restrict__ amax, int* __restrict__ rmll, int* __restrict__ rmlr,
    float4* __restrict__ nodes, uint32_t* __restrict__ warn) {
  __shared__ float4 smin[kRefitCH], smax[kRefitCH];
  __shared__ int sarr[kRefitCH], srml[kRefitCH], srmr[kRefitCH];
  __shared__ int2 esc[2 * kRefitCH];
  __shared__ int esc_n;
  const int nin = t - 1, base = blockIdx.x * kRefitCH, tid = threadIdx.x;
  sarr[tid] = 0;
  if (tid == 0) esc_n = 0;
  __syncthreads();
  auto is_local = [&](int q) {
    if (q >= nin) return false;
    const int2 rg = ranges[q];
    return rg.x >= base && rg.y < base + kRefitCH;
  };
  uint32_t w = 0;
  const int leaf = base + tid;
  if (leaf < t) {
    int stk[kClimbStack];
    int top = 0;
    stk[top++] = nin + leaf;
    int guard = 0;
    while (top > 0) {
      if (++guard > 4 * kRefitCH) {
        w |= kRefitOverflow;
        break;
      }
      const int x = stk[--top];
      const int k = cnt[x];
      if (k > kMaxClaims) w |= kRefitOverflow;
      const int kk = min(k, kMaxClaims);
      for (int s = 0; s < kk; ++s) {
        const int q = list[kMaxClaims * (int64_t)x + s];
        if (is_local(q)) {
          const int sq = q - base;
          if (__hip_atomic_fetch_add(&sarr[sq], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 1) {
            const int l = lefts[q], r = rights[q];
            f3 lmn, lmx, rmn, rmx;
            int a, b;
            if (l >= nin) {
              lmn = xyz(amin[l]); lmx = xyz(amax[l]); a = l;
            } else {
              lmn = xyz(smin[l - base]); lmx = xyz(smax[l - base]); a = max(srml[l - base], srmr[l - base]);
            }
            if (r >= nin) {
              rmn = xyz(amin[r]); rmx = xyz(amax[r]); b = r;
            } else {
              rmn = xyz(smin[r - base]); rmx = xyz(smax[r - base]); b = max(srml[r - base], srmr[r - base]);
            }
            smin[sq] = f4(vmin(lmn, rmn), 0.f);
            smax[sq] = f4(vmax(lmx, rmx), 0.f);
            srml[sq] = a;
            srmr[sq] = b;
            write_node(nodes, q, l, r, a, b, lmn, lmx, rmn, rmx);
            if (top < kClimbStack) stk[top++] = q; else w |= kRefitOverflow;
          }
        } else {
          const int e = atomicAdd(&esc_n, 1);
          if (e < 2 * kRefitCH) esc[e] = make_int2(x, q); else w |= kRefitOverflow;
        }
      }
    }
  }
  __syncthreads();
  {
    const int q = base + tid;
    if (q < nin && sarr[tid] >= 2 && is_local(q)) {
      amin[q] = smin[tid];
      amax[q] = smax[tid];
      rmll[q] = srml[tid];
      rmlr[q] = srmr[tid];
    }
  }
  __threadfence();
  __syncthreads();
  const int ne = min(esc_n, 2 * kRefitCH);
  for (int e = tid; e < ne; e += kRefitCH) {
    int stk[kClimbStack];
    int top = 0;
    const int q0 = esc[e].y;
    if (atomic_add_acq_rel(&visited[q0], 1) == 1) {
      union_global(q0, nin, lefts, rights, amin, amax, rmll, rmlr, nodes);
      stk[top++] = q0;
    }
    int guard = 0;
    while (top > 0) {
      if (++guard > 2 * t) {
        w |= CM_WARN_REFIT_LIMIT;
        break;
      }
      const int x = stk[--top];
      const int k = cnt[x];
      if (k > kMaxClaims) w |= kRefitOverflow;
      const int kk = min(k, kMaxClaims);
      for (int s = 0; s < kk; ++s) {
        const int q = list[kMaxClaims * (int64_t)x + s];
        if (atomic_add_acq_rel(&visited[q], 1) == 1) {
          union_global(q, nin, lefts, rights, amin, amax, rmll, rmlr, nodes);
          if (top < kClimbStack) stk[top++] = q; else w |= kRefitOverflow;
        }
      }
    }
  }
  if (w) atomicOr(warn, w);
}

// Fallback when a node has more than two claimants or a climb stack filled:
// one block recomputes every internal node, children first, in passes.
__global__ void __launch_bounds__(1024) k_refit_fallback(int t, const int* __restrict__ lefts,
                                                         const int* __restrict__ rights,
                                                         int* __restrict__ done,
                                                         float4* __restrict__ amin,
                                                         float4* __restrict__ amax,
                                                         int* __restrict__ rmll,
                                                         int* __restrict__ rmlr,
                                                         float4* __restrict__ nodes,
                                                         const uint32_t* __restrict__ warn) {
  if (!(*warn & kRefitOverflow)) return;
  const int nin = t - 1;
  for (int q = threadIdx.x; q < nin; q += blockDim.x) done[q] = 0;
  __syncthreads();
  for (int pass = 1; pass <= nin; ++pass) {
    int changed = 0;
    for (int q = threadIdx.x; q < nin; q += blockDim.x) {
      if (done[q]) continue;
      const int l = lefts[q], r = rights[q];
      const bool rl = l >= nin || (done[l] != 0 && done[l] < pass);
      const bool rr = r >= nin || (done[r] != 0 && done[r] < pass);
      if (rl && rr) {
        union_global(q, nin, lefts, rights, amin, amax, rmll, rmlr, nodes);
        done[q] = pass;
        changed = 1;
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
}

// ---- traversal ----------------------------------------------------------------
constexpr int kLdsStack = 24;  // per-thread stack entries kept in LDS, the rest private

__global__ void __launch_bounds__(kBlock) k_traverse(int t, const float4* __restrict__ nodes,
                                                     const float4* __restrict__ amin,
                                                     const float4* __restrict__ amax, int cap,
                                                     int2* __restrict__ chunk,
                                                     int* __restrict__ counts,
                                                     uint32_t* __restrict__ warn) {
  __shared__ int sstk[kLdsStack][kBlock];
  const int leaf = blockIdx.x * kBlock + threadIdx.x;
  if (leaf >= t) return;
  const int nin = t - 1;
  const int me = nin + leaf;
  const f3 mn = xyz(amin[me]), mx = xyz(amax[me]);
  int2* out = chunk + (int64_t)leaf * cap;
  int ncol = 0;
  uint32_t w = 0;
  int pstk[63 - kLdsStack];
  int top = 0;  // 63 usable entries, as the reference's int[64] with its -1 sentinel
  int node = 0;
  const long long limit = 64ll * t + 4096;
  long long visits = 0;
  while (true) {
    if (++visits > limit) {
      w |= CM_WARN_TRAVERSAL_LIMIT;
      break;
    }
    const float4* nd = nodes + 4 * (int64_t)node;
    const float4 r0 = nd[0], r1 = nd[1], r2 = nd[2], r3 = nd[3];
    const int l = __float_as_int(r0.x), r = __float_as_int(r0.y);
    const int rml = __float_as_int(r0.z), rmr = __float_as_int(r0.w);
    const bool ol = rml > me && overlap(mn, mx, mk3(r1.x, r1.y, r1.z), mk3(r1.w, r2.x, r2.y));
    const bool orr = rmr > me && overlap(mn, mx, mk3(r2.z, r2.w, r3.x), mk3(r3.y, r3.z, r3.w));
    // Past the cap the reference keeps walking but drops every further pair, so
    // stopping at the first dropped pair yields the same list.
    if (ol && l >= nin) {
      if (ncol == cap) { w |= CM_WARN_PAIR_CAP; break; }
      out[ncol++] = make_int2(leaf, l - nin);
    }
    if (orr && r >= nin) {
      if (ncol == cap) { w |= CM_WARN_PAIR_CAP; break; }
      out[ncol++] = make_int2(leaf, r - nin);
    }
    const bool tl = ol && l < nin, tr = orr && r < nin;
    if (!tl && !tr) {
      if (top == 0) break;
      --top;
      node = top < kLdsStack ? sstk[top][threadIdx.x] : pstk[top - kLdsStack];
    } else {
      node = tl ? l : r;
      if (tl && tr) {
        if (top < 63) {
          if (top < kLdsStack) sstk[top][threadIdx.x] = r; else pstk[top - kLdsStack] = r;
          ++top;
        } else {
          w |= CM_WARN_STACK_OVERFLOW;
        }
      }
    }
  }
  counts[leaf] = ncol;
  if (w) atomicOr(warn, w);
}

// packed node records from separate arrays (unit-level pair query)
__global__ void __launch_bounds__(kBlock) k_pack_nodes(int t, const int* __restrict__ lefts,
                                                       const int* __restrict__ rights,
                                                       const float4* __restrict__ amin,
                                                       const float4* __restrict__ amax,
                                                       const int* __restrict__ rmll,
                                                       const int* __restrict__ rmlr,
                                                       float4* __restrict__ nodes) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= t - 1) return;
  const int l = lefts[q], r = rights[q];
  write_node(nodes, q, l, r, rmll[q], rmlr[q], xyz(amin[l]), xyz(amax[l]), xyz(amin[r]),
             xyz(amax[r]));
}

// one wave per leaf copies its chunk to the compacted list
__global__ void __launch_bounds__(kBlock) k_compact_pairs(int t, int cap,
                                                          const int2* __restrict__ chunk,
                                                          const int* __restrict__ counts,
                                                          const int* __restrict__ offsets,
                                                          int2* __restrict__ pairs) {
  const int leaf = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (leaf >= t) return;
  const int c = counts[leaf], o = offsets[leaf];
  const int2* src = chunk + (int64_t)leaf * cap;
  for (int k = lane; k < c; k += 64) pairs[o + k] = src[k];
}

__global__ void __launch_bounds__(kBlock) k_gather_tris(int t, const uint32_t* __restrict__ perm,
                                                        const int4* __restrict__ tin,
                                                        int4* __restrict__ tout, int fix,
                                                        const float4* __restrict__ amin,
                                                        const float4* __restrict__ amax,
                                                        float4* __restrict__ tmin,
                                                        float4* __restrict__ tmax) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= t) return;
  const uint32_t s = perm[i];
  tout[i] = tin[s];
  if (fix) {
    tmin[i] = amin[t - 1 + s];
    tmax[i] = amax[t - 1 + s];
  }
}
__global__ void __launch_bounds__(kBlock) k_put_leaves(int t, const float4* __restrict__ tmin,
                                                       const float4* __restrict__ tmax,
                                                       float4* __restrict__ amin,
                                                       float4* __restrict__ amax) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= t) return;
  amin[t - 1 + i] = tmin[i];
  amax[t - 1 + i] = tmax[i];
}
__global__ void __launch_bounds__(kBlock) k_copy_i4(int64_t n, const int4* __restrict__ s,
                                                    int4* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) d[i] = s[i];
}

}  // namespace

// ---- launchers --------------------------------------------------------------
int launch_tri_update(int n, int t, const float4* x, const float4* p, const int4* tri, float th,
                      int bits, int normalize, float4* amin, float4* amax, uint64_t* keys,
                      uint32_t* vals, float4* centroid_tmp, float4* box_tmp, hipStream_t s) {
  (void)n;
  if (t <= 0) return CM_OK;
  const int nb = nblocks(t);
  const bool norm = bits == 30 && normalize;
  hipLaunchKernelGGL(k_tri_update, dim3(nb), dim3(kBlock), 0, s, t, x, p, tri, th, bits, amin,
                     amax, keys, vals, norm ? centroid_tmp : nullptr, norm ? box_tmp : nullptr);
  if (norm) {
    hipLaunchKernelGGL(k_box_reduce, dim3(1), dim3(kBlock), 0, s, box_tmp, nb);
    hipLaunchKernelGGL(k_keys30, dim3(nb), dim3(kBlock), 0, s, t, centroid_tmp, box_tmp, nb, keys);
  }
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_build(const LbvhBufs& b, const uint64_t* keys, int midpoint, hipStream_t s) {
  const int t = b.t;
  if (t < 2) return CM_OK;
  hipLaunchKernelGGL(k_build_init, dim3(nblocks(2 * t - 1)), dim3(kBlock), 0, s, t, b.claim_cnt,
                     b.claim_hi, b.visited, b.warn);
  hipLaunchKernelGGL(k_build, dim3(nblocks(t - 1)), dim3(kBlock), 0, s, t, keys, midpoint, b.lefts,
                     b.rights, b.claim_cnt, b.claim_hi, b.claim_list, b.ranges, b.warn);
  hipLaunchKernelGGL(k_parents, dim3(nblocks(2 * t - 1)), dim3(kBlock), 0, s, t, b.claim_cnt,
                     b.claim_hi, b.parents, b.warn);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_refit(const LbvhBufs& b, hipStream_t s) {
  const int t = b.t;
  if (t < 2) return CM_OK;
  hipLaunchKernelGGL(k_refit_chunk, dim3((t + kRefitCH - 1) / kRefitCH), dim3(kRefitCH), 0, s, t,
                     b.lefts, b.rights, b.ranges, b.claim_cnt, b.claim_list, b.visited,
                     b.amin, b.amax, b.rmll, b.rmlr, b.nodes, b.warn);
  hipLaunchKernelGGL(k_refit_fallback, dim3(1), dim3(1024), 0, s, t, b.lefts, b.rights, b.done,
                     b.amin, b.amax, b.rmll, b.rmlr, b.nodes, b.warn);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_pairs(const LbvhBufs& b, int cap, int2* chunk, int* counts, int* offsets, int2* pairs,
                 int* d_npairs, void* scan_temp, hipStream_t s) {
  const int t = b.t;
  if (t < 2) return fill_i32(d_npairs, 0, 1, s);
  hipLaunchKernelGGL(k_traverse, dim3(nblocks(t)), dim3(kBlock), 0, s, t, b.nodes, b.amin, b.amax,
                     cap, chunk, counts, b.warn);
  CM_LAUNCH_CHECK();
  int rc = scan_exclusive(counts, offsets, t, nullptr, d_npairs, scan_temp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compact_pairs, dim3(nblocks((int64_t)t * 64)), dim3(kBlock), 0, s, t, cap,
                     chunk, counts, offsets, pairs);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_pack_nodes(const LbvhBufs& b, hipStream_t s) {
  const int t = b.t;
  if (t < 2) return CM_OK;
  hipLaunchKernelGGL(k_pack_nodes, dim3(nblocks(t - 1)), dim3(kBlock), 0, s, t, b.lefts, b.rights,
                     b.amin, b.amax, b.rmll, b.rmlr, b.nodes);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_gather_tris(int t, const uint32_t* perm, const int4* tri_in, int4* tri_out,
                       int fix_leaves, float4* amin, float4* amax, float4* tmp_min,
                       float4* tmp_max, hipStream_t s) {
  if (t <= 0) return CM_OK;
  hipLaunchKernelGGL(k_gather_tris, dim3(nblocks(t)), dim3(kBlock), 0, s, t, perm, tri_in, tri_out,
                     fix_leaves, amin, amax, tmp_min, tmp_max);
  if (fix_leaves)
    hipLaunchKernelGGL(k_put_leaves, dim3(nblocks(t)), dim3(kBlock), 0, s, t, tmp_min, tmp_max,
                       amin, amax);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_copy_i4(int64_t n, const int4* src, int4* dst, hipStream_t s) {
  if (n <= 0) return CM_OK;
  hipLaunchKernelGGL(k_copy_i4, dim3(nblocks(n)), dim3(kBlock), 0, s, n, src, dst);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

size_t lbvh_build_temp_bytes(int t) {
  const size_t m = t > 0 ? 2 * (size_t)t - 1 : 0;
  const size_t tt = (size_t)(t > 0 ? t : 1);
  return 2 * align_up(sizeof(int) * m) + align_up(sizeof(int) * kMaxClaims * m) +
         2 * align_up(sizeof(int) * tt) +
         align_up(sizeof(int2) * tt) + align_up(4 * sizeof(float4) * tt);
}
size_t lbvh_pairs_temp_bytes(int t, int cap) {
  return align_up(sizeof(int2) * (size_t)t * cap) + 2 * align_up(sizeof(int) * (size_t)t) +
         align_up(4 * sizeof(float4) * (size_t)t) + scan_temp_bytes(t);
}

}  // namespace cm

// ---- unit-level C ABI ---------------------------------------------------------
using namespace cm;

extern "C" int cm_tri_update(int32_t n, int32_t t, const void* d_x, const void* d_p,
                             const void* d_tri, float thickness, int32_t morton_bits,
                             int32_t morton_normalize, void* d_aabb_min, void* d_aabb_max,
                             uint64_t* d_keys, void* stream) {
  if (t < 1 || n < 1 || !d_x || !d_p || !d_tri || !d_aabb_min || !d_aabb_max || !d_keys ||
      (morton_bits != 30 && morton_bits != 63)) {
    set_error("cm_tri_update: invalid argument");
    return CM_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const bool norm = morton_bits == 30 && morton_normalize;
  float4* cen = nullptr;
  float4* box = nullptr;
  uint32_t* vals = nullptr;
  CM_HIP(hipMallocAsync((void**)&vals, sizeof(uint32_t) * (size_t)t, s));
  if (norm) {
    CM_HIP(hipMallocAsync((void**)&cen, sizeof(float4) * (size_t)t, s));
    CM_HIP(hipMallocAsync((void**)&box, sizeof(float4) * (2 * (size_t)nblocks(t) + 2), s));
  }
  int rc = launch_tri_update(n, t, (const float4*)d_x, (const float4*)d_p, (const int4*)d_tri,
                             thickness, morton_bits, morton_normalize, (float4*)d_aabb_min,
                             (float4*)d_aabb_max, d_keys, vals, cen, box, s);
  hipFreeAsync(vals, s);
  if (cen) hipFreeAsync(cen, s);
  if (box) hipFreeAsync(box, s);
  return rc;
}

extern "C" size_t cm_lbvh_sort_temp_bytes(int32_t t) { return radix_sort_temp_bytes(t); }

extern "C" int cm_lbvh_sort(int32_t t, int32_t key_bits, uint64_t* d_keys, uint32_t* d_vals,
                            void* d_temp, size_t temp_bytes, void* stream) {
  if (t < 0 || key_bits < 1 || key_bits > 64 || (t > 1 && (!d_keys || !d_vals || !d_temp))) {
    set_error("cm_lbvh_sort: invalid argument");
    return CM_ERR_ARG;
  }
  if (temp_bytes < radix_sort_temp_bytes(t)) {
    set_error("cm_lbvh_sort: temp buffer too small");
    return CM_ERR_ARG;
  }
  return radix_sort(d_keys, d_vals, t, key_bits, d_temp, (hipStream_t)stream);
}

extern "C" size_t cm_lbvh_build_temp_bytes(int32_t t) { return lbvh_build_temp_bytes(t); }

static void carve_build(int t, void* d_temp, LbvhBufs& b) {
  const size_t m = 2 * (size_t)t - 1;
  char* tp = (char*)d_temp;
  b.claim_cnt = (int*)tp;
  tp += align_up(sizeof(int) * m);
  b.claim_hi = (int*)tp;
  tp += align_up(sizeof(int) * m);
  b.claim_list = (int*)tp;
  tp += align_up(sizeof(int) * kMaxClaims * m);
  b.visited = (int*)tp;
  tp += align_up(sizeof(int) * (size_t)t);
  b.done = (int*)tp;
  tp += align_up(sizeof(int) * (size_t)t);
  b.ranges = (int2*)tp;
  tp += align_up(sizeof(int2) * (size_t)t);
  b.nodes = (float4*)tp;
}

extern "C" int cm_lbvh_build(int32_t t, const uint64_t* d_keys, int32_t midpoint_split,
                             int32_t* d_lefts, int32_t* d_rights, int32_t* d_parents,
                             uint32_t* d_warn, void* d_temp, size_t temp_bytes, void* stream) {
  if (t < 2 || !d_keys || !d_lefts || !d_rights || !d_parents || !d_warn || !d_temp ||
      temp_bytes < lbvh_build_temp_bytes(t)) {
    set_error("cm_lbvh_build: invalid argument");
    return CM_ERR_ARG;
  }
  LbvhBufs b;
  b.t = t;
  b.lefts = d_lefts;
  b.rights = d_rights;
  b.parents = d_parents;
  b.warn = d_warn;
  carve_build(t, d_temp, b);
  return launch_build(b, d_keys, midpoint_split, (hipStream_t)stream);
}

// d_temp must be the same buffer that the preceding cm_lbvh_build used (claim records).
extern "C" int cm_lbvh_refit(int32_t t, const int32_t* d_lefts, const int32_t* d_rights,
                             const int32_t* d_parents, void* d_aabb_min, void* d_aabb_max,
                             int32_t* d_rmll, int32_t* d_rmlr, uint32_t* d_warn, void* d_temp,
                             size_t temp_bytes, void* stream) {
  (void)d_parents;
  if (t < 2 || !d_lefts || !d_rights || !d_aabb_min || !d_aabb_max || !d_rmll || !d_rmlr ||
      !d_warn || !d_temp || temp_bytes < lbvh_build_temp_bytes(t)) {
    set_error("cm_lbvh_refit: invalid argument");
    return CM_ERR_ARG;
  }
  LbvhBufs b;
  b.t = t;
  b.lefts = (int*)d_lefts;
  b.rights = (int*)d_rights;
  b.amin = (float4*)d_aabb_min;
  b.amax = (float4*)d_aabb_max;
  b.rmll = d_rmll;
  b.rmlr = d_rmlr;
  b.warn = d_warn;
  carve_build(t, d_temp, b);
  return launch_refit(b, (hipStream_t)stream);
}

extern "C" size_t cm_lbvh_pairs_temp_bytes(int32_t t, int32_t cap) {
  return lbvh_pairs_temp_bytes(t, cap);
}

extern "C" int cm_lbvh_pairs(int32_t t, const int32_t* d_lefts, const int32_t* d_rights,
                             const void* d_aabb_min, const void* d_aabb_max,
                             const int32_t* d_rmll, const int32_t* d_rmlr, int32_t cap,
                             int32_t* d_pairs, int32_t* d_n_pairs, uint32_t* d_warn,
                             void* d_temp, size_t temp_bytes, void* stream) {
  if (t < 2 || cap < 1 || !d_lefts || !d_rights || !d_aabb_min || !d_aabb_max || !d_rmll ||
      !d_rmlr || !d_pairs || !d_n_pairs || !d_warn || !d_temp ||
      temp_bytes < lbvh_pairs_temp_bytes(t, cap)) {
    set_error("cm_lbvh_pairs: invalid argument");
    return CM_ERR_ARG;
  }
  LbvhBufs b;
  b.t = t;
  b.lefts = (int*)d_lefts;
  b.rights = (int*)d_rights;
  b.amin = (float4*)d_aabb_min;
  b.amax = (float4*)d_aabb_max;
  b.rmll = (int*)d_rmll;
  b.rmlr = (int*)d_rmlr;
  b.warn = d_warn;
  char* tp = (char*)d_temp;
  int2* chunk = (int2*)tp;
  tp += align_up(sizeof(int2) * (size_t)t * cap);
  int* counts = (int*)tp;
  tp += align_up(sizeof(int) * (size_t)t);
  int* offsets = (int*)tp;
  tp += align_up(sizeof(int) * (size_t)t);
  b.nodes = (float4*)tp;
  tp += align_up(4 * sizeof(float4) * (size_t)t);
  int rc = launch_pack_nodes(b, (hipStream_t)stream);
  if (rc) return rc;
  return launch_pairs(b, cap, chunk, counts, offsets, (int2*)d_pairs, d_n_pairs, tp,
                      (hipStream_t)stream);
}
