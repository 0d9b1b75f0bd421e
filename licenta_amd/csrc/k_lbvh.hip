// Morton/LBVH stages of the substep as gfx950 kernels.
//
//  k_tri_update   <- DeformableUtils::_UpdateTrianglesContinuous
//                    (LicentaV2/LicentaV2/Physics/CudaDeformableUtils.cu:120-151)
//  k_build        <- CudaBVH::_GenerateBVH2 / FindRange / FindSplit
//                    (Collision/CudaBVH.cu:163-242), Delta/sgn (Core/CudaUtils.cu:338-362)
//  k_refit_local,
//  k_refit_top    <- CudaBVH::ComputeTreeAABBs (Collision/CudaBVH.cu:128-149)
//  k_traverse     <- CudaBVH::_GetAABBCollisions (Collision/CudaBVH.cu:23-113)
//
// Layout: node AABBs are float4 min / float4 max arrays of 2T-1 entries, internal
// nodes [0, T-1) then leaves [T-1, 2T-1), as in the reference. Parent writes are
// made deterministic with claim records (count, highest claimant): parents[c] =
// highest claimant, which is what the reference's last writer in thread order
// produces. The refit pulls each node from its two children, so a node shared by
// several parents (midpoint split on tied codes) is still the union of its two
// children.
#include "cm_internal.hpp"

#include <algorithm>
#include <climits>

namespace cm {

namespace {

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
    if (vals) vals[i] = (uint32_t)i;
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

// ---- per-cloth centroid boxes of a batch -------------------------------------------
// Floats are mapped to unsigned integers of the same order, so the box is a pair of
// integer atomic min/max per axis: exact, and independent of the order of arrival.
__device__ __forceinline__ uint32_t ford(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float funord(uint32_t e) {
  return __uint_as_float((e & 0x80000000u) ? (e & 0x7fffffffu) : ~e);
}

__global__ void __launch_bounds__(kBlock) k_cloth_box_init(int ncloth, uint32_t* __restrict__ cbox) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < 6 * ncloth) cbox[i] = (i % 6) < 3 ? 0xffffffffu : 0u;  // min xyz, max xyz
}

__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) {
  for (int o = 32; o >= 1; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) {
  for (int o = 32; o >= 1; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

// centroid box of each cloth; a wave whose triangles all belong to one cloth
// reduces first and issues one atomic per bound
__global__ void __launch_bounds__(kBlock) k_cloth_box(int t, const float4* __restrict__ cen,
                                                      const int* __restrict__ slot_cloth,
                                                      uint32_t* __restrict__ cbox) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < t;
  const int c = valid ? slot_cloth[i] : -1;
  const f3 p = valid ? xyz(cen[i]) : mk3(0, 0, 0);
  uint32_t e[6];
  e[0] = valid ? ford(p.x) : 0xffffffffu;
  e[1] = valid ? ford(p.y) : 0xffffffffu;
  e[2] = valid ? ford(p.z) : 0xffffffffu;
  e[3] = valid ? ford(p.x) : 0u;
  e[4] = valid ? ford(p.y) : 0u;
  e[5] = valid ? ford(p.z) : 0u;
  const int c0 = __shfl(c, 0, 64);
  const int cmax = (int)wave_max_u((uint32_t)(c + 1)) - 1, cmin = (int)wave_min_u((uint32_t)(c + 1)) - 1;
  const bool uniform = c0 >= 0 && cmin == c0 && cmax == c0;
  if (uniform) {
    for (int a = 0; a < 3; ++a) e[a] = wave_min_u(e[a]);
    for (int a = 3; a < 6; ++a) e[a] = wave_max_u(e[a]);
    if ((threadIdx.x & 63) == 0)
      for (int a = 0; a < 6; ++a)
        if (a < 3) atomicMin(&cbox[6 * c0 + a], e[a]); else atomicMax(&cbox[6 * c0 + a], e[a]);
  } else if (valid) {
    for (int a = 0; a < 6; ++a)
      if (a < 3) atomicMin(&cbox[6 * c + a], e[a]); else atomicMax(&cbox[6 * c + a], e[a]);
  }
}

__global__ void __launch_bounds__(kBlock) k_keys30_cloth(int t, const float4* __restrict__ cen,
                                                         const int* __restrict__ slot_cloth,
                                                         const uint32_t* __restrict__ cbox,
                                                         uint64_t* __restrict__ keys) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= t) return;
  const int c = slot_cloth[i];
  const uint32_t* bx = cbox + 6 * c;
  const f3 lo = mk3(funord(bx[0]), funord(bx[1]), funord(bx[2]));
  const f3 hi = mk3(funord(bx[3]), funord(bx[4]), funord(bx[5]));
  keys[i] = ((uint64_t)c << 30) | morton30(xyz(cen[i]), lo, hi);
}

// ---- Karras build -----------------------------------------------------------
// Core/CudaUtils.cu:338-351. In a batch, equal keys belong to one cloth and the
// index tie-break uses cloth-local indices, so each cloth's subtree is the tree its
// own keys would build.
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int i, int j, int maxj,
                                     const int* __restrict__ tstart) {
  if (j < 0 || j > maxj) return -1;
  const uint64_t x = k[i] ^ k[j];
  if (x) return clz64(x);
  const int base = tstart ? tstart[(int)(k[i] >> 30)] : 0;
  return clz64((uint64_t)(i - base) ^ (uint64_t)(j - base)) + 64;
}
__device__ __forceinline__ int sgn(int a) { return signbit((float)a) ? -1 : 1; }

__global__ void __launch_bounds__(kBlock) k_build_init(int t, int* __restrict__ cnt,
                                                       int* __restrict__ hi,
                                                       int* __restrict__ ntop, int ncloth,
                                                       int* __restrict__ croot) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < 2 * t - 1) {
    cnt[i] = 0;
    hi[i] = -1;
  }
  if (i < ncloth) croot[i] = -1;
  if (i == 0) *ntop = 0;  // deferred-node count of the refit
}

__global__ void __launch_bounds__(kBlock) k_build(int t, const uint64_t* __restrict__ k,
                                                  int midpoint, int* __restrict__ lefts,
                                                  int* __restrict__ rights, int* __restrict__ cnt,
                                                  int* __restrict__ hi, int2* __restrict__ ranges,
                                                  uint32_t* __restrict__ warn,
                                                  const int* __restrict__ tstart,
                                                  int* __restrict__ croot) {
  const int id = blockIdx.x * kBlock + threadIdx.x;
  const int nin = t - 1;
  if (id >= nin) return;
  // range (Collision/CudaBVH.cu:190-212)
  const int d = sgn(delta(k, id, id + 1, nin, tstart) - delta(k, id, id - 1, nin, tstart));
  const int dmin = delta(k, id, id - d, nin, tstart);
  int lmax = 2;
  while (delta(k, id, id + lmax * d, nin, tstart) > dmin) lmax <<= 1;
  int l = 0;
  for (int s = lmax / 2; s >= 1; s /= 2)
    if (delta(k, id, id + (l + s) * d, nin, tstart) > dmin) l += s;
  const int j = id + l * d;
  const int first = min(id, j), last = max(id, j);
  ranges[id] = make_int2(first, last);
  if (tstart) {  // the node spanning exactly one cloth is that cloth's root
    const int c = (int)(k[first] >> 30);
    if (first == tstart[c] && last == tstart[c + 1] - 1) croot[c] = id;
  }
  // split (Collision/CudaBVH.cu:214-242)
  const uint64_t fc = k[first], lc = k[last];
  int split;
  if (midpoint && fc == lc) {
    split = (first + last) >> 1;
  } else {
    const int com = midpoint ? clz64(fc ^ lc) : delta(k, first, last, nin, tstart);
    split = first;
    int step = last - first;
    do {
      step = (step + 1) >> 1;
      const int ns = split + step;
      if (ns < last) {
        const int pref = midpoint ? clz64(fc ^ k[ns]) : delta(k, first, ns, nin, tstart);
        if (pref > com) split = ns;
      }
    } while (step > 1);
  }
  const int left = split == first ? nin + split : split;
  const int right = split + 1 == last ? nin + split + 1 : split + 1;
  lefts[id] = left;
  rights[id] = right;
  const int o1 = atomicAdd(&cnt[left], 1);
  atomicMax(&hi[left], id);
  const int o2 = atomicAdd(&cnt[right], 1);
  atomicMax(&hi[right], id);
  if (o1 > 0 || o2 > 0) atomicOr(warn, (uint32_t)CM_WARN_BROKEN_TREE);
}

__global__ void __launch_bounds__(kBlock) k_parents(int t, const int* __restrict__ cnt,
                                                    const int* __restrict__ hi,
                                                    int* __restrict__ parents) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= 2 * t - 1) return;
  if (cnt[c] > 0) parents[c] = hi[c];
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

constexpr int kRefitCH = 1024;  // leaves (and node ids) per refit workgroup
constexpr int kTopCap = 2048;   // deferred nodes the top pass keeps in LDS
constexpr int kTopList = 8192;  // pending deferred nodes the top pass lists in LDS

// Refit, pass 1: one workgroup per chunk of 1024 consecutive leaves. Internal node
// q belongs to the chunk holding index q (a Karras node's index lies in its key
// range). Nodes whose range lies inside the chunk are finished here: in each
// synchronised pass a node whose children are leaves of the chunk or nodes finished
// in an earlier pass takes the union of its children (Collision/CudaBVH.cu:140-144),
// all in LDS. Every other node of the chunk (its range crosses the chunk, or a child
// is unfinished) is appended to the deferred list for k_refit_top. A node is the
// union of its two children however many parents claimed it.
__global__ void __launch_bounds__(kRefitCH) k_refit_local(
    int t, const int* __restrict__ lefts, const int* __restrict__ rights,
    const int2* __restrict__ ranges, float4* __restrict__ amin, float4* __restrict__ amax,
    int* __restrict__ rmll, int* __restrict__ rmlr, float4* __restrict__ nodes,
    int* __restrict__ state, int* __restrict__ top_list, int* __restrict__ top_index,
    int* __restrict__ ntop, int* __restrict__ gdone) {
  __shared__ float4 lmin[kRefitCH], lmax[kRefitCH];  // leaf boxes of the chunk
  __shared__ float4 nmin[kRefitCH], nmax[kRefitCH];  // node boxes of the chunk
  __shared__ int nrm[kRefitCH];                      // max(rmll, rmlr) of each node
  __shared__ int ndone[kRefitCH];                    // pass that finished the node, 0 = not
  const int nin = t - 1, base = blockIdx.x * kRefitCH, tid = threadIdx.x;
  if (base + tid < t) {
    lmin[tid] = amin[nin + base + tid];
    lmax[tid] = amax[nin + base + tid];
  }
  const int q = base + tid;
  int l = 0, r = 0;
  bool pending = false;
  if (q < nin) {
    const int2 rg = ranges[q];
    if (rg.x >= base && rg.y < base + kRefitCH) {
      l = lefts[q];
      r = rights[q];
      // children must be leaves or nodes of this chunk
      const int lo = l >= nin ? l - nin : l, ro = r >= nin ? r - nin : r;
      pending = lo >= base && lo < base + kRefitCH && ro >= base && ro < base + kRefitCH;
    }
  }
  ndone[tid] = 0;
  __syncthreads();
  bool fin = false;
  f3 lmn = mk3(0, 0, 0), lmx = lmn, rmn = lmn, rmx = lmn;
  int a = 0, b = 0;
  for (int pass = 1;; ++pass) {
    bool changed = false;
    if (pending) {
      const bool rl = l >= nin || (ndone[l - base] != 0 && ndone[l - base] < pass);
      const bool rr = r >= nin || (ndone[r - base] != 0 && ndone[r - base] < pass);
      if (rl && rr) {
        if (l >= nin) {
          lmn = xyz(lmin[l - nin - base]); lmx = xyz(lmax[l - nin - base]); a = l;
        } else {
          lmn = xyz(nmin[l - base]); lmx = xyz(nmax[l - base]); a = nrm[l - base];
        }
        if (r >= nin) {
          rmn = xyz(lmin[r - nin - base]); rmx = xyz(lmax[r - nin - base]); b = r;
        } else {
          rmn = xyz(nmin[r - base]); rmx = xyz(nmax[r - base]); b = nrm[r - base];
        }
        nmin[tid] = f4(vmin(lmn, rmn), 0.f);
        nmax[tid] = f4(vmax(lmx, rmx), 0.f);
        nrm[tid] = max(a, b);
        ndone[tid] = pass;
        pending = false;
        fin = true;
        changed = true;
      }
    }
    if (!__syncthreads_or(changed)) break;
  }
  if (q >= nin) return;
  if (fin) {
    amin[q] = nmin[tid];
    amax[q] = nmax[tid];
    rmll[q] = a;
    rmlr[q] = b;
    write_node(nodes, q, l, r, a, b, lmn, lmx, rmn, rmx);
    state[q] = 1;
  } else {
    state[q] = 0;
    const int j = atomicAdd(ntop, 1);
    top_list[j] = q;
    top_index[q] = j;
    gdone[j] = 0;  // pending for the later passes
  }
}

// child c of a deferred node: a leaf or a node finished by k_refit_local is final
// in global memory (kind 0); otherwise it is deferred node top_index[c] (kind 1)
__device__ __forceinline__ void top_child(int c, int nin, const int* __restrict__ state,
                                          const int* __restrict__ top_index,
                                          const float4* __restrict__ amin,
                                          const float4* __restrict__ amax,
                                          const int* __restrict__ rmll,
                                          const int* __restrict__ rmlr, bool& deferred,
                                          int& ref, f3& mn, f3& mx) {
  deferred = c < nin && state[c] == 0;
  if (deferred) {
    ref = top_index[c];
  } else {
    mn = xyz(amin[c]);
    mx = xyz(amax[c]);
    ref = c >= nin ? c : max(rmll[c], rmlr[c]);
  }
}

// a finished deferred node j (node qq): its box and rightmost leaf in the deferred
// scratch (for its parent) and in the tree
__device__ __forceinline__ void put_deferred(int j, int qq, int ll, int rr, int a, int b, f3 lmn,
                                             f3 lmx, f3 rmn, f3 rmx, float4* __restrict__ gmin,
                                             float4* __restrict__ gmax, int* __restrict__ grm,
                                             float4* __restrict__ amin, float4* __restrict__ amax,
                                             int* __restrict__ rmll, int* __restrict__ rmlr,
                                             float4* __restrict__ nodes) {
  gmin[j] = f4(vmin(lmn, rmn), 0.f);
  gmax[j] = f4(vmax(lmx, rmx), 0.f);
  grm[j] = max(a, b);
  amin[qq] = gmin[j];
  amax[qq] = gmax[j];
  rmll[qq] = a;
  rmlr[qq] = b;
  write_node(nodes, qq, ll, rr, a, b, lmn, lmx, rmn, rmx);
}

constexpr int kSuper = 64 * kRefitCH;  // leaves per workgroup of the middle pass
constexpr int kMidCap = 8192;          // deferred nodes one middle workgroup holds

// Refit, middle pass (large trees: more deferred nodes than the top pass holds in
// LDS): one workgroup per range of kSuper leaves finishes the deferred nodes whose
// key range lies inside it, with the synchronised passes of k_refit_local over the
// deferred scratch. A deferred child outside the range (possible only in a broken
// tree) leaves its parent to the top pass. Finished nodes are marked gdone = -pass;
// the top pass treats them as finished before its first pass.
__global__ void __launch_bounds__(1024) k_refit_mid(
    int t, const int* __restrict__ lefts, const int* __restrict__ rights,
    const int2* __restrict__ ranges, float4* __restrict__ amin, float4* __restrict__ amax,
    int* __restrict__ rmll, int* __restrict__ rmlr, float4* __restrict__ nodes,
    const int* __restrict__ state, const int* __restrict__ top_list,
    const int* __restrict__ top_index, const int* __restrict__ ntop, float4* __restrict__ gmin,
    float4* __restrict__ gmax, int* __restrict__ grm, int* __restrict__ gdone) {
  __shared__ int own[kMidCap];
  __shared__ int nown;
  const int nin = t - 1, tid = threadIdx.x;
  const int nt = *ntop;
  if (nt <= kTopCap) return;  // the top pass keeps a short list in LDS
  const int sbase = blockIdx.x * kSuper, send = sbase + kSuper;
  if (tid == 0) nown = 0;
  __syncthreads();
  for (int j = tid; j < nt; j += 1024) {
    const int2 rg = ranges[top_list[j]];
    if (rg.x >= sbase && rg.y < send) {
      const int k = atomicAdd(&nown, 1);
      if (k < kMidCap) own[k] = j;
    }
  }
  __syncthreads();
  const int no = nown;
  if (no > kMidCap) return;  // too many: all of them are left to the top pass
  auto inside = [&](int c) {
    const int2 rg = ranges[c];
    return rg.x >= sbase && rg.y < send;
  };
  for (int pass = 1;; ++pass) {
    bool changed = false;
    for (int k = tid; k < no; k += 1024) {
      const int j = own[k];
      if (gdone[j] != 0) continue;
      const int qq = top_list[j], ll = lefts[qq], rr = rights[qq];
      bool lt, rt;
      int a, b;
      f3 lmn, lmx, rmn, rmx;
      top_child(ll, nin, state, top_index, amin, amax, rmll, rmlr, lt, a, lmn, lmx);
      top_child(rr, nin, state, top_index, amin, amax, rmll, rmlr, rt, b, rmn, rmx);
      if (lt && !(inside(ll) && gdone[a] < 0 && -gdone[a] < pass)) continue;
      if (rt && !(inside(rr) && gdone[b] < 0 && -gdone[b] < pass)) continue;
      if (lt) { lmn = xyz(gmin[a]); lmx = xyz(gmax[a]); a = grm[a]; }
      if (rt) { rmn = xyz(gmin[b]); rmx = xyz(gmax[b]); b = grm[b]; }
      put_deferred(j, qq, ll, rr, a, b, lmn, lmx, rmn, rmx, gmin, gmax, grm, amin, amax, rmll,
                   rmlr, nodes);
      __threadfence_block();
      gdone[j] = -pass;
      changed = true;
    }
    if (!__syncthreads_or(changed)) break;
  }
}

// Refit, pass 2: one workgroup finishes the deferred nodes with the same
// synchronised passes. Up to kTopCap deferred nodes keep their boxes in LDS and
// their children in registers (two per thread); beyond that the boxes live in
// global scratch and node data is re-read each pass. Clears the list count.
__global__ void __launch_bounds__(1024) k_refit_top(
    int t, const int* __restrict__ lefts, const int* __restrict__ rights,
    float4* __restrict__ amin, float4* __restrict__ amax, int* __restrict__ rmll,
    int* __restrict__ rmlr, float4* __restrict__ nodes, const int* __restrict__ state,
    const int* __restrict__ top_list, const int* __restrict__ top_index, int* __restrict__ ntop,
    float4* __restrict__ gmin, float4* __restrict__ gmax, int* __restrict__ grm,
    int* __restrict__ gdone, uint32_t* __restrict__ warn) {
  __shared__ float4 smin[kTopCap], smax[kTopCap];
  __shared__ int srm[kTopCap], sdone[kTopCap];
  __shared__ int plist[kTopList];
  __shared__ int npend;
  const int nin = t - 1, tid = threadIdx.x;
  const int nt = *ntop;
  bool left_over = false;
  if (nt <= kTopCap) {
    constexpr int kPer = kTopCap / 1024;
    int q[kPer], l[kPer], r[kPer], la[kPer], ra[kPer];
    bool own[kPer], ltop[kPer], rtop[kPer];
    f3 lmn[kPer], lmx[kPer], rmn[kPer], rmx[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int j = tid + k * 1024;
      own[k] = j < nt;
      if (own[k]) {
        q[k] = top_list[j];
        l[k] = lefts[q[k]];
        r[k] = rights[q[k]];
        top_child(l[k], nin, state, top_index, amin, amax, rmll, rmlr, ltop[k], la[k], lmn[k],
                  lmx[k]);
        top_child(r[k], nin, state, top_index, amin, amax, rmll, rmlr, rtop[k], ra[k], rmn[k],
                  rmx[k]);
        sdone[j] = 0;
      }
    }
    __syncthreads();
    for (int pass = 1;; ++pass) {
      bool changed = false;
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int j = tid + k * 1024;
        if (!own[k] || sdone[j] != 0) continue;
        const bool rl = !ltop[k] || (sdone[la[k]] != 0 && sdone[la[k]] < pass);
        const bool rr = !rtop[k] || (sdone[ra[k]] != 0 && sdone[ra[k]] < pass);
        if (!(rl && rr)) continue;
        if (ltop[k]) {
          const int c = la[k];
          lmn[k] = xyz(smin[c]); lmx[k] = xyz(smax[c]); la[k] = srm[c];
        }
        if (rtop[k]) {
          const int c = ra[k];
          rmn[k] = xyz(smin[c]); rmx[k] = xyz(smax[c]); ra[k] = srm[c];
        }
        smin[j] = f4(vmin(lmn[k], rmn[k]), 0.f);
        smax[j] = f4(vmax(lmx[k], rmx[k]), 0.f);
        srm[j] = max(la[k], ra[k]);
        sdone[j] = pass;
        changed = true;
      }
      if (!__syncthreads_or(changed)) break;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int j = tid + k * 1024;
      if (!own[k]) continue;
      if (sdone[j] == 0) {
        left_over = true;
        continue;
      }
      amin[q[k]] = smin[j];
      amax[q[k]] = smax[j];
      rmll[q[k]] = la[k];
      rmlr[q[k]] = ra[k];
      write_node(nodes, q[k], l[k], r[k], la[k], ra[k], lmn[k], lmx[k], rmn[k], rmx[k]);
    }
  } else {
    // gdone: 0 pending (set by k_refit_local), < 0 finished by k_refit_mid, > 0 the
    // pass of this kernel that finished it. The pending nodes are listed in LDS when
    // they fit, so the passes skip the finished ones.
    if (tid == 0) npend = 0;
    __syncthreads();
    for (int j = tid; j < nt; j += 1024)
      if (gdone[j] == 0) {
        const int k = atomicAdd(&npend, 1);
        if (k < kTopList) plist[k] = j;
      }
    __syncthreads();
    const int np = npend;
    const bool listed = np <= kTopList;
    const int nloop = listed ? np : nt;
    for (int pass = 1;; ++pass) {
      bool changed = false;
      for (int k = tid; k < nloop; k += 1024) {
        const int j = listed ? plist[k] : k;
        if (gdone[j] != 0) continue;
        const int qq = top_list[j], ll = lefts[qq], rr = rights[qq];
        bool lt, rt;
        int a, b;
        f3 lmn, lmx, rmn, rmx;
        top_child(ll, nin, state, top_index, amin, amax, rmll, rmlr, lt, a, lmn, lmx);
        top_child(rr, nin, state, top_index, amin, amax, rmll, rmlr, rt, b, rmn, rmx);
        if (lt && !(gdone[a] != 0 && gdone[a] < pass)) continue;
        if (rt && !(gdone[b] != 0 && gdone[b] < pass)) continue;
        if (lt) { lmn = xyz(gmin[a]); lmx = xyz(gmax[a]); a = grm[a]; }
        if (rt) { rmn = xyz(gmin[b]); rmx = xyz(gmax[b]); b = grm[b]; }
        put_deferred(j, qq, ll, rr, a, b, lmn, lmx, rmn, rmx, gmin, gmax, grm, amin, amax, rmll,
                     rmlr, nodes);
        __threadfence_block();
        gdone[j] = pass;
        changed = true;
      }
      if (!__syncthreads_or(changed)) break;
    }
    for (int k = tid; k < nloop; k += 1024)
      if (gdone[listed ? plist[k] : k] == 0) left_over = true;
  }
  if (__syncthreads_or(left_over) && tid == 0) atomicOr(warn, (uint32_t)CM_WARN_REFIT_LIMIT);
  if (tid == 0) *ntop = 0;
}

// ---- traversal ----------------------------------------------------------------
constexpr int kTravBlock = 128;  // 2 waves
constexpr int kStack = 63;       // usable entries, as the reference's int[64] with its -1 sentinel
// The first kLdsStack entries of a lane's stack live in LDS (8 KiB per workgroup, so
// LDS leaves room for 8 waves per SIMD); deeper entries, rare, go to the lane's
// global spill row of kStack - kLdsStack ints.
constexpr int kLdsStack = 16;
constexpr int kSpill = kStack - kLdsStack;
static_assert(kSpill == kTravSpill, "spill row");

__device__ __forceinline__ void put_pair(int2* p, int a, int b) {
  // a plain store: a leaf's chunk fills whole lines in L2 before they go out (streamed
  // stores left each 8-byte pair a partial-line write; the 512-cloth batch traversal
  // took 2.2 ms with them, 1.36 ms without)
  *(long long*)p = (long long)(uint32_t)a | ((long long)b << 32);
}

// State of one leaf's DFS between two visits: the node to visit next, the stack
// depth, the pairs written and the visits made.
struct DfsState {
  int node, top, ncol;
  long long visits;
};

// DFS of one leaf (Collision/CudaBVH.cu:23-113) from state st: its pairs in the
// reference's order into out (at most cap), stack entry k in sstk[k][col]. Returns
// true when the walk ended (stack empty, the first dropped pair, or the visit limit);
// false when `budget` visits had been made first, with st at the next visit.
// The stack's first kLds entries are sstk[k][col], the rest spill[k - kLds].
template <int kRow, int kLds>
__device__ __forceinline__ bool dfs_run(int nin, int me, int leaf, f3 mn, f3 mx,
                                        const float4* __restrict__ nodes, int cap,
                                        int2* __restrict__ out, long long limit,
                                        long long budget, DfsState& st, uint32_t& w,
                                        int (*sstk)[kRow], int col, int* __restrict__ spill) {
  int node = st.node, top = st.top, ncol = st.ncol;
  long long visits = st.visits;
  bool done = true;
  while (true) {
    if (visits >= budget) {
      done = false;
      break;
    }
    if (++visits > limit) {
      w |= CM_WARN_TRAVERSAL_LIMIT;
      break;
    }
    // One 64-byte record per node: children, rightmost leaves, both child boxes.
    // Both tests use & so that all four loads are issued together.
    const float4* nd = nodes + 4 * (int64_t)node;
    const float4 r0 = nd[0], r1 = nd[1], r2 = nd[2], r3 = nd[3];
    const int l = __float_as_int(r0.x), r = __float_as_int(r0.y);
    const int rml = __float_as_int(r0.z), rmr = __float_as_int(r0.w);
    const bool ol = (rml > me) & overlap(mn, mx, mk3(r1.x, r1.y, r1.z), mk3(r1.w, r2.x, r2.y));
    const bool orr = (rmr > me) & overlap(mn, mx, mk3(r2.z, r2.w, r3.x), mk3(r3.y, r3.z, r3.w));
    const bool el = ol & (l >= nin), er = orr & (r >= nin);
    if (el | er) {
      // Past the cap the reference keeps walking but drops every further pair, so
      // stopping at the first dropped pair yields the same list.
      if (ncol + (int)el + (int)er > cap) {
        if (el && ncol < cap) put_pair(out + ncol++, leaf, l - nin);
        else if (er && ncol < cap) put_pair(out + ncol++, leaf, r - nin);
        w |= CM_WARN_PAIR_CAP;
        break;
      }
      if (el) put_pair(out + ncol++, leaf, l - nin);
      if (er) put_pair(out + ncol++, leaf, r - nin);
    }
    const bool tl = ol & (l < nin), tr = orr & (r < nin);
    if (tl | tr) {
      node = tl ? l : r;
      if (tl & tr) {
        if (top < kStack) {
          if constexpr (kLds < kStack) {
            if (top < kLds) sstk[top][col] = r;
            else spill[top - kLds] = r;
          } else {
            sstk[top][col] = r;
          }
          ++top;
        } else {
          w |= CM_WARN_STACK_OVERFLOW;  // the push is dropped, as in the reference
        }
      }
    } else {
      if (top == 0) break;
      --top;
      if constexpr (kLds < kStack) node = top < kLds ? sstk[top][col] : spill[top - kLds];
      else node = sstk[top][col];
    }
  }
  st.node = node;
  st.top = top;
  st.ncol = ncol;
  st.visits = visits;
  return done;
}

// A leaf whose walk outlasts the first pass's visit budget leaves its state in a
// record of kRec ints: leaf, pairs written, stack depth, next node, cap, visits, then
// the kStack stack entries. The second pass finishes it with a whole wave.
constexpr int kRec = kTravRec;
static_assert(kRec >= 6 + kStack, "record holds the stack");

// First pass: one lane per leaf, at most `budget` visits (no budget: budget <= 0 or
// no record buffer). A leaf still walking then takes a record slot, if one is left,
// and counts zero pairs here; past the slots it walks on to the end.
__global__ void __launch_bounds__(kTravBlock) k_traverse(int t, const float4* __restrict__ nodes,
                                                         const float4* __restrict__ amin,
                                                         const float4* __restrict__ amax, int cap,
                                                         int2* __restrict__ chunk,
                                                         int* __restrict__ counts,
                                                         uint32_t* __restrict__ warn,
                                                         const int* __restrict__ slot_cloth,
                                                         const int* __restrict__ croot,
                                                         const int* __restrict__ ccap,
                                                         int* __restrict__ tile_sum,
                                                         int xchunk, int budget,
                                                         int* __restrict__ rec, int rcap,
                                                         int* __restrict__ hctl,
                                                         int* __restrict__ spill) {
  // Entry k of every lane's stack is one row, so a wave's pushes and pops at equal
  // depths hit 64 consecutive banks.
  __shared__ int sstk[kLdsStack][kTravBlock];
  // Workgroups are dealt to the 8 XCDs in turn. Each XCD takes every eighth chunk of
  // `xchunk` consecutive tiles (Morton-ordered leaves), so its L2 serves compact
  // subtrees while the dense regions of the cloth, whose leaves visit many more
  // nodes, are shared evenly between the XCDs. The grid is padded to whole rounds.
  const int j = blockIdx.x / kXcds, xcd = blockIdx.x % kXcds;
  const int tile = ((j / xchunk) * kXcds + xcd) * xchunk + j % xchunk;
  if (tile >= (t + kTravBlock - 1) / kTravBlock) return;
  const int leaf = tile * kTravBlock + threadIdx.x;
  int ncol = 0;
  const int stride = cap;  // every leaf's chunk holds the body's cap (a cloth's may be less)
  uint32_t w = 0;
  const int nin = t - 1;
  const int me = nin + leaf;
  DfsState st{0, 0, 0, 0};  // root (a batch: the root of the leaf's cloth, its cap)
  bool walk = leaf < t;
  if (walk && croot) {
    const int c = slot_cloth[leaf];
    st.node = croot[c];
    cap = ccap[c];
    walk = st.node >= 0;  // a one-triangle cloth has no pairs
  }
  f3 mn = mk3(0, 0, 0), mx = mk3(0, 0, 0);
  int2* out = chunk + (int64_t)leaf * stride;
  int* sp = spill + (int64_t)leaf * kSpill;
  const long long limit = 64ll * t + 4096;
  bool done = true;
  if (walk) {
    mn = xyz(amin[me]);
    mx = xyz(amax[me]);
    const bool budgeted = budget > 0 && rec;
    done = dfs_run<kTravBlock, kLdsStack>(nin, me, leaf, mn, mx, nodes, cap, out, limit,
                                          budgeted ? (long long)budget : LLONG_MAX, st, w, sstk,
                                          threadIdx.x, sp);
  }
  // Leaves still walking claim record slots with one atomic per wave (thousands of
  // lanes on one counter would queue at its L2 channel).
  const int lane = threadIdx.x & 63;
  const uint64_t want = __ballot(!done);
  bool handed = false;
  if (want) {
    int base = 0;
    if (lane == 0) base = atomicAdd(hctl, __popcll(want));
    base = __shfl(base, 0, 64);
    if (!done) {
      const int slot = base + __popcll(want & ((1ull << lane) - 1));
      if (slot < rcap) {
        int* R = rec + (int64_t)slot * kRec;
        R[0] = leaf;
        R[1] = st.ncol;
        R[2] = st.top;
        R[3] = st.node;
        R[4] = cap;
        R[5] = (int)st.visits;
        for (int k = 0; k < st.top; ++k)
          R[6 + k] = k < kLdsStack ? sstk[k][threadIdx.x] : sp[k - kLdsStack];
        handed = true;
      } else {
        dfs_run<kTravBlock, kLdsStack>(nin, me, leaf, mn, mx, nodes, cap, out, limit, LLONG_MAX,
                                       st, w, sstk, threadIdx.x, sp);
      }
    }
  }
  if (walk) ncol = handed ? 0 : st.ncol;
  if (leaf < t) counts[leaf] = ncol;
  // one warning atomic per wave: two thirds of the leaves of a draped cloth hit the cap
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) w |= __shfl_xor(w, o, 64);
  if (lane == 0 && w) atomicOr(warn, w);
  // pairs of the tile (its leaves' counts follow in the compaction)
  __shared__ int wsum[kTravBlock / 64];
  int v = ncol;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int q = 0; q < kTravBlock / 64; ++q) tot += wsum[q];
    tile_sum[tile] = tot;
  }
}

// ---- second pass: one wave finishes one handed-over leaf -------------------------
// The reference's DFS visits nodes in preorder and emits a node's left leaf child
// before its right one, so a pair's place in the leaf's list is the preorder place of
// the node that emits it. The wave expands up to 64 pending nodes at once, in any
// order, and tags every pair with a 64-bit key that sorts in that preorder:
//   bits 63..58  rank of the pending subtree handed over (0 = the next node to visit,
//                then the stack entries from the top down: they are visited in that
//                order, each subtree before the next)
//   bits 57..8   the path below that subtree's root, one bit per level (0 = left)
//   bits  7..1   the depth of that path
//   bit   0      0 for the left leaf child, 1 for the right one
// Sorting (rank, path padded with zeros, depth, side) is preorder: a node comes before
// its descendants (same padded path, smaller depth) and a left subtree before the
// right one. Pending nodes are kept sorted by key (the smallest on top), so every pair
// below the top pending node's key is final. The wave stops once more than the
// remaining cap of pairs are final: the first `cap` are the leaf's list and one more
// means the reference dropped pairs (CM_WARN_PAIR_CAP). A push the reference would
// drop (stack full) is mirrored by tracking each pending node's stack depth: a node
// whose children are both walked at depth kStack loses its right child, and
// CM_WARN_STACK_OVERFLOW is raised when such a node precedes the point where the
// reference's walk stopped. A leaf that overflows this pass's buffers or path bits,
// or could reach the visit limit, is walked by lane 0 alone from its record.
constexpr int kPend = 512;     // pending nodes per leaf
constexpr int kHits = 256;     // pairs kept per leaf
constexpr int kPathBits = 50;
constexpr int kResumeBlocks = 4096;
constexpr int kResumeMaxCap = 96;  // kHits holds cap + 1 kept pairs and a round's 128 new ones
static_assert(kResumeMaxCap + 1 + 128 <= kHits, "hit buffer");

__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
  int s = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(s, o, 64);
    if (lane >= o) s += y;
  }
  total = __shfl(s, 63, 64);
  return s - v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// More records than this and each is finished by one lane (k_traverse_lanes) instead
// of a whole wave: a wave per record pays off for the long tail of a few heavy
// leaves, while with many records the lanes' serial walks keep more of them going
// at once.
constexpr int kLaneModeMin = 16384;

__global__ void __launch_bounds__(64) k_traverse_resume(int t, const float4* __restrict__ nodes,
                                                        const float4* __restrict__ amin,
                                                        const float4* __restrict__ amax,
                                                        int2* __restrict__ chunk, int stride,
                                                        int* __restrict__ counts,
                                                        uint32_t* __restrict__ warn,
                                                        int* __restrict__ tile_sum,
                                                        const int* __restrict__ rec, int rcap,
                                                        int* __restrict__ hctl) {
  __shared__ int pe[kPend];  // node | stack depth << 26
  __shared__ uint64_t pk[kPend];
  __shared__ uint64_t hk[kHits];
  __shared__ int hp[kHits];
  __shared__ int fstk[kStack][1];
  __shared__ unsigned long long s_ovf, s_stop;
  const int lane = threadIdx.x;
  const int nin = t - 1;
  const long long limit = 64ll * t + 4096;
  const int nres = min(hctl[0], rcap);
  if (nres > kLaneModeMin) return;  // k_traverse_lanes takes them
  uint32_t wacc = 0;
  for (int rid = blockIdx.x; rid < nres; rid += gridDim.x) {
    const int* R = rec + (int64_t)rid * kRec;
    const int leaf = R[0], ncol0 = R[1], top0 = R[2], cur = R[3], cap = R[4];
    const long long visits0 = R[5];
    const int me = nin + leaf;
    const f3 mn = xyz(amin[me]), mx = xyz(amax[me]);
    const int capr = cap - ncol0;  // pairs still to find (0: only whether one more exists)
    int2* out = chunk + (int64_t)leaf * stride;
    // pending: the stack entries bottom-up, then the next node on top
    const int m = top0 + 1;
    for (int k = lane; k < m; k += 64) {
      pe[k] = (k < top0 ? R[6 + k] : cur) | (k << 26);  // visited at stack depth k
      pk[k] = (uint64_t)(top0 - k) << 58;
    }
    if (lane == 0) {
      s_ovf = ~0ull;
      s_stop = ~0ull;
    }
    __syncthreads();
    int np = m, nh = 0;
    long long expanded = 0;
    bool fallback = cap > kResumeMaxCap;
    while (np > 0 && !fallback) {
      const int take = min(64, np);
      const int base = np - take;
      const bool act = lane < take;
      int node = 0, sd = 0;
      uint64_t key = 0;
      if (act) {
        const int e = pe[np - 1 - lane];
        node = e & 0x3ffffff;
        sd = e >> 26;
        key = pk[np - 1 - lane];
      }
      __syncthreads();
      expanded += take;
      int nc = 0, ne = 0, ce[2] = {0, 0}, ep[2] = {0, 0};
      uint64_t ck[2] = {0, 0}, ek[2] = {0, 0};
      bool deep = false;
      if (act) {
        const float4* nd = nodes + 4 * (int64_t)node;
        const float4 r0 = nd[0], r1 = nd[1], r2 = nd[2], r3 = nd[3];
        const int l = __float_as_int(r0.x), r = __float_as_int(r0.y);
        const int rml = __float_as_int(r0.z), rmr = __float_as_int(r0.w);
        const bool ol = (rml > me) & overlap(mn, mx, mk3(r1.x, r1.y, r1.z), mk3(r1.w, r2.x, r2.y));
        const bool orr =
            (rmr > me) & overlap(mn, mx, mk3(r2.z, r2.w, r3.x), mk3(r3.y, r3.z, r3.w));
        if (ol & (l >= nin)) {
          ek[ne] = key;
          ep[ne++] = l - nin;
        }
        if (orr & (r >= nin)) {
          ek[ne] = key | 1ull;
          ep[ne++] = r - nin;
        }
        const bool tl = ol & (l < nin), tr = orr & (r < nin);
        if (tl | tr) {
          const int d = (int)((key >> 1) & 127ull);
          if (d + 1 > kPathBits) deep = true;
          const uint64_t kl = (key & ~0xffull) | ((uint64_t)(d + 1) << 1);
          const uint64_t kr = kl | (1ull << (57 - d));
          const bool both = tl & tr;
          if (tl) {
            ce[nc] = l | ((both && sd < kStack ? sd + 1 : sd) << 26);
            ck[nc++] = kl;
          }
          if (tr) {
            if (both && sd >= kStack) {
              atomicMin(&s_ovf, (unsigned long long)key);
            } else {
              ce[nc] = r | (sd << 26);
              ck[nc++] = kr;
            }
          }
        }
      }
      int totc, tote;
      const int exc = wave_excl_scan(nc, lane, totc);
      const int exe = wave_excl_scan(ne, lane, tote);
      if (__any(deep) || base + totc > kPend || visits0 + expanded > limit) {
        fallback = true;
        break;
      }
      if (nh + tote > kHits) {
        // keep the capr + 1 smallest keys: a larger one can only move further back
        // once the pending subtrees add theirs, so it is never listed nor the stop
        int rk[kHits / 64];
        uint64_t kk[kHits / 64];
        int kp[kHits / 64];
#pragma unroll
        for (int u = 0; u < kHits / 64; ++u) {
          const int i = lane + 64 * u;
          rk[u] = kHits;
          if (i < nh) {
            kk[u] = hk[i];
            kp[u] = hp[i];
            int rank = 0;
            for (int j2 = 0; j2 < nh; ++j2) rank += hk[j2] < kk[u];
            rk[u] = rank;
          }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kHits / 64; ++u)
          if (rk[u] <= capr) {
            hk[rk[u]] = kk[u];
            hp[rk[u]] = kp[u];
          }
        nh = min(nh, capr + 1);
        __syncthreads();
      }
      for (int q = 0; q < nc; ++q) {
        const int at = base + totc - 1 - (exc + q);
        pe[at] = ce[q];
        pk[at] = ck[q];
      }
      for (int q = 0; q < ne; ++q) {
        hk[nh + exe + q] = ek[q];
        hp[nh + exe + q] = ep[q];
      }
      __syncthreads();
      np = base + totc;
      nh += tote;
      if (nh > capr) {
        const uint64_t bound = np > 0 ? pk[np - 1] : ~0ull;
        int below = 0;
        for (int i = lane; i < nh; i += 64) below += hk[i] < bound;
        if (wave_sum(below) > capr) break;
      }
    }
    int nout;
    uint32_t w = 0;
    if (fallback) {
      __syncthreads();
      nout = 0;
      if (lane == 0) {
        for (int k = 0; k < top0; ++k) fstk[k][0] = R[6 + k];
        DfsState st{cur, top0, ncol0, visits0};
        dfs_run<1, kStack>(nin, me, leaf, mn, mx, nodes, cap, out, limit, LLONG_MAX, st, w, fstk,
                           0, nullptr);
        nout = st.ncol;
      }
    } else {
      // each pair's place is the number of smaller keys (keys are distinct)
      for (int i = lane; i < nh; i += 64) {
        const uint64_t ki = hk[i];
        int rank = 0;
        for (int j2 = 0; j2 < nh; ++j2) rank += hk[j2] < ki;
        if (rank < capr) put_pair(out + ncol0 + rank, leaf, hp[i]);
        if (rank == capr) s_stop = ki & ~1ull;  // the node where the reference stopped
      }
      __syncthreads();
      nout = ncol0 + min(nh, capr);
      if (nh > capr) w |= CM_WARN_PAIR_CAP;
      if (s_ovf != ~0ull && s_ovf < s_stop) w |= CM_WARN_STACK_OVERFLOW;
    }
    if (lane == 0) {
      counts[leaf] = nout;
      atomicAdd(&tile_sum[leaf / kTravBlock], nout);
      wacc |= w;
    }
    __syncthreads();
  }
  // the workgroup's warnings in one atomic (the record count is cleared by the next
  // kernel on the stream, once every workgroup here has read it)
  if (lane == 0 && wacc) atomicOr(warn, wacc);
}

// Second pass when there are more than kLaneModeMin records: one lane per record
// walks on serially from the record's state (next node, stack, pairs written,
// visits), exactly as the first pass would have without a budget.
__global__ void __launch_bounds__(kTravBlock) k_traverse_lanes(
    int t, const float4* __restrict__ nodes, const float4* __restrict__ amin,
    const float4* __restrict__ amax, int2* __restrict__ chunk, int stride,
    int* __restrict__ counts, uint32_t* __restrict__ warn, int* __restrict__ tile_sum,
    const int* __restrict__ rec, int rcap, const int* __restrict__ hctl,
    int* __restrict__ spill) {
  __shared__ int sstk[kLdsStack][kTravBlock];
  const int nres = min(hctl[0], rcap);
  if (nres <= kLaneModeMin) return;  // k_traverse_resume takes them
  const int nin = t - 1;
  const long long limit = 64ll * t + 4096;
  uint32_t w = 0;
  for (int rid = blockIdx.x * kTravBlock + threadIdx.x; rid < nres;
       rid += gridDim.x * kTravBlock) {
    const int* R = rec + (int64_t)rid * kRec;
    const int leaf = R[0], top0 = R[2];
    const int me = nin + leaf;
    int* sp = spill + (int64_t)leaf * kSpill;
    for (int k = 0; k < top0; ++k) {
      if (k < kLdsStack) sstk[k][threadIdx.x] = R[6 + k];
      else sp[k - kLdsStack] = R[6 + k];
    }
    DfsState st{R[3], top0, R[1], (long long)R[5]};
    dfs_run<kTravBlock, kLdsStack>(nin, me, leaf, xyz(amin[me]), xyz(amax[me]), nodes, R[4],
                                   chunk + (int64_t)leaf * stride, limit, LLONG_MAX, st, w, sstk,
                                   threadIdx.x, sp);
    counts[leaf] = st.ncol;
    atomicAdd(&tile_sum[leaf / kTravBlock], st.ncol);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) w |= __shfl_xor(w, o, 64);
  if ((threadIdx.x & 63) == 0 && w) atomicOr(warn, w);
}

__global__ void k_clear_records(int* __restrict__ hctl) {
  if (threadIdx.x == 0) hctl[0] = 0;
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

// One wave per leaf copies its chunk to the compacted list. The leaf's offset is its
// tile's prefix plus the counts of the tile's earlier leaves (at most 127, summed by
// the wave).
// One workgroup per traversal tile of kTravBlock leaves: the first wave scans the
// leaves' counts, then the workgroup writes the tile's pairs in order, kBlock
// consecutive output slots at a time (each thread finds its slot's leaf by binary
// search over the tile's prefixes), so reads and writes are contiguous runs.
__global__ void __launch_bounds__(kBlock) k_compact_pairs(int t, int cap,
                                                          const int2* __restrict__ chunk,
                                                          const int* __restrict__ counts,
                                                          const int* __restrict__ tile_pre,
                                                          int2* __restrict__ pairs,
                                                          int* __restrict__ hctl) {
  static_assert(kTravBlock == 128, "two leaves per lane of the first wave");
  // the traversal's hand-over record count, read by its second pass, starts the next
  // substep at zero
  if (hctl && blockIdx.x == 0 && threadIdx.x == 0) hctl[0] = 0;
  __shared__ int pre[kTravBlock + 1];
  const int tile = blockIdx.x, tid = threadIdx.x;
  const int l0 = tile * kTravBlock;
  if (tid < 64) {
    const int i0 = l0 + 2 * tid, i1 = i0 + 1;
    const int c0 = i0 < t ? counts[i0] : 0, c1 = i1 < t ? counts[i1] : 0;
    int s = c0 + c1;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(s, o, 64);
      if (tid >= o) s += y;
    }
    const int ex = s - (c0 + c1);
    pre[2 * tid] = ex;
    pre[2 * tid + 1] = ex + c0;
    if (tid == 63) pre[kTravBlock] = s;
  }
  __syncthreads();
  const int total = pre[kTravBlock];
  const int out0 = tile_pre[tile];
  for (int k = tid; k < total; k += kBlock) {
    int lo = 0, hi = kTravBlock - 1;  // the last leaf whose prefix is <= k holds slot k
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    pairs[out0 + k] = chunk[(int64_t)(l0 + lo) * cap + (k - pre[lo])];
  }
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

int launch_tri_update_batch(const LbvhBufs& b, const float4* x, const float4* p, const int4* tri,
                            float th, uint64_t* keys, uint32_t* vals, float4* centroid_tmp,
                            hipStream_t s) {
  const int t = b.t;
  if (t <= 0) return CM_OK;
  hipLaunchKernelGGL(k_cloth_box_init, dim3(nblocks(6 * b.ncloth)), dim3(kBlock), 0, s, b.ncloth,
                     b.cbox);
  hipLaunchKernelGGL(k_tri_update, dim3(nblocks(t)), dim3(kBlock), 0, s, t, x, p, tri, th, 30,
                     b.amin, b.amax, keys, vals, centroid_tmp, nullptr);
  hipLaunchKernelGGL(k_cloth_box, dim3(nblocks(t)), dim3(kBlock), 0, s, t, centroid_tmp,
                     b.slot_cloth, b.cbox);
  hipLaunchKernelGGL(k_keys30_cloth, dim3(nblocks(t)), dim3(kBlock), 0, s, t, centroid_tmp,
                     b.slot_cloth, b.cbox, keys);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_build(const LbvhBufs& b, const uint64_t* keys, int midpoint, hipStream_t s) {
  const int t = b.t;
  if (t < 2) return CM_OK;
  hipLaunchKernelGGL(k_build_init, dim3(nblocks(std::max(2 * t - 1, b.ncloth))), dim3(kBlock), 0,
                     s, t, b.claim_cnt, b.claim_hi, b.ntop, b.ncloth, b.croot);
  hipLaunchKernelGGL(k_build, dim3(nblocks(t - 1)), dim3(kBlock), 0, s, t, keys, midpoint, b.lefts,
                     b.rights, b.claim_cnt, b.claim_hi, b.ranges, b.warn,
                     b.ncloth ? b.tstart : nullptr, b.croot);
  hipLaunchKernelGGL(k_parents, dim3(nblocks(2 * t - 1)), dim3(kBlock), 0, s, t, b.claim_cnt,
                     b.claim_hi, b.parents);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_refit(const LbvhBufs& b, hipStream_t s) {
  const int t = b.t;
  if (t < 2) return CM_OK;
  hipLaunchKernelGGL(k_refit_local, dim3((t + kRefitCH - 1) / kRefitCH), dim3(kRefitCH), 0, s, t,
                     b.lefts, b.rights, b.ranges, b.amin, b.amax, b.rmll, b.rmlr, b.nodes, b.state,
                     b.top_list, b.top_index, b.ntop, b.tdone);
  if (t > 2 * kSuper)  // smaller trees defer at most a few hundred nodes
    hipLaunchKernelGGL(k_refit_mid, dim3((t + kSuper - 1) / kSuper), dim3(1024), 0, s, t, b.lefts,
                       b.rights, b.ranges, b.amin, b.amax, b.rmll, b.rmlr, b.nodes, b.state,
                       b.top_list, b.top_index, b.ntop, b.tmin, b.tmax, b.trm, b.tdone);
  hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(1024), 0, s, t, b.lefts, b.rights, b.amin, b.amax,
                     b.rmll, b.rmlr, b.nodes, b.state, b.top_list, b.top_index, b.ntop, b.tmin,
                     b.tmax, b.trm, b.tdone, b.warn);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

// both traversal passes (the second only with a budget and record buffers)
static int launch_traverse(const LbvhBufs& b, int cap, int2* chunk, int* counts, int* offsets,
                           hipStream_t s) {
  const int t = b.t;
  const int nt = (t + kTravBlock - 1) / kTravBlock;
  const int xchunk = b.trav_chunk > 0 ? b.trav_chunk : 4;
  const int round = kXcds * xchunk;
  const bool two = b.trav_budget > 0 && b.trec && b.hctl && b.rcap > 0;
  hipLaunchKernelGGL(k_traverse, dim3((nt + round - 1) / round * round), dim3(kTravBlock), 0, s, t,
                     b.nodes, b.amin, b.amax,
                     cap, chunk, counts, b.warn, b.ncloth ? b.slot_cloth : nullptr,
                     b.ncloth ? b.croot : nullptr, b.ccap, offsets, xchunk,
                     two ? b.trav_budget : 0, two ? b.trec : nullptr, b.rcap, b.hctl, b.spill);
  if (two) {  // one of the two finishes the records, the other returns at once
    hipLaunchKernelGGL(k_traverse_resume, dim3(std::min(b.rcap, kResumeBlocks)), dim3(64), 0, s,
                       t, b.nodes, b.amin, b.amax, chunk, cap, counts, b.warn, offsets, b.trec,
                       b.rcap, b.hctl);
    if (b.rcap > kLaneModeMin)
      hipLaunchKernelGGL(k_traverse_lanes,
                         dim3(std::min((b.rcap + kTravBlock - 1) / kTravBlock, 8192)),
                         dim3(kTravBlock), 0, s, t, b.nodes, b.amin, b.amax, chunk, cap, counts,
                         b.warn, offsets, b.trec, b.rcap, b.hctl, b.spill);
  }
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_pairs(const LbvhBufs& b, int cap, int2* chunk, int* counts, int* offsets, int2* pairs,
                 int* d_npairs, void* scan_temp, hipStream_t s) {
  const int t = b.t;
  if (t < 2) return fill_i32(d_npairs, 0, 1, s);
  // offsets[0, nt): pair count of each tile of kTravBlock leaves; offsets[nt, 2nt): the
  // tiles' exclusive prefixes (2 * nt <= t)
  const int nt = (t + kTravBlock - 1) / kTravBlock;
  int rc = launch_traverse(b, cap, chunk, counts, offsets, s);
  if (rc) return rc;
  rc = scan_tiles(nullptr, 1, nt, offsets, offsets + nt, 0, (int64_t)t * cap, 0, d_npairs,
                  nullptr, scan_temp, s);
  if (rc) return rc;
  const bool two = b.trav_budget > 0 && b.trec && b.hctl && b.rcap > 0;
  hipLaunchKernelGGL(k_compact_pairs, dim3(nt), dim3(kBlock), 0, s, t, cap, chunk, counts,
                     offsets + nt, pairs, two ? b.hctl : nullptr);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_traverse_only(const LbvhBufs& b, int cap, int2* chunk, int* counts, int* offsets,
                         hipStream_t s) {
  if (b.t < 2) return CM_OK;
  const int rc = launch_traverse(b, cap, chunk, counts, offsets, s);
  if (rc || !(b.trav_budget > 0 && b.trec && b.hctl && b.rcap > 0)) return rc;
  hipLaunchKernelGGL(k_clear_records, dim3(1), dim3(64), 0, s, b.hctl);
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

// Build/refit scratch carved from one buffer; with b == nullptr only the size.
size_t carve_build(int t, void* d_temp, LbvhBufs* b) {
  const size_t m = t > 0 ? 2 * (size_t)t - 1 : 0;
  const size_t tt = (size_t)(t > 0 ? t : 1);
  char* tp = (char*)d_temp;
  size_t off = 0;
  auto take = [&](auto** dst, size_t bytes) {
    if (b) *dst = (std::remove_reference_t<decltype(**dst)>*)(tp + off);
    off += align_up(bytes);
  };
  LbvhBufs dummy;
  LbvhBufs* o = b ? b : &dummy;
  take(&o->claim_cnt, sizeof(int) * m);
  take(&o->claim_hi, sizeof(int) * m);
  take(&o->state, sizeof(int) * tt);
  take(&o->top_list, sizeof(int) * tt);
  take(&o->top_index, sizeof(int) * tt);
  take(&o->ntop, sizeof(int));
  take(&o->ranges, sizeof(int2) * tt);
  take(&o->nodes, 4 * sizeof(float4) * tt);
  take(&o->tmin, sizeof(float4) * tt);
  take(&o->tmax, sizeof(float4) * tt);
  take(&o->trm, sizeof(int) * tt);
  take(&o->tdone, sizeof(int) * tt);
  return off;
}

size_t lbvh_build_temp_bytes(int t) { return carve_build(t, nullptr, nullptr); }
size_t lbvh_pairs_temp_bytes(int t, int cap) {
  return align_up(sizeof(int2) * (size_t)t * cap) + 2 * align_up(sizeof(int) * (size_t)t) +
         align_up(4 * sizeof(float4) * (size_t)t) + align_up(sizeof(int) * kSpill * (size_t)t) +
         scan_temp_bytes(t);
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
  if (norm) {
    CM_HIP(hipMallocAsync((void**)&cen, sizeof(float4) * (size_t)t, s));
    CM_HIP(hipMallocAsync((void**)&box, sizeof(float4) * (2 * (size_t)nblocks(t) + 2), s));
  }
  int rc = launch_tri_update(n, t, (const float4*)d_x, (const float4*)d_p, (const int4*)d_tri,
                             thickness, morton_bits, morton_normalize, (float4*)d_aabb_min,
                             (float4*)d_aabb_max, d_keys, nullptr, cen, box, s);
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
  int rc = radix_sort_clear(d_temp, t, (hipStream_t)stream);
  if (rc) return rc;
  return radix_sort(d_keys, d_vals, t, key_bits, d_temp, (hipStream_t)stream);
}

extern "C" size_t cm_lbvh_build_temp_bytes(int32_t t) { return lbvh_build_temp_bytes(t); }


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
  carve_build(t, d_temp, &b);
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
  carve_build(t, d_temp, &b);
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
  b.spill = (int*)tp;
  tp += align_up(sizeof(int) * kSpill * (size_t)t);
  int rc = launch_pack_nodes(b, (hipStream_t)stream);
  if (rc) return rc;
  rc = scan_temp_clear(tp, t, (hipStream_t)stream);
  if (rc) return rc;
  return launch_pairs(b, cap, chunk, counts, offsets, (int2*)d_pairs, d_n_pairs, tp,
                      (hipStream_t)stream);
}
