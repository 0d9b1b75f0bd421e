// Narrow phase of the substep: candidate expansion, feature swept-AABB filter,
// continuous collision tests and ordered compaction, plus the per-vertex record
// lists the solver gathers contact corrections from.
//
//  k_cand_count / k_cand_write <- DeformableUtils::CreateTriangleTests
//        (LicentaV2/LicentaV2/Collision/CudaCollisionDetection.cu:6-189) fused with
//        CudaPrimitiveTests::FeatureAABBTests (Collision/CudaPrimitiveTests.cu:427-490);
//        vertex and edge swept AABBs are recomputed from p and x on the fly
//        (same values as Collision/FeatureList.cu:134-162), which removes the
//        reference's N x N edge map.
//  k_ccd_* <- CudaPrimitiveTests::CCDTriangleTests (Collision/CudaPrimitiveTests.cu:109-238)
//        with Physics::FindCoplanarityTimes (Physics/CudaCubicSolver.cu:5-153)
//  compaction keeps candidate order, like CubWrap::SelectFlagged (Core/CubWrappers.cuh:24-50)
//  dynamic CSR: records (contact, role) grouped by vertex in record order, the order
//        the reference's stable sort + run-length encode gives
//        (Physics/CudaPBD.cu:669-725, 1448-1475).
#include "cm_internal.hpp"

namespace cm {

namespace {

__device__ __forceinline__ int bit(int f, int b) { return (f >> b) & 1; }

// Items are processed in tiles, one tile of 64 * R consecutive items per wave: lane
// l takes items base + l + 64 * j, j < R, so every load is coalesced and a tile needs
// no workgroup barrier. The pass over all candidates uses R = 8 (fewer tile sums to
// scan); the passes over pairs and over the short list of near candidates use R = 4
// (more waves share the work).
constexpr int kBig = 8, kSmall = 4;
constexpr int kWavesPerBlock = kBlock / 64;

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Offsets of a tile's items within the tile, in item order: v[j] is the (packed)
// count of item base + lane + 64 * j; off[j] receives the sum of the counts of all
// earlier items of the tile.
template <int R>
__device__ __forceinline__ void tile_offsets(const int (&v)[R], int (&off)[R]) {
  const int lane = threadIdx.x & 63;
  int inc[R];
#pragma unroll
  for (int j = 0; j < R; ++j) inc[j] = v[j];
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int u = __shfl_up(inc[j], o, 64);
      if (lane >= o) inc[j] += u;
    }
  }
  int carry = 0;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    off[j] = carry + inc[j] - v[j];
    carry += __shfl(inc[j], 63, 64);
  }
}

__device__ __forceinline__ void vert_aabb(const float4* __restrict__ p,
                                          const float4* __restrict__ x, int i, float th, f3& mn,
                                          f3& mx) {
  const f3 a = xyz(p[i]), b = xyz(x[i]);
  mn = subs(vmin(a, b), th);
  mx = adds(vmax(a, b), th);
}

// EE slot table: (edge of tri1, edge of tri2) as vertex index pairs into {v1,v2,v3}
// and the ownership bits that gate them (Collision/CudaCollisionDetection.cu:139-187).
__constant__ int kEE1[9][2] = {{0, 1}, {0, 2}, {1, 2}, {0, 1}, {0, 2}, {1, 2}, {0, 1}, {0, 2}, {1, 2}};
__constant__ int kEE2[9][2] = {{0, 1}, {0, 1}, {0, 1}, {0, 2}, {0, 2}, {0, 2}, {1, 2}, {1, 2}, {1, 2}};
__constant__ int kEEB1[9] = {2, 0, 1, 2, 0, 1, 2, 0, 1};
__constant__ int kEEB2[9] = {2, 2, 2, 0, 0, 0, 1, 1, 1};

// Feature mask of one pair (bits 0-5: VF slots, 6-14: EE slots).
__device__ __forceinline__ unsigned pair_mask(int2 pr, const int4* __restrict__ tri,
                                              const float4* __restrict__ x,
                                              const float4* __restrict__ p, float th, int ee_on) {
  const int4 t1 = tri[pr.x], t2 = tri[pr.y];
  const int v1[3] = {t1.x, t1.y, t1.z}, v2[3] = {t2.x, t2.y, t2.z};
  f3 mn1[3], mx1[3], mn2[3], mx2[3];
  for (int a = 0; a < 3; ++a) {
    vert_aabb(p, x, v1[a], th, mn1[a], mx1[a]);
    vert_aabb(p, x, v2[a], th, mn2[a], mx2[a]);
  }
  unsigned m = 0;
  // VF: vertices of tri1 against tri2, then vertices of tri2 against tri1
  const f3 tmn2 = min3(mn2[0], mn2[1], mn2[2]), tmx2 = max3(mx2[0], mx2[1], mx2[2]);
  const f3 tmn1 = min3(mn1[0], mn1[1], mn1[2]), tmx1 = max3(mx1[0], mx1[1], mx1[2]);
  for (int s = 0; s < 3; ++s) {
    if (bit(t1.w, 5 - s) && overlap(mn1[s], mx1[s], tmn2, tmx2)) m |= 1u << s;
    if (bit(t2.w, 5 - s) && overlap(mn2[s], mx2[s], tmn1, tmx1)) m |= 1u << (3 + s);
  }
  if (ee_on) {
    for (int s = 0; s < 9; ++s) {
      if (!(bit(t1.w, kEEB1[s]) & bit(t2.w, kEEB2[s]))) continue;
      const int a0 = kEE1[s][0], a1 = kEE1[s][1], b0 = kEE2[s][0], b1 = kEE2[s][1];
      const f3 amn = vmin(mn1[a0], mn1[a1]), amx = vmax(mx1[a0], mx1[a1]);
      const f3 bmn = vmin(mn2[b0], mn2[b1]), bmx = vmax(mx2[b0], mx2[b1]);
      if (overlap(amn, amx, bmn, bmx)) m |= 1u << (6 + s);
    }
  }
  return m;
}

// VF and EE candidate counts of a mask, packed low / high
__device__ __forceinline__ int mask_counts(unsigned m) {
  return __popc(m & 63u) | (__popc(m >> 6) << 16);
}

// pass 1: feature mask per pair, candidate counts per tile of pairs
__global__ void __launch_bounds__(kBlock) k_cand_count(
    const int2* __restrict__ pairs, const int* __restrict__ d_npairs, int64_t pair_cap,
    const int4* __restrict__ tri, const float4* __restrict__ x, const float4* __restrict__ p,
    float th, int ee_on, uint16_t* __restrict__ mask, int* __restrict__ tile_sum) {
  const int np = (int)min((int64_t)*d_npairs, pair_cap);
  const int ntiles = (np + (64 * kSmall) - 1) / (64 * kSmall);
  const int lane = threadIdx.x & 63;
  for (int tile = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); tile < ntiles;
       tile += gridDim.x * kWavesPerBlock) {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kSmall; ++j) {
      const int k = tile * (64 * kSmall) + j * 64 + lane;
      if (k < np) {
        const unsigned m = pair_mask(pairs[k], tri, x, p, th, ee_on);
        mask[k] = (uint16_t)m;
        c += mask_counts(m);
      }
    }
    c = wave_sum(c);
    if (lane == 0) tile_sum[tile] = c;
  }
}

// pass 2 (after the scan of the tile counts): candidates at their ordered positions
__global__ void __launch_bounds__(kBlock) k_cand_write(
    const int2* __restrict__ pairs, const int* __restrict__ d_npairs, int64_t pair_cap,
    const int4* __restrict__ tri, const uint16_t* __restrict__ mask,
    const int* __restrict__ tile_pre, int4* __restrict__ cand_vf, int4* __restrict__ cand_ee,
    int64_t vf_cap, int64_t ee_cap, uint32_t* __restrict__ warn) {
  const int np = (int)min((int64_t)*d_npairs, pair_cap);
  const int ntiles = (np + (64 * kSmall) - 1) / (64 * kSmall);
  const int lane = threadIdx.x & 63;
  bool over = false;
  for (int tile = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); tile < ntiles;
       tile += gridDim.x * kWavesPerBlock) {
    unsigned m[kSmall];
    int c[kSmall], off[kSmall];
#pragma unroll
    for (int j = 0; j < kSmall; ++j) {
      const int k = tile * (64 * kSmall) + j * 64 + lane;
      m[j] = k < np ? mask[k] : 0u;
      c[j] = mask_counts(m[j]);
    }
    tile_offsets<kSmall>(c, off);
    const int pvf = tile_pre[2 * tile], pee = tile_pre[2 * tile + 1];
#pragma unroll
    for (int j = 0; j < kSmall; ++j) {
      if (!m[j]) continue;
      const int k = tile * (64 * kSmall) + j * 64 + lane;
      const int2 pr = pairs[k];
      const int4 t1 = tri[pr.x], t2 = tri[pr.y];
      const int v1[3] = {t1.x, t1.y, t1.z}, v2[3] = {t2.x, t2.y, t2.z};
      int64_t o = (int64_t)pvf + (off[j] & 0xffff);
      for (int s = 0; s < 6; ++s) {
        if (!(m[j] & (1u << s))) continue;
        const int4 q = s < 3 ? make_int4(v1[s], t2.x, t2.y, t2.z)
                             : make_int4(v2[s - 3], t1.x, t1.y, t1.z);
        if (o < vf_cap) cand_vf[o] = q; else over = true;
        ++o;
      }
      o = (int64_t)pee + (off[j] >> 16);
      for (int s = 0; s < 9; ++s) {
        if (!(m[j] & (1u << (6 + s)))) continue;
        const int4 q = make_int4(v1[kEE1[s][0]], v1[kEE1[s][1]], v2[kEE2[s][0]], v2[kEE2[s][1]]);
        if (o < ee_cap) cand_ee[o] = q; else over = true;
        ++o;
      }
    }
  }
  if (over) atomicOr(warn, (uint32_t)CM_WARN_CANDIDATE_CAP);
}

// ---- CCD (FP32, same arithmetic as the oracle) -------------------------------
constexpr float kCubicEps = 1.f / 1000000;  // Physics/CudaCubicSolver.cuh:12
constexpr float kUtilEps = 0.00000001f;     // Core/CudaUtils.cu:4

__device__ __forceinline__ float triple_scalar(f3 x21, f3 x31, f3 x41, f3 v21, f3 v31, f3 v41,
                                               float t) {
  return dot(cross(add(x21, mul(v21, t)), add(x31, mul(v31, t))), add(x41, mul(v41, t)));
}

// The coplanarity cubic of Physics/CudaCubicSolver.cu:5-153: relative positions and
// velocities of vertices 2-4 against vertex 1, and the critical points in [0, 1]
// that split [0, 1] into cc + 1 intervals (ends 0, [cp1], [cp2], 1).
struct Cubic {
  f3 x21, x31, x41, v21, v31, v41;
  float cp1, cp2;
  int cc;
};

__device__ __forceinline__ void cubic_setup(f3 x1, f3 x2, f3 x3, f3 x4, f3 v1, f3 v2, f3 v3,
                                            f3 v4, Cubic& q) {
  const f3 x21 = sub(x2, x1), x31 = sub(x3, x1), x41 = sub(x4, x1);
  const f3 v21 = sub(v2, v1), v31 = sub(v3, v1), v41 = sub(v4, v1);
  const f3 v21v31 = cross(v21, v31), x21v31 = cross(x21, v31), v21x31 = cross(v21, x31),
           x21x31 = cross(x21, x31);
  const float v1v2v3 = dot(v21v31, v41), x1v2v3 = dot(x21v31, v41), v1x2v3 = dot(v21x31, v41),
              v1v2x3 = dot(v21v31, x41), x1x2v3 = dot(x21x31, v41), x1v2x3 = dot(x21v31, x41),
              v1x2x3 = dot(v21x31, x41);
  const float a = v1v2v3;
  const float b = x1v2v3 + v1x2v3 + v1v2x3;
  const float c = x1x2v3 + x1v2x3 + v1x2x3;
  const float ad = 3 * a, bd = 2 * b;
  const float dd = 4 * (b * b - ad * c);
  const float sq = sqrtf(dd);
  const float da = 2 * ad;
  const float t1 = ((-bd + sq) / da);
  const float t2 = ((-bd - sq) / da);
  float cp1 = 0.f, cp2 = 0.f;
  int cc = 0;
  bool have1 = false;
  if (dd >= kCubicEps) {
    if (between(t1, 0.f, 1.f)) { cp1 = t1; have1 = true; ++cc; }
    if (between(t2, 0.f, 1.f)) { cp2 = t2; ++cc; }
  } else if (fabsf(dd) < kCubicEps) {
    if (between(t1, 0.f, 1.f)) { cp1 = t1; have1 = true; ++cc; }
  }
  if (cc == 1 && !have1) cp1 = cp2;
  q.x21 = x21; q.x31 = x31; q.x41 = x41;
  q.v21 = v21; q.v31 = v31; q.v41 = v41;
  q.cp1 = cp1; q.cp2 = cp2; q.cc = cc;
}

// end of interval i (its start is the end of interval i - 1, or 0)
__device__ __forceinline__ float cubic_end(float cp1, float cp2, int cc, int i) {
  return i == cc ? 1.f : (i == 0 ? cp1 : cp2);
}

__device__ __forceinline__ float cubic_f(const Cubic& q, float t) {
  return triple_scalar(q.x21, q.x31, q.x41, q.v21, q.v31, q.v41, t);
}

struct Hit {
  float w[4];
  float t;
  f3 n;
};

__device__ __forceinline__ bool test_vf(f3 x4, f3 x1, f3 x2, f3 x3, Hit& h, float th) {
  const f3 x13 = sub(x1, x3), x23 = sub(x2, x3), x43 = sub(x4, x3);
  const f3 crs = cross(sub(x2, x1), sub(x3, x1));
  const f3 n = normalize(crs);
  if (fabsf(dot(x43, n)) <= th) {
    const float dtmp = dot(x13, x23);
    const float m11 = dot(x13, x13), m12 = dtmp, m21 = dtmp, m22 = dot(x23, x23);
    const float vr1 = dot(x13, x43), vr2 = dot(x23, x43);
    const float w2 = (m11 * vr2 - m21 * vr1) / (m11 * m22 - m21 * m12);
    const float w1 = (vr1 - m12 * w2) / m11;
    const float w3 = 1.f - w1 - w2;
    const float delta = th / (len(x13) + len(x23) + len(sub(x1, x2)));
    const float lo = 0.f, hi = 1 + delta;
    if (between(w1, lo, hi) && between(w2, lo, hi) && between(w3, lo, hi)) {
      h.w[0] = 1.f;
      h.w[1] = -w1;
      h.w[2] = -w2;
      h.w[3] = -w3;
      h.n = n;
      h.t = 0.f;
      return true;
    }
  }
  return false;
}

__device__ f3 project_on_line(f3 p, f3 a, f3 b) {
  const f3 ab = sub(b, a), ap = sub(p, a);
  return add(a, mul(ab, dot(ap, ab) / dot(ab, ab)));
}
__device__ f3 clamp_on_line(f3 p, f3 a, f3 b) {
  const float d1 = len(sub(p, a)), d2 = len(sub(p, b));
  if (fabsf(d1 + d2 - len(sub(a, b))) < kUtilEps) return p;
  return d1 < d2 ? a : b;
}

__device__ __forceinline__ bool test_ee(f3 x1, f3 x2, f3 x3, f3 x4, Hit& h, float th) {
  const f3 x21 = sub(x2, x1), x31 = sub(x3, x1), x43 = sub(x4, x3);
  const f3 crs = cross(x21, x43);
  if (len(crs) < 0.000000001f) {  // TestEdgeDegenerate (Collision/CudaPrimitiveTests.cu:408-425)
    const f3 m1 = add(x1, divs(x21, 2.f)), m2 = add(x3, divs(x43, 2.f));
    if (len(sub(m1, m2)) <= th) {
      h.w[0] = h.w[1] = h.w[2] = h.w[3] = 0.5f;
      h.n = normalize(sub(m1, m2));
      return true;
    }
    return false;
  }
  if (is_zero(crs)) return false;
  const float dtmp = dot(neg(x21), x43);
  const float m11 = dot(x21, x21), m12 = dtmp, m21 = dtmp, m22 = dot(x43, x43);
  const float vr1 = dot(x21, x31), vr2 = dot(neg(x43), x31);
  float b = (m11 * vr2 - m21 * vr1) / (m11 * m22 - m21 * m12);
  float a = (vr1 - m12 * b) / m11;
  if (between(a, 0.f, 1.f) && between(b, 0.f, 1.f)) {
    const f3 cp1 = add(x1, mul(x21, a));
    const f3 cp2 = add(x3, mul(x43, a));
    if (len(sub(cp1, cp2)) <= th) {
      h.w[0] = 1 - a;
      h.w[1] = a;
      h.w[2] = -(1 - b);
      h.w[3] = -b;
      const f3 dir = sub(add(x1, mul(x21, a)), add(x3, mul(x43, b)));
      h.n = normalize(is_zero(dir) ? crs : dir);
      return true;
    }
  } else {
    const float ac = fclamp(a, 0.f, 1.f), bc = fclamp(b, 0.f, 1.f);
    const float d1 = fabsf(a - ac) * len(x21);
    const float d2 = fabsf(b - bc) * len(x43);
    a = ac;
    b = bc;
    f3 cp1, cp2;
    if (d1 > d2) {
      cp1 = add(x1, mul(x21, a));
      cp2 = clamp_on_line(project_on_line(cp1, x3, x4), x3, x4);
    } else {
      cp2 = add(x3, mul(x43, b));
      cp1 = clamp_on_line(project_on_line(cp2, x1, x2), x1, x2);
    }
    if (len(sub(cp1, cp2)) <= th) {
      h.w[0] = 1 - a;
      h.w[1] = a;
      h.w[2] = -(1 - b);
      h.w[3] = -b;
      h.n = normalize(sub(cp1, cp2));
      h.t = 0.f;
      return true;
    }
  }
  return false;
}

// ---- CCD in passes ------------------------------------------------------------
// The reference tests each candidate (Collision/CudaPrimitiveTests.cu:128-238): it
// finds the coplanarity roots in [0, 1] in interval order (a root at an interval
// start when |f| < EPS there, then a secant root when f changes sign), advances all
// four positions cumulatively by root * dt, and runs the VF / EE proximity test
// after each root until one hits. Here:
//   k_ccd_plan    per candidate: cubic, f at the interval ends, plan word, the secant
//                 root of each interval with a sign change, then the roots in order
//                 with the same cumulative advance and only the box-gap rejection:
//                 a near flag per candidate, near counts per tile. Plan word,
//                 critical points and roots (interval i of candidate id in
//                 roots[3 * id + i]) are stored for near candidates only, the only
//                 ones the later passes read.
//   k_near_list   ids of the near candidates in order
//   k_ccd_test    per near candidate: the full tests; hit flag and hit counts
//   k_ccd_write   hits at their ordered positions in the VF / EE contact lists
// Plan word: bits 0-1 cc, bits 2-4 root at the start of interval i, bits 5-7 sign
// change on interval i, bit 8 nothing to test (degenerate candidate).
constexpr int kPlanDead = 1 << 8;
constexpr float kNoRoot = -15.f;  // Physics/CudaCubicSolver.cu:152

// Vertex ids of candidate c in the cubic's order (x1, x2, x3, x4): VF candidates are
// (point, triangle) and the cubic takes the triangle first; EE as stored.
__device__ __forceinline__ int4 cubic_ids(bool isvf, int4 c) {
  return isvf ? make_int4(c.y, c.z, c.w, c.x) : c;
}
__device__ __forceinline__ bool sane(bool isvf, int4 c) {
  return isvf ? !(c.x == c.y || c.x == c.z || c.x == c.w)                    // VFSanity
              : !(c.x == c.z || c.x == c.w || c.y == c.z || c.y == c.w);   // EESanity
}

__device__ __forceinline__ int4 cand_of(int id, int nvf, const int4* __restrict__ cand_vf,
                                        const int4* __restrict__ cand_ee, bool& isvf) {
  isvf = id < nvf;
  return isvf ? cand_vf[id] : cand_ee[id - nvf];
}

// Secant searches (Physics/CudaCubicSolver.cu:131-153, 20 steps) on the intervals
// with a sign change, from x0 = start, x1 = end, f0 = f(x0), f1 = f(x1). The
// reference evaluates f at both points every step; f(x0) of a step is f(x1) of the
// previous one, and the first step's two values are the ones the plan computed at
// the interval's ends, all from the same inputs, so they are carried over unchanged.
// A search converges when |xc - x1| < EPS (root xc) or gives up after 20 steps (the
// reference returns -15). A lane runs its searches one after the other in one loop,
// so the wave waits for the lane with the most steps in all, not for the slowest
// lane of every interval in turn.
__device__ __forceinline__ void secants(const Cubic& q, int pl, const float (&fe)[4],
                                        float (&r)[3]) {
  auto next = [&](int i) {  // first interval >= i with a sign change, or 3
    while (i <= q.cc && !(pl & (1 << (5 + i)))) ++i;
    return i <= q.cc ? i : 3;
  };
  auto start = [&](int i, float& x0, float& x1, float& f0, float& f1) {
    x0 = i == 0 ? 0.f : (i == 1 ? q.cp1 : q.cp2);
    x1 = cubic_end(q.cp1, q.cp2, q.cc, i);
    f0 = i == 0 ? fe[0] : (i == 1 ? fe[1] : fe[2]);
    f1 = i == 0 ? fe[1] : (i == 1 ? fe[2] : fe[3]);
  };
  int i = next(0), k = 0;
  float x0 = 0.f, x1 = 0.f, f0 = 0.f, f1 = 0.f;
  if (i < 3) start(i, x0, x1, f0, f1);
  while (i < 3) {
    const float xc = (x0 * f1 - x1 * f0) / (f1 - f0);
    const bool conv = fabsf(xc - x1) < kCubicEps;
    if (conv) {
      r[0] = i == 0 ? xc : r[0];
      r[1] = i == 1 ? xc : r[1];
      r[2] = i == 2 ? xc : r[2];
    }
    if (conv || ++k == 20) {
      i = next(i + 1);
      k = 0;
      if (i < 3) start(i, x0, x1, f0, f1);
      continue;
    }
    x0 = x1;
    f0 = f1;
    x1 = xc;
    f1 = cubic_f(q, x1);
  }
}

// Conservative rejections before the proximity tests. A VF hit puts the point within
// th of the triangle (plane distance <= th, projection inside); an EE hit puts two
// points of the segments (or their midpoints) within th. So a box gap larger than
// th plus a rounding margin means the full test returns false.
__device__ __forceinline__ float box_gap2(f3 amn, f3 amx, f3 bmn, f3 bmx) {
  const float gx = fmaxf(fmaxf(bmn.x - amx.x, amn.x - bmx.x), 0.f);
  const float gy = fmaxf(fmaxf(bmn.y - amx.y, amn.y - bmx.y), 0.f);
  const float gz = fmaxf(fmaxf(bmn.z - amx.z, amn.z - bmx.z), 0.f);
  return gx * gx + gy * gy + gz * gz;
}
__device__ __forceinline__ float reach2(float th) {
  const float r = th * 1.0001f + 2e-5f;
  return r * r;
}
__device__ __forceinline__ bool vf_far(f3 p, f3 a, f3 b, f3 c, float th) {
  return box_gap2(p, p, min3(a, b, c), max3(a, b, c)) > reach2(th);
}
__device__ __forceinline__ bool ee_far(f3 a, f3 b, f3 c, f3 d, float th) {
  return box_gap2(vmin(a, b), vmax(a, b), vmin(c, d), vmax(c, d)) > reach2(th);
}

// The reference's per-candidate test loop over the planned roots; true on a hit
// (contact data in h).
__device__ __forceinline__ bool ccd_tests(bool isvf, int4 c, int pl, float2 cp, int id,
                                          const float* __restrict__ roots,
                                          const float4* __restrict__ x,
                                          const float4* __restrict__ v, float th, float dt,
                                          int dcd, Hit& h) {
  h.t = -1.f;  // Physics/Structs.h:122-128
  h.w[0] = h.w[1] = h.w[2] = h.w[3] = -1.f;
  h.n = mk3(-1.f, -1.f, -1.f);
  if (pl & kPlanDead) return false;
  const int4 q4 = cubic_ids(isvf, c);
  const f3 o1 = xyz(x[q4.x]), o2 = xyz(x[q4.y]), o3 = xyz(x[q4.z]), o4 = xyz(x[q4.w]);
  if (dcd)  // Collision/CudaPrimitiveTests.cu:41-94: the tests at x; t as they leave it
    return isvf ? !vf_far(o4, o1, o2, o3, th) && test_vf(o4, o1, o2, o3, h, th)
                : !ee_far(o1, o2, o3, o4, th) && test_ee(o1, o2, o3, o4, h, th);
  const f3 u1 = xyz(v[q4.x]), u2 = xyz(v[q4.y]), u3 = xyz(v[q4.z]), u4 = xyz(v[q4.w]);
  f3 p1 = o1, p2 = o2, p3 = o3, p4 = o4;  // advanced cumulatively by each root
  // one root: advance (Collision/CudaPrimitiveTests.cu:164-171), then test
  auto at_root = [&](float root) {
    const float t = root * dt;
    if (isvf) {  // the reference advances x4 (the point) first; the order is immaterial
      p4 = add(p4, mul(u4, t));
      p1 = add(p1, mul(u1, t));
      p2 = add(p2, mul(u2, t));
      p3 = add(p3, mul(u3, t));
      if (vf_far(p4, p1, p2, p3, th) || !test_vf(p4, p1, p2, p3, h, th)) return false;
    } else {
      p1 = add(p1, mul(u1, t));
      p2 = add(p2, mul(u2, t));
      p3 = add(p3, mul(u3, t));
      p4 = add(p4, mul(u4, t));
      if (ee_far(p1, p2, p3, p4, th) || !test_ee(p1, p2, p3, p4, h, th)) return false;
    }
    h.t = t;
    return true;
  };
  const int cc = pl & 3;
  float ea = 0.f;
  for (int i = 0; i <= cc; ++i) {
    if ((pl & (1 << (2 + i))) && at_root(ea)) return true;
    if (pl & (1 << (5 + i))) {
      const float r = roots[3 * (int64_t)id + i];
      if (r != kNoRoot && at_root(r)) return true;
    }
    ea = cubic_end(cp.x, cp.y, cc, i);
  }
  return false;
}

// One candidate of k_ccd_plan: the plan of its roots and whether any root survives
// the box-gap rejection after the same cumulative advance as ccd_tests (false: the
// full tests cannot hit). Stores plan word, critical points and roots when near.
__device__ __forceinline__ bool plan_near(bool isvf, int4 c, int id, const float4* __restrict__ x,
                                          const float4* __restrict__ v, float th, float dt,
                                          int dcd, int* __restrict__ plan,
                                          float2* __restrict__ cps, float* __restrict__ roots) {
  if (!sane(isvf, c)) return false;
  const int4 q4 = cubic_ids(isvf, c);
  const f3 o1 = xyz(x[q4.x]), o2 = xyz(x[q4.y]), o3 = xyz(x[q4.z]), o4 = xyz(x[q4.w]);
  if (dcd) {  // one test at x, no cubic
    const bool nr = isvf ? !vf_far(o4, o1, o2, o3, th) : !ee_far(o1, o2, o3, o4, th);
    if (nr) {
      plan[id] = 0;
      cps[id] = make_float2(0.f, 0.f);
    }
    return nr;
  }
  const f3 u1 = xyz(v[q4.x]), u2 = xyz(v[q4.y]), u3 = xyz(v[q4.z]), u4 = xyz(v[q4.w]);
  Cubic q;
  cubic_setup(o1, o2, o3, o4, u1, u2, u3, u4, q);
  // f at the interval ends: fe[i] at the start of interval i, fe[i + 1] at its end
  int pl = q.cc;
  float fe[4];
  fe[0] = cubic_f(q, 0.f);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i > q.cc) break;
    fe[i + 1] = cubic_f(q, cubic_end(q.cp1, q.cp2, q.cc, i));
    if (fabsf(fe[i]) < kCubicEps) pl |= 1 << (2 + i);
    if (fe[i] * fe[i + 1] < 0.f) pl |= 1 << (5 + i);
  }
  // the secant root of each interval with a sign change, with the cubic in registers
  float r[3] = {kNoRoot, kNoRoot, kNoRoot};
  secants(q, pl, fe, r);
  f3 p1 = o1, p2 = o2, p3 = o3, p4 = o4;
  auto near_at = [&](float root) {
    const float t = root * dt;
    if (isvf) {
      p4 = add(p4, mul(u4, t));
      p1 = add(p1, mul(u1, t));
      p2 = add(p2, mul(u2, t));
      p3 = add(p3, mul(u3, t));
      return !vf_far(p4, p1, p2, p3, th);
    }
    p1 = add(p1, mul(u1, t));
    p2 = add(p2, mul(u2, t));
    p3 = add(p3, mul(u3, t));
    p4 = add(p4, mul(u4, t));
    return !ee_far(p1, p2, p3, p4, th);
  };
  bool nr = false;
  float ea = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    if (i > q.cc) break;
    if ((pl & (1 << (2 + i))) && near_at(ea)) { nr = true; break; }
    if ((pl & (1 << (5 + i))) && r[i] != kNoRoot && near_at(r[i])) { nr = true; break; }
    ea = cubic_end(q.cp1, q.cp2, q.cc, i);
  }
  if (nr) {
    plan[id] = pl;
    cps[id] = make_float2(q.cp1, q.cp2);
#pragma unroll
    for (int i = 0; i < 3; ++i)
      if (i <= q.cc && (pl & (1 << (5 + i)))) roots[3 * (int64_t)id + i] = r[i];
  }
  return nr;
}

// pass 3: near flag per candidate (VF then EE ids), near counts per tile of ids
__global__ void __launch_bounds__(kBlock) k_ccd_plan(
    const int4* __restrict__ cand_vf, const int4* __restrict__ cand_ee,
    const int* __restrict__ d_nvf, const int* __restrict__ d_nee, int64_t vf_cap,
    int64_t ee_cap, const float4* __restrict__ x, const float4* __restrict__ v, float th,
    float dt, int dcd, int* __restrict__ plan, float2* __restrict__ cps,
    float* __restrict__ roots, int* __restrict__ d_ntotal, uint8_t* __restrict__ near_flag,
    int* __restrict__ tile_sum) {
  const int nvf = (int)min((int64_t)*d_nvf, vf_cap), nee = (int)min((int64_t)*d_nee, ee_cap);
  const int total = nvf + nee;
  if (blockIdx.x == 0 && threadIdx.x == 0) *d_ntotal = total;
  const int ntiles = (total + (64 * kBig) - 1) / (64 * kBig);
  const int lane = threadIdx.x & 63;
  for (int tile = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); tile < ntiles;
       tile += gridDim.x * kWavesPerBlock) {
    int c = 0;
    for (int j = 0; j < kBig; ++j) {
      const int id = tile * (64 * kBig) + j * 64 + lane;
      if (id >= total) break;
      bool isvf;
      const int4 q = cand_of(id, nvf, cand_vf, cand_ee, isvf);
      const bool nr = plan_near(isvf, q, id, x, v, th, dt, dcd, plan, cps, roots);
      near_flag[id] = (uint8_t)nr;
      c += nr;
    }
    c = wave_sum(c);
    if (lane == 0) tile_sum[tile] = c;
  }
}

// pass 4: the near candidates' ids in order
__global__ void __launch_bounds__(kBlock) k_near_list(const int* __restrict__ d_ntotal,
                                                      const uint8_t* __restrict__ near_flag,
                                                      const int* __restrict__ tile_pre,
                                                      int* __restrict__ near_ids) {
  // lane l takes the kBig consecutive ids from tile * 64 * kBig + l * kBig (one 8-byte
  // load of their flags), so the ids leave in order: lane by lane, then within a lane
  static_assert(kBig == 8, "one 8-byte load of near flags per lane");
  const int total = *d_ntotal;
  const int ntiles = (total + (64 * kBig) - 1) / (64 * kBig);
  const int lane = threadIdx.x & 63;
  for (int tile = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); tile < ntiles;
       tile += gridDim.x * kWavesPerBlock) {
    const int b0 = tile * (64 * kBig) + lane * kBig;
    uint32_t lo = 0, hi = 0;
    if (b0 + kBig <= total) {
      const uint2 u = *reinterpret_cast<const uint2*>(near_flag + b0);
      lo = u.x;
      hi = u.y;
    } else {
#pragma unroll
      for (int j = 0; j < kBig; ++j) {
        const uint32_t f = b0 + j < total ? near_flag[b0 + j] : 0u;
        if (j < 4) lo |= f << (8 * j);
        else hi |= f << (8 * (j - 4));
      }
    }
    const int cnt = __popc(lo & 0x01010101u) + __popc(hi & 0x01010101u);
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(inc, o, 64);
      if (lane >= o) inc += u;
    }
    int at = tile_pre[tile] + inc - cnt;
#pragma unroll
    for (int j = 0; j < kBig; ++j) {
      const uint32_t f = ((j < 4 ? lo : hi) >> (8 * (j & 3))) & 1u;
      if (f) near_ids[at++] = b0 + j;
    }
  }
}

// pass 5: full tests of the near candidates; hit flag per near index, hit counts
// (VF low, EE high) per tile of near indices
__global__ void __launch_bounds__(kBlock) k_ccd_test(
    const int4* __restrict__ cand_vf, const int4* __restrict__ cand_ee,
    const int* __restrict__ d_nvf, int64_t vf_cap, const int* __restrict__ d_nnear,
    const int* __restrict__ near_ids, const int* __restrict__ plan,
    const float2* __restrict__ cps, const float* __restrict__ roots,
    const float4* __restrict__ x, const float4* __restrict__ v, float th, float dt, int dcd,
    uint8_t* __restrict__ hit_flag, int* __restrict__ tile_sum, float4* __restrict__ hit_data,
    int64_t hit_cap) {
  const int nvf = (int)min((int64_t)*d_nvf, vf_cap), nn = *d_nnear;
  const int ntiles = (nn + (64 * kSmall) - 1) / (64 * kSmall);
  const int lane = threadIdx.x & 63;
  for (int tile = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); tile < ntiles;
       tile += gridDim.x * kWavesPerBlock) {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kSmall; ++j) {
      const int k = tile * (64 * kSmall) + j * 64 + lane;
      if (k < nn) {
        const int id = near_ids[k];
        bool isvf;
        const int4 q = cand_of(id, nvf, cand_vf, cand_ee, isvf);
        Hit h;
        const bool hit = ccd_tests(isvf, q, plan[id], cps[id], id, roots, x, v, th, dt, dcd, h);
        hit_flag[k] = (uint8_t)hit;
        if (hit) c += isvf ? 1 : (1 << 16);
        if (hit && k < hit_cap) {  // kept for k_ccd_write
          hit_data[2 * (int64_t)k] = make_float4(h.w[0], h.w[1], h.w[2], h.w[3]);
          hit_data[2 * (int64_t)k + 1] = make_float4(h.t, h.n.x, h.n.y, h.n.z);
        }
      }
    }
    c = wave_sum(c);
    if (lane == 0) tile_sum[tile] = c;
  }
}

// pass 6 (after the scan of the tile counts): each hit is written at its ordered
// position among the VF or EE contacts, with the weights, time and normal k_ccd_test
// kept for it (recomputed for near indices at or past hit_cap).
__global__ void __launch_bounds__(kBlock) k_ccd_write(
    const int4* __restrict__ cand_vf, const int4* __restrict__ cand_ee,
    const int* __restrict__ d_nvf, int64_t vf_cap, const int* __restrict__ d_nnear,
    const int* __restrict__ near_ids, const int* __restrict__ plan,
    const float2* __restrict__ cps, const float* __restrict__ roots,
    const uint8_t* __restrict__ hit_flag, const int* __restrict__ tile_pre, int64_t vf_out_cap,
    int64_t ee_out_cap, const float4* __restrict__ x, const float4* __restrict__ v, float th,
    float dt, int dcd, int4* __restrict__ vf_v, float4* __restrict__ vf_w,
    float4* __restrict__ vf_tn, int4* __restrict__ ee_v, float4* __restrict__ ee_w,
    float4* __restrict__ ee_tn, uint32_t* __restrict__ warn,
    const float4* __restrict__ hit_data, int64_t hit_cap) {
  const int nvf = (int)min((int64_t)*d_nvf, vf_cap), nn = *d_nnear;
  const int ntiles = (nn + (64 * kSmall) - 1) / (64 * kSmall);
  const int lane = threadIdx.x & 63;
  bool over = false;
  for (int tile = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); tile < ntiles;
       tile += gridDim.x * kWavesPerBlock) {
    int id[kSmall], c[kSmall], off[kSmall];
#pragma unroll
    for (int j = 0; j < kSmall; ++j) {
      const int k = tile * (64 * kSmall) + j * 64 + lane;
      const bool hit = k < nn && hit_flag[k];
      id[j] = hit ? near_ids[k] : -1;
      c[j] = hit ? (id[j] < nvf ? 1 : (1 << 16)) : 0;
    }
    tile_offsets<kSmall>(c, off);
    const int pvf = tile_pre[2 * tile], pee = tile_pre[2 * tile + 1];
#pragma unroll
    for (int j = 0; j < kSmall; ++j) {
      if (id[j] < 0) continue;
      bool isvf;
      const int4 q = cand_of(id[j], nvf, cand_vf, cand_ee, isvf);
      const int64_t o = isvf ? (int64_t)pvf + (off[j] & 0xffff) : (int64_t)pee + (off[j] >> 16);
      if (o >= (isvf ? vf_out_cap : ee_out_cap)) {
        over = true;
        continue;
      }
      const int64_t k = (int64_t)tile * (64 * kSmall) + j * 64 + lane;
      float4 w4, tn;
      if (k < hit_cap) {
        w4 = hit_data[2 * k];
        tn = hit_data[2 * k + 1];
      } else {
        Hit h;
        ccd_tests(isvf, q, plan[id[j]], cps[id[j]], id[j], roots, x, v, th, dt, dcd, h);
        w4 = make_float4(h.w[0], h.w[1], h.w[2], h.w[3]);
        tn = make_float4(h.t, h.n.x, h.n.y, h.n.z);
      }
      if (isvf) { vf_v[o] = q; vf_w[o] = w4; vf_tn[o] = tn; }
      else { ee_v[o] = q; ee_w[o] = w4; ee_tn[o] = tn; }
    }
  }
  if (over) atomicOr(warn, (uint32_t)CM_WARN_CANDIDATE_CAP);
}

__global__ void __launch_bounds__(kBlock) k_dyn_count(const int4* __restrict__ vfv,
                                                      const int4* __restrict__ eev,
                                                      const int* __restrict__ d_nvf,
                                                      const int* __restrict__ d_nee,
                                                      int* __restrict__ dcount, int n) {
  const int nvf = *d_nvf, nee = *d_nee;
  if (blockIdx.x == 0 && threadIdx.x == 0) dcount[n] = 0;  // long-list counter of k_dyn_sort
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < nvf + nee; c += gridDim.x * kBlock) {
    const int4 q = c < nvf ? vfv[c] : eev[c - nvf];
    atomicAdd(&dcount[q.x], 1);
    atomicAdd(&dcount[q.y], 1);
    atomicAdd(&dcount[q.z], 1);
    atomicAdd(&dcount[q.w], 1);
  }
}

__global__ void __launch_bounds__(kBlock) k_dyn_fill(const int4* __restrict__ vfv,
                                                     const int4* __restrict__ eev,
                                                     const int* __restrict__ d_nvf,
                                                     const int* __restrict__ d_nee,
                                                     const int* __restrict__ prefix,
                                                     int* __restrict__ dcount,
                                                     uint32_t* __restrict__ drec) {
  // each vertex's count runs down to zero as its records are placed (in any order:
  // k_dyn_sort orders them), which leaves the counts cleared for the next substep
  const int nvf = *d_nvf, nee = *d_nee;
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < nvf + nee; c += gridDim.x * kBlock) {
    const int4 q = c < nvf ? vfv[c] : eev[c - nvf];
    const int vv[4] = {q.x, q.y, q.z, q.w};
    for (int r = 0; r < 4; ++r)
      drec[prefix[vv[r]] + atomicSub(&dcount[vv[r]], 1) - 1] = 4u * (uint32_t)c + r;
  }
}

__device__ __forceinline__ void put_payload(uint32_t r, int k, int nvf,
                                            const int4* __restrict__ vf_v,
                                            const float4* __restrict__ vf_w,
                                            const int4* __restrict__ ee_v,
                                            const float4* __restrict__ ee_w,
                                            RecPayload* __restrict__ out) {
  const int ci = (int)(r >> 2), role = (int)(r & 3u);
  const bool isvf = ci < nvf;
  int4 ids = isvf ? vf_v[ci] : ee_v[ci - nvf];
  const float4 w = isvf ? vf_w[ci] : ee_w[ci - nvf];
  ids.x |= (role << 29) | (isvf ? (int)0x80000000u : 0);
  RecPayload rp;
  rp.ids = ids;
  rp.w = w;
  out[k] = rp;
}

constexpr int kShortList = 32;   // longer record lists are ordered by a workgroup
constexpr int kLongLds = 4096;   // records a workgroup ranks in LDS

// per vertex: order its records by record index. Short lists: insertion sort by
// the vertex's thread in its own LDS column (one load per record, then LDS only);
// longer ones go to a work list for k_dyn_sort_long.
__global__ void __launch_bounds__(kBlock) k_dyn_sort(int n, const int* __restrict__ prefix,
                                                     uint32_t* __restrict__ drec,
                                                     int* __restrict__ nlong,
                                                     int* __restrict__ long_list,
                                                     const int* __restrict__ d_nvf,
                                                     const int4* __restrict__ vf_v,
                                                     const float4* __restrict__ vf_w,
                                                     const int4* __restrict__ ee_v,
                                                     const float4* __restrict__ ee_w,
                                                     RecPayload* __restrict__ rec) {
  __shared__ uint32_t sk[kShortList][kBlock];  // column per thread: no bank conflicts
  const int tid = threadIdx.x;
  const int v = blockIdx.x * kBlock + tid;
  if (v >= n) return;
  const int b = prefix[v], e = prefix[v + 1];
  const int k = e - b;
  if (k > kShortList) {
    long_list[atomicAdd(nlong, 1)] = v;
    return;
  }
  for (int j = 0; j < k; ++j) sk[j][tid] = drec[b + j];
  for (int i = 1; i < k; ++i) {
    const uint32_t key = sk[i][tid];
    int j = i - 1;
    while (j >= 0 && sk[j][tid] > key) {
      sk[j + 1][tid] = sk[j][tid];
      --j;
    }
    sk[j + 1][tid] = key;
  }
  const int nvf = *d_nvf;
  for (int j = 0; j < k; ++j) {
    const uint32_t r = sk[j][tid];
    drec[b + j] = r;
    put_payload(r, b + j, nvf, vf_v, vf_w, ee_v, ee_w, rec);
  }
}

// one workgroup per long list: record indices are unique, so the rank of a record
// is the number of smaller records in its list
__global__ void __launch_bounds__(kBlock) k_dyn_sort_long(const int* __restrict__ prefix,
                                                          const int* __restrict__ nlong,
                                                          const int* __restrict__ long_list,
                                                          uint32_t* __restrict__ drec,
                                                          uint32_t* __restrict__ tmp,
                                                          const int* __restrict__ d_nvf,
                                                          const int4* __restrict__ vf_v,
                                                          const float4* __restrict__ vf_w,
                                                          const int4* __restrict__ ee_v,
                                                          const float4* __restrict__ ee_w,
                                                          RecPayload* __restrict__ rec) {
  __shared__ uint32_t sl[kLongLds];
  const int nl = *nlong;
  const int nvf = *d_nvf;
  for (int w = blockIdx.x; w < nl; w += gridDim.x) {
    const int v = long_list[w];
    const int b = prefix[v], k = prefix[v + 1] - b;
    const uint32_t* src = drec + b;
    if (k <= kLongLds) {
      for (int j = threadIdx.x; j < k; j += kBlock) sl[j] = src[j];
      __syncthreads();
      for (int j = threadIdx.x; j < k; j += kBlock) {
        const uint32_t key = sl[j];
        int r = 0;
        for (int m = 0; m < k; ++m) r += sl[m] < key;
        drec[b + r] = key;
        put_payload(key, b + r, nvf, vf_v, vf_w, ee_v, ee_w, rec);
      }
    } else {
      for (int j = threadIdx.x; j < k; j += kBlock) {
        const uint32_t key = src[j];
        int r = 0;
        for (int m = 0; m < k; ++m) r += src[m] < key;
        tmp[b + r] = key;
      }
      __syncthreads();
      for (int j = threadIdx.x; j < k; j += kBlock) {
        drec[b + j] = tmp[b + j];
        put_payload(tmp[b + j], b + j, nvf, vf_v, vf_w, ee_v, ee_w, rec);
      }
    }
    __syncthreads();
  }
}

// ---- sorted CSR: the records keyed by vertex id through the stable radix sort --------
// record r = 4c + role of contact c (VF then EE) is keyed by that vertex; a stable
// sort leaves each vertex's records in increasing r, the order k_dyn_sort gives
__global__ void __launch_bounds__(kBlock) k_rec_keys(const int4* __restrict__ vfv,
                                                     const int4* __restrict__ eev,
                                                     const int* __restrict__ d_nvf,
                                                     const int* __restrict__ d_nee,
                                                     uint4* __restrict__ keys,
                                                     uint4* __restrict__ vals,
                                                     int* __restrict__ d_nrec) {
  const int nvf = *d_nvf, nee = *d_nee;
  if (blockIdx.x == 0 && threadIdx.x == 0) *d_nrec = 4 * (nvf + nee);
  for (int c = blockIdx.x * kBlock + threadIdx.x; c < nvf + nee; c += gridDim.x * kBlock) {
    const int4 q = c < nvf ? vfv[c] : eev[c - nvf];
    keys[c] = make_uint4((uint32_t)q.x, (uint32_t)q.y, (uint32_t)q.z, (uint32_t)q.w);
    const uint32_t r = 4u * (uint32_t)c;
    vals[c] = make_uint4(r, r + 1u, r + 2u, r + 3u);
  }
}

// From the sorted keys and values: dprefix[v] = the number of keys below v (binary
// search, v in [0, n]), and the records and their payloads in sorted order. vals
// may be drec itself (an even number of sort passes).
__global__ void __launch_bounds__(kBlock) k_rec_place(int n, const uint32_t* __restrict__ keys,
                                                      const uint32_t* vals,
                                                      const int* __restrict__ d_nrec,
                                                      const int* __restrict__ d_nvf,
                                                      const int4* __restrict__ vf_v,
                                                      const float4* __restrict__ vf_w,
                                                      const int4* __restrict__ ee_v,
                                                      const float4* __restrict__ ee_w,
                                                      int* __restrict__ dprefix, uint32_t* drec,
                                                      RecPayload* __restrict__ rec) {
  const int R = *d_nrec;
  const int nvf = *d_nvf;
  const int m = max(R, n + 1);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < m; i += gridDim.x * kBlock) {
    if (i <= n) {
      int lo = 0, hi = R;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < (uint32_t)i) lo = mid + 1;
        else hi = mid;
      }
      dprefix[i] = lo;
    }
    if (i < R) {
      const uint32_t r = vals[i];
      drec[i] = r;
      put_payload(r, i, nvf, vf_v, vf_w, ee_v, ee_w, rec);
    }
  }
}

__global__ void k_copy_i32(const int* __restrict__ s, int* __restrict__ d, int n) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) d[i] = s[i];
}


}  // namespace

static int grid_for(int64_t cap) {
  int64_t g = (cap + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

// workgroups for tiles of 64 * R items over up to `items` items
static int tile_grid(int64_t items, int R) {
  const int64_t g = (items + kBlock * R - 1) / (kBlock * R);
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 2048));
}

int launch_narrow(const NarrowBufs& nb, const int4* tri, const float4* x, const float4* p,
                  const float4* v, float th, float dt, int ee_on, int dcd, hipStream_t s) {
  // candidates: masks and tile counts, tile prefixes (and the clamped totals), writes
  const int ptiles = (int)((nb.pair_cap + 64 * kSmall - 1) / (64 * kSmall));
  hipLaunchKernelGGL(k_cand_count, dim3(tile_grid(nb.pair_cap, kSmall)), dim3(kBlock), 0, s, nb.pairs,
                     nb.d_npairs, nb.pair_cap, tri, x, p, th, ee_on, nb.mask, nb.tile_sum);
  CM_LAUNCH_CHECK();
  int rc = scan_tiles(nb.d_npairs, 64 * kSmall, ptiles, nb.tile_sum, nb.tile_pre, 1, nb.vf_cap,
                      nb.ee_cap, nb.d_ncand_vf, nb.d_ncand_ee, nb.scan_temp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cand_write, dim3(tile_grid(nb.pair_cap, kSmall)), dim3(kBlock), 0, s, nb.pairs,
                     nb.d_npairs, nb.pair_cap, tri, nb.mask, nb.tile_pre, nb.cand_vf, nb.cand_ee,
                     nb.vf_cap, nb.ee_cap, nb.warn);
  // CCD: plan with secant roots and near flags, near list, tests with tile hit
  // counts, tile prefixes, writes
  const int64_t ccap = nb.vf_cap + nb.ee_cap;
  hipLaunchKernelGGL(k_ccd_plan, dim3(tile_grid(ccap, kBig)), dim3(kBlock), 0, s, nb.cand_vf,
                     nb.cand_ee, nb.d_ncand_vf, nb.d_ncand_ee, nb.vf_cap, nb.ee_cap, x, v, th, dt,
                     dcd, nb.cplan, nb.ccp, nb.croot, nb.d_ntotal, nb.hit_flag, nb.tile_sum);
  CM_LAUNCH_CHECK();
  const int ctiles = (int)((ccap + 64 * kBig - 1) / (64 * kBig));
  rc = scan_tiles(nb.d_ntotal, 64 * kBig, ctiles, nb.tile_sum, nb.tile_pre, 0, ccap, 0,
                  nb.d_nnear, nullptr, nb.scan_temp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_near_list, dim3(tile_grid(ccap, kBig)), dim3(kBlock), 0, s, nb.d_ntotal,
                     nb.hit_flag, nb.tile_pre, nb.near_ids);
  // the record payloads (32 B each) are free until k_rec_place refills them: they keep
  // each hit's weights, time and normal from k_ccd_test for k_ccd_write
  const int64_t hit_cap = 4 * (nb.vf_out_cap + nb.ee_out_cap);
  hipLaunchKernelGGL(k_ccd_test, dim3(tile_grid(ccap, kSmall)), dim3(kBlock), 0, s, nb.cand_vf,
                     nb.cand_ee, nb.d_ncand_vf, nb.vf_cap, nb.d_nnear, nb.near_ids, nb.cplan,
                     nb.ccp, nb.croot, x, v, th, dt, dcd, nb.hit_flag, nb.tile_sum,
                     (float4*)nb.rec, hit_cap);
  CM_LAUNCH_CHECK();
  const int ntiles_near = (int)((ccap + 64 * kSmall - 1) / (64 * kSmall));
  rc = scan_tiles(nb.d_nnear, 64 * kSmall, ntiles_near, nb.tile_sum, nb.tile_pre, 1, nb.vf_out_cap,
                  nb.ee_out_cap, nb.d_nvf, nb.d_nee, nb.scan_temp, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ccd_write, dim3(tile_grid(ccap, kSmall)), dim3(kBlock), 0, s, nb.cand_vf,
                     nb.cand_ee, nb.d_ncand_vf, nb.vf_cap, nb.d_nnear, nb.near_ids, nb.cplan,
                     nb.ccp, nb.croot, nb.hit_flag, nb.tile_pre, nb.vf_out_cap, nb.ee_out_cap, x,
                     v, th, dt, dcd, nb.vf_v, nb.vf_w, nb.vf_tn, nb.ee_v, nb.ee_w, nb.ee_tn,
                     nb.warn, (const float4*)nb.rec, hit_cap);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_ccd_plan_only(const NarrowBufs& nb, const float4* x, const float4* v, float th,
                         float dt, int dcd, hipStream_t s) {
  const int64_t ccap = nb.vf_cap + nb.ee_cap;
  hipLaunchKernelGGL(k_ccd_plan, dim3(tile_grid(ccap, kBig)), dim3(kBlock), 0, s, nb.cand_vf,
                     nb.cand_ee, nb.d_ncand_vf, nb.d_ncand_ee, nb.vf_cap, nb.ee_cap, x, v, th, dt,
                     dcd, nb.cplan, nb.ccp, nb.croot, nb.d_ntotal, nb.hit_flag, nb.tile_sum);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_dynamic_csr(const NarrowBufs& nb, hipStream_t s) {
  const int gc = grid_for(nb.vf_out_cap + nb.ee_out_cap);
  if (nb.sorted_csr) {
    // every record keyed by its vertex, one stable radix sort, then the prefixes and
    // payloads from the sorted keys
    const int64_t rcap = 4 * (nb.vf_out_cap + nb.ee_out_cap);
    uint32_t* keys = nb.skey;
    uint32_t* kalt = nb.skey + rcap;
    hipLaunchKernelGGL(k_rec_keys, dim3(gc), dim3(kBlock), 0, s, nb.vf_v, nb.ee_v, nb.d_nvf,
                       nb.d_nee, (uint4*)keys, (uint4*)nb.drec, nb.d_nrec);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < nb.n) ++bits;
    int in_alt = 0;
    int rc = radix_sort_dev(keys, nb.drec, kalt, nb.dtmp, (int)rcap, nb.d_nrec, bits,
                            nb.sort_temp, &in_alt, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rec_place, dim3(grid_for(std::max<int64_t>(rcap, nb.n + 1))),
                       dim3(kBlock), 0, s, nb.n, in_alt ? kalt : keys, in_alt ? nb.dtmp : nb.drec,
                       nb.d_nrec, nb.d_nvf, nb.vf_v, nb.vf_w, nb.ee_v, nb.ee_w, nb.dprefix,
                       nb.drec, nb.rec);
    CM_LAUNCH_CHECK();
    return CM_OK;
  }
  // per-vertex counts, a scan, atomic placement, then each vertex's list ordered;
  // dcount[0, n) is zero here: cleared at creation, then by every k_dyn_fill
  hipLaunchKernelGGL(k_dyn_count, dim3(gc), dim3(kBlock), 0, s, nb.vf_v, nb.ee_v, nb.d_nvf,
                     nb.d_nee, nb.dcount, nb.n);
  CM_LAUNCH_CHECK();
  int rc = scan_exclusive(nb.dcount, nb.dprefix, nb.n, nullptr, nb.dprefix + nb.n, nb.scan_temp,
                          s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dyn_fill, dim3(gc), dim3(kBlock), 0, s, nb.vf_v, nb.ee_v, nb.d_nvf,
                     nb.d_nee, nb.dprefix, nb.dcount, nb.drec);
  hipLaunchKernelGGL(k_dyn_sort, dim3(nblocks(nb.n)), dim3(kBlock), 0, s, nb.n, nb.dprefix,
                     nb.drec, nb.dcount + nb.n, nb.dlong, nb.d_nvf, nb.vf_v, nb.vf_w, nb.ee_v,
                     nb.ee_w, nb.rec);
  hipLaunchKernelGGL(k_dyn_sort_long, dim3(512), dim3(kBlock), 0, s, nb.dprefix, nb.dcount + nb.n,
                     nb.dlong, nb.drec, nb.dtmp, nb.d_nvf, nb.vf_v, nb.vf_w, nb.ee_v, nb.ee_w,
                     nb.rec);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

}  // namespace cm
