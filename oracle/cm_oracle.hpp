// ORACLE — CPU restatement of the reference's GPU cloth substep (PBD + Morton/LBVH
// self-collision + CCD narrow phase). This is test infrastructure: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as
// the checker or the reported CPU baseline. The product (licenta_amd) never calls it.
//
// Every function cites the reference file:line it restates (paths relative to
// LicentaV2/LicentaV2/ in Tureyz/Licenta). Templated on Real: Real=float is the
// FP32 path; Real=double runs the continuous dynamics in FP64 while the discrete
// stages (Morton, sort, LBVH, pairs, CCD) run in FP32 on a float shadow state
// computed with the FP32 arithmetic, so integer stages see identical FP32 inputs.
//
// Defined choices where the reference is racy or undefined (each is pinned by a
// test in tests/test_oracle_*.py):
//  * feature ownership: the lowest triangle index in the initial order wins
//    (reference: atomicExch race, Collision/FeatureList.cu:84-132);
//  * parents[c] written by several internal nodes: the highest node id wins
//    (sequential order of Collision/CudaBVH.cu:186-187); unclaimed entries keep
//    their previous value as in the reference;
//  * refit: every internal node gets the union of its two children
//    (Collision/CudaBVH.cu:140-144) in children-first order, so a node shared by
//    two parents (midpoint split on tied codes) is well defined;
//  * traversal stack: 64 entries like Collision/CudaBVH.cu:38; a push into a full
//    stack is dropped and CM_WARN_STACK_OVERFLOW raised (reference: overflow UB);
//  * FindCriticalPoints with only the second root in [0,1]: that root is used as
//    the critical point (reference reads an uninitialised value,
//    Physics/CudaCubicSolver.cu:31-34, 110-114);
//  * vector length: sqrt(x*x+y*y+z*z) in place of CUDA's norm3df
//    (Core/CudaUtils.cu:212-220), parity unpinned at the last ulp;
//  * vertex normals: the per-vertex sum runs over the carried normal, then the
//    incident triangles by ascending index (reference: float atomics,
//    Physics/CudaDeformableUtils.cu:230-240);
//  * friction: an EE contact's entries carry its own four vertex ids (reference:
//    left unwritten, Physics/CudaPBD.cu:1810-1813).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../include/clothmi.h"

// Loops over independent items (vertices, constraints, triangles, leaves, pairs,
// candidates) run on OpenMP threads; every result is written to its own index and
// joined in index order, so the output does not depend on the thread count.
#ifdef _OPENMP
#include <omp.h>
#define CM_OMP_FOR _Pragma("omp parallel for schedule(static)")
inline int omp_threads() { return omp_get_max_threads(); }
inline int omp_tid() { return omp_get_thread_num(); }
inline int omp_nthreads() { return omp_get_num_threads(); }
#else
#define CM_OMP_FOR
inline int omp_threads() { return 1; }
inline int omp_tid() { return 0; }
inline int omp_nthreads() { return 1; }
#endif

namespace orc {

// ---------------------------------------------------------------------------
// float3 arithmetic, same operation order as Core/CudaUtils.cu:32-107,252-276
// ---------------------------------------------------------------------------
template <class R>
struct V3 {
  R x, y, z;
};
template <class R>
inline V3<R> mk(R x, R y, R z) {
  return V3<R>{x, y, z};
}
template <class R>
inline V3<R> operator+(V3<R> a, V3<R> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class R>
inline V3<R> operator-(V3<R> a, V3<R> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class R>
inline V3<R> operator-(V3<R> a) {
  return {-a.x, -a.y, -a.z};
}
template <class R>
inline V3<R> mul(V3<R> a, R s) {  // operator*(float3,float) and operator*(float,float3)
  return {a.x * s, a.y * s, a.z * s};
}
template <class R>
inline V3<R> divs(V3<R> a, R s) {  // operator/(float3,float)
  return {a.x / s, a.y / s, a.z / s};
}
template <class R>
inline V3<R> subs(V3<R> a, R s) {  // operator-(float3,float)
  return {a.x - s, a.y - s, a.z - s};
}
template <class R>
inline V3<R> adds(V3<R> a, R s) {
  return {a.x + s, a.y + s, a.z + s};
}
template <class R>
inline R dot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R>
inline R len(V3<R> a) {
  return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}
template <class R>
inline V3<R> normalize(V3<R> a) {
  return divs(a, len(a));
}
template <class R>
inline V3<R> vmin(V3<R> a, V3<R> b) {
  return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)};
}
template <class R>
inline V3<R> vmax(V3<R> a, V3<R> b) {
  return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)};
}
template <class R>
inline V3<R> min3(V3<R> a, V3<R> b, V3<R> c) {  // Core/CudaUtils.cu:232-235
  return vmin(a, vmin(b, c));
}
template <class R>
inline V3<R> max3(V3<R> a, V3<R> b, V3<R> c) {  // Core/CudaUtils.cu:247-250
  return vmax(a, vmax(b, c));
}
template <class R>
inline bool is_zero(V3<R> a) {  // Core/CudaUtils.cu:139-142
  return a.x == R(0) && a.y == R(0) && a.z == R(0);
}
template <class R>
inline R fclamp(R v, R lo, R hi) {  // Core/CudaUtils.cu:180-183
  return std::fmin(std::fmax(lo, v), hi);
}
template <class R>
inline bool between(R v, R lo, R hi) {  // Core/CudaUtils.cu:170-173
  return v >= lo && v <= hi;
}
template <class T, class R>
inline V3<T> cvt(V3<R> a) {
  return {T(a.x), T(a.y), T(a.z)};
}
// Core/CudaUtils.cu:364-371 (inclusive overlap)
inline bool aabb_overlap(V3<float> amin, V3<float> amax, V3<float> bmin, V3<float> bmax) {
  bool a = amax.x < bmin.x || amin.x > bmax.x;
  bool b = amax.y < bmin.y || amin.y > bmax.y;
  bool c = amax.z < bmin.z || amin.z > bmax.z;
  return !(a || b || c);
}

struct Tri {
  int32_t v1, v2, v3, feat;  // Physics/Structs.h:176-188 (m_v1..m_v3, m_assignedFeatures)
};

// ---------------------------------------------------------------------------
// Morton codes (Core/CudaUtils.cu:278-325): 21-bit per-axis quantisation of the
// raw coordinate clamped to [0,1], or 10-bit quantisation inside a box.
// ---------------------------------------------------------------------------
inline uint64_t spread21(uint64_t v) {  // bit k of v -> bit 3k
  uint64_t r = 0;
  for (int k = 0; k < 21; ++k) r |= ((v >> k) & 1ull) << (3 * k);
  return r;
}
inline uint64_t quant21(float c) {
  float s = std::fmin(std::fmax(c * 2097152.f, 0.0f), 2097151.0f);
  return (uint64_t)s;
}
inline uint64_t morton63(V3<float> c) {
  return (spread21(quant21(c.x)) << 2) + (spread21(quant21(c.y)) << 1) + spread21(quant21(c.z));
}
inline uint64_t quant10(float u) {
  float s = std::fmin(std::fmax(u * 1024.f, 0.0f), 1023.0f);
  return (uint64_t)s;
}
inline float unit_coord(float c, float lo, float hi) {
  return hi > lo ? (c - lo) / (hi - lo) : 0.0f;
}
inline uint64_t morton30(V3<float> c, V3<float> lo, V3<float> hi) {
  return (spread21(quant10(unit_coord(c.x, lo.x, hi.x))) << 2) +
         (spread21(quant10(unit_coord(c.y, lo.y, hi.y))) << 1) +
         spread21(quant10(unit_coord(c.z, lo.z, hi.z)));
}

inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }  // __clzll(0) == 64

// Core/CudaUtils.cu:338-351
inline int delta(const uint64_t* k, int i, int j, int maxj) {
  if (j < 0 || j > maxj) return -1;
  uint64_t x = k[i] ^ k[j];
  return x ? clz64(x) : clz64((uint64_t)i ^ (uint64_t)j) + 64;
}
// Core/CudaUtils.cu:353-362: signbit((float)a) ? -1 : 1, so sgn(0) = +1
inline int sgn(int a) { return std::signbit((float)a) ? -1 : 1; }

// Collision/CudaBVH.cu:190-212
inline void find_range(const uint64_t* k, int count, int id, int& first, int& last) {
  int d = sgn(delta(k, id, id + 1, count) - delta(k, id, id - 1, count));
  int dmin = delta(k, id, id - d, count);
  int lmax = 2;
  while (delta(k, id, id + lmax * d, count) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (delta(k, id, id + (l + t) * d, count) > dmin) l += t;
  int j = id + l * d;
  first = std::min(id, j);
  last = std::max(id, j);
}
// Collision/CudaBVH.cu:214-242 (midpoint on equal end codes), or the
// index-augmented Karras split when midpoint == false.
inline int find_split(const uint64_t* k, int count, int first, int last, bool midpoint) {
  uint64_t fc = k[first], lc = k[last];
  int com;
  if (midpoint) {
    if (fc == lc) return (first + last) >> 1;
    com = clz64(fc ^ lc);
  } else {
    com = delta(k, first, last, count);
  }
  int split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    int ns = split + step;
    if (ns < last) {
      int pref = midpoint ? clz64(fc ^ k[ns]) : delta(k, first, ns, count);
      if (pref > com) split = ns;
    }
  } while (step > 1);
  return split;
}

// ---------------------------------------------------------------------------
// LBVH stages on plain arrays (FP32 AABBs).
// ---------------------------------------------------------------------------
// Collision/CudaBVH.cu:163-188 in thread-id order.
inline void lbvh_build(int t, const uint64_t* keys, bool midpoint, int32_t* lefts,
                       int32_t* rights, int32_t* parents, uint32_t* warn) {
  const int nin = t - 1;
  std::vector<int> claims(2 * (size_t)t - 1, 0);
  CM_OMP_FOR
  for (int id = 0; id < nin; ++id) {
    int first, last;
    find_range(keys, nin, id, first, last);
    int split = find_split(keys, nin, first, last, midpoint);
    lefts[id] = split == first ? nin + split : split;
    rights[id] = split + 1 == last ? nin + split + 1 : split + 1;
  }
  // parents in node order: on a broken tree the last claimant wins, as in one pass
  for (int id = 0; id < nin; ++id) {
    const int l = lefts[id], r = rights[id];
    parents[l] = id;
    parents[r] = id;
    if (++claims[l] > 1 || ++claims[r] > 1) {
      if (warn) *warn |= CM_WARN_BROKEN_TREE;
    }
  }
}

// Collision/CudaBVH.cu:128-149, restated children-first.
inline void lbvh_refit(int t, const int32_t* lefts, const int32_t* rights, V3<float>* amin,
                       V3<float>* amax, int32_t* rmll, int32_t* rmlr) {
  const int nin = t - 1;
  std::vector<uint8_t> state(nin, 0);  // 0 new, 1 on stack, 2 done
  std::vector<int> stack;
  for (int root = 0; root < nin; ++root) {
    if (state[root] == 2) continue;
    stack.push_back(root);
    while (!stack.empty()) {
      int p = stack.back();
      if (state[p] == 2) {
        stack.pop_back();
        continue;
      }
      int l = lefts[p], r = rights[p];
      bool ready = true;
      if (l < nin && state[l] != 2) {
        ready = false;
        stack.push_back(l);
      }
      if (r < nin && state[r] != 2) {
        ready = false;
        stack.push_back(r);
      }
      if (!ready) {
        state[p] = 1;
        continue;
      }
      stack.pop_back();
      amin[p] = vmin(amin[l], amin[r]);
      amax[p] = vmax(amax[l], amax[r]);
      rmll[p] = l >= nin ? l : std::max(rmll[l], rmlr[l]);
      rmlr[p] = r >= nin ? r : std::max(rmll[r], rmlr[r]);
      state[p] = 2;
    }
  }
}

// Collision/CudaBVH.cu:23-113 for one leaf: pairs (leaf, j) in DFS order.
inline int lbvh_leaf_pairs(int t, int leaf, const int32_t* lefts, const int32_t* rights,
                           const V3<float>* amin, const V3<float>* amax, const int32_t* rmll,
                           const int32_t* rmlr, int cap, int32_t* out, uint32_t* warn,
                           bool* capped) {
  const int nin = t - 1;
  const int me = nin + leaf;
  int ncol = 0;
  int stack[64];
  stack[0] = -1;
  int top = 1;
  int node = 0;
  const long long visit_limit = 64ll * t + 4096;
  long long visits = 0;
  do {
    if (++visits > visit_limit) {
      if (warn) *warn |= CM_WARN_TRAVERSAL_LIMIT;
      break;
    }
    int l = lefts[node], r = rights[node];
    bool ol = false, orr = false;
    if (rmll[node] > me && aabb_overlap(amin[me], amax[me], amin[l], amax[l])) ol = true;
    if (rmlr[node] > me && aabb_overlap(amin[me], amax[me], amin[r], amax[r])) orr = true;
    // The reference keeps walking past the cap (Collision/CudaBVH.cu:67-93) but
    // drops every further pair; stopping at the first dropped one gives the same list.
    if (ol && l >= nin) {
      if (ncol == cap) {
        if (capped) *capped = true;
        break;
      }
      out[2 * ncol] = leaf;
      out[2 * ncol + 1] = l - nin;
      ++ncol;
    }
    if (orr && r >= nin) {
      if (ncol == cap) {
        if (capped) *capped = true;
        break;
      }
      out[2 * ncol] = leaf;
      out[2 * ncol + 1] = r - nin;
      ++ncol;
    }
    bool tl = ol && l < nin, tr = orr && r < nin;
    if (!tl && !tr) {
      node = stack[--top];
    } else {
      node = tl ? l : r;
      if (tl && tr) {
        if (top < 64) {
          stack[top++] = r;
        } else if (warn) {
          *warn |= CM_WARN_STACK_OVERFLOW;
        }
      }
    }
  } while (top != 0);
  return ncol;
}

// ---------------------------------------------------------------------------
// CCD narrow phase (Physics/CudaCubicSolver.cu, Collision/CudaPrimitiveTests.cu),
// always FP32.
// ---------------------------------------------------------------------------
typedef V3<float> F3;
static const float kCubicEps = 1.f / 1000000;  // Physics/CudaCubicSolver.cuh:12
static const float kUtilEps = 0.00000001f;     // Core/CudaUtils.cu:4

inline float triple_scalar(F3 x21, F3 x31, F3 x41, F3 v21, F3 v31, F3 v41, float t) {
  // Physics/CudaCubicSolver.cu:126-129
  return dot(cross(x21 + mul(v21, t), x31 + mul(v31, t)), x41 + mul(v41, t));
}
inline float secant(F3 x21, F3 x31, F3 x41, F3 v21, F3 v31, F3 v41, float a, float b,
                    int maxit) {  // Physics/CudaCubicSolver.cu:131-153
  float x0 = a, x1 = b;
  for (int i = 0; i < maxit; ++i) {
    float f0 = triple_scalar(x21, x31, x41, v21, v31, v41, x0);
    float f1 = triple_scalar(x21, x31, x41, v21, v31, v41, x1);
    float xc = (x0 * f1 - x1 * f0) / (f1 - f0);
    if (std::fabs(xc - x1) < kCubicEps) return xc;
    x0 = x1;
    x1 = xc;
  }
  return -15.f;
}
inline int coplanarity_times(F3 x1, F3 x2, F3 x3, F3 x4, F3 v1, F3 v2, F3 v3, F3 v4,
                             float* times) {  // Physics/CudaCubicSolver.cu:5-67
  int n = 0;
  F3 x21 = x2 - x1, x31 = x3 - x1, x41 = x4 - x1;
  F3 v21 = v2 - v1, v31 = v3 - v1, v41 = v4 - v1;
  // critical points (restated inline to expose both roots)
  F3 v21v31 = cross(v21, v31), x21v31 = cross(x21, v31), v21x31 = cross(v21, x31),
     x21x31 = cross(x21, x31);
  float v1v2v3 = dot(v21v31, v41), x1v2v3 = dot(x21v31, v41), v1x2v3 = dot(v21x31, v41),
        v1v2x3 = dot(v21v31, x41), x1x2v3 = dot(x21x31, v41), x1v2x3 = dot(x21v31, x41),
        v1x2x3 = dot(v21x31, x41);
  float a = v1v2v3;
  float b = x1v2v3 + v1x2v3 + v1v2x3;
  float c = x1x2v3 + x1v2x3 + v1x2x3;
  float ad = 3 * a, bd = 2 * b;
  float dd = 4 * (b * b - ad * c);
  float sq = std::sqrt(dd);
  float da = 2 * ad;
  float t1 = ((-bd + sq) / da);
  float t2 = ((-bd - sq) / da);
  float cp1 = 0.f, cp2 = 0.f;
  int ccount = 0;
  bool have1 = false;
  if (dd >= kCubicEps) {
    if (between(t1, 0.f, 1.f)) {
      cp1 = t1;
      have1 = true;
      ++ccount;
    }
    if (between(t2, 0.f, 1.f)) {
      cp2 = t2;
      ++ccount;
    }
  } else if (std::fabs(dd) < kCubicEps) {
    if (between(t1, 0.f, 1.f)) {
      cp1 = t1;
      have1 = true;
      ++ccount;
    }
  }
  if (ccount == 1 && !have1) cp1 = cp2;
  float ends[4];
  int ne = 0;
  ends[ne++] = 0.f;
  if (ccount == 1) {
    ends[ne++] = cp1;
  } else if (ccount == 2) {
    ends[ne++] = cp1;
    ends[ne++] = cp2;
  }
  ends[ne++] = 1.f;
  for (int i = 0; i < ne - 1; ++i) {
    float ea = ends[i], eb = ends[i + 1];
    float fa = triple_scalar(x21, x31, x41, v21, v31, v41, ea);
    float fb = triple_scalar(x21, x31, x41, v21, v31, v41, eb);
    if (std::fabs(fa) < kCubicEps) times[n++] = ea;
    if (fa * fb < 0.f) {
      float tt = secant(x21, x31, x41, v21, v31, v41, ea, eb, 20);
      if (tt != -15) times[n++] = tt;
    }
  }
  return n;
}

// Collision/CudaPrimitiveTests.cu:348-406
inline bool test_vf(F3 x4, F3 x1, F3 x2, F3 x3, cm_contact& c, float th) {
  F3 x13 = x1 - x3, x23 = x2 - x3, x43 = x4 - x3;
  F3 crs = cross(x2 - x1, x3 - x1);
  F3 n = normalize(crs);
  if (std::fabs(dot(x43, n)) <= th) {
    float dtmp = dot(x13, x23);
    float m11 = dot(x13, x13), m12 = dtmp, m21 = dtmp, m22 = dot(x23, x23);
    float vr1 = dot(x13, x43), vr2 = dot(x23, x43);
    float w2 = (m11 * vr2 - m21 * vr1) / (m11 * m22 - m21 * m12);
    float w1 = (vr1 - m12 * w2) / m11;
    float w3 = 1.f - w1 - w2;
    float delta = th / (len(x13) + len(x23) + len(x1 - x2));
    const float lo = 0.f, hi = 1 + delta;
    if (between(w1, lo, hi) && between(w2, lo, hi) && between(w3, lo, hi)) {
      c.w[0] = 1.f;
      c.w[1] = -w1;
      c.w[2] = -w2;
      c.w[3] = -w3;
      c.n[0] = n.x;
      c.n[1] = n.y;
      c.n[2] = n.z;
      c.t = 0.f;
      return true;
    }
  }
  return false;
}
// Core/CudaUtils.cu:149-168
inline F3 project_on_line(F3 p, F3 a, F3 b) {
  F3 ab = b - a, ap = p - a;
  return a + mul(ab, dot(ap, ab) / dot(ab, ab));
}
inline F3 clamp_on_line(F3 p, F3 a, F3 b) {
  float d1 = len(p - a), d2 = len(p - b);
  if (std::fabs(d1 + d2 - len(a - b)) < kUtilEps) return p;
  return d1 < d2 ? a : b;
}
// Collision/CudaPrimitiveTests.cu:408-425
inline bool test_edge_degenerate(F3 x1, F3 x3, F3 x21, F3 x43, cm_contact& c, float th) {
  F3 m1 = x1 + divs(x21, 2.f), m2 = x3 + divs(x43, 2.f);
  if (len(m1 - m2) <= th) {
    c.w[0] = c.w[1] = c.w[2] = c.w[3] = 0.5f;
    F3 n = normalize(m1 - m2);
    c.n[0] = n.x;
    c.n[1] = n.y;
    c.n[2] = n.z;
    return true;
  }
  return false;
}
// Collision/CudaPrimitiveTests.cu:252-346 (cp2 uses `a`, as at :291)
inline bool test_ee(F3 x1, F3 x2, F3 x3, F3 x4, cm_contact& c, float th, float eps) {
  F3 x21 = x2 - x1, x31 = x3 - x1, x43 = x4 - x3;
  F3 crs = cross(x21, x43);
  if (len(crs) < eps) return test_edge_degenerate(x1, x3, x21, x43, c, th);
  if (is_zero(crs)) return false;
  float dtmp = dot(-x21, x43);
  float m11 = dot(x21, x21), m12 = dtmp, m21 = dtmp, m22 = dot(x43, x43);
  float vr1 = dot(x21, x31), vr2 = dot(-x43, x31);
  float b = (m11 * vr2 - m21 * vr1) / (m11 * m22 - m21 * m12);
  float a = (vr1 - m12 * b) / m11;
  if (between(a, 0.f, 1.f) && between(b, 0.f, 1.f)) {
    F3 cp1 = x1 + mul(x21, a);
    F3 cp2 = x3 + mul(x43, a);
    if (len(cp1 - cp2) <= th) {
      c.w[0] = 1 - a;
      c.w[1] = a;
      c.w[2] = -(1 - b);
      c.w[3] = -b;
      F3 dir = (x1 + mul(x21, a)) - (x3 + mul(x43, b));
      F3 n = normalize(is_zero(dir) ? crs : dir);
      c.n[0] = n.x;
      c.n[1] = n.y;
      c.n[2] = n.z;
      return true;
    }
  } else {
    float ac = fclamp(a, 0.f, 1.f), bc = fclamp(b, 0.f, 1.f);
    float d1 = std::fabs(a - ac) * len(x21);
    float d2 = std::fabs(b - bc) * len(x43);
    a = ac;
    b = bc;
    F3 cp1, cp2;
    if (d1 > d2) {
      cp1 = x1 + mul(x21, a);
      cp2 = clamp_on_line(project_on_line(cp1, x3, x4), x3, x4);
    } else {
      cp2 = x3 + mul(x43, b);
      cp1 = clamp_on_line(project_on_line(cp2, x1, x2), x1, x2);
    }
    if (len(cp1 - cp2) <= th) {
      c.w[0] = 1 - a;
      c.w[1] = a;
      c.w[2] = -(1 - b);
      c.w[3] = -b;
      F3 n = normalize(cp1 - cp2);
      c.n[0] = n.x;
      c.n[1] = n.y;
      c.n[2] = n.z;
      c.t = 0.f;
      return true;
    }
  }
  return false;
}

// Collision/CudaPrimitiveTests.cu:128-184 (VF) and 187-238 (EE) on prev positions
// x and velocities v, candidate already filled with vertex ids.
inline bool ccd_vf(cm_contact& c, const F3* x, const F3* v, float th, float dt) {
  int p = c.v[0], a = c.v[1], b = c.v[2], d = c.v[3];
  if (p == a || p == b || p == d) return false;  // VFSanity :240-243
  F3 x4 = x[p], x1 = x[a], x2 = x[b], x3 = x[d];
  F3 v4 = v[p], v1 = v[a], v2 = v[b], v3 = v[d];
  float times[10];
  int nt = coplanarity_times(x1, x2, x3, x4, v1, v2, v3, v4, times);
  for (int i = 0; i < nt; ++i) {
    float t = times[i] * dt;
    x4 = x4 + mul(v4, t);
    x1 = x1 + mul(v1, t);
    x2 = x2 + mul(v2, t);
    x3 = x3 + mul(v3, t);
    if (test_vf(x4, x1, x2, x3, c, th)) {
      c.t = t;
      return true;
    }
  }
  return false;
}
inline bool ccd_ee(cm_contact& c, const F3* x, const F3* v, float th, float dt) {
  int p1 = c.v[0], p2 = c.v[1], q1 = c.v[2], q2 = c.v[3];
  if (p1 == q1 || p1 == q2 || p2 == q1 || p2 == q2) return false;  // EESanity :245-248
  F3 x1 = x[p1], x2 = x[p2], x3 = x[q1], x4 = x[q2];
  F3 v1 = v[p1], v2 = v[p2], v3 = v[q1], v4 = v[q2];
  float times[10];
  int nt = coplanarity_times(x1, x2, x3, x4, v1, v2, v3, v4, times);
  for (int i = 0; i < nt; ++i) {
    float t = times[i] * dt;
    x1 = x1 + mul(v1, t);
    x2 = x2 + mul(v2, t);
    x3 = x3 + mul(v3, t);
    x4 = x4 + mul(v4, t);
    if (test_ee(x1, x2, x3, x4, c, th, 0.000000001f)) {
      c.t = t;
      return true;
    }
  }
  return false;
}

// Collision/CudaPrimitiveTests.cu:41-94 (DCDTestVFs, DCDTestEEsBridson): the same
// sanity checks and proximity tests once, at the positions x; the contact's time is
// what the test leaves (0, or the candidate's -1 where TestEE does not set it).
inline bool dcd_vf(cm_contact& c, const F3* x, float th) {
  int p = c.v[0], a = c.v[1], b = c.v[2], d = c.v[3];
  if (p == a || p == b || p == d) return false;  // VFSanity :240-243
  return test_vf(x[p], x[a], x[b], x[d], c, th);
}
inline bool dcd_ee(cm_contact& c, const F3* x, float th) {
  int p1 = c.v[0], p2 = c.v[1], q1 = c.v[2], q2 = c.v[3];
  if (p1 == q1 || p1 == q2 || p2 == q1 || p2 == q2) return false;  // EESanity :245-248
  return test_ee(x[p1], x[p2], x[q1], x[q2], c, th, 0.000000001f);
}

inline cm_contact make_contact(int a, int b, int c, int d) {  // Physics/Structs.h:112-128
  cm_contact r;
  r.v[0] = a;
  r.v[1] = b;
  r.v[2] = c;
  r.v[3] = d;
  r.w[0] = r.w[1] = r.w[2] = r.w[3] = -1.f;
  r.t = -1.f;
  r.n[0] = r.n[1] = r.n[2] = -1.f;
  return r;
}

// Candidate (pair, slot) -> contact and ownership gate
// (Collision/CudaCollisionDetection.cu:101-187). slot 0..5 VF, 6..14 EE.
inline bool candidate(const Tri& t1, const Tri& t2, int slot, cm_contact& c) {
  auto bit = [](int f, int b) { return (f >> b) & 1; };
  switch (slot) {
    case 0: c = make_contact(t1.v1, t2.v1, t2.v2, t2.v3); return bit(t1.feat, 5);
    case 1: c = make_contact(t1.v2, t2.v1, t2.v2, t2.v3); return bit(t1.feat, 4);
    case 2: c = make_contact(t1.v3, t2.v1, t2.v2, t2.v3); return bit(t1.feat, 3);
    case 3: c = make_contact(t2.v1, t1.v1, t1.v2, t1.v3); return bit(t2.feat, 5);
    case 4: c = make_contact(t2.v2, t1.v1, t1.v2, t1.v3); return bit(t2.feat, 4);
    case 5: c = make_contact(t2.v3, t1.v1, t1.v2, t1.v3); return bit(t2.feat, 3);
    case 6: c = make_contact(t1.v1, t1.v2, t2.v1, t2.v2); return bit(t1.feat, 2) & bit(t2.feat, 2);
    case 7: c = make_contact(t1.v1, t1.v3, t2.v1, t2.v2); return bit(t1.feat, 0) & bit(t2.feat, 2);
    case 8: c = make_contact(t1.v2, t1.v3, t2.v1, t2.v2); return bit(t1.feat, 1) & bit(t2.feat, 2);
    case 9: c = make_contact(t1.v1, t1.v2, t2.v1, t2.v3); return bit(t1.feat, 2) & bit(t2.feat, 0);
    case 10: c = make_contact(t1.v1, t1.v3, t2.v1, t2.v3); return bit(t1.feat, 0) & bit(t2.feat, 0);
    case 11: c = make_contact(t1.v2, t1.v3, t2.v1, t2.v3); return bit(t1.feat, 1) & bit(t2.feat, 0);
    case 12: c = make_contact(t1.v1, t1.v2, t2.v2, t2.v3); return bit(t1.feat, 2) & bit(t2.feat, 1);
    case 13: c = make_contact(t1.v1, t1.v3, t2.v2, t2.v3); return bit(t1.feat, 0) & bit(t2.feat, 1);
    default: c = make_contact(t1.v2, t1.v3, t2.v2, t2.v3); return bit(t1.feat, 1) & bit(t2.feat, 1);
  }
}

// Feature swept-AABB filter (Collision/CudaPrimitiveTests.cu:447-490 with the
// particle/edge AABBs of Collision/FeatureList.cu:134-162).
inline void particle_aabb(const F3* p, const F3* x, int i, float th, F3& mn, F3& mx) {
  mn = subs(vmin(p[i], x[i]), th);
  mx = adds(vmax(p[i], x[i]), th);
}
inline bool feature_filter(const cm_contact& c, bool vf, const F3* p, const F3* x, float th) {
  F3 mn[4], mx[4];
  for (int k = 0; k < 4; ++k) particle_aabb(p, x, c.v[k], th, mn[k], mx[k]);
  if (vf) {
    F3 tmn = min3(mn[1], mn[2], mn[3]), tmx = max3(mx[1], mx[2], mx[3]);
    return aabb_overlap(mn[0], mx[0], tmn, tmx);
  }
  int a0 = std::min(c.v[0], c.v[1]), a1 = std::max(c.v[0], c.v[1]);
  int b0 = std::min(c.v[2], c.v[3]), b1 = std::max(c.v[2], c.v[3]);
  F3 amn, amx, bmn, bmx, t0, t1, u0, u1;
  particle_aabb(p, x, a0, th, t0, t1);
  particle_aabb(p, x, a1, th, u0, u1);
  amn = vmin(t0, u0);
  amx = vmax(t1, u1);
  particle_aabb(p, x, b0, th, t0, t1);
  particle_aabb(p, x, b1, th, u0, u1);
  bmn = vmin(t0, u0);
  bmx = vmax(t1, u1);
  return aabb_overlap(amn, amx, bmn, bmx);
}

// ---------------------------------------------------------------------------
// Topology (Physics/CudaPBD.cu:39-370, Collision/FeatureList.cu:12-132).
// ---------------------------------------------------------------------------
struct Topology {
  int n = 0, t = 0;
  std::vector<int32_t> sid1, sid2;  // distance constraints
  std::vector<float> sl0, skp;
  std::vector<int32_t> bid[4];  // bend constraints p1..p4
  std::vector<float> bkp, bphi0;
  std::vector<int32_t> prefix;     // N+1 CSR over correction slots
  std::vector<int32_t> soff[2];    // slot of each stretch end
  std::vector<int32_t> boff[4];    // slot of each bend vertex
  std::vector<Tri> tri;            // initial triangles with feature bits
  std::vector<int32_t> lra_fixed;  // pinned vertex ids
  std::vector<float> lra_l0;       // N * F
};

inline float glm_distance(const float* xyz, int a, int b) {  // glm::distance in float
  float dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1],
        dz = xyz[3 * a + 2] - xyz[3 * b + 2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

inline float kp_of(float k, int iters) {  // Physics/CudaPBD.cu:990, 1029
  return 1.f - std::pow(1.f - k, 1.f / (float)iters);
}

// Physics/CudaPBD.cu:289-370: p1 = last shared vertex of tri1 in its order,
// p2 = the other shared one, p3 = tri1's free vertex, p4 = tri2's free vertex.
inline void bend_quad(const int t1[3], const int t2[3], int& p1, int& p2, int& p3, int& p4) {
  p1 = p2 = p3 = p4 = -1;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (t1[i] == t2[j]) {
        p1 = t1[i];
        break;
      }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      if (t1[i] == t2[j] && t1[i] != p1) {
        p2 = t1[i];
        break;
      }
  for (int i = 0; i < 3; ++i)
    if (t1[i] != p1 && t1[i] != p2) {
      p3 = t1[i];
      break;
    }
  for (int i = 0; i < 3; ++i)
    if (t2[i] != p1 && t2[i] != p2) {
      p4 = t2[i];
      break;
    }
}

inline float dihedral_rest(const float* xyz, int p1, int p2, int p3, int p4) {
  // Physics/CudaPBD.cu:348-366 (the value the reference computes, then discards)
  auto v = [&](int a, int b) {
    return mk(xyz[3 * a] - xyz[3 * b], xyz[3 * a + 1] - xyz[3 * b + 1],
              xyz[3 * a + 2] - xyz[3 * b + 2]);
  };
  F3 p21 = v(p2, p1), p31 = v(p3, p1), p41 = v(p4, p1);
  F3 a = normalize(cross(p21, p31)), b = normalize(cross(p21, p41));
  // clamped: the unit-vector dot can round past +-1 (the reference's acos is
  // unclamped, but it discards the value)
  return std::acos(fclamp(dot(a, b), -1.f, 1.f));
}

inline void build_topology(Topology& tp, const cm_params& prm, const float* xyz, int n,
                           const uint32_t* tri, int t, const uint8_t* pinned,
                           const int32_t* dist_pairs, int n_dist, const float* k_dist,
                           const float* k_bend) {
  tp.n = n;
  tp.t = t;
  const int I = prm.solver_iterations;
  // Edge walk: first sighting -> distance constraint, second -> bend quad.
  struct Key {
    int a, b;
    bool operator==(const Key& o) const { return a == o.a && b == o.b; }
  };
  struct H {
    size_t operator()(const Key& k) const { return (size_t)k.a * 1000003u ^ (size_t)k.b; }
  };
  std::unordered_map<Key, int, H> open;  // unordered edge -> first triangle index
  std::vector<int32_t> s1, s2;
  for (int i = 0; i < t; ++i) {
    int tv[3] = {(int)tri[3 * i], (int)tri[3 * i + 1], (int)tri[3 * i + 2]};
    for (int j = 0; j < 3; ++j) {
      int id1 = tv[j % 3], id2 = tv[(j + 1) % 3];
      Key k{std::min(id1, id2), std::max(id1, id2)};
      auto it = open.find(k);
      if (it != open.end()) {
        int f = it->second;
        int tf[3] = {(int)tri[3 * f], (int)tri[3 * f + 1], (int)tri[3 * f + 2]};
        int p1, p2, p3, p4;
        bend_quad(tf, tv, p1, p2, p3, p4);
        tp.bid[0].push_back(p1);
        tp.bid[1].push_back(p2);
        tp.bid[2].push_back(p3);
        tp.bid[3].push_back(p4);
        float kb = k_bend ? k_bend[tp.bkp.size()] : prm.k_bend;
        tp.bkp.push_back(kp_of(kb, I));
        tp.bphi0.push_back((prm.ref_compat & CM_RC_BEND_ABS) ? 3.14159265359f
                                                                : dihedral_rest(xyz, p1, p2, p3, p4));
        open.erase(it);
      } else {
        open[k] = i;
        s1.push_back(id1);
        s2.push_back(id2);
      }
    }
  }
  if (dist_pairs) {
    s1.assign(n_dist, 0);
    s2.assign(n_dist, 0);
    for (int i = 0; i < n_dist; ++i) {
      s1[i] = dist_pairs[2 * i];
      s2[i] = dist_pairs[2 * i + 1];
    }
  }
  tp.sid1 = s1;
  tp.sid2 = s2;
  const int S = (int)s1.size(), B = (int)tp.bkp.size();
  tp.sl0.resize(S);
  tp.skp.resize(S);
  for (int i = 0; i < S; ++i) {
    tp.sl0[i] = glm_distance(xyz, s1[i], s2[i]);
    tp.skp[i] = kp_of(k_dist ? k_dist[i] : prm.k_stretch, I);
  }
  // CSR slots (Physics/CudaPBD.cu:146-233): stretch ends first, then bends.
  std::vector<int> cnt(n, 0);
  for (int i = 0; i < S; ++i) {
    cnt[s1[i]]++;
    cnt[s2[i]]++;
  }
  for (int i = 0; i < B; ++i)
    for (int r = 0; r < 4; ++r) cnt[tp.bid[r][i]]++;
  tp.prefix.assign(n + 1, 0);
  for (int i = 1; i <= n; ++i) tp.prefix[i] = tp.prefix[i - 1] + cnt[i - 1];
  std::fill(cnt.begin(), cnt.end(), 0);
  tp.soff[0].resize(S);
  tp.soff[1].resize(S);
  for (int i = 0; i < S; ++i) {
    tp.soff[0][i] = tp.prefix[s1[i]] + cnt[s1[i]]++;
    tp.soff[1][i] = tp.prefix[s2[i]] + cnt[s2[i]]++;
  }
  for (int r = 0; r < 4; ++r) tp.boff[r].resize(B);
  for (int i = 0; i < B; ++i)
    for (int r = 0; r < 4; ++r) tp.boff[r][i] = tp.prefix[tp.bid[r][i]] + cnt[tp.bid[r][i]]++;
  // Triangles and feature ownership, lowest index wins (Collision/FeatureList.cu:84-132).
  tp.tri.resize(t);
  std::vector<uint8_t> vown(n, 0);
  std::unordered_map<Key, int, H> eown;
  auto claim_edge = [&](int a, int b) {
    Key k{std::min(a, b), std::max(a, b)};
    if (eown.count(k)) return false;
    eown[k] = 1;
    return true;
  };
  for (int i = 0; i < t; ++i) {
    Tri tr{(int)tri[3 * i], (int)tri[3 * i + 1], (int)tri[3 * i + 2], 0};
    int f = 0;
    if (!vown[tr.v1]) { vown[tr.v1] = 1; f |= 1 << 5; }
    if (!vown[tr.v2]) { vown[tr.v2] = 1; f |= 1 << 4; }
    if (!vown[tr.v3]) { vown[tr.v3] = 1; f |= 1 << 3; }
    if (claim_edge(tr.v1, tr.v2)) f |= 1 << 2;
    if (claim_edge(tr.v2, tr.v3)) f |= 1 << 1;
    if (claim_edge(tr.v3, tr.v1)) f |= 1 << 0;
    tr.feat = f;
    tp.tri[i] = tr;
  }
  // LRA rest lengths (Physics/CudaPBD.cu:60-81).
  tp.lra_fixed.clear();
  for (int i = 0; i < n; ++i)
    if (pinned[i]) tp.lra_fixed.push_back(i);
  const int F = (int)tp.lra_fixed.size();
  tp.lra_l0.assign((size_t)n * F, 0.f);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < F; ++j)
      if (!pinned[i])
        tp.lra_l0[(size_t)i * F + j] = glm_distance(xyz, i, tp.lra_fixed[j]) * (1.f + prm.strain_limit);
}

// ---------------------------------------------------------------------------
// Deterministic reduction order used for the damping sums: chunks of 256
// consecutive elements reduced by a halving tree, chunk results added in order.
// ---------------------------------------------------------------------------
template <class T, class F>
inline T tree_sum(int n, F get) {
  T total{};
  bool first = true;
  std::vector<T> s(256);
  for (int base = 0; base < n; base += 256) {
    for (int j = 0; j < 256; ++j) s[j] = (base + j < n) ? get(base + j) : T{};
    for (int stride = 128; stride >= 1; stride >>= 1)
      for (int j = 0; j < stride; ++j) s[j] = s[j] + s[j + stride];
    total = first ? s[0] : total + s[0];
    first = false;
  }
  return total;
}

// ---------------------------------------------------------------------------
// The simulation object (Physics/CudaDeformableBody.cpp:30-285, 550-627).
// ---------------------------------------------------------------------------
template <class R>
struct Sim {
  typedef V3<R> RV;
  cm_params prm;
  Topology tp;
  int n = 0, t = 0;
  std::vector<uint8_t> pinned;
  std::vector<R> w;       // inverse masses
  std::vector<R> mass;    // masses for damping
  std::vector<RV> x, p, v;
  std::vector<Tri> tri;   // current (sorted) triangle order
  // collision buffers (FP32)
  std::vector<uint64_t> keys;
  std::vector<uint32_t> perm;
  std::vector<F3> amin, amax;
  std::vector<int32_t> lefts, rights, parents, rmll, rmlr;
  std::vector<int32_t> pairs;  // 2 * P
  std::vector<cm_contact> vf, ee;
  // CCD inputs of the last substep: the VF / EE candidates that passed the
  // ownership gate and the feature swept-AABB filter, in candidate order
  std::vector<cm_contact> cand_vf, cand_ee;
  uint32_t warn = 0;
  // differentiable rollout (no reference counterpart): with `record` set, step()
  // keeps each substep's entering state, the positions at the start of every Jacobi
  // iteration and after the last, and the substep's contact set
  struct Tape {
    std::vector<RV> x, v;
    std::vector<std::vector<RV>> pk;
    std::vector<cm_contact> vf, ee;
    // contact activation decisions the forward took, one byte per (iteration,
    // contact); the adjoint differentiates the substep with these decisions
    std::vector<uint8_t> act;
  };
  bool record = false;
  std::vector<Tape> tape;
  // Contact activation decisions of the last solver stage, one byte per (iteration,
  // contact) with the VF contacts first: act_mode 1 records them, 2 replays the
  // ones given (a run in another precision on the same contact set); act_mismatch
  // is set when the given decisions do not fit the contact set.
  int act_mode = 0;
  std::vector<uint8_t> act;
  bool act_mismatch = false;
  std::vector<uint8_t> gflag;  // ground contacts of the last solver set-up
  std::vector<RV> vnrm;        // vertex normals (carried with CM_F_VERTEX_NORMALS)
  bool sphere_on = false;
  F3 sphere_c{0, 0, 0};
  float sphere_r = 0.f;

  void create(const cm_params& P, const float* xyz, int n_, const uint32_t* tris, int t_,
              const uint8_t* pin, const float* inv_mass, const int32_t* dist_pairs, int n_dist,
              const float* k_dist, const float* k_bend) {
    prm = P;
    n = n_;
    t = t_;
    pinned.resize(n);
    for (int i = 0; i < n; ++i) pinned[i] = pin ? pin[i] : (inv_mass[i] == 0.f);
    build_topology(tp, prm, xyz, n, tris, t, pinned.data(), dist_pairs, n_dist, k_dist, k_bend);
    w.resize(n);
    mass.resize(n);
    x.resize(n);
    p.resize(n);
    v.resize(n);
    for (int i = 0; i < n; ++i) {
      w[i] = R(inv_mass[i]);
      mass[i] = pinned[i] ? R(prm.pinned_mass) : (inv_mass[i] > 0 ? R(1.f / inv_mass[i]) : R(prm.pinned_mass));
      x[i] = p[i] = mk<R>(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
      v[i] = mk<R>(0, 0, 0);
    }
    tri = tp.tri;
    keys.assign(t, 0);
    perm.assign(t, 0);
    amin.assign(t > 0 ? 2 * (size_t)t - 1 : 0, F3{0, 0, 0});
    amax = amin;
    lefts.assign(t > 1 ? t - 1 : 0, 0);
    rights = lefts;
    rmll = lefts;
    rmlr = lefts;
    parents.assign(t > 0 ? 2 * (size_t)t - 1 : 0, 0);
    vnrm.assign(n, mk<R>(0, 0, 0));
    if (prm.flags & CM_F_VERTEX_NORMALS) vertex_normals(x, false, false);
    sphere_on = (prm.flags & CM_F_SPHERE) != 0;
    sphere_c = mk(prm.sphere_center[0], prm.sphere_center[1], prm.sphere_center[2]);
    sphere_r = prm.sphere_radius;
  }

  // ---- stage 1: prediction (Physics/CudaPBD.cu:610-630, 1992-2012, 411-458) ----
  template <class T>
  void predict(std::vector<V3<T>>& xs, std::vector<V3<T>>& vs, std::vector<V3<T>>& ps) const {
    const T dt = T(prm.timestep);
    const V3<T> g = mk<T>(prm.gravity[0], prm.gravity[1], prm.gravity[2]);
    const V3<T> dtg = mul(g, dt);
    for (int i = 0; i < n; ++i) vs[i] = vs[i] + dtg;
    if (prm.flags & CM_F_DAMPING) damp<T>(ps, vs);
    for (int i = 0; i < n; ++i) ps[i] = xs[i] + mul(vs[i], dt);
  }

  // Physics/CudaPBD.cu:411-458 with the sums in the tree order above; ps are the
  // positions the reference passes (m_dPositions from the previous substep).
  template <class T>
  void damp(const std::vector<V3<T>>& ps, std::vector<V3<T>>& vs) const {
    const T kg = T(prm.global_vel_damp);
    for (int i = 0; i < n; ++i) vs[i] = mul(vs[i], T(1) - kg);
    std::vector<T> m(n);
    for (int i = 0; i < n; ++i) m[i] = T(mass[i]);
    T msum = tree_sum<T>(n, [&](int i) { return m[i]; });
    V3<T> xm = tree_sum<V3<T>>(n, [&](int i) { return mul(ps[i], m[i]); });
    V3<T> vm = tree_sum<V3<T>>(n, [&](int i) { return mul(vs[i], m[i]); });
    V3<T> xcm = divs(xm, msum), vcm = divs(vm, msum);
    V3<T> L = tree_sum<V3<T>>(n, [&](int i) { return cross(ps[i] - xcm, mul(vs[i], m[i])); });
    V3<T> i1 = tree_sum<V3<T>>(n, [&](int i) {
      V3<T> r = ps[i] - xcm;
      return mul(mk(r.x * r.x, r.y * r.y, r.z * r.z), m[i]);
    });
    V3<T> i2 = tree_sum<V3<T>>(n, [&](int i) {
      V3<T> r = ps[i] - xcm;
      return mul(-mk(r.x * r.y, r.x * r.z, r.y * r.z), m[i]);
    });
    // symmetric inertia tensor, columns as Physics/CudaPBD.cu:447-449
    T a00 = i1.y + i1.z, a01 = i2.x, a02 = i2.y;
    T a11 = i1.x + i1.z, a12 = i2.z, a22 = i1.x + i1.y;
    // m[c][r]: column-major cofactor inverse of the 3x3 matrix
    T m00 = a00, m01 = a01, m02 = a02, m10 = a01, m11 = a11, m12 = a12, m20 = a02, m21 = a12,
      m22 = a22;
    T inv_det = T(1) / (+m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02) +
                        m20 * (m01 * m12 - m11 * m02));
    T i00 = +(m11 * m22 - m21 * m12) * inv_det, i10 = -(m10 * m22 - m20 * m12) * inv_det,
      i20 = +(m10 * m21 - m20 * m11) * inv_det, i01 = -(m01 * m22 - m21 * m02) * inv_det,
      i11 = +(m00 * m22 - m20 * m02) * inv_det, i21 = -(m00 * m21 - m20 * m01) * inv_det,
      i02 = +(m01 * m12 - m11 * m02) * inv_det, i12 = -(m00 * m12 - m10 * m02) * inv_det,
      i22 = +(m00 * m11 - m10 * m01) * inv_det;
    // omega = Inverse * L (column-major matrix times vector)
    V3<T> om = mk(i00 * L.x + i10 * L.y + i20 * L.z, i01 * L.x + i11 * L.y + i21 * L.z,
                  i02 * L.x + i12 * L.y + i22 * L.z);
    const T kd = T(prm.k_damp);
    for (int i = 0; i < n; ++i) {
      V3<T> r = ps[i] - xcm;
      V3<T> dv = vcm + cross(om, r) - vs[i];
      vs[i] = vs[i] + mul(dv, kd);
    }
  }

  // ---- stages 2-5: collision pipeline on FP32 state (xf prev, pf predicted, vel) ----
  void collide(const std::vector<F3>& xf, const std::vector<F3>& pf, const std::vector<F3>& vel) {
    pairs.clear();
    vf_clear();
    if (!(prm.flags & CM_F_SELF_COLLISION) || t < 2) return;
    const float th = prm.thickness;
    const int nin = t - 1;
    // stage 2: swept leaf AABBs + keys in current triangle order
    // (Physics/CudaDeformableUtils.cu:120-151)
    std::vector<F3> cen(t);
    CM_OMP_FOR
    for (int i = 0; i < t; ++i) {
      const Tri& tr = tri[i];
      F3 mn = subs(min3(pf[tr.v1], pf[tr.v2], pf[tr.v3]), th);
      F3 mx = adds(max3(pf[tr.v1], pf[tr.v2], pf[tr.v3]), th);
      F3 pmn = subs(min3(xf[tr.v1], xf[tr.v2], xf[tr.v3]), th);
      F3 pmx = adds(max3(xf[tr.v1], xf[tr.v2], xf[tr.v3]), th);
      F3 smn = vmin(mn, pmn), smx = vmax(mx, pmx);
      amin[nin + i] = smn;
      amax[nin + i] = smx;
      cen[i] = smn + divs(smx - smn, 2.f);
    }
    compute_keys(cen);
    // stage 3: stable sort of (key, triangle) (Core/CubWrappers.cuh:141-156)
    for (int i = 0; i < t; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::vector<uint64_t> k2(t);
    std::vector<Tri> t2(t);
    for (int i = 0; i < t; ++i) {
      k2[i] = keys[perm[i]];
      t2[i] = tri[perm[i]];
    }
    keys.swap(k2);
    tri.swap(t2);
    if (!(prm.ref_compat & CM_RC_STALE_LEAF_AABB)) {
      std::vector<F3> lm(t), lx(t);
      for (int i = 0; i < t; ++i) {
        lm[i] = amin[nin + perm[i]];
        lx[i] = amax[nin + perm[i]];
      }
      for (int i = 0; i < t; ++i) {
        amin[nin + i] = lm[i];
        amax[nin + i] = lx[i];
      }
    }
    lbvh_build(t, keys.data(), (prm.ref_compat & CM_RC_MIDPOINT_SPLIT) != 0, lefts.data(),
               rights.data(), parents.data(), &warn);
    lbvh_refit(t, lefts.data(), rights.data(), amin.data(), amax.data(), rmll.data(), rmlr.data());
    // stage 4: traversal -> pairs (Collision/CudaBVH.cu:23-113), compaction in leaf order
    const int cap = std::min(prm.pair_cap, t);  // Physics/CudaDeformableBody.cpp:233
    {
      // per leaf in parallel, concatenated in leaf order
      std::vector<int32_t> buf((size_t)t * 2 * cap);
      std::vector<int> cnt(t);
      uint32_t w = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : w)
      for (int i = 0; i < t; ++i) {
        bool capped = false;
        uint32_t wl = 0;
        cnt[i] = lbvh_leaf_pairs(t, i, lefts.data(), rights.data(), amin.data(), amax.data(),
                                 rmll.data(), rmlr.data(), cap, buf.data() + (size_t)i * 2 * cap,
                                 &wl, &capped);
        w |= wl | (capped ? (uint32_t)CM_WARN_PAIR_CAP : 0u);
      }
      warn |= w;
      for (int i = 0; i < t; ++i)
        pairs.insert(pairs.end(), buf.begin() + (size_t)i * 2 * cap,
                     buf.begin() + (size_t)i * 2 * cap + 2 * cnt[i]);
    }
    // stage 5: candidates, feature filter, CCD, compaction
    // (Collision/CudaCollisionDetection.cu:6-189, 223-283); threads take contiguous
    // pair ranges and their lists are joined in range order
    const size_t P = pairs.size() / 2;
    const bool ee_on = (prm.flags & CM_F_EE_CONTACTS) != 0;
    const int nth = omp_threads();
    std::vector<std::vector<cm_contact>> pvf(nth), pee(nth);
#pragma omp parallel num_threads(nth)
    {
      const int id = omp_tid(), nt = omp_nthreads();
      const size_t lo = P * id / nt, hi = P * (id + 1) / nt;
      for (size_t k = lo; k < hi; ++k) {
        const Tri& t1 = tri[pairs[2 * k]];
        const Tri& tt2 = tri[pairs[2 * k + 1]];
        for (int s = 0; s < 6; ++s) {
          cm_contact c;
          if (candidate(t1, tt2, s, c) && feature_filter(c, true, pf.data(), xf.data(), th))
            pvf[id].push_back(c);
        }
      }
      if (ee_on)
        for (size_t k = lo; k < hi; ++k) {
          const Tri& t1 = tri[pairs[2 * k]];
          const Tri& tt2 = tri[pairs[2 * k + 1]];
          for (int s = 6; s < 15; ++s) {
            cm_contact c;
            if (candidate(t1, tt2, s, c) && feature_filter(c, false, pf.data(), xf.data(), th))
              pee[id].push_back(c);
          }
        }
    }
    std::vector<cm_contact>& vfc = cand_vf;
    std::vector<cm_contact>& eec = cand_ee;
    vfc.clear();
    eec.clear();
    for (int id = 0; id < nth; ++id) {
      vfc.insert(vfc.end(), pvf[id].begin(), pvf[id].end());
      eec.insert(eec.end(), pee[id].begin(), pee[id].end());
    }
    const float dt = prm.timestep;
    const bool dcd = (prm.flags & CM_F_DCD) != 0;  // HandleCollisionsDiscrete's tests
    std::vector<uint8_t> hvf(vfc.size()), hee(eec.size());
#pragma omp parallel for schedule(dynamic, 512)
    for (long k = 0; k < (long)vfc.size(); ++k)
      hvf[k] = dcd ? dcd_vf(vfc[k], xf.data(), th) : ccd_vf(vfc[k], xf.data(), vel.data(), th, dt);
#pragma omp parallel for schedule(dynamic, 512)
    for (long k = 0; k < (long)eec.size(); ++k)
      hee[k] = dcd ? dcd_ee(eec[k], xf.data(), th) : ccd_ee(eec[k], xf.data(), vel.data(), th, dt);
    for (size_t k = 0; k < vfc.size(); ++k)
      if (hvf[k]) vf.push_back(vfc[k]);
    for (size_t k = 0; k < eec.size(); ++k)
      if (hee[k]) ee.push_back(eec[k]);
  }
  void vf_clear() {
    vf.clear();
    ee.clear();
    cand_vf.clear();
    cand_ee.clear();
  }

  void compute_keys(const std::vector<F3>& cen) {
    if (prm.morton_bits == 30) {
      F3 lo = cen[0], hi = cen[0];
      if (prm.morton_normalize) {
        for (int i = 1; i < t; ++i) {
          lo = vmin(lo, cen[i]);
          hi = vmax(hi, cen[i]);
        }
      } else {
        lo = F3{0, 0, 0};
        hi = F3{1, 1, 1};
      }
      CM_OMP_FOR
      for (int i = 0; i < t; ++i) keys[i] = morton30(cen[i], lo, hi);
    } else {
      CM_OMP_FOR
      for (int i = 0; i < t; ++i) keys[i] = morton63(cen[i]);
    }
  }

  // ---- stage 6 helpers: constraint projections (Physics/CudaPBD.cu:986-1415) ----
  void stretch(int s, const std::vector<RV>& ps, RV& c1, RV& c2) const {
    const int a = tp.sid1[s], b = tp.sid2[s];
    const R w1 = w[a], w2 = w[b], ws = w1 + w2;
    c1 = c2 = mk<R>(0, 0, 0);
    if (ws == 0) return;
    RV diff = ps[a] - ps[b];
    R l = len(diff);
    if (l <= R(1e-8f)) return;
    RV nn = divs(diff, l);
    RV tt = divs(mul(mul(nn, l - R(tp.sl0[s])), R(tp.skp[s])), ws);
    c1 = mul(tt, -w1);
    c2 = mul(tt, w2);
  }
  void bend(int bi, const std::vector<RV>& ps, RV c[4]) const {
    if ((prm.flags & CM_F_BEND_FP64) && sizeof(R) < sizeof(double)) {
      V3<double> cd[4];
      bend_impl<double>(bi, ps, cd);
      for (int r = 0; r < 4; ++r) c[r] = cvt<R>(cd[r]);
    } else {
      bend_impl<R>(bi, ps, c);
    }
  }
  // Physics/CudaPBD.cu:1024-1147 evaluated in arithmetic T on the state's values
  // v / s per component: FP32 divides like the reference; the FP64 form (CM_F_BEND_FP64,
  // this project's own precision choice) multiplies by one reciprocal per length.
  template <class T>
  static V3<T> bdiv(V3<T> v, T s) {
    if constexpr (std::is_same<T, double>::value) return mul(v, T(1) / s);
    else return divs(v, s);
  }
  template <class T>
  void bend_impl(int bi, const std::vector<RV>& ps, V3<T> c[4]) const {
    typedef V3<T> TV;
    for (int r = 0; r < 4; ++r) c[r] = mk<T>(0, 0, 0);
    const int i1 = tp.bid[0][bi], i2 = tp.bid[1][bi], i3 = tp.bid[2][bi], i4 = tp.bid[3][bi];
    const T kp = T(tp.bkp[bi]), phi0 = T(tp.bphi0[bi]);
    const T w1 = T(w[i1]), w2 = T(w[i2]), w3 = T(w[i3]), w4 = T(w[i4]);
    TV p2 = cvt<T>(ps[i2]), p3 = cvt<T>(ps[i3]), p4 = cvt<T>(ps[i4]);
    if (!(prm.ref_compat & CM_RC_BEND_ABS)) {
      TV p1 = cvt<T>(ps[i1]);
      p2 = p2 - p1;
      p3 = p3 - p1;
      p4 = p4 - p1;
    }
    TV c23 = cross(p2, p3), c24 = cross(p2, p4);
    T l23 = len(c23), l24 = len(c24);
    const T eps = T(1e-8f);
    if (l23 <= eps || l24 <= eps) return;
    TV n1 = bdiv(c23, l23), n2 = bdiv(c24, l24);
    T d = fclamp(dot(n1, n2), T(-1), T(1));
    T phi = std::acos(d);
    const T PIf = T(3.14159265359f);
    if (d == T(-1)) {
      phi = PIf;
      // the n/100 push of :1077-1078 only matters when phi0 != pi, i.e. outside the
      // reference's own rest angle; the p1-relative variant takes the limit of the
      // smooth formula there (sqrt(1 - d^2) = 0: no correction)
      if (phi == phi0 || !(prm.ref_compat & CM_RC_BEND_ABS)) return;
      c[0] = divs(n1, T(100));
      c[1] = divs(n2, T(100));
      return;
    }
    if (d == T(1)) {
      phi = 0;
      if (phi == phi0) return;
    }
    TV q3 = bdiv(cross(p2, n2) + mul(cross(n1, p2), d), l23);
    TV q4 = bdiv(cross(p2, n1) + mul(cross(n2, p2), d), l24);
    TV q2 = -bdiv(cross(p3, n2) + mul(cross(n1, p3), d), l23) -
            bdiv(cross(p4, n1) + mul(cross(n2, p4), d), l24);
    TV q1 = -q2 - q3 - q4;
    T term = std::sqrt(T(1) - d * d) * (phi - phi0) * kp;
    T s1 = dot(q1, q1), s2 = dot(q2, q2), s3 = dot(q3, q3), s4 = dot(q4, q4);
    T sf = w1 * s1 + w2 * s2 + w3 * s3 + w4 * s4;
    if (sf <= eps) return;
    T ts = term / sf;
    c[0] = mul(q1, -(w1 * ts));
    c[1] = mul(q2, -(w2 * ts));
    c[2] = mul(q3, -(w3 * ts));
    c[3] = mul(q4, -(w4 * ts));
  }
  // Contact projections (Physics/CudaPBD.cu:1212-1311 VF, :1313-1415 EE) on the
  // contact's four positions p and inverse masses w. act < 0: the constraint is
  // active when its value cc = |n| - thickness is negative (`s = c < 0 ? s : 0`);
  // act = 0 / 1 forces the decision (replay of another precision's decisions).
  // Returns whether it was active.
  static bool contact_vf_core(const cm_contact& k, const RV p4[4], const R w4[4], R th, RV c[4],
                              int act = -1) {
    R b0 = fclamp(R(-k.w[1]), R(0), R(1)), b1 = fclamp(R(-k.w[2]), R(0), R(1)),
      b2 = fclamp(R(-k.w[3]), R(0), R(1));
    RV pp = p4[0], p0 = p4[1], p1 = p4[2], p2 = p4[3];
    R ww = w4[0], w0 = w4[1], w1 = w4[2], w2 = w4[3];
    RV q = mul(p0, b0) + mul(p1, b1) + mul(p2, b2);
    RV nn = pp - q;
    R l = len(nn);
    nn = divs(nn, l);
    R cc = l - th;
    RV g = nn, g0 = mul(-nn, b0), g1 = mul(-nn, b1), g2 = mul(-nn, b2);
    R s = ww + w0 * b0 * b0 + w1 * b1 * b1 + w2 * b2 * b2;
    for (int r = 0; r < 4; ++r) c[r] = mk<R>(0, 0, 0);
    const bool on = act < 0 ? cc < 0 : act > 0;
    if (s == 0) return on;
    s = cc / s;
    s = on ? s : R(0);
    c[0] = mul(g, -s * ww);
    c[1] = mul(g0, -s * w0);
    c[2] = mul(g1, -s * w1);
    c[3] = mul(g2, -s * w2);
    return on;
  }
  static bool contact_ee_core(const cm_contact& k, const RV p4[4], const R w4[4], R th, RV c[4],
                              int act = -1) {
    R b0 = R(k.w[0]), b1 = R(k.w[1]), b2 = R(-k.w[2]), b3 = R(-k.w[3]);
    RV p0 = p4[0], p1 = p4[1], p2 = p4[2], p3 = p4[3];
    R w0 = w4[0], w1 = w4[1], w2 = w4[2], w3 = w4[3];
    RV q0 = mul(p0, b0) + mul(p1, b1), q1 = mul(p2, b2) + mul(p3, b3);
    RV nn = q0 - q1;
    R l = len(nn);
    nn = divs(nn, l);
    R cc = l - th;
    RV g0 = mul(nn, b0), g1 = mul(nn, b1), g2 = mul(-nn, b2), g3 = mul(-nn, b3);
    R s = w0 * b0 * b0 + w1 * b1 * b1 + w2 * b2 * b2 + w3 * b3 * b3;
    for (int r = 0; r < 4; ++r) c[r] = mk<R>(0, 0, 0);
    const bool on = act < 0 ? cc < 0 : act > 0;
    if (s == 0) return on;
    s = cc / s;
    s = on ? s : R(0);
    c[0] = mul(g0, -s * w0);
    c[1] = mul(g1, -s * w1);
    c[2] = mul(g2, -s * w2);
    c[3] = mul(g3, -s * w3);
    return on;
  }
  bool contact_vf(const cm_contact& k, const std::vector<RV>& ps, RV c[4], int act = -1) const {
    RV p4[4];
    R w4[4];
    for (int r = 0; r < 4; ++r) {
      p4[r] = ps[k.v[r]];
      w4[r] = w[k.v[r]];
    }
    return contact_vf_core(k, p4, w4, R(prm.thickness), c, act);
  }
  bool contact_ee(const cm_contact& k, const std::vector<RV>& ps, RV c[4], int act = -1) const {
    RV p4[4];
    R w4[4];
    for (int r = 0; r < 4; ++r) {
      p4[r] = ps[k.v[r]];
      w4[r] = w[k.v[r]];
    }
    return contact_ee_core(k, p4, w4, R(prm.thickness), c, act);
  }

  // Per-vertex sum of the dynamic corrections (Physics/CudaPBD.cu:711-720 sort by
  // vertex + run-length encoding, then _ApplyDynamicCorrections :1448-1475): the
  // records (vertex rec_id[k], value rec[k]) are visited per vertex in record order;
  // only non-zero values are added and counted.
  static void accumulate_records(const std::vector<uint32_t>& rec_id,
                                 const std::vector<uint32_t>& order, const std::vector<RV>& rec,
                                 std::vector<RV>& acc, std::vector<int>& cnt) {
    size_t k = 0;
    while (k < order.size()) {
      uint32_t vid = rec_id[order[k]];
      RV a = mk<R>(0, 0, 0);
      int c = 0;
      size_t e = k;
      while (e < order.size() && rec_id[order[e]] == vid) {
        const RV& val = rec[order[e]];
        if (!is_zero(val)) {
          a = a + val;
          ++c;
        }
        ++e;
      }
      acc[vid] = acc[vid] + a;
      cnt[vid] += c;
      k = e;
    }
  }
  static void record_order(const std::vector<uint32_t>& rec_id, std::vector<uint32_t>& order) {
    order.resize(rec_id.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return rec_id[a] < rec_id[b]; });
  }

  // Long-range attachments of one vertex (Physics/CudaPBD.cu:1656-1688): for every
  // flagged pinned vertex j, -(|p - p_j| - l0_j) along (p - p_j)/|p - p_j|, one count each.
  static void lra_vertex(RV pi, int F, const int32_t* fixed_ids, const float* l0,
                         const uint8_t* flags, const std::vector<RV>& ps, RV& a, int& c) {
    a = mk<R>(0, 0, 0);
    c = 0;
    for (int j = 0; j < F; ++j)
      if (flags[j]) {
        RV diff = pi - ps[fixed_ids[j]];
        R l = len(diff);
        RV nn = divs(diff, l);
        a = a + mul(nn, -(l - R(l0[j])));
        ++c;
      }
  }

  // Core/.../CudaDeformableUtils RaySphere (Collision/CudaCollisionDetection.cu:529-567)
  static int ray_sphere(RV a, RV b, RV c, R rad, RV& qc) {
    RV dir = b - a;
    if (is_zero(dir)) return 0;
    R l = len(dir);
    dir = divs(dir, l);
    RV m = a - c;
    R bb = dot(m, dir);
    R cc = dot(m, m) - rad * rad;
    if (cc > R(0) && bb > R(0)) return 0;
    R del = bb * bb - cc;
    if (del < 0) return 0;
    R tt = -bb - std::sqrt(del);
    if (tt < R(0)) {
      qc = a + mul(dir, tt);
      return -1;
    }
    if (tt > l) return 0;
    qc = a + mul(dir, tt);
    return 1;
  }

  // ---- stage 6: solver (Physics/CudaPBD.cu:632-660) ----
  void solve() {
    const R th = R(prm.thickness);
    // ground (Physics/CudaPBD.cu:1477-1507)
    gflag.assign(n, 0);
    std::vector<RV> gqc(n), gnc(n);
    if (prm.flags & CM_F_GROUND) {
      const R gh = R(prm.ground_height);
      for (int i = 0; i < n; ++i) {
        RV pos = p[i], prev = x[i];
        if (pos.y <= gh + th) {
          RV nc = mk<R>(0, 1, 0);
          RV p0 = mk<R>(1, gh + th, 1);
          RV l = pos - prev;
          R d = dot(p0 - pos, nc) / dot(l, nc);
          gqc[i] = mul(l, d) + pos;
          gnc[i] = nc;
          gflag[i] = 1;
        }
      }
    }
    // LRA (Physics/CudaPBD.cu:1571-1592)
    const int F = (int)tp.lra_fixed.size();
    std::vector<uint8_t> lflag;
    if ((prm.flags & CM_F_LRA) && F) {
      lflag.assign((size_t)n * F, 0);
      for (int i = 0; i < n; ++i) {
        if (pinned[i]) continue;
        for (int j = 0; j < F; ++j)
          if (len(p[i] - p[tp.lra_fixed[j]]) > R(tp.lra_l0[(size_t)i * F + j])) lflag[(size_t)i * F + j] = 1;
      }
    }
    // sphere (Physics/CudaPBD.cu:1509-1569)
    if (sphere_on) {
      const RV sc = cvt<R>(sphere_c);
      const R rad = R(float(sphere_r + prm.thickness));
      for (int i = 0; i < n; ++i) {
        RV qc{0, 0, 0};
        int ret = ray_sphere(x[i], p[i], sc, rad, qc);
        if (ret == -1) {
          RV q = sc + mul(normalize(qc - sc), rad);
          RV nc = normalize(q - sc);
          x[i] = q + mul(nc, R(0.0001f));
          p[i] = q;
        } else if (ret == 1) {
          RV nc = normalize(qc - sc);
          p[i] = qc + mul(nc, R(0.0001f));
        }
      }
    }
    // dynamic records sorted stably by vertex (Physics/CudaPBD.cu:669-725)
    const size_t C = vf.size() + ee.size();
    std::vector<uint32_t> rec_id(4 * C), order(4 * C);
    for (size_t c = 0; c < C; ++c) {
      const cm_contact& k = c < vf.size() ? vf[c] : ee[c - vf.size()];
      for (int r = 0; r < 4; ++r) rec_id[4 * c + r] = (uint32_t)k.v[r];
    }
    record_order(rec_id, order);
    if (act_mode == 2 && act.size() != (size_t)prm.solver_iterations * C) {
      act_mismatch = true;  // decisions of another contact set: evaluate them here
      act_mode = 0;
    }
    // a recorded substep keeps the decisions it took (evaluated ones are recorded)
    const int saved_mode = act_mode;
    if (act_mode == 1 || (act_mode == 0 && record)) {
      act.assign((size_t)prm.solver_iterations * C, 0);
      act_mode = 1;
    }
    const int S = (int)tp.sid1.size(), B = (int)tp.bkp.size();
    std::vector<RV> slots(tp.prefix[n]);
    std::vector<RV> acc(n);
    std::vector<int> cnt(n);
    std::vector<RV> rec(4 * C);
    std::vector<std::vector<RV>>* pk_out = record ? &tape.back().pk : nullptr;
    for (int it = 0; it < prm.solver_iterations; ++it) {
      if (pk_out) pk_out->push_back(p);
      std::fill(acc.begin(), acc.end(), mk<R>(0, 0, 0));
      std::fill(cnt.begin(), cnt.end(), 0);
      // ground corrections (Physics/CudaPBD.cu:1594-1629)
      for (int i = 0; i < n; ++i)
        if (gflag[i]) {
          R cv = dot(p[i] - gqc[i], gnc[i]);
          if (cv < 0) {
            acc[i] = acc[i] + mul(gnc[i], -cv);
            cnt[i] += 1;
          }
        }
      // LRA corrections (Physics/CudaPBD.cu:1656-1688)
      if (!lflag.empty())
        for (int i = 0; i < n; ++i) {
          RV a;
          int c;
          lra_vertex(p[i], F, tp.lra_fixed.data(), tp.lra_l0.data() + (size_t)i * F,
                     lflag.data() + (size_t)i * F, p, a, c);
          acc[i] = acc[i] + a;
          cnt[i] += c;
        }
      // static constraints into slots, then per-vertex gather (:942-962, :1417-1446)
      CM_OMP_FOR
      for (int s = 0; s < S; ++s) {
        RV c1, c2;
        stretch(s, p, c1, c2);
        slots[tp.soff[0][s]] = c1;
        slots[tp.soff[1][s]] = c2;
      }
      CM_OMP_FOR
      for (int b = 0; b < B; ++b) {
        RV c[4];
        bend(b, p, c);
        for (int r = 0; r < 4; ++r) slots[tp.boff[r][b]] = c[r];
      }
      CM_OMP_FOR
      for (int i = 0; i < n; ++i) {
        RV a = mk<R>(0, 0, 0);
        for (int k = tp.prefix[i]; k < tp.prefix[i + 1]; ++k) a = a + slots[k];
        acc[i] = acc[i] + a;
        cnt[i] += tp.prefix[i + 1] - tp.prefix[i];
      }
      // dynamic constraints (:1190-1415, :1448-1475)
      if (C) {
        uint8_t* act_it = act_mode ? act.data() + (size_t)it * C : nullptr;
        CM_OMP_FOR
        for (long c = 0; c < (long)C; ++c) {
          RV cc[4];
          const int force = act_mode == 2 ? (int)act_it[c] : -1;
          bool on;
          if ((size_t)c < vf.size())
            on = contact_vf(vf[c], p, cc, force);
          else
            on = contact_ee(ee[c - vf.size()], p, cc, force);
          if (act_mode == 1) act_it[c] = on ? 1 : 0;
          for (int r = 0; r < 4; ++r) rec[4 * c + r] = cc[r];
        }
        accumulate_records(rec_id, order, rec, acc, cnt);
      }
      // FinalCorrectionStep (:1690-1698)
      CM_OMP_FOR
      for (int i = 0; i < n; ++i) p[i] = p[i] + divs(acc[i], R(cnt[i]));
    }
    if (pk_out) pk_out->push_back(p);
    if (record) tape.back().act = act;
    act_mode = saved_mode;
  }

  // ---- stage 7: integration + world clamp (Physics/CudaPBD.cu:1916-1932,
  // Physics/CudaDeformableUtils.cu:260-269) ----
  void integrate() {
    const R dt = R(prm.timestep);
    for (int i = 0; i < n; ++i) {
      if (pinned[i]) {
        v[i] = mk<R>(0, 0, 0);
        continue;
      }
      v[i] = divs(p[i] - x[i], dt);
      x[i] = p[i];
    }
    if (prm.flags & CM_F_WORLD_CLAMP) {
      const RV lo = mk<R>(prm.world_min[0], prm.world_min[1], prm.world_min[2]);
      const RV hi = mk<R>(prm.world_max[0], prm.world_max[1], prm.world_max[2]);
      for (int i = 0; i < n; ++i) {
        p[i].x = std::fmin(std::fmax(p[i].x, lo.x), hi.x);
        p[i].y = std::fmin(std::fmax(p[i].y, lo.y), hi.y);
        p[i].z = std::fmin(std::fmax(p[i].z, lo.z), hi.z);
      }
    }
  }

  // ---- time-of-contact rollback (Physics/CudaPBD.cu:727-768, FillTOC :2014-2046,
  // _RevertToTOC :2048-2072; its call at Physics/CudaDeformableBody.cpp:583 is
  // commented out) ----
  // A vertex in contacts takes the smallest time t of its contacts (minimum started
  // at 100) and moves to (1 - t) x + t p (Core/CudaUtils.cu:396-399).
  void revert_toc() {
    std::vector<float> tmin(n, 100.f);
    std::vector<uint8_t> hit(n, 0);
    for (const std::vector<cm_contact>* cs : {&vf, &ee})
      for (const cm_contact& c : *cs)
        for (int k = 0; k < 4; ++k) {
          tmin[c.v[k]] = std::fmin(tmin[c.v[k]], c.t);
          hit[c.v[k]] = 1;
        }
    for (int i = 0; i < n; ++i)
      if (hit[i]) p[i] = mul(x[i], R(1) - R(tmin[i])) + mul(p[i], R(tmin[i]));
  }

  // ---- vertex normals (Physics/CudaDeformableUtils.cu:190-249) ----
  // Each triangle's face normal (CudaUtils::FaceNormal = cross(b-a, c-a)), a unit one
  // in a substep (:135) and the raw cross product at creation (:54, used by
  // Physics/CudaDeformableBody.cpp:269), is added to its three vertices' normals,
  // which are then normalised. The reference adds with float atomics onto the
  // previous substep's normal (never cleared); here the previous normal (carry) comes
  // first, then the triangles in ascending input index.
  void vertex_normals(const std::vector<RV>& pos, bool carry, bool unit = true) {
    std::vector<RV> acc(n, mk<R>(0, 0, 0));
    if (carry) acc = vnrm;
    for (size_t j = 0; j < tp.tri.size(); ++j) {
      const Tri& q = tp.tri[j];
      const RV cr = cross(pos[q.v2] - pos[q.v1], pos[q.v3] - pos[q.v1]);
      const RV fn = unit ? normalize(cr) : cr;
      acc[q.v1] = acc[q.v1] + fn;
      acc[q.v2] = acc[q.v2] + fn;
      acc[q.v3] = acc[q.v3] + fn;
    }
    for (int i = 0; i < n; ++i) vnrm[i] = normalize(acc[i]);
  }

  // ---- stage 8: friction on the integrated velocities (Physics/CudaPBD.cu:770-833,
  // called at Physics/CudaDeformableBody.cpp:596-602) ----
  void friction() {
    const R kf = R(prm.k_friction);
    // _DampVelRigid (Physics/CudaPBD.cu:1879-1912) on the ground flags; its launch of
    // 255-thread blocks over ceil(N/256) blocks (:778) stops short of the last
    // ceil(N/256) vertices. Its sphere branch reads flags the sphere set-up never sets
    // (:1531, :1538).
    const int rigid_n =
        (prm.ref_compat & CM_RC_RIGID_FRICTION_TAIL) ? std::min(n, (n + 255) / 256 * 255) : n;
    const RV up = mk<R>(0, 1, 0);
    for (int i = 0; i < rigid_n; ++i) {
      if (gflag.empty() || !gflag[i]) continue;
      const RV vn = mul(up, dot(v[i], up));  // CudaUtils::tangentNormal
      const RV a = (v[i] - vn) + mul(vn, R(1) - kf);
      v[i] = a - mul(up, R(2) * dot(a, up));  // CudaUtils::reflect
    }
    // _ApplyFriction (:1725-1743): the contact normal once per contact vertex, VF
    // contacts then EE; sorted by vertex (stable) and summed per vertex in entry
    // order (:1846-1877). The EE entries' vertex ids are the contact's own here; the
    // reference leaves them unwritten (:1810-1813).
    std::vector<RV> acc(n, mk<R>(0, 0, 0));
    std::vector<int> cnt(n, 0);
    for (const std::vector<cm_contact>* cs : {&vf, &ee})
      for (const cm_contact& c : *cs)
        for (int k = 0; k < 4; ++k) {
          acc[c.v[k]] = acc[c.v[k]] + mk<R>(c.n[0], c.n[1], c.n[2]);
          ++cnt[c.v[k]];
        }
    // FinalFrictionStep (:1700-1723)
    for (int i = 0; i < n; ++i) {
      const RV s = acc[i];
      if (cnt[i] == 0 || (s.x == R(0) && s.y == R(0) && s.z == R(0))) continue;
      const RV fn = normalize(s);
      const RV vn = mul(fn, dot(v[i], fn));
      const RV a = mul(vn, R(std::pow(R(1) - kf, R(cnt[i])))) + (v[i] - vn);
      v[i] = a - mul(fn, R(2) * dot(a, fn));
    }
  }

  // One substep (Physics/CudaDeformableBody.cpp:550-627).
  void step() {
    if (record) {
      tape.emplace_back();
      tape.back().x = x;
      tape.back().v = v;
    }
    std::vector<F3> xf(n), pf(n), vf_(n);
    for (int i = 0; i < n; ++i) {
      xf[i] = cvt<float>(x[i]);
      pf[i] = cvt<float>(p[i]);
      vf_[i] = cvt<float>(v[i]);
    }
    predict<R>(x, v, p);
    if (prm.flags & CM_F_VERTEX_NORMALS) vertex_normals(p, true);
    if (sizeof(R) == sizeof(float)) {
      for (int i = 0; i < n; ++i) {
        pf[i] = cvt<float>(p[i]);
        vf_[i] = cvt<float>(v[i]);
      }
    } else {
      predict<float>(xf, vf_, pf);  // FP32 shadow for the discrete stages
    }
    collide(xf, pf, vf_);
    if (record) {
      tape.back().vf = vf;
      tape.back().ee = ee;
    }
    if (prm.flags & CM_F_REVERT_TOC) revert_toc();
    solve();
    integrate();
    if (prm.flags & CM_F_FRICTION) friction();
  }

  // ---- adjoint of the recorded rollout (no reference counterpart) ----
  // The differentiable substep: v' = v + dt g, p0 = x + dt v', the Jacobi iterations
  // over stretch, bend and the substep's (frozen) VF / EE contacts, then x = p, v =
  // (p - x) / dt for free vertices (pinned: x kept, v = 0). Ground, LRA, sphere,
  // damping and the world clamp are not part of it. Counts are piecewise constant.
  // Vector-Jacobian products below are written out by hand in R.
  void stretch_vjp(int s, const std::vector<RV>& ps, const std::vector<RV>& a,
                   std::vector<RV>& la, R& gkp) const {
    const int ia = tp.sid1[s], ib = tp.sid2[s];
    const R w1 = w[ia], w2 = w[ib], ws = w1 + w2;
    if (ws == 0) return;
    const RV d = ps[ia] - ps[ib];
    const R l = len(d);
    if (l <= R(1e-8f)) return;
    const RV n = divs(d, l);
    const R l0 = R(tp.sl0[s]), kp = R(tp.skp[s]);
    // c1 = -w1 f kp / ws, c2 = w2 f kp / ws with f = (l - l0) n
    const RV g = mul(a[ib], w2) - mul(a[ia], w1);  // dL/d(f kp / ws)
    const RV f = mul(n, l - l0);
    gkp += dot(g, f) / ws;
    const RV gf = mul(g, kp / ws);
    // df/dd = I - (l0 / l)(I - n n^T)
    const RV gd = gf - mul(gf - mul(n, dot(n, gf)), l0 / l);
    la[ia] = la[ia] + gd;
    la[ib] = la[ib] - gd;
  }

  template <class T>
  void bend_vjp_impl(int bi, const std::vector<RV>& ps, const std::vector<RV>& a,
                     std::vector<RV>& la, R& gkp) const {
    typedef V3<T> TV;
    const int id[4] = {tp.bid[0][bi], tp.bid[1][bi], tp.bid[2][bi], tp.bid[3][bi]};
    const T kp = T(tp.bkp[bi]), phi0 = T(tp.bphi0[bi]);
    const T wv[4] = {T(w[id[0]]), T(w[id[1]]), T(w[id[2]]), T(w[id[3]])};
    const bool abs_mode = (prm.ref_compat & CM_RC_BEND_ABS) != 0;
    TV P2 = cvt<T>(ps[id[1]]), P3 = cvt<T>(ps[id[2]]), P4 = cvt<T>(ps[id[3]]);
    if (!abs_mode) {
      const TV p1 = cvt<T>(ps[id[0]]);
      P2 = P2 - p1;
      P3 = P3 - p1;
      P4 = P4 - p1;
    }
    const TV c23 = cross(P2, P3), c24 = cross(P2, P4);
    const T l23 = len(c23), l24 = len(c24);
    const T eps = T(1e-8f);
    if (l23 <= eps || l24 <= eps) return;
    // n1, n2 and the q_k as the forward evaluates them (bend_impl): near d = +-1 the
    // reverse sweep amplifies any difference in them
    const TV n1 = bdiv(c23, l23), n2 = bdiv(c24, l24);
    const T dd = dot(n1, n2);
    const T d = fclamp(dd, T(-1), T(1));
    if (d == T(-1) || d == T(1)) return;  // the branches with constant output
    const T phi = std::acos(d);
    const TV A = cross(P3, n2) + mul(cross(n1, P3), d), B = cross(P4, n1) + mul(cross(n2, P4), d);
    const TV C = cross(P2, n2) + mul(cross(n1, P2), d), D = cross(P2, n1) + mul(cross(n2, P2), d);
    const TV q3 = bdiv(C, l23), q4 = bdiv(D, l24);
    const TV q2 = -bdiv(A, l23) - bdiv(B, l24);
    const TV q1 = -q2 - q3 - q4;
    const TV q[4] = {q1, q2, q3, q4};
    const T r = std::sqrt(T(1) - d * d), e = phi - phi0;
    const T term = r * e * kp;
    T sf = 0;
    for (int k = 0; k < 4; ++k) sf += wv[k] * dot(q[k], q[k]);
    if (sf <= eps) return;
    const T ts = term / sf;
    // reverse sweep: c_k = -w_k ts q_k
    TV qb[4];
    T tsb = 0;
    for (int k = 0; k < 4; ++k) {
      const TV ak = cvt<T>(a[id[k]]);
      qb[k] = mul(ak, -wv[k] * ts);
      tsb -= wv[k] * dot(q[k], ak);
    }
    const T termb = tsb / sf, sfb = -tsb * term / (sf * sf);
    for (int k = 0; k < 4; ++k) qb[k] = qb[k] + mul(q[k], T(2) * wv[k] * sfb);
    gkp += R(termb * r * e);
    const T rb = termb * e * kp, eb = termb * r * kp;
    T db = -eb / r - rb * d / r;  // phi = acos(d), r = sqrt(1 - d^2)
    // q1 = -q2 - q3 - q4
    qb[1] = qb[1] - qb[0];
    qb[2] = qb[2] - qb[0];
    qb[3] = qb[3] - qb[0];
    // q2 = -A / l23 - B / l24, q3 = C / l23, q4 = D / l24
    const T i23 = T(1) / l23, i24 = T(1) / l24;
    const TV Ab = mul(qb[1], -i23), Bb = mul(qb[1], -i24);
    const TV Cb = mul(qb[2], i23), Db = mul(qb[3], i24);
    T l23b = (dot(qb[1], A) - dot(qb[2], C)) * i23 * i23;
    T l24b = (dot(qb[1], B) - dot(qb[3], D)) * i24 * i24;
    TV P2b = mk<T>(0, 0, 0), P3b = mk<T>(0, 0, 0), P4b = mk<T>(0, 0, 0);
    TV n1b = mk<T>(0, 0, 0), n2b = mk<T>(0, 0, 0);
    // X = U x N + (M x U) d, for the four (U, N, M) = (P3,n2,n1), (P4,n1,n2), (P2,n2,n1), (P2,n1,n2)
    auto back = [&](const TV& Xb, const TV& U, const TV& N, const TV& M, TV& Ub, TV& Nb, TV& Mb) {
      Ub = Ub + cross(N, Xb);
      Nb = Nb + cross(Xb, U);
      const TV t = mul(Xb, d);
      Mb = Mb + cross(U, t);
      Ub = Ub + cross(t, M);
      db += dot(Xb, cross(M, U));
    };
    back(Ab, P3, n2, n1, P3b, n2b, n1b);
    back(Bb, P4, n1, n2, P4b, n1b, n2b);
    back(Cb, P2, n2, n1, P2b, n2b, n1b);
    back(Db, P2, n1, n2, P2b, n1b, n2b);
    // d = clamp(n1 . n2) (interior here)
    n1b = n1b + mul(n2, db);
    n2b = n2b + mul(n1, db);
    // n = c / l, l = |c|
    l23b -= dot(n1b, c23) * i23 * i23;
    l24b -= dot(n2b, c24) * i24 * i24;
    const TV c23b = mul(n1b, i23) + mul(n1, l23b), c24b = mul(n2b, i24) + mul(n2, l24b);
    // c23 = P2 x P3, c24 = P2 x P4
    P2b = P2b + cross(P3, c23b) + cross(P4, c24b);
    P3b = P3b + cross(c23b, P2);
    P4b = P4b + cross(c24b, P2);
    la[id[1]] = la[id[1]] + cvt<R>(P2b);
    la[id[2]] = la[id[2]] + cvt<R>(P3b);
    la[id[3]] = la[id[3]] + cvt<R>(P4b);
    if (!abs_mode) la[id[0]] = la[id[0]] - cvt<R>(P2b + P3b + P4b);
  }
  void bend_vjp(int bi, const std::vector<RV>& ps, const std::vector<RV>& a, std::vector<RV>& la,
                R& gkp) const {
    if ((prm.flags & CM_F_BEND_FP64) && sizeof(R) < sizeof(double))
      bend_vjp_impl<double>(bi, ps, a, la, gkp);
    else
      bend_vjp_impl<R>(bi, ps, a, la, gkp);
  }

  // VF and EE contacts: c_r = k_r g(nn), g(nn) = nn - th nn / |nn| while |nn| < th,
  // with fixed weights; nn = sum_r m_r p_r. Roles whose record is zero were neither
  // summed nor counted, so their cotangents are left out.
  // active: the forward's activation decision for this contact and iteration (its
  // value cc = |nn| - thickness was negative, or the decision was replayed)
  void contact_vjp(const cm_contact& k, bool isvf, const std::vector<RV>& ps,
                   const std::vector<RV>& a, const bool used[4], bool active,
                   std::vector<RV>& la) const {
    R m[4], kr[4];
    const int* v = k.v;
    if (isvf) {
      const R b0 = fclamp(R(-k.w[1]), R(0), R(1)), b1 = fclamp(R(-k.w[2]), R(0), R(1)),
              b2 = fclamp(R(-k.w[3]), R(0), R(1));
      m[0] = 1; m[1] = -b0; m[2] = -b1; m[3] = -b2;
      const R S = w[v[0]] + w[v[1]] * b0 * b0 + w[v[2]] * b1 * b1 + w[v[3]] * b2 * b2;
      if (S == 0) return;
      kr[0] = -w[v[0]] / S; kr[1] = w[v[1]] * b0 / S; kr[2] = w[v[2]] * b1 / S;
      kr[3] = w[v[3]] * b2 / S;
    } else {
      const R b0 = R(k.w[0]), b1 = R(k.w[1]), b2 = R(-k.w[2]), b3 = R(-k.w[3]);
      m[0] = b0; m[1] = b1; m[2] = -b2; m[3] = -b3;
      const R S = w[v[0]] * b0 * b0 + w[v[1]] * b1 * b1 + w[v[2]] * b2 * b2 + w[v[3]] * b3 * b3;
      if (S == 0) return;
      kr[0] = -w[v[0]] * b0 / S; kr[1] = -w[v[1]] * b1 / S; kr[2] = w[v[2]] * b2 / S;
      kr[3] = w[v[3]] * b3 / S;
    }
    RV nn = mk<R>(0, 0, 0);
    for (int r = 0; r < 4; ++r) nn = nn + mul(ps[v[r]], m[r]);
    const R l = len(nn), th = R(prm.thickness);
    if (!active) return;
    const RV n = divs(nn, l);
    RV G = mk<R>(0, 0, 0);
    for (int r = 0; r < 4; ++r)
      if (used[r]) G = G + mul(a[v[r]], kr[r]);
    // dg/dnn = I - (th / l)(I - n n^T)
    const RV gn = G - mul(G - mul(n, dot(n, G)), th / l);
    for (int r = 0; r < 4; ++r) la[v[r]] = la[v[r]] + mul(gn, m[r]);
  }

  // Backward sweep over the tape: lx / lv = dL/dx, dL/dv after the last substep in,
  // dL/dx, dL/dv before the first out; dL/dkp per distance and bend constraint added
  // to gs / gb.
  void adjoint(std::vector<RV>& lx, std::vector<RV>& lv, std::vector<R>& gs,
               std::vector<R>& gb) const {
    const R dt = R(prm.timestep);
    const int S = (int)tp.sid1.size(), Bn = (int)tp.bkp.size();
    gs.assign(S, R(0));
    gb.assign(Bn, R(0));
    std::vector<RV> lp(n), a(n), la(n);
    for (int s = (int)tape.size() - 1; s >= 0; --s) {
      const Tape& tpe = tape[s];
      const int I = (int)tpe.pk.size() - 1;
      for (int i = 0; i < n; ++i) {
        if (pinned[i]) {
          lp[i] = mk<R>(0, 0, 0);  // x kept, v = 0
        } else {
          lp[i] = lx[i] + divs(lv[i], dt);
          lx[i] = -divs(lv[i], dt);
        }
      }
      const size_t C = tpe.vf.size() + tpe.ee.size();
      std::vector<uint8_t> used(4 * C);
      for (int it = I - 1; it >= 0; --it) {
        const std::vector<RV>& ps = tpe.pk[it];
        std::vector<int> cnt(n, 0);
        for (int i = 0; i < n; ++i) cnt[i] = tp.prefix[i + 1] - tp.prefix[i];
        const bool have_act = tpe.act.size() == (size_t)I * C;
        std::vector<uint8_t> on(C);
        for (size_t c = 0; c < C; ++c) {
          RV cc[4];
          const bool isvf = c < tpe.vf.size();
          const cm_contact& k = isvf ? tpe.vf[c] : tpe.ee[c - tpe.vf.size()];
          const int force = have_act ? (int)tpe.act[(size_t)it * C + c] : -1;
          on[c] = isvf ? contact_vf(k, ps, cc, force) : contact_ee(k, ps, cc, force);
          for (int r = 0; r < 4; ++r) {
            used[4 * c + r] = !is_zero(cc[r]);
            cnt[k.v[r]] += used[4 * c + r];
          }
        }
        for (int i = 0; i < n; ++i) {
          a[i] = divs(lp[i], R(cnt[i]));
          la[i] = lp[i];
        }
        for (int j = 0; j < S; ++j) stretch_vjp(j, ps, a, la, gs[j]);
        for (int j = 0; j < Bn; ++j) bend_vjp(j, ps, a, la, gb[j]);
        for (size_t c = 0; c < C; ++c) {
          const bool isvf = c < tpe.vf.size();
          const cm_contact& k = isvf ? tpe.vf[c] : tpe.ee[c - tpe.vf.size()];
          bool u[4];
          for (int r = 0; r < 4; ++r) u[r] = used[4 * c + r] != 0;
          contact_vjp(k, isvf, ps, a, u, on[c] != 0, la);
        }
        lp.swap(la);
      }
      // p0 = x + dt v', v' = v + dt g
      for (int i = 0; i < n; ++i) {
        lx[i] = lx[i] + lp[i];
        lv[i] = mul(lp[i], dt);
      }
    }
  }
};

}  // namespace orc
