// Vector arithmetic shared by host and device code of libclothmi.
//
// Each operation keeps the evaluation order of the reference's float3 operators
// (LicentaV2/LicentaV2/Core/CudaUtils.cu:32-107, 252-276) so that, compiled
// with -ffp-contract=off and correctly rounded division/sqrt, FP32 results are
// reproducible bit for bit across host and device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CM_HD __host__ __device__ __forceinline__

namespace cm {

struct f3 {
  float x, y, z;
};

CM_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
CM_HD f3 xyz(float4 a) { return f3{a.x, a.y, a.z}; }
CM_HD float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
CM_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
CM_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CM_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
CM_HD f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
CM_HD f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
CM_HD f3 subs(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
CM_HD f3 adds(f3 a, float s) { return f3{a.x + s, a.y + s, a.z + s}; }
CM_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CM_HD f3 cross(f3 a, f3 b) {
  return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
CM_HD float len(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
CM_HD f3 normalize(f3 a) { return divs(a, len(a)); }
CM_HD f3 vmin(f3 a, f3 b) { return f3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
CM_HD f3 vmax(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
CM_HD f3 min3(f3 a, f3 b, f3 c) { return vmin(a, vmin(b, c)); }
CM_HD f3 max3(f3 a, f3 b, f3 c) { return vmax(a, vmax(b, c)); }
CM_HD bool is_zero(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
// mirror image of a about the plane with unit normal nn (Core/CudaUtils.cu:262-265)
CM_HD f3 reflect(f3 a, f3 nn) { return sub(a, mul(nn, 2.f * dot(a, nn))); }
// a = tangential + normal part along unit nn (Core/CudaUtils.cu:267-271)
CM_HD void tangent_normal(f3 a, f3 nn, f3& tang, f3& norm) {
  norm = mul(nn, dot(a, nn));
  tang = sub(a, norm);
}
CM_HD float fclamp(float v, float lo, float hi) { return fminf(fmaxf(lo, v), hi); }
CM_HD bool between(float v, float lo, float hi) { return v >= lo && v <= hi; }

// inclusive AABB overlap (Core/CudaUtils.cu:364-371)
CM_HD bool overlap(f3 amin, f3 amax, f3 bmin, f3 bmax) {
  bool a = amax.x < bmin.x || amin.x > bmax.x;
  bool b = amax.y < bmin.y || amin.y > bmax.y;
  bool c = amax.z < bmin.z || amin.z > bmax.z;
  return !(a || b || c);
}

// ---- Morton codes ---------------------------------------------------------
// 21-bit spread: bit k of v moves to bit 3k (shift-and-mask form).
CM_HD uint64_t spread21(uint64_t v) {
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}
CM_HD uint64_t quant21(float c) {
  return (uint64_t)fminf(fmaxf(c * 2097152.f, 0.0f), 2097151.0f);
}
// Core/CudaUtils.cu:310-325: raw coordinates clamped to [0,1], 21 bits per axis.
CM_HD uint64_t morton63(f3 c) {
  return (spread21(quant21(c.x)) << 2) + (spread21(quant21(c.y)) << 1) + spread21(quant21(c.z));
}
CM_HD uint64_t quant10(float u) { return (uint64_t)fminf(fmaxf(u * 1024.f, 0.0f), 1023.0f); }
CM_HD float unit_coord(float c, float lo, float hi) { return hi > lo ? (c - lo) / (hi - lo) : 0.0f; }
// 10 bits per axis inside the box [lo, hi] (same bit layout as morton63 >> 33).
CM_HD uint64_t morton30(f3 c, f3 lo, f3 hi) {
  return (spread21(quant10(unit_coord(c.x, lo.x, hi.x))) << 2) +
         (spread21(quant10(unit_coord(c.y, lo.y, hi.y))) << 1) +
         spread21(quant10(unit_coord(c.z, lo.z, hi.z)));
}

CM_HD int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

}  // namespace cm
