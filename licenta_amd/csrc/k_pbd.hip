// Position-based-dynamics stages of the substep as gfx950 kernels.
//
//  k_predict      <- CudaPBD::ApplyExternalForces + ExplicitEuler
//                    (LicentaV2/LicentaV2/Physics/CudaPBD.cu:1992-2012, 610-630)
//  k_damp_*       <- CudaPBD::DampVelocities (Physics/CudaPBD.cu:411-458, 1934-1990);
//                    sums are block trees over 256 consecutive vertices, block
//                    results added in block order (a fixed, reproducible order)
//  k_setup        <- _CreateGroundConstraints / _CreateLRAConstraints (:1477-1507, :1571-1592)
//  k_sphere       <- _CreateSphereConstraints + RaySphere (:1509-1569,
//                    Collision/CudaCollisionDetection.cu:529-567)
//  k_solve        <- one Jacobi iteration of PBDStepSolver (:648-659): ground, LRA,
//                    static (stretch :986-1022, bend :1024-1147), dynamic (VF :1212-1311,
//                    EE :1313-1415) corrections and FinalCorrectionStep (:1690-1698)
//                    in one launch: a workgroup evaluates the slots of a tile of up to
//                    128 consecutive vertices in parallel into LDS, then each vertex sums its slots in
//                    the reference's slot order, so the per-vertex sums are the
//                    reference's gathers (:1417-1475) without the global slot buffer.
//  k_integrate    <- _FinalUpdate (:1916-1932) + _WorldConstraints
//                    (Physics/CudaDeformableUtils.cu:260-269)
//  k_vertex_normals <- DeformableUtils::UpdateVertexNormals
//                    (Physics/CudaDeformableUtils.cu:190-249) with the unit face normals
//                    of _UpdateTrianglesContinuous (:135)
//  k_revert_toc   <- CudaPBD::RevertToTOC (:727-768, FillTOC :2014-2046, _RevertToTOC
//                    :2048-2072) over the per-vertex contact records
//  k_friction     <- CudaPBD::ApplyFriction (:770-833): _DampVelRigid (:1879-1912),
//                    then _ApplyFriction / _AccumulateFriction / FinalFrictionStep
//                    (:1725-1877, :1700-1723) over the per-vertex contact records
//
// Positions are float4 with the inverse mass in .w, so one 16-byte load gives a
// neighbour's position and weight.
#include "cm_internal.hpp"

namespace cm {

namespace {

__global__ void __launch_bounds__(kBlock) k_predict(int n, const float4* __restrict__ x,
                                                    float4* __restrict__ v,
                                                    float4* __restrict__ p, float dt, f3 g) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = x[i];
  const f3 dtg = mul(g, dt);
  const f3 vi = add(xyz(v[i]), dtg);
  v[i] = f4(vi, 0.f);
  p[i] = f4(add(xyz(xi), mul(vi, dt)), xi.w);
}

// ---- damping ------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void block_tree(float (&vals)[K], float (*s)[kBlock], float* out) {
  for (int k = 0; k < K; ++k) s[k][threadIdx.x] = vals[k];
  __syncthreads();
  for (int stride = kBlock / 2; stride >= 1; stride >>= 1) {
    if (threadIdx.x < stride)
      for (int k = 0; k < K; ++k) s[k][threadIdx.x] = s[k][threadIdx.x] + s[k][threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < K; ++k) out[k] = s[k][0];
}

// v' = (v + dt g)(1 - kg); partial sums of m, p m, v' m
__global__ void __launch_bounds__(kBlock) k_damp_sum1(int n, const float4* __restrict__ p,
                                                      float4* __restrict__ v,
                                                      const float* __restrict__ mass, f3 dtg,
                                                      float kg, float* __restrict__ partials) {
  __shared__ float s[7][kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  float vals[7] = {0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    const f3 vi = mul(add(xyz(v[i]), dtg), 1.f - kg);
    v[i] = f4(vi, 0.f);
    const float m = mass[i];
    const f3 pm = mul(xyz(p[i]), m), vm = mul(vi, m);
    vals[0] = m;
    vals[1] = pm.x; vals[2] = pm.y; vals[3] = pm.z;
    vals[4] = vm.x; vals[5] = vm.y; vals[6] = vm.z;
  }
  block_tree<7>(vals, s, partials + 7 * blockIdx.x);
}

// single thread: block sums in order -> xcm, vcm (state[0..5])
__global__ void k_damp_mid(int nb, const float* __restrict__ partials,
                           float* __restrict__ state) {
  if (threadIdx.x != 0) return;
  float t[7];
  for (int k = 0; k < 7; ++k) t[k] = partials[k];
  for (int b = 1; b < nb; ++b)
    for (int k = 0; k < 7; ++k) t[k] = t[k] + partials[7 * b + k];
  const f3 xcm = divs(mk3(t[1], t[2], t[3]), t[0]);
  const f3 vcm = divs(mk3(t[4], t[5], t[6]), t[0]);
  state[0] = xcm.x; state[1] = xcm.y; state[2] = xcm.z;
  state[3] = vcm.x; state[4] = vcm.y; state[5] = vcm.z;
}

// partial sums of L = sum r x (v' m), i1 = sum (r.x^2, r.y^2, r.z^2) m,
// i2 = -sum (r.x r.y, r.x r.z, r.y r.z) m
__global__ void __launch_bounds__(kBlock) k_damp_sum2(int n, const float4* __restrict__ p,
                                                      const float4* __restrict__ v,
                                                      const float* __restrict__ mass,
                                                      const float* __restrict__ state,
                                                      float* __restrict__ partials) {
  __shared__ float s[9][kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  float vals[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < n) {
    const f3 xcm = mk3(state[0], state[1], state[2]);
    const float m = mass[i];
    const f3 r = sub(xyz(p[i]), xcm);
    const f3 L = cross(r, mul(xyz(v[i]), m));
    const f3 a = mul(mk3(r.x * r.x, r.y * r.y, r.z * r.z), m);
    const f3 b = mul(neg(mk3(r.x * r.y, r.x * r.z, r.y * r.z)), m);
    vals[0] = L.x; vals[1] = L.y; vals[2] = L.z;
    vals[3] = a.x; vals[4] = a.y; vals[5] = a.z;
    vals[6] = b.x; vals[7] = b.y; vals[8] = b.z;
  }
  block_tree<9>(vals, s, partials + 9 * blockIdx.x);
}

// single thread: omega = I^-1 L (state[6..8])
__global__ void k_damp_final(int nb, const float* __restrict__ partials,
                             float* __restrict__ state) {
  if (threadIdx.x != 0) return;
  float t[9];
  for (int k = 0; k < 9; ++k) t[k] = partials[k];
  for (int b = 1; b < nb; ++b)
    for (int k = 0; k < 9; ++k) t[k] = t[k] + partials[9 * b + k];
  const f3 L = mk3(t[0], t[1], t[2]), i1 = mk3(t[3], t[4], t[5]), i2 = mk3(t[6], t[7], t[8]);
  const float a00 = i1.y + i1.z, a01 = i2.x, a02 = i2.y;
  const float a11 = i1.x + i1.z, a12 = i2.z, a22 = i1.x + i1.y;
  const float m00 = a00, m01 = a01, m02 = a02, m10 = a01, m11 = a11, m12 = a12, m20 = a02,
              m21 = a12, m22 = a22;
  const float inv_det = 1.f / (+m00 * (m11 * m22 - m21 * m12) - m10 * (m01 * m22 - m21 * m02) +
                               m20 * (m01 * m12 - m11 * m02));
  const float i00 = +(m11 * m22 - m21 * m12) * inv_det, i10 = -(m10 * m22 - m20 * m12) * inv_det,
              i20 = +(m10 * m21 - m20 * m11) * inv_det, i01 = -(m01 * m22 - m21 * m02) * inv_det,
              i11 = +(m00 * m22 - m20 * m02) * inv_det, i21 = -(m00 * m21 - m20 * m01) * inv_det,
              i02 = +(m01 * m12 - m11 * m02) * inv_det, i12 = -(m00 * m12 - m10 * m02) * inv_det,
              i22 = +(m00 * m11 - m10 * m01) * inv_det;
  state[6] = i00 * L.x + i10 * L.y + i20 * L.z;
  state[7] = i01 * L.x + i11 * L.y + i21 * L.z;
  state[8] = i02 * L.x + i12 * L.y + i22 * L.z;
}

__global__ void __launch_bounds__(kBlock) k_damp_apply(int n, const float4* __restrict__ x,
                                                       float4* __restrict__ p,
                                                       float4* __restrict__ v,
                                                       const float* __restrict__ state, float kd,
                                                       float dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const f3 xcm = mk3(state[0], state[1], state[2]), vcm = mk3(state[3], state[4], state[5]),
           om = mk3(state[6], state[7], state[8]);
  const f3 r = sub(xyz(p[i]), xcm);
  const f3 vi = xyz(v[i]);
  const f3 dv = sub(add(vcm, cross(om, r)), vi);
  const f3 vn = add(vi, mul(dv, kd));
  v[i] = f4(vn, 0.f);
  const float4 xi = x[i];
  p[i] = f4(add(xyz(xi), mul(vn, dt)), xi.w);
}

// ---- solver set-up ------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_setup(int n, const float4* __restrict__ x,
                                                  const float4* __restrict__ p, float gh,
                                                  float th, int ground_on,
                                                  float4* __restrict__ ground, int F,
                                                  const int* __restrict__ fixed,
                                                  const float* __restrict__ l0,
                                                  const uint8_t* __restrict__ pinned,
                                                  uint8_t* __restrict__ lflag) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const f3 pos = xyz(p[i]);
  if (ground_on) {
    float4 gq = make_float4(0.f, 0.f, 0.f, 0.f);
    if (pos.y <= gh + th) {
      const f3 nc = mk3(0.f, 1.f, 0.f);
      const f3 p0 = mk3(1.f, gh + th, 1.f);
      const f3 l = sub(pos, xyz(x[i]));
      const float d = dot(sub(p0, pos), nc) / dot(l, nc);
      gq = f4(add(mul(l, d), pos), 1.f);
    }
    ground[i] = gq;
  }
  for (int j = 0; j < F; ++j) {
    uint8_t f = 0;
    if (!pinned[i]) {
      const f3 q = xyz(p[fixed[j]]);
      f = len(sub(pos, q)) > l0[(int64_t)i * F + j] ? 1 : 0;
    }
    lflag[(int64_t)i * F + j] = f;
  }
}

__global__ void __launch_bounds__(kBlock) k_sphere(int n, float4* __restrict__ x,
                                                   float4* __restrict__ p, f3 c, float rad) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 xi = x[i], pi = p[i];
  const f3 a = xyz(xi), b = xyz(pi);
  f3 dir = sub(b, a);
  if (is_zero(dir)) return;
  const float l = len(dir);
  dir = divs(dir, l);
  const f3 m = sub(a, c);
  const float bb = dot(m, dir);
  const float cc = dot(m, m) - rad * rad;
  if (cc > 0.f && bb > 0.f) return;
  const float del = bb * bb - cc;
  if (del < 0) return;
  const float tt = -bb - sqrtf(del);
  if (tt < 0.f) {
    const f3 qc = add(a, mul(dir, tt));
    const f3 q = add(c, mul(normalize(sub(qc, c)), rad));
    const f3 nc = normalize(sub(q, c));
    x[i] = f4(add(q, mul(nc, 0.0001f)), xi.w);
    p[i] = f4(q, pi.w);
    return;
  }
  if (tt > l) return;
  const f3 qc = add(a, mul(dir, tt));
  const f3 nc = normalize(sub(qc, c));
  p[i] = f4(add(qc, mul(nc, 0.0001f)), pi.w);
}

// ---- constraint corrections for one vertex role ---------------------------------
__device__ __forceinline__ f3 stretch_corr(int2 ids, float2 l0kp, int role,
                                           const float4* __restrict__ p) {
  const float4 a = p[ids.x], b = p[ids.y];
  const float w1 = a.w, w2 = b.w, ws = w1 + w2;
  if (ws == 0) return mk3(0, 0, 0);
  const f3 diff = sub(xyz(a), xyz(b));
  const float l = len(diff);
  if (l <= 1e-8f) return mk3(0, 0, 0);
  const f3 nn = divs(diff, l);
  const f3 tt = divs(mul(mul(nn, l - l0kp.x), l0kp.y), ws);
  return role == 0 ? mul(tt, -w1) : mul(tt, w2);
}

// vector arithmetic in T (float or double) with the operation order of cm_math.hpp
template <class T>
struct tv {
  T x, y, z;
};
template <class T> __device__ __forceinline__ tv<T> tadd(tv<T> a, tv<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> __device__ __forceinline__ tv<T> tsub(tv<T> a, tv<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> __device__ __forceinline__ tv<T> tneg(tv<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> __device__ __forceinline__ tv<T> tmul(tv<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> __device__ __forceinline__ tv<T> tdiv(tv<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> __device__ __forceinline__ T tdot(tv<T> a, tv<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ __forceinline__ tv<T> tcross(tv<T> a, tv<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> __device__ __forceinline__ T tsqrt(T v);
template <> __device__ __forceinline__ float tsqrt<float>(float v) { return sqrtf(v); }
template <> __device__ __forceinline__ double tsqrt<double>(double v) { return sqrt(v); }
template <class T> __device__ __forceinline__ T tacos(T v);
template <> __device__ __forceinline__ float tacos<float>(float v) { return acosf(v); }
template <> __device__ __forceinline__ double tacos<double>(double v) { return acos(v); }
template <class T> __device__ __forceinline__ T tlen(tv<T> a) { return tsqrt<T>(a.x * a.x + a.y * a.y + a.z * a.z); }
template <class T> __device__ __forceinline__ tv<T> tfrom(float4 a) { return {(T)a.x, (T)a.y, (T)a.z}; }
// v / s per component for the bend: FP32 divides like the reference; the FP64 form
// (CM_F_BEND_FP64, this project's own precision choice) multiplies by one
// reciprocal, the same operations as the oracle's FP64 bend.
template <class T> __device__ __forceinline__ tv<T> bdiv(tv<T> v, T s);
template <> __device__ __forceinline__ tv<float> bdiv<float>(tv<float> v, float s) { return tdiv(v, s); }
template <> __device__ __forceinline__ tv<double> bdiv<double>(tv<double> v, double s) {
  return tmul(v, 1.0 / s);
}
template <class T> __device__ __forceinline__ f3 tto(tv<T> a) { return mk3((float)a.x, (float)a.y, (float)a.z); }

// Physics/CudaPBD.cu:1024-1147 for one vertex role, evaluated in T
template <class T>
__device__ f3 bend_corr(int4 ids, float2 kpphi0, int role, int abs_mode,
                        const float4* __restrict__ p) {
  const float4 a1 = p[ids.x], a2 = p[ids.y], a3 = p[ids.z], a4 = p[ids.w];
  const T w1 = (T)a1.w, w2 = (T)a2.w, w3 = (T)a3.w, w4 = (T)a4.w;
  tv<T> p2 = tfrom<T>(a2), p3 = tfrom<T>(a3), p4 = tfrom<T>(a4);
  if (!abs_mode) {
    const tv<T> p1 = tfrom<T>(a1);
    p2 = tsub(p2, p1);
    p3 = tsub(p3, p1);
    p4 = tsub(p4, p1);
  }
  const T kp = (T)kpphi0.x, phi0 = (T)kpphi0.y;
  const tv<T> c23 = tcross(p2, p3), c24 = tcross(p2, p4);
  const T l23 = tlen(c23), l24 = tlen(c24);
  const T eps = (T)1e-8f;
  if (l23 <= eps || l24 <= eps) return mk3(0, 0, 0);
  const tv<T> n1 = bdiv<T>(c23, l23), n2 = bdiv<T>(c24, l24);
  const T d = fmin(fmax((T)-1, tdot(n1, n2)), (T)1);
  T phi = tacos<T>(d);
  const T PIf = (T)3.14159265359f;
  if (d == (T)-1) {
    phi = PIf;
    // the n/100 push of :1077-1078 is kept for the reference's absolute form only;
    // the p1-relative variant takes the smooth limit (no correction)
    if (phi == phi0 || !abs_mode) return mk3(0, 0, 0);
    if (role == 0) return tto(tdiv(n1, (T)100));
    if (role == 1) return tto(tdiv(n2, (T)100));
    return mk3(0, 0, 0);
  }
  if (d == (T)1) {
    phi = (T)0;
    if (phi == phi0) return mk3(0, 0, 0);
  }
  const tv<T> q3 = bdiv<T>(tadd(tcross(p2, n2), tmul(tcross(n1, p2), d)), l23);
  const tv<T> q4 = bdiv<T>(tadd(tcross(p2, n1), tmul(tcross(n2, p2), d)), l24);
  const tv<T> q2 = tsub(tneg(bdiv<T>(tadd(tcross(p3, n2), tmul(tcross(n1, p3), d)), l23)),
                        bdiv<T>(tadd(tcross(p4, n1), tmul(tcross(n2, p4), d)), l24));
  const tv<T> q1 = tsub(tsub(tneg(q2), q3), q4);
  const T term = tsqrt<T>((T)1 - d * d) * (phi - phi0) * kp;
  const T s1 = tdot(q1, q1), s2 = tdot(q2, q2), s3 = tdot(q3, q3), s4 = tdot(q4, q4);
  const T sf = w1 * s1 + w2 * s2 + w3 * s3 + w4 * s4;
  if (sf <= eps) return mk3(0, 0, 0);
  const T ts = term / sf;
  switch (role) {
    case 0: return tto(tmul(q1, -(w1 * ts)));
    case 1: return tto(tmul(q2, -(w2 * ts)));
    case 2: return tto(tmul(q3, -(w3 * ts)));
    default: return tto(tmul(q4, -(w4 * ts)));
  }
}

// stretch_corr for both roles at once (same operations as stretch_corr), from the
// two positions already loaded
__device__ __forceinline__ void stretch_both(float4 a, float4 b, float2 l0kp, f3& c0, f3& c1) {
  c0 = c1 = mk3(0, 0, 0);
  const float w1 = a.w, w2 = b.w, ws = w1 + w2;
  if (ws == 0) return;
  const f3 diff = sub(xyz(a), xyz(b));
  const float l = len(diff);
  if (l <= 1e-8f) return;
  const f3 nn = divs(diff, l);
  const f3 tt = divs(mul(mul(nn, l - l0kp.x), l0kp.y), ws);
  c0 = mul(tt, -w1);
  c1 = mul(tt, w2);
}

// bend_corr for all four roles at once (same operations as bend_corr), from the four
// positions already loaded
template <class T>
__device__ __forceinline__ void bend_all(float4 a1, float4 a2, float4 a3, float4 a4,
                                         float2 kpphi0, int abs_mode, f3 c[4]) {
  c[0] = c[1] = c[2] = c[3] = mk3(0, 0, 0);
  const T w1 = (T)a1.w, w2 = (T)a2.w, w3 = (T)a3.w, w4 = (T)a4.w;
  tv<T> p2 = tfrom<T>(a2), p3 = tfrom<T>(a3), p4 = tfrom<T>(a4);
  if (!abs_mode) {
    const tv<T> p1 = tfrom<T>(a1);
    p2 = tsub(p2, p1);
    p3 = tsub(p3, p1);
    p4 = tsub(p4, p1);
  }
  const T kp = (T)kpphi0.x, phi0 = (T)kpphi0.y;
  const tv<T> c23 = tcross(p2, p3), c24 = tcross(p2, p4);
  const T l23 = tlen(c23), l24 = tlen(c24);
  const T eps = (T)1e-8f;
  if (l23 <= eps || l24 <= eps) return;
  const tv<T> n1 = bdiv<T>(c23, l23), n2 = bdiv<T>(c24, l24);
  const T d = fmin(fmax((T)-1, tdot(n1, n2)), (T)1);
  T phi = tacos<T>(d);
  const T PIf = (T)3.14159265359f;
  if (d == (T)-1) {
    phi = PIf;
    if (phi == phi0 || !abs_mode) return;
    c[0] = tto(tdiv(n1, (T)100));
    c[1] = tto(tdiv(n2, (T)100));
    return;
  }
  if (d == (T)1) {
    phi = (T)0;
    if (phi == phi0) return;
  }
  const tv<T> q3 = bdiv<T>(tadd(tcross(p2, n2), tmul(tcross(n1, p2), d)), l23);
  const tv<T> q4 = bdiv<T>(tadd(tcross(p2, n1), tmul(tcross(n2, p2), d)), l24);
  const tv<T> q2 = tsub(tneg(bdiv<T>(tadd(tcross(p3, n2), tmul(tcross(n1, p3), d)), l23)),
                        bdiv<T>(tadd(tcross(p4, n1), tmul(tcross(n2, p4), d)), l24));
  const tv<T> q1 = tsub(tsub(tneg(q2), q3), q4);
  const T term = tsqrt<T>((T)1 - d * d) * (phi - phi0) * kp;
  const T s1 = tdot(q1, q1), s2 = tdot(q2, q2), s3 = tdot(q3, q3), s4 = tdot(q4, q4);
  const T sf = w1 * s1 + w2 * s2 + w3 * s3 + w4 * s4;
  if (sf <= eps) return;
  const T ts = term / sf;
  c[0] = tto(tmul(q1, -(w1 * ts)));
  c[1] = tto(tmul(q2, -(w2 * ts)));
  c[2] = tto(tmul(q3, -(w3 * ts)));
  c[3] = tto(tmul(q4, -(w4 * ts)));
}

__device__ __forceinline__ f3 vf_corr(float4 A, float4 B0, float4 B1, float4 B2, float4 w,
                                      int role, float th) {
  const float b0 = fclamp(-w.y, 0.f, 1.f), b1 = fclamp(-w.z, 0.f, 1.f),
              b2 = fclamp(-w.w, 0.f, 1.f);
  const float ww = A.w, w0 = B0.w, w1 = B1.w, w2 = B2.w;
  const f3 q = add(add(mul(xyz(B0), b0), mul(xyz(B1), b1)), mul(xyz(B2), b2));
  f3 nn = sub(xyz(A), q);
  const float l = len(nn);
  nn = divs(nn, l);
  const float cc = l - th;
  float s = ww + w0 * b0 * b0 + w1 * b1 * b1 + w2 * b2 * b2;
  if (s == 0) return mk3(0, 0, 0);
  s = cc / s;
  s = cc < 0 ? s : 0.f;
  switch (role) {
    case 0: return mul(nn, -s * ww);
    case 1: return mul(mul(neg(nn), b0), -s * w0);
    case 2: return mul(mul(neg(nn), b1), -s * w1);
    default: return mul(mul(neg(nn), b2), -s * w2);
  }
}

__device__ __forceinline__ f3 ee_corr(float4 P0, float4 P1, float4 P2, float4 P3, float4 w,
                                      int role, float th) {
  const float b0 = w.x, b1 = w.y, b2 = -w.z, b3 = -w.w;
  const float w0 = P0.w, w1 = P1.w, w2 = P2.w, w3 = P3.w;
  const f3 q0 = add(mul(xyz(P0), b0), mul(xyz(P1), b1));
  const f3 q1 = add(mul(xyz(P2), b2), mul(xyz(P3), b3));
  f3 nn = sub(q0, q1);
  const float l = len(nn);
  nn = divs(nn, l);
  const float cc = l - th;
  float s = w0 * b0 * b0 + w1 * b1 * b1 + w2 * b2 * b2 + w3 * b3 * b3;
  if (s == 0) return mk3(0, 0, 0);
  s = cc / s;
  s = cc < 0 ? s : 0.f;
  switch (role) {
    case 0: return mul(mul(nn, b0), -s * w0);
    case 1: return mul(mul(nn, b1), -s * w1);
    case 2: return mul(mul(neg(nn), b2), -s * w2);
    default: return mul(mul(neg(nn), b3), -s * w3);
  }
}

struct SolveArgs {
  int n, F, abs_mode, ground_on, bend_fp64;
  float th;
  const float4* ground;
  const int* lra_fixed;
  const float* lra_l0;
  const uint8_t* lra_flag;
  const int* vprefix;
  const uint32_t* vslot;
  const int2* sid;
  const float2* s_l0kp;
  const int4* bid;
  const float2* b_kpphi0;
  const int* dprefix;
  const RecPayload* rec;  // dynamic records in per-vertex order
  // vertex tiles: tstart[ntiles + 1] vertex ranges, eprefix[ntiles + 1] entry ranges
  const int* tstart;
  const int* eprefix;
  const TileEntry* entries;
};

// one static slot's correction: slot word = bend flag (bit 31), role (bits 29-30),
// constraint index (bits 0-28)
__device__ __forceinline__ f3 slot_corr(const SolveArgs& a, uint32_t s,
                                        const float4* __restrict__ pin) {
  const int idx = (int)(s & 0x1fffffffu);
  const int role = (int)((s >> 29) & 3u);
  if (s & 0x80000000u)
    return a.bend_fp64 ? bend_corr<double>(a.bid[idx], a.b_kpphi0[idx], role, a.abs_mode, pin)
                       : bend_corr<float>(a.bid[idx], a.b_kpphi0[idx], role, a.abs_mode, pin);
  return stretch_corr(a.sid[idx], a.s_l0kp[idx], role, pin);
}

// one dynamic record's correction (payload: contact ids with role and VF bit, weights)
// from the record's four positions already loaded
__device__ __forceinline__ f3 rec_corr(const RecPayload& rp, const float4 (&P)[4], float th) {
  const int role = (rp.ids.x >> 29) & 3;
  return rp.ids.x < 0 ? vf_corr(P[0], P[1], P[2], P[3], rp.w, role, th)
                      : ee_corr(P[0], P[1], P[2], P[3], rp.w, role, th);
}

constexpr int kSolveV = 64;  // vertices per workgroup (slot-parallel form)
constexpr int kLds = kTileSlots;  // slot corrections staged in LDS per pass
constexpr int kBatch = 4;         // constraint entries / records per thread in flight

__device__ __forceinline__ void lds_put(float* cx, float* cy, float* cz, uint32_t off, f3 c) {
  cx[off] = c.x;
  cy[off] = c.y;
  cz[off] = c.z;
}

// One Jacobi iteration over a range of consecutive vertices. The range's static
// corrections are staged in LDS: in the tiled form (kTiles) each constraint that
// touches the tile is evaluated once, from the tile's entry list, and its
// corrections go to the slots of the tile's vertices; in the slot-parallel form
// every slot re-evaluates its constraint, kLds slots per pass. The range's dynamic
// records follow the same way. Each vertex then adds its own entries in slot
// order, so every per-vertex sum has the reference's order and rounding.
template <bool kTiles>
__global__ void __launch_bounds__(kBlock) k_solve(SolveArgs a, const float4* __restrict__ pin,
                                                  float4* __restrict__ pout) {
  __shared__ float cx[kLds], cy[kLds], cz[kLds];
  int v0, v1, tile = 0;
  if (kTiles) {
    // each XCD takes a contiguous range of tiles, so the positions that neighbouring
    // tiles share are fetched into one L2
    tile = xcd_block(blockIdx.x, gridDim.x);
    v0 = a.tstart[tile];
    v1 = a.tstart[tile + 1];
  } else {
    v0 = blockIdx.x * kSolveV;
    v1 = min(v0 + kSolveV, a.n);
  }
  const int tid = threadIdx.x;
  const int i = v0 + tid;
  const bool mine = i < v1;  // tiles hold at most kBlock vertices
  float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
  f3 total = mk3(0, 0, 0);
  int cnt = 0;
  if (mine) {
    pi = pin[i];
    if (a.ground_on) {
      const float4 gq = a.ground[i];
      if (gq.w != 0.f) {
        const f3 nc = mk3(0.f, 1.f, 0.f);
        const float cv = dot(sub(xyz(pi), xyz(gq)), nc);
        if (cv < 0) {
          total = add(total, mul(nc, -cv));
          cnt += 1;
        }
      }
    }
    if (a.F) {
      f3 acc = mk3(0, 0, 0);
      int c = 0;
      for (int j = 0; j < a.F; ++j) {
        if (a.lra_flag[(int64_t)i * a.F + j]) {
          const f3 diff = sub(xyz(pi), xyz(pin[a.lra_fixed[j]]));
          const float l = len(diff);
          const f3 nn = divs(diff, l);
          acc = add(acc, mul(nn, -(l - a.lra_l0[(int64_t)i * a.F + j])));
          ++c;
        }
      }
      total = add(total, acc);
      cnt += c;
    }
  }
  // static slots
  {
    const int s0 = a.vprefix[v0], s1 = a.vprefix[v1];
    int k = mine ? a.vprefix[i] : 0;
    const int e = mine ? a.vprefix[i + 1] : 0;
    f3 acc = mk3(0, 0, 0);
    for (int c0 = s0; c0 < s1; c0 += kLds) {
      const int c1 = min(c0 + kLds, s1);
      if (kTiles) {  // a tile's slots fit one pass (c0 == s0)
        const int eb = a.eprefix[tile], ee = a.eprefix[tile + 1];
        // kBatch entries per thread at a time: their loads, then their positions'
        // loads, are in flight together before any of them is evaluated
        for (int q0 = eb + tid; q0 < ee; q0 += kBlock * kBatch) {
          TileEntry te[kBatch];
          float4 P[kBatch][4];
#pragma unroll
          for (int u = 0; u < kBatch; ++u)
            if (q0 + u * kBlock < ee) te[u] = a.entries[q0 + u * kBlock];
#pragma unroll
          for (int u = 0; u < kBatch; ++u) {
            if (q0 + u * kBlock >= ee) continue;
            P[u][0] = pin[te[u].ids.x];
            P[u][1] = pin[te[u].ids.y];
            if (te[u].ids.z >= 0) {
              P[u][2] = pin[te[u].ids.z];
              P[u][3] = pin[te[u].ids.w];
            }
          }
#pragma unroll
          for (int u = 0; u < kBatch; ++u) {
            if (q0 + u * kBlock >= ee) continue;
            const TileEntry& t = te[u];
            if (t.ids.z < 0) {
              f3 c0v, c1v;
              stretch_both(P[u][0], P[u][1], make_float2(t.a, t.b), c0v, c1v);
              if ((t.s01 & 0xffffu) != 0xffffu) lds_put(cx, cy, cz, t.s01 & 0xffffu, c0v);
              if ((t.s01 >> 16) != 0xffffu) lds_put(cx, cy, cz, t.s01 >> 16, c1v);
            } else {
              f3 cr[4];
              if (a.bend_fp64)
                bend_all<double>(P[u][0], P[u][1], P[u][2], P[u][3], make_float2(t.a, t.b),
                                 a.abs_mode, cr);
              else
                bend_all<float>(P[u][0], P[u][1], P[u][2], P[u][3], make_float2(t.a, t.b),
                                a.abs_mode, cr);
              if ((t.s01 & 0xffffu) != 0xffffu) lds_put(cx, cy, cz, t.s01 & 0xffffu, cr[0]);
              if ((t.s01 >> 16) != 0xffffu) lds_put(cx, cy, cz, t.s01 >> 16, cr[1]);
              if ((t.s23 & 0xffffu) != 0xffffu) lds_put(cx, cy, cz, t.s23 & 0xffffu, cr[2]);
              if ((t.s23 >> 16) != 0xffffu) lds_put(cx, cy, cz, t.s23 >> 16, cr[3]);
            }
            // evaluate the batch's entries one after the other (bounded registers)
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      } else {
        for (int q = c0 + tid; q < c1; q += kBlock)
          lds_put(cx, cy, cz, q - c0, slot_corr(a, a.vslot[q], pin));
      }
      __syncthreads();
      const int kend = min(e, c1);
      // LDS reads batched ahead of the dependent adds; the add order is unchanged
      for (; k + 4 <= kend; k += 4) {
        const int q = k - c0;
        const f3 e0 = mk3(cx[q], cy[q], cz[q]), e1 = mk3(cx[q + 1], cy[q + 1], cz[q + 1]);
        const f3 e2 = mk3(cx[q + 2], cy[q + 2], cz[q + 2]), e3 = mk3(cx[q + 3], cy[q + 3], cz[q + 3]);
        acc = add(add(add(add(acc, e0), e1), e2), e3);
      }
      for (; k < kend; ++k) acc = add(acc, mk3(cx[k - c0], cy[k - c0], cz[k - c0]));
      __syncthreads();
    }
    if (mine) {
      total = add(total, acc);
      cnt += e - (int)a.vprefix[i];
    }
  }
  // dynamic records: only non-zero corrections are summed and counted
  if (a.dprefix) {
    const int r0 = a.dprefix[v0], r1 = a.dprefix[v1];
    if (r1 > r0) {
      int k = mine ? a.dprefix[i] : 0;
      const int e = mine ? a.dprefix[i + 1] : 0;
      const bool any = mine && e > k;
      f3 acc = mk3(0, 0, 0);
      int c = 0;
      for (int c0 = r0; c0 < r1; c0 += kLds) {
        const int c1 = min(c0 + kLds, r1);
        for (int q0 = c0 + tid; q0 < c1; q0 += kBlock * kBatch) {
          RecPayload rp[kBatch];
          float4 P[kBatch][4];
#pragma unroll
          for (int u = 0; u < kBatch; ++u)
            if (q0 + u * kBlock < c1) rp[u] = a.rec[q0 + u * kBlock];
#pragma unroll
          for (int u = 0; u < kBatch; ++u) {
            if (q0 + u * kBlock >= c1) continue;
            P[u][0] = pin[rp[u].ids.x & 0x1fffffff];
            P[u][1] = pin[rp[u].ids.y];
            P[u][2] = pin[rp[u].ids.z];
            P[u][3] = pin[rp[u].ids.w];
          }
#pragma unroll
          for (int u = 0; u < kBatch; ++u)
            if (q0 + u * kBlock < c1)
              lds_put(cx, cy, cz, q0 + u * kBlock - c0, rec_corr(rp[u], P[u], a.th));
        }
        __syncthreads();
        const int kend = min(e, c1);
        for (; k + 4 <= kend; k += 4) {
          const int q = k - c0;
          f3 ev[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) ev[u] = mk3(cx[q + u], cy[q + u], cz[q + u]);
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (!is_zero(ev[u])) {
              acc = add(acc, ev[u]);
              ++c;
            }
        }
        for (; k < kend; ++k) {
          const f3 val = mk3(cx[k - c0], cy[k - c0], cz[k - c0]);
          if (!is_zero(val)) {
            acc = add(acc, val);
            ++c;
          }
        }
        __syncthreads();
      }
      if (any) {
        total = add(total, acc);
        cnt += c;
      }
    }
  }
  if (mine) pout[i] = f4(add(xyz(pi), divs(total, (float)cnt)), pi.w);
}

__global__ void __launch_bounds__(kBlock) k_integrate(int n, const float4* __restrict__ pfin,
                                                      float4* __restrict__ x,
                                                      float4* __restrict__ v,
                                                      float4* __restrict__ p,
                                                      const uint8_t* __restrict__ pinned,
                                                      float dt, int clamp, f3 lo, f3 hi) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float4 pi = pfin[i];
  if (pinned[i]) {
    v[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  } else {
    const float4 xi = x[i];
    v[i] = f4(divs(sub(xyz(pi), xyz(xi)), dt), 0.f);
    x[i] = make_float4(pi.x, pi.y, pi.z, xi.w);
  }
  if (clamp) {
    pi.x = fminf(fmaxf(pi.x, lo.x), hi.x);
    pi.y = fminf(fmaxf(pi.y, lo.y), hi.y);
    pi.z = fminf(fmaxf(pi.z, lo.z), hi.z);
  }
  p[i] = pi;
}


// Friction on the integrated velocities. First the ground step of _DampVelRigid
// (:1879-1912): a vertex whose solver set-up flagged a ground contact keeps its
// tangential velocity and has its normal velocity reflected and scaled by 1 - kf
// (ground normal +y). The reference launches that kernel with 255-thread blocks over
// ceil(N/256) blocks (:778), so it reaches only vertices below rigid_n; its sphere
// branch never runs, since the sphere set-up leaves every flag clear (:1531, :1538).
// Then the contacts: the reference writes each contact's normal once per contact
// vertex, sorts the entries by vertex (stable) and sums each vertex's run in entry
// order (:1745-1795, :1846-1877). The vertex's record list here is that run in the
// same order (records ascend by contact, VF before EE), so the sum is the same. The
// reference leaves the vertex ids of its EE entries unwritten (:1810-1813), so they
// land on whatever ids the buffer held before; here an EE contact counts for its own
// four vertices. With the sum's direction fn and count c, FinalFrictionStep
// (:1700-1723) keeps the tangential velocity along fn and reflects the normal part
// scaled by (1 - kf)^c.
__global__ void __launch_bounds__(kBlock) k_friction(
    int n, int rigid_n, const float4* __restrict__ ground, const int* __restrict__ dprefix,
    const uint32_t* __restrict__ drec, const int* __restrict__ d_nvf,
    const float4* __restrict__ vf_tn, const float4* __restrict__ ee_tn, float4* __restrict__ v,
    float kf) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 vi = v[i];
  f3 vel = xyz(vi);
  if (ground && i < rigid_n && ground[i].w != 0.f) {
    const f3 nc = mk3(0.f, 1.f, 0.f);
    f3 vt, vn;
    tangent_normal(vel, nc, vt, vn);
    vel = reflect(add(vt, mul(vn, 1.f - kf)), nc);
  }
  if (dprefix) {
    const int b = dprefix[i], e = dprefix[i + 1];
    if (e > b) {
      const int nvf = *d_nvf;
      f3 acc = mk3(0.f, 0.f, 0.f);
      for (int k = b; k < e; ++k) {
        const int c = (int)(drec[k] >> 2);
        const float4 tn = c < nvf ? vf_tn[c] : ee_tn[c - nvf];
        acc = add(acc, mk3(tn.y, tn.z, tn.w));
      }
      if (!is_zero(acc)) {
        const f3 fn = normalize(acc);
        f3 vt, vn;
        tangent_normal(vel, fn, vt, vn);
        vel = reflect(add(mul(vn, powf(1.f - kf, (float)(e - b))), vt), fn);
      }
    }
  }
  v[i] = f4(vel, vi.w);
}

// Time-of-contact rollback between the narrow phase and the solver (the reference's
// call at Physics/CudaDeformableBody.cpp:583 is commented out; CM_F_REVERT_TOC turns
// it on). A vertex in contacts takes the smallest contact time t of its contacts
// (the reference starts its minimum at 100) and moves to (1 - t) x + t p
// (CudaUtils::AdvancePositionInTimePos, Core/CudaUtils.cu:396-399). The contact t is
// the time of contact in seconds, used as the blend weight as the reference does.
__global__ void __launch_bounds__(kBlock) k_revert_toc(int n, const int* __restrict__ dprefix,
                                                       const uint32_t* __restrict__ drec,
                                                       const int* __restrict__ d_nvf,
                                                       const float4* __restrict__ vf_tn,
                                                       const float4* __restrict__ ee_tn,
                                                       const float4* __restrict__ x,
                                                       float4* __restrict__ p) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int b = dprefix[i], e = dprefix[i + 1];
  if (e == b) return;
  const int nvf = *d_nvf;
  float tmin = 100.f;
  for (int k = b; k < e; ++k) {
    const int c = (int)(drec[k] >> 2);
    tmin = fminf(tmin, (c < nvf ? vf_tn[c] : ee_tn[c - nvf]).x);
  }
  const float4 pi = p[i];
  p[i] = f4(add(mul(xyz(x[i]), 1.f - tmin), mul(xyz(pi), tmin)), pi.w);
}

// Vertex normals. The reference writes each triangle's face normal once per corner
// and adds them into the per-vertex normals with float atomics, then normalises in
// place (Physics/CudaDeformableUtils.cu:190-249); the per-vertex normals are never
// cleared, so each substep adds to the previous substep's unit normal. The face
// normals are unit ones in a substep (:135) and the raw cross products at creation
// (:54, used by Physics/CudaDeformableBody.cpp:269). Here each vertex gathers its
// incident triangles (vertex ids in their original order, listed by ascending
// triangle index) and adds their face normals (unit = 1: normalised) to the carried
// normal (carry = 1) or to zero (carry = 0) in that fixed order.
__global__ void __launch_bounds__(kBlock) k_vertex_normals(int n, const int* __restrict__ vt_prefix,
                                                           const int4* __restrict__ vt_tri,
                                                           const float4* __restrict__ pos,
                                                           float4* __restrict__ nrm, int carry,
                                                           int unit) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  f3 acc = carry ? xyz(nrm[i]) : mk3(0.f, 0.f, 0.f);
  for (int k = vt_prefix[i], e = vt_prefix[i + 1]; k < e; ++k) {
    const int4 q = vt_tri[k];
    const f3 a = xyz(pos[q.x]);
    const f3 fn = cross(sub(xyz(pos[q.y]), a), sub(xyz(pos[q.z]), a));
    acc = add(acc, unit ? normalize(fn) : fn);
  }
  nrm[i] = f4(normalize(acc), 0.f);
}

// ---- adjoint of the differentiable substep (no reference counterpart) ------------
// The substep as a function of (x, v, kp): v' = v + dt g, p0 = x + dt v', I Jacobi
// iterations p_{k+1} = p_k + (sum of the corrections at p_k) / count_k over the
// distance and bend constraints and the substep's frozen VF / EE contacts, then
// x' = p_I, v' = (p_I - x) / dt for free vertices (pinned: x' = x, v' = 0). Counts
// are piecewise constant. Cotangents run in FP32; stiffness gradients sum in FP64.

// x' = p_I, v' = (p_I - x) / dt; pinned: x' = x, v' = 0
__global__ void __launch_bounds__(kBlock) k_adj_integrate(int n, const uint8_t* __restrict__ pinned,
                                                          float4* __restrict__ lx,
                                                          const float4* __restrict__ lv,
                                                          float4* __restrict__ lp, float dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (pinned[i]) {
    lp[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const f3 gv = divs(xyz(lv[i]), dt);
  lp[i] = f4(add(xyz(lx[i]), gv), 0.f);
  lx[i] = f4(neg(gv), 0.f);
}

// p0 = x + dt v', v' = v + dt g
__global__ void __launch_bounds__(kBlock) k_adj_predict(int n, float4* __restrict__ lx,
                                                        float4* __restrict__ lv,
                                                        const float4* __restrict__ lp, float dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const f3 g = xyz(lp[i]);
  lx[i] = f4(add(xyz(lx[i]), g), 0.f);
  lv[i] = f4(mul(g, dt), 0.f);
}

// The dynamic part of every iteration's Jacobi counts, all iterations at once (the
// substep's iterates are on the tape): record q adds one to dcnt[it * n + v], v its
// vertex, when its correction at the iterate pk[it] is non-zero, as k_solve counts it.
// Integer sums, so the order of the atomics does not matter.
__global__ void __launch_bounds__(kBlock) k_adj_counts(int n, int iters,
                                                       const int* __restrict__ dprefix,
                                                       const RecPayload* __restrict__ rec,
                                                       const float4* __restrict__ pk,
                                                       int* __restrict__ dcnt, float th) {
  const int nrec = dprefix[n];
  const int64_t total = (int64_t)nrec * iters;
  for (int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * kBlock) {
    const int it = (int)(k / nrec), q = (int)(k - (int64_t)it * nrec);
    const RecPayload rp = rec[q];
    const float4* ps = pk + (int64_t)it * n;
    const int role = (rp.ids.x >> 29) & 3;
    const int id0 = rp.ids.x & 0x1fffffff;
    const float4 P[4] = {ps[id0], ps[rp.ids.y], ps[rp.ids.z], ps[rp.ids.w]};
    if (!is_zero(rec_corr(rp, P, th))) {
      const int v = role == 0 ? id0 : (role == 1 ? rp.ids.y : (role == 2 ? rp.ids.z : rp.ids.w));
      atomicAdd(&dcnt[(int64_t)it * n + v], 1);
    }
  }
}

// per vertex: the iteration's count (static slots plus the non-zero dynamic records
// counted by k_adj_counts; dcnt null: no contacts) and av = lp / count
__global__ void __launch_bounds__(kBlock) k_adj_scale(int n, const int* __restrict__ vprefix,
                                                      const int* __restrict__ dcnt,
                                                      const float4* __restrict__ lp,
                                                      float4* __restrict__ av) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int cnt = vprefix[i + 1] - vprefix[i] + (dcnt ? dcnt[i] : 0);
  av[i] = cnt ? f4(divs(xyz(lp[i]), (float)cnt), 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Distance constraint: c0 = -w1 t, c1 = w2 t with t = f kp / ws, f = (l - l0) n.
// g0 / g1: cotangents of c0 / c1. Returns the cotangent of p_a (p_b gets its
// negative); gkp receives dL/dkp.
__device__ __forceinline__ f3 stretch_vjp(float4 A, float4 B, float4 g0, float4 g1, float l0,
                                          float kp, float& gkp) {
  gkp = 0.f;
  const float w1 = A.w, w2 = B.w, ws = w1 + w2;
  if (ws == 0) return mk3(0, 0, 0);
  const f3 d = sub(xyz(A), xyz(B));
  const float l = len(d);
  if (l <= 1e-8f) return mk3(0, 0, 0);
  const f3 nn = divs(d, l);
  const f3 g = sub(mul(xyz(g1), w2), mul(xyz(g0), w1));
  gkp = dot(g, mul(nn, l - l0)) / ws;
  const f3 gf = mul(g, kp / ws);
  // df/dd = I - (l0 / l)(I - n n^T)
  return sub(gf, mul(sub(gf, mul(nn, dot(nn, gf))), l0 / l));
}

// Bend constraint (the dihedral correction of bend_all), reverse mode in T: out[k]
// receives the cotangent of vertex k's position, gkp dL/dkp. The branches whose
// output is constant (degenerate triangles, d = +-1) give zero.
template <class T>
__device__ void bend_vjp(const float4 (&P)[4], const float4 (&G)[4], float kpf, float phi0f,
                         int abs_mode, f3 out[4], float& gkp) {
  out[0] = out[1] = out[2] = out[3] = mk3(0, 0, 0);
  gkp = 0.f;
  typedef tv<T> V;
  const T wv[4] = {(T)P[0].w, (T)P[1].w, (T)P[2].w, (T)P[3].w};
  V P2 = tfrom<T>(P[1]), P3 = tfrom<T>(P[2]), P4 = tfrom<T>(P[3]);
  if (!abs_mode) {
    const V p1 = tfrom<T>(P[0]);
    P2 = tsub(P2, p1);
    P3 = tsub(P3, p1);
    P4 = tsub(P4, p1);
  }
  const T kp = (T)kpf, phi0 = (T)phi0f;
  const V c23 = tcross(P2, P3), c24 = tcross(P2, P4);
  const T l23 = tlen(c23), l24 = tlen(c24);
  const T eps = (T)1e-8f;
  if (l23 <= eps || l24 <= eps) return;
  const V n1 = bdiv<T>(c23, l23), n2 = bdiv<T>(c24, l24);
  const T d = fmin(fmax((T)-1, tdot(n1, n2)), (T)1);
  if (d == (T)-1 || d == (T)1) return;
  const T phi = tacos<T>(d);
  const V A = tadd(tcross(P3, n2), tmul(tcross(n1, P3), d));
  const V B = tadd(tcross(P4, n1), tmul(tcross(n2, P4), d));
  const V C = tadd(tcross(P2, n2), tmul(tcross(n1, P2), d));
  const V D = tadd(tcross(P2, n1), tmul(tcross(n2, P2), d));
  const T i23 = (T)1 / l23, i24 = (T)1 / l24;
  // n1, n2 and the q_k as bend_all evaluates them: near d = +-1 the reverse sweep
  // amplifies any difference in them
  V q[4];
  q[2] = bdiv<T>(C, l23);
  q[3] = bdiv<T>(D, l24);
  q[1] = tsub(tneg(bdiv<T>(A, l23)), bdiv<T>(B, l24));
  q[0] = tsub(tsub(tneg(q[1]), q[2]), q[3]);
  const T r = tsqrt<T>((T)1 - d * d), e = phi - phi0;
  const T term = r * e * kp;
  T sf = 0;
  for (int k = 0; k < 4; ++k) sf += wv[k] * tdot(q[k], q[k]);
  if (sf <= eps) return;
  const T ts = term / sf;
  // c_k = -w_k ts q_k
  V qb[4];
  T tsb = 0;
  for (int k = 0; k < 4; ++k) {
    const V ak = tfrom<T>(G[k]);
    qb[k] = tmul(ak, -wv[k] * ts);
    tsb -= wv[k] * tdot(q[k], ak);
  }
  const T termb = tsb / sf, sfb = -tsb * term / (sf * sf);
  for (int k = 0; k < 4; ++k) qb[k] = tadd(qb[k], tmul(q[k], (T)2 * wv[k] * sfb));
  gkp = (float)(termb * r * e);
  const T rb = termb * e * kp, eb = termb * r * kp;
  T db = -eb / r - rb * d / r;  // phi = acos(d), r = sqrt(1 - d^2)
  // q1 = -q2 - q3 - q4
  qb[1] = tsub(qb[1], qb[0]);
  qb[2] = tsub(qb[2], qb[0]);
  qb[3] = tsub(qb[3], qb[0]);
  // q2 = -A / l23 - B / l24, q3 = C / l23, q4 = D / l24
  const V Ab = tmul(qb[1], -i23), Bb = tmul(qb[1], -i24);
  const V Cb = tmul(qb[2], i23), Db = tmul(qb[3], i24);
  T l23b = (tdot(qb[1], A) - tdot(qb[2], C)) * i23 * i23;
  T l24b = (tdot(qb[1], B) - tdot(qb[3], D)) * i24 * i24;
  V P2b = {0, 0, 0}, P3b = {0, 0, 0}, P4b = {0, 0, 0}, n1b = {0, 0, 0}, n2b = {0, 0, 0};
  // X = U x N + (M x U) d for (U, N, M) = (P3, n2, n1), (P4, n1, n2), (P2, n2, n1), (P2, n1, n2)
  auto back = [&](const V& Xb, const V& U, const V& N, const V& M, V& Ub, V& Nb, V& Mb) {
    Ub = tadd(Ub, tcross(N, Xb));
    Nb = tadd(Nb, tcross(Xb, U));
    const V t = tmul(Xb, d);
    Mb = tadd(Mb, tcross(U, t));
    Ub = tadd(Ub, tcross(t, M));
    db += tdot(Xb, tcross(M, U));
  };
  back(Ab, P3, n2, n1, P3b, n2b, n1b);
  back(Bb, P4, n1, n2, P4b, n1b, n2b);
  back(Cb, P2, n2, n1, P2b, n2b, n1b);
  back(Db, P2, n1, n2, P2b, n1b, n2b);
  // d = n1 . n2 (interior of the clamp)
  n1b = tadd(n1b, tmul(n2, db));
  n2b = tadd(n2b, tmul(n1, db));
  // n = c / l, l = |c|
  l23b -= tdot(n1b, c23) * i23 * i23;
  l24b -= tdot(n2b, c24) * i24 * i24;
  const V c23b = tadd(tmul(n1b, i23), tmul(n1, l23b));
  const V c24b = tadd(tmul(n2b, i24), tmul(n2, l24b));
  // c23 = P2 x P3, c24 = P2 x P4
  P2b = tadd(P2b, tadd(tcross(P3, c23b), tcross(P4, c24b)));
  P3b = tadd(P3b, tcross(c23b, P2));
  P4b = tadd(P4b, tcross(c24b, P2));
  out[1] = tto(P2b);
  out[2] = tto(P3b);
  out[3] = tto(P4b);
  if (!abs_mode) out[0] = tto(tneg(tadd(tadd(P2b, P3b), P4b)));
}

// One record of a frozen contact: c_r = k_r g(nn) for the contact's roles r, with
// nn = sum_r m_r p_r and g(nn) = nn - th nn / |nn| while |nn| < th (fixed weights).
// Roles whose correction is zero at P were neither summed nor counted, so their
// cotangents are left out. Returns the cotangent of this record's vertex.
__device__ __forceinline__ f3 rec_vjp(const RecPayload& rp, const float4 (&P)[4],
                                      const float4 (&G)[4], float th) {
  const int role = (rp.ids.x >> 29) & 3;
  const bool isvf = rp.ids.x < 0;
  const float4 w = rp.w;
  bool used[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    used[r] = !is_zero(isvf ? vf_corr(P[0], P[1], P[2], P[3], w, r, th)
                            : ee_corr(P[0], P[1], P[2], P[3], w, r, th));
  if (!(used[0] || used[1] || used[2] || used[3])) return mk3(0, 0, 0);
  float m[4], kr[4];
  if (isvf) {
    const float b0 = fclamp(-w.y, 0.f, 1.f), b1 = fclamp(-w.z, 0.f, 1.f),
                b2 = fclamp(-w.w, 0.f, 1.f);
    m[0] = 1.f; m[1] = -b0; m[2] = -b1; m[3] = -b2;
    const float S = P[0].w + P[1].w * b0 * b0 + P[2].w * b1 * b1 + P[3].w * b2 * b2;
    kr[0] = -P[0].w / S; kr[1] = P[1].w * b0 / S; kr[2] = P[2].w * b1 / S;
    kr[3] = P[3].w * b2 / S;
  } else {
    const float b0 = w.x, b1 = w.y, b2 = -w.z, b3 = -w.w;
    m[0] = b0; m[1] = b1; m[2] = -b2; m[3] = -b3;
    const float S = P[0].w * b0 * b0 + P[1].w * b1 * b1 + P[2].w * b2 * b2 + P[3].w * b3 * b3;
    kr[0] = -P[0].w * b0 / S; kr[1] = -P[1].w * b1 / S; kr[2] = P[2].w * b2 / S;
    kr[3] = P[3].w * b3 / S;
  }
  f3 nn = mk3(0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r) nn = add(nn, mul(xyz(P[r]), m[r]));
  const float l = len(nn);
  const f3 nh = divs(nn, l);
  f3 g = mk3(0, 0, 0);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (used[r]) g = add(g, mul(xyz(G[r]), kr[r]));
  // dg/dnn = I - (th / l)(I - n n^T)
  const f3 gn = sub(g, mul(sub(g, mul(nh, dot(nh, g))), th / l));
  return mul(gn, m[role]);
}

struct AdjArgs {
  int abs_mode, bend_fp64;
  float th;
  const int* vprefix;
  const int* tstart;
  const int* eprefix;
  const TileEntry* entries;
  const int* ecid;  // per entry: constraint index, bit 31 set for a bend
  const int* dprefix;
  const RecPayload* rec;
  double* gks;  // dL/dkp per distance constraint
  double* gkb;  // dL/dkp per bend constraint
};

// One reverse Jacobi iteration over a vertex tile, in the layout of k_solve<true>:
// every constraint touching the tile is differentiated once, its vertex cotangents
// go to the tile's slots, and each vertex adds its slots in slot order, then its
// dynamic records in record order: la = lp + sum over slots and records. The tile
// holding a constraint's role-0 vertex adds its dL/dkp (one owner per constraint).
__global__ void __launch_bounds__(kBlock) k_adj_iter(AdjArgs a, const float4* __restrict__ ps,
                                                     const float4* __restrict__ av,
                                                     const float4* __restrict__ lp,
                                                     float4* __restrict__ la) {
  __shared__ float cx[kLds], cy[kLds], cz[kLds];
  const int tile = xcd_block(blockIdx.x, gridDim.x);  // contiguous tiles per XCD
  const int v0 = a.tstart[tile], v1 = a.tstart[tile + 1];
  const int tid = threadIdx.x;
  const int i = v0 + tid;
  const bool mine = i < v1;
  f3 total = mk3(0, 0, 0);
  {
    const int s0 = a.vprefix[v0];
    const int eb = a.eprefix[tile], ee = a.eprefix[tile + 1];
    for (int q = eb + tid; q < ee; q += kBlock) {
      const TileEntry t = a.entries[q];
      const int cid = a.ecid[q];
      float gkp;
      if (t.ids.z < 0) {
        const float4 P0 = ps[t.ids.x], P1 = ps[t.ids.y];
        const float4 G0 = av[t.ids.x], G1 = av[t.ids.y];
        const f3 gd = stretch_vjp(P0, P1, G0, G1, t.a, t.b, gkp);
        if ((t.s01 & 0xffffu) != 0xffffu) {
          lds_put(cx, cy, cz, t.s01 & 0xffffu, gd);
          a.gks[cid] += (double)gkp;
        }
        if ((t.s01 >> 16) != 0xffffu) lds_put(cx, cy, cz, t.s01 >> 16, neg(gd));
      } else {
        const float4 P[4] = {ps[t.ids.x], ps[t.ids.y], ps[t.ids.z], ps[t.ids.w]};
        const float4 G[4] = {av[t.ids.x], av[t.ids.y], av[t.ids.z], av[t.ids.w]};
        f3 o[4];
        if (a.bend_fp64)
          bend_vjp<double>(P, G, t.a, t.b, a.abs_mode, o, gkp);
        else
          bend_vjp<float>(P, G, t.a, t.b, a.abs_mode, o, gkp);
        if ((t.s01 & 0xffffu) != 0xffffu) {
          lds_put(cx, cy, cz, t.s01 & 0xffffu, o[0]);
          a.gkb[cid & 0x7fffffff] += (double)gkp;
        }
        if ((t.s01 >> 16) != 0xffffu) lds_put(cx, cy, cz, t.s01 >> 16, o[1]);
        if ((t.s23 & 0xffffu) != 0xffffu) lds_put(cx, cy, cz, t.s23 & 0xffffu, o[2]);
        if ((t.s23 >> 16) != 0xffffu) lds_put(cx, cy, cz, t.s23 >> 16, o[3]);
      }
    }
    __syncthreads();
    if (mine) {
      const int e = a.vprefix[i + 1];
      for (int k = a.vprefix[i]; k < e; ++k)
        total = add(total, mk3(cx[k - s0], cy[k - s0], cz[k - s0]));
    }
    __syncthreads();
  }
  if (a.dprefix) {
    const int r0 = a.dprefix[v0], r1 = a.dprefix[v1];
    int k = mine ? a.dprefix[i] : 0;
    const int e = mine ? a.dprefix[i + 1] : 0;
    for (int c0 = r0; c0 < r1; c0 += kLds) {
      const int c1 = min(c0 + kLds, r1);
      for (int q = c0 + tid; q < c1; q += kBlock) {
        const RecPayload rp = a.rec[q];
        const int id0 = rp.ids.x & 0x1fffffff;
        const float4 P[4] = {ps[id0], ps[rp.ids.y], ps[rp.ids.z], ps[rp.ids.w]};
        const float4 G[4] = {av[id0], av[rp.ids.y], av[rp.ids.z], av[rp.ids.w]};
        lds_put(cx, cy, cz, q - c0, rec_vjp(rp, P, G, a.th));
      }
      __syncthreads();
      const int kend = min(e, c1);
      for (; k < kend; ++k) total = add(total, mk3(cx[k - c0], cy[k - c0], cz[k - c0]));
      __syncthreads();
    }
  }
  if (mine) la[i] = f4(add(xyz(lp[i]), total), 0.f);
}

// Tape of a recorded substep's contact set: slot `slot` of ctl ([0] records used, [1]
// overflow flag, [2 + slot] the slot's first record or -1) takes the next dprefix[n]
// records of the arena. The cursor advances even past the capacity, so the host
// learns what a segment needs.
__global__ void k_tape_claim(int n, const int* __restrict__ dprefix, int64_t cap,
                             int* __restrict__ ctl, int slot) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int off = ctl[0], nrec = dprefix[n];
  const bool fits = (int64_t)off + nrec <= cap;
  ctl[2 + slot] = fits ? off : -1;
  if (!fits) ctl[1] = 1;
  ctl[0] = off + nrec;
}

__global__ void __launch_bounds__(kBlock) k_tape_copy(int n, const int* __restrict__ dprefix,
                                                      const RecPayload* __restrict__ rec,
                                                      RecPayload* __restrict__ arena,
                                                      const int* __restrict__ ctl, int slot) {
  const int off = ctl[2 + slot];
  if (off < 0) return;
  const int nrec = dprefix[n];
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < nrec; k += gridDim.x * kBlock)
    arena[(int64_t)off + k] = rec[k];
}

}  // namespace

int launch_predict(const PbdBufs& b, const PbdParams& pp, hipStream_t s) {
  const int nb = nblocks(b.n);
  if (pp.flags & CM_F_DAMPING) {
    hipLaunchKernelGGL(k_damp_sum1, dim3(nb), dim3(kBlock), 0, s, b.n, b.p, b.v, b.mass,
                       mul(pp.g, pp.dt), pp.kg, b.damp_partials);
    hipLaunchKernelGGL(k_damp_mid, dim3(1), dim3(64), 0, s, nb, b.damp_partials, b.damp_state);
    hipLaunchKernelGGL(k_damp_sum2, dim3(nb), dim3(kBlock), 0, s, b.n, b.p, b.v, b.mass,
                       b.damp_state, b.damp_partials);
    hipLaunchKernelGGL(k_damp_final, dim3(1), dim3(64), 0, s, nb, b.damp_partials, b.damp_state);
    hipLaunchKernelGGL(k_damp_apply, dim3(nb), dim3(kBlock), 0, s, b.n, b.x, b.p, b.v,
                       b.damp_state, pp.kd, pp.dt);
  } else {
    hipLaunchKernelGGL(k_predict, dim3(nb), dim3(kBlock), 0, s, b.n, b.x, b.v, b.p, pp.dt, pp.g);
  }
  CM_LAUNCH_CHECK();
  return CM_OK;
}

static SolveArgs solve_args(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb) {
  const int ground_on = (pp.flags & CM_F_GROUND) ? 1 : 0;
  const int F = (pp.flags & CM_F_LRA) ? b.F : 0;
  SolveArgs a;
  a.n = b.n;
  a.F = F;
  a.abs_mode = (pp.ref_compat & CM_RC_BEND_ABS) ? 1 : 0;
  a.bend_fp64 = (pp.flags & CM_F_BEND_FP64) ? 1 : 0;
  a.ground_on = ground_on;
  a.th = pp.th;
  a.ground = b.ground;
  a.lra_fixed = b.lra_fixed;
  a.lra_l0 = b.lra_l0;
  a.lra_flag = b.lra_flag;
  a.vprefix = b.vprefix;
  a.vslot = b.vslot;
  a.tstart = b.tstart;
  a.eprefix = b.eprefix;
  a.entries = b.entries;
  a.sid = b.sid;
  a.s_l0kp = b.s_l0kp;
  a.bid = b.bid;
  a.b_kpphi0 = b.b_kpphi0;
  const bool coll = (pp.flags & CM_F_SELF_COLLISION) && nb.dprefix;
  a.dprefix = coll ? nb.dprefix : nullptr;
  a.rec = nb.rec;
  return a;
}

// `iters` Jacobi iterations ping-ponging p -> p2 -> p ...
static int launch_iterations(const PbdBufs& b, const SolveArgs& a, int iters, hipStream_t s) {
  float4* src = b.p;
  float4* dst = b.p2;
  for (int it = 0; it < iters; ++it) {
    if (b.ntiles > 0)
      hipLaunchKernelGGL(k_solve<true>, dim3(b.ntiles), dim3(kBlock), 0, s, a, src, dst);
    else
      hipLaunchKernelGGL(k_solve<false>, dim3((b.n + kSolveV - 1) / kSolveV), dim3(kBlock), 0, s,
                         a, src, dst);
    std::swap(src, dst);
  }
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_solver(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, hipStream_t s) {
  const int g = nblocks(b.n);
  const int ground_on = (pp.flags & CM_F_GROUND) ? 1 : 0;
  const int F = (pp.flags & CM_F_LRA) ? b.F : 0;
  if (ground_on || F)  // otherwise the set-up writes nothing
    hipLaunchKernelGGL(k_setup, dim3(g), dim3(kBlock), 0, s, b.n, b.x, b.p, pp.gh, pp.th,
                       ground_on, b.ground, F, b.lra_fixed, b.lra_l0, b.pinned, b.lra_flag);
  if (pp.flags & CM_F_SPHERE)
    hipLaunchKernelGGL(k_sphere, dim3(g), dim3(kBlock), 0, s, b.n, b.x, b.p, pp.sc, pp.sr + pp.th);
  CM_LAUNCH_CHECK();
  return launch_iterations(b, solve_args(b, pp, nb), pp.iters, s);
}

int launch_solve_only(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, int iters,
                      hipStream_t s) {
  return launch_iterations(b, solve_args(b, pp, nb), iters, s);
}

int launch_integrate(const PbdBufs& b, const PbdParams& pp, hipStream_t s) {
  return launch_integrate_from(b, pp, (pp.iters & 1) ? b.p2 : b.p, s);
}

int launch_integrate_from(const PbdBufs& b, const PbdParams& pp, const float4* pfin,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_integrate, dim3(nblocks(b.n)), dim3(kBlock), 0, s, b.n, pfin, b.x, b.v,
                     b.p, b.pinned, pp.dt, (pp.flags & CM_F_WORLD_CLAMP) ? 1 : 0, pp.wmin,
                     pp.wmax);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_vertex_normals(int n, const int* vt_prefix, const int4* vt_tri, const float4* pos,
                          float4* nrm, int carry, int unit, hipStream_t s) {
  hipLaunchKernelGGL(k_vertex_normals, dim3(nblocks(n)), dim3(kBlock), 0, s, n, vt_prefix, vt_tri,
                     pos, nrm, carry, unit);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_revert_toc(const PbdBufs& b, const NarrowBufs& nb, hipStream_t s) {
  hipLaunchKernelGGL(k_revert_toc, dim3(nblocks(b.n)), dim3(kBlock), 0, s, b.n, nb.dprefix,
                     nb.drec, nb.d_nvf, nb.vf_tn, nb.ee_tn, b.x, b.p);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_friction(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, hipStream_t s) {
  const int g = nblocks(b.n);
  const int rigid_n = (pp.ref_compat & CM_RC_RIGID_FRICTION_TAIL) ? std::min(b.n, g * (kBlock - 1))
                                                                   : b.n;
  const float4* ground = (pp.flags & CM_F_GROUND) ? b.ground : nullptr;
  const bool coll = (pp.flags & CM_F_SELF_COLLISION) && nb.dprefix;
  hipLaunchKernelGGL(k_friction, dim3(g), dim3(kBlock), 0, s, b.n, rigid_n, ground,
                     coll ? nb.dprefix : nullptr, nb.drec, nb.d_nvf, nb.vf_tn, nb.ee_tn, b.v,
                     pp.kf);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_solver_rec(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, float4* pk,
                      hipStream_t s) {
  const int g = nblocks(b.n);
  const int ground_on = (pp.flags & CM_F_GROUND) ? 1 : 0;
  const int F = (pp.flags & CM_F_LRA) ? b.F : 0;
  if (ground_on || F)  // otherwise the set-up writes nothing
    hipLaunchKernelGGL(k_setup, dim3(g), dim3(kBlock), 0, s, b.n, b.x, b.p, pp.gh, pp.th,
                       ground_on, b.ground, F, b.lra_fixed, b.lra_l0, b.pinned, b.lra_flag);
  if (pp.flags & CM_F_SPHERE)
    hipLaunchKernelGGL(k_sphere, dim3(g), dim3(kBlock), 0, s, b.n, b.x, b.p, pp.sc, pp.sr + pp.th);
  CM_LAUNCH_CHECK();
  CM_HIP(hipMemcpyAsync(pk, b.p, sizeof(float4) * b.n, hipMemcpyDeviceToDevice, s));
  const SolveArgs a = solve_args(b, pp, nb);
  for (int it = 0; it < pp.iters; ++it) {
    const float4* src = pk + (int64_t)it * b.n;
    float4* dst = pk + (int64_t)(it + 1) * b.n;
    if (b.ntiles > 0)
      hipLaunchKernelGGL(k_solve<true>, dim3(b.ntiles), dim3(kBlock), 0, s, a, src, dst);
    else
      hipLaunchKernelGGL(k_solve<false>, dim3((b.n + kSolveV - 1) / kSolveV), dim3(kBlock), 0, s,
                         a, src, dst);
  }
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_tape_records(int n, const int* dprefix, const RecPayload* rec, RecPayload* arena,
                        int64_t cap, int* ctl, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_tape_claim, dim3(1), dim3(64), 0, s, n, dprefix, cap, ctl, slot);
  hipLaunchKernelGGL(k_tape_copy, dim3(1024), dim3(kBlock), 0, s, n, dprefix, rec, arena, ctl,
                     slot);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

int launch_adj_substep(const PbdBufs& b, const PbdParams& pp, const AdjBufs& ab, const float4* pk,
                       const int* dprefix, const RecPayload* rec, hipStream_t s) {
  if (b.ntiles <= 0) {
    set_error("adjoint: needs the tiled Jacobi layout");
    return CM_ERR_ARG;
  }
  const int g = nblocks(b.n);
  hipLaunchKernelGGL(k_adj_integrate, dim3(g), dim3(kBlock), 0, s, b.n, b.pinned, ab.lx, ab.lv,
                     ab.lp, pp.dt);
  AdjArgs a;
  a.abs_mode = (pp.ref_compat & CM_RC_BEND_ABS) ? 1 : 0;
  a.bend_fp64 = (pp.flags & CM_F_BEND_FP64) ? 1 : 0;
  a.th = pp.th;
  a.vprefix = b.vprefix;
  a.tstart = b.tstart;
  a.eprefix = b.eprefix;
  a.entries = b.entries;
  a.ecid = ab.ecid;
  a.dprefix = dprefix;
  a.rec = rec;
  a.gks = ab.gks;
  a.gkb = ab.gkb;
  if (dprefix) {
    CM_HIP(hipMemsetAsync(ab.dcnt, 0, sizeof(int) * (size_t)pp.iters * b.n, s));
    hipLaunchKernelGGL(k_adj_counts, dim3(1024), dim3(kBlock), 0, s, b.n, pp.iters, dprefix, rec,
                       pk, ab.dcnt, pp.th);
  }
  float4* src = ab.lp;
  float4* dst = ab.la;
  for (int it = pp.iters - 1; it >= 0; --it) {
    const float4* ps = pk + (int64_t)it * b.n;
    hipLaunchKernelGGL(k_adj_scale, dim3(g), dim3(kBlock), 0, s, b.n, b.vprefix,
                       dprefix ? ab.dcnt + (int64_t)it * b.n : nullptr, src, ab.av);
    hipLaunchKernelGGL(k_adj_iter, dim3(b.ntiles), dim3(kBlock), 0, s, a, ps, ab.av, src, dst);
    std::swap(src, dst);
  }
  hipLaunchKernelGGL(k_adj_predict, dim3(g), dim3(kBlock), 0, s, b.n, ab.lx, ab.lv, src, pp.dt);
  CM_LAUNCH_CHECK();
  return CM_OK;
}

}  // namespace cm
