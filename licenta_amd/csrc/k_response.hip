// Impulse response of the reference's GPU path. No reference configuration reaches
// it: it runs from SolveCollisions inside HandleCollisionsContinuous, whose call is
// commented out in FixedUpdate (Physics/CudaDeformableBody.cpp:605-613). Unit-level
// entry point on caller buffers:
//
//  cm_impulses <- SolveCollisions (Physics/CudaDeformableBody.cpp:525-548):
//     CreateImpulses (Collision/CudaCollisionResponse.cu:11-139): one impulse per
//       contact vertex (_CreateImpulses / _CreateImpulseVF / _CreateImpulseEE /
//       _CreateImpulse :142-231), a stable sort of the impulses by vertex id, the
//       runs of equal ids, and AccumulateImpulses (:233-256): each vertex's impulse
//       is the mean of its run, summed in sorted order;
//     ApplyImpulses (:258-297): v += impulse / m, p = x + v dt for free vertices.
//
// The stable LSD radix sort of the LBVH keys the impulses by vertex id with the
// impulse index as value, which is the order of the reference's CUB sort; each run's
// first position then sums the run.
#include "cm_internal.hpp"

namespace cm {
namespace {

__device__ __forceinline__ f3 vec3(const float* a) { return mk3(a[0], a[1], a[2]); }

// _CreateImpulse for contact c: record 4c + k is vertex v[k]'s impulse. A contact
// with a vertex id outside [0, n) gives no impulse (its records get key n, which
// the accumulation skips).
__global__ void __launch_bounds__(kBlock) k_imp_create(
    int n, int nvf, int nee, const cm_contact* __restrict__ vf, const cm_contact* __restrict__ ee,
    const float4* __restrict__ x, const float4* __restrict__ v, float stiffness, float m,
    float dt, float th, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
    float4* __restrict__ imp) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nvf + nee) return;
  const bool isvf = c < nvf;
  const cm_contact ct = isvf ? vf[c] : ee[c - nvf];
  bool valid = true;
#pragma unroll
  for (int k = 0; k < 4; ++k) valid = valid && ct.v[k] >= 0 && ct.v[k] < n;
  if (!valid) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      keys[4 * c + k] = (uint64_t)n;
      vals[4 * c + k] = (uint32_t)(4 * c + k);
      imp[4 * c + k] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    return;
  }
  const float w1 = ct.w[0], w2 = ct.w[1], w3 = ct.w[2], w4 = ct.w[3];
  const f3 u1 = xyz(v[ct.v[0]]), u2 = xyz(v[ct.v[1]]), u3 = xyz(v[ct.v[2]]), u4 = xyz(v[ct.v[3]]);
  f3 cv1, cv2;
  if (isvf) {  // _CreateImpulseVF (:160-170)
    cv1 = mul(u1, w1);
    cv2 = add(add(mul(u2, w2), mul(u3, w3)), mul(u4, w4));
  } else {  // _CreateImpulseEE (:172-182)
    cv1 = add(mul(u1, w1), mul(u2, w2));
    cv2 = add(mul(u3, w3), mul(u4, w4));
  }
  const f3 vr = add(cv1, cv2);
  const f3 nn = vec3(ct.n);
  const f3 xs = add(add(add(mul(xyz(x[ct.v[0]]), w1), mul(xyz(x[ct.v[1]]), w2)),
                        mul(xyz(x[ct.v[2]]), w3)),
                    mul(xyz(x[ct.v[3]]), w4));
  const float d = th - dot(xs, nn);
  const float vn = dot(vr, nn);
  const float dts = (0.1f * d) / dt;
  float i = 0.f;
  if (vn < dts) i = -fminf(dt * stiffness * d, m * (dts - vn));
  if (vn < 0) i += m * vn * 0.5f;
  i = fclamp(i, -m * 0.1f * th, m * 0.1f * th);
  const float ip = (2 * i) / (w1 * w1 + w2 * w2 + w3 * w3 + w4 * w4);
  const f3 ipv = mul(nn, ip);
  const float w[4] = {w1, w2, w3, w4};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int r = 4 * c + k;
    keys[r] = (uint64_t)(uint32_t)ct.v[k];
    vals[r] = (uint32_t)r;
    imp[r] = f4(mul(ipv, w[k]), 0.f);
  }
}

__global__ void __launch_bounds__(kBlock) k_imp_zero(int n, float4* __restrict__ acc) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) acc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// AccumulateImpulses: the first position of each run of equal vertex ids sums the run
// in sorted order and writes its mean (keys >= n: contacts with invalid ids)
__global__ void __launch_bounds__(kBlock) k_imp_accumulate(int n, int R,
                                                           const uint64_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals,
                                                           const float4* __restrict__ imp,
                                                           float4* __restrict__ acc) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= R || (q > 0 && keys[q] == keys[q - 1])) return;
  const uint64_t key = keys[q];
  if (key >= (uint64_t)n) return;
  f3 s = mk3(0.f, 0.f, 0.f);
  int e = q;
  for (; e < R && keys[e] == key; ++e) s = add(s, xyz(imp[vals[e]]));
  acc[key] = f4(divs(s, (float)(e - q)), 0.f);
}

// ApplyImpulses (:276-297)
__global__ void __launch_bounds__(kBlock) k_imp_apply(int n, const uint8_t* __restrict__ fixed,
                                                      const float4* __restrict__ acc,
                                                      const float4* __restrict__ x,
                                                      float4* __restrict__ v, float4* __restrict__ p,
                                                      float m, float dt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || fixed[i]) return;
  const float4 v4 = v[i], p4 = p[i];
  const f3 nv = add(xyz(v4), divs(xyz(acc[i]), m));
  v[i] = f4(nv, v4.w);
  p[i] = f4(add(xyz(x[i]), mul(nv, dt)), p4.w);
}

struct ImpTemp {
  uint64_t* keys;
  uint32_t* vals;
  float4* imp;
  float4* acc;
  void* sort;
  size_t bytes;
};

ImpTemp imp_layout(char* base, int n, int64_t R) {
  ImpTemp t;
  size_t o = 0;
  t.keys = (uint64_t*)(base + o);
  o = align_up(o + sizeof(uint64_t) * (size_t)R);
  t.vals = (uint32_t*)(base + o);
  o = align_up(o + sizeof(uint32_t) * (size_t)R);
  t.imp = (float4*)(base + o);
  o = align_up(o + sizeof(float4) * (size_t)R);
  t.acc = (float4*)(base + o);
  o = align_up(o + sizeof(float4) * (size_t)n);
  t.sort = base + o;
  o += radix_sort_temp_bytes((int)R);
  t.bytes = o;
  return t;
}

// bits that hold every key in [0, m)
int key_bits_for(int64_t m) {
  int b = 1;
  while (b < 62 && ((int64_t)1 << b) < m) ++b;
  return b;
}

}  // namespace
}  // namespace cm

using namespace cm;

extern "C" size_t cm_impulses_temp_bytes(int32_t n, int32_t n_contacts) {
  if (n < 1 || n_contacts < 0 || (int64_t)n_contacts * 4 >= (int64_t)1 << 31) return 0;
  return imp_layout(nullptr, n, 4 * (int64_t)n_contacts).bytes;
}

extern "C" int cm_impulses(int32_t n, const void* d_x, void* d_v, void* d_p,
                           const uint8_t* d_fixed, const cm_contact* d_vf, int32_t n_vf,
                           const cm_contact* d_ee, int32_t n_ee, float stiffness,
                           float vertex_mass, float dt, float thickness, void* d_temp,
                           size_t temp_bytes, void* stream) {
  if (n < 1 || n_vf < 0 || n_ee < 0 || (int64_t)(n_vf + (int64_t)n_ee) * 4 >= (int64_t)1 << 31 ||
      !d_x || !d_v || !d_p || !d_fixed || (n_vf && !d_vf) || (n_ee && !d_ee) || !d_temp ||
      !(dt > 0.f)) {
    set_error("cm_impulses: invalid argument");
    return CM_ERR_ARG;
  }
  const int nc = n_vf + n_ee;
  const int64_t R = 4 * (int64_t)nc;
  if (temp_bytes < imp_layout(nullptr, n, R).bytes) {
    set_error("cm_impulses: temp buffer too small");
    return CM_ERR_ARG;
  }
  const ImpTemp t = imp_layout((char*)d_temp, n, R);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_imp_zero, dim3(nblocks(n)), dim3(kBlock), 0, s, n, t.acc);
  if (nc > 0) {
    hipLaunchKernelGGL(k_imp_create, dim3(nblocks(nc)), dim3(kBlock), 0, s, n, n_vf, n_ee, d_vf,
                       d_ee, (const float4*)d_x, (const float4*)d_v, stiffness, vertex_mass, dt,
                       thickness, t.keys, t.vals, t.imp);
    CM_LAUNCH_CHECK();
    int rc = radix_sort_clear(t.sort, (int)R, s);
    if (!rc) rc = radix_sort(t.keys, t.vals, (int)R, key_bits_for(n + 1), t.sort, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_imp_accumulate, dim3(nblocks(R)), dim3(kBlock), 0, s, n, (int)R, t.keys,
                       t.vals, t.imp, t.acc);
  }
  hipLaunchKernelGGL(k_imp_apply, dim3(nblocks(n)), dim3(kBlock), 0, s, n, d_fixed, t.acc,
                     (const float4*)d_x, (float4*)d_v, (float4*)d_p, vertex_mass, dt);
  CM_LAUNCH_CHECK();
  return CM_OK;
}
