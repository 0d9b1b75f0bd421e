// Mass-spring cloth of the reference's GPU path. No reference configuration reaches
// it: FixedUpdate's call of ClothInternalDynamics is commented out
// (Physics/CudaDeformableBody.cpp:605) and so is the code that builds the spring
// network it reads (:99-152). Unit-level entry points on caller buffers:
//
//  cm_ms_grid_springs <- the grid spring network (:107-152) with AddSpring (:288-311)
//  cm_ms_step         <- ClothInternalDynamics (:327-357): ClothEngineStep
//                        (Physics/CudaMassSpring.cu:22-88), then AdjustSprings
//                        (:95-148, its kernel _AdjustSprings :121-148)
//
// A spring whose end ids fall outside [0, n) is skipped (the reference assumes its
// own network). The reference updates positions in place while other threads read
// their neighbours' (a race whose outcome depends on the schedule). Here every read of
// another vertex's position sees the value it had when the kernel started:
// k_ms_engine reads a snapshot of the positions taken before it, and each
// adjustment iteration reads a snapshot taken before that iteration. A vertex's own
// position changes as its thread walks its springs, as in the reference.
#include <cmath>
#include <vector>

#include "cm_internal.hpp"

namespace cm {
namespace {

__global__ void __launch_bounds__(kBlock) k_ms_snapshot(int n, const float4* __restrict__ pos,
                                                        float4* __restrict__ snap) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) snap[i] = pos[i];
}

// ClothEngineStep: gravity, damping against the previous position, the pull of
// every stretched spring listed at the vertex, and a Verlet step
__global__ void __launch_bounds__(kBlock) k_ms_engine(
    int n, const float4* __restrict__ snap, float4* __restrict__ pos, float4* __restrict__ prev,
    float4* __restrict__ vel, const float* __restrict__ mass, const uint8_t* __restrict__ fixed,
    const int* __restrict__ aid, const int* __restrict__ bid, const float* __restrict__ ks,
    const float* __restrict__ l0, const int* __restrict__ sinfo, const int* __restrict__ sids,
    f3 g, float kd, float dt) {
  const int id = blockIdx.x * kBlock + threadIdx.x;
  if (id >= n || fixed[id]) return;
  const float m = mass[id];
  const float4 p4 = snap[id], q4 = prev[id];
  const f3 pi = xyz(p4), qi = xyz(q4);
  f3 force = sub(mul(g, m), mul(sub(pi, qi), kd));
  for (int k = sids[id]; k < sids[id + 1]; ++k) {
    const int s = sinfo[k];
    if ((unsigned)aid[s] >= (unsigned)n || (unsigned)bid[s] >= (unsigned)n) continue;
    const f3 dir = sub(xyz(snap[aid[s]]), xyz(snap[bid[s]]));
    if (len(dir) > l0[s]) force = sub(force, mul(sub(dir, mul(normalize(dir), l0[s])), ks[s]));
  }
  const f3 accel = divs(force, m);
  const f3 np = add(sub(mul(pi, 2.f), qi), mul(mul(accel, dt), dt));
  prev[id] = f4(pi, q4.w);
  pos[id] = f4(np, p4.w);
  vel[id] = f4(divs(sub(np, pi), dt), vel[id].w);
}

// one _AdjustSprings iteration: a vertex whose non-bend spring is stretched beyond
// max_def moves towards the other end by the excess (half of it when that end is
// free), then its velocity follows from the previous position
__global__ void __launch_bounds__(kBlock) k_ms_adjust(
    int n, const float4* __restrict__ snap, float4* __restrict__ pos,
    const float4* __restrict__ prev, float4* __restrict__ vel, const uint8_t* __restrict__ fixed,
    const int* __restrict__ bid, const float* __restrict__ ks, const float* __restrict__ l0,
    const int* __restrict__ sinfo, const int* __restrict__ sids, float k_bend, float max_def,
    float dt) {
  const int id = blockIdx.x * kBlock + threadIdx.x;
  if (id >= n || fixed[id]) return;
  const float4 p4 = snap[id];
  f3 p = xyz(p4);
  for (int k = sids[id]; k < sids[id + 1]; ++k) {
    const int s = sinfo[k];
    const int b = bid[s];
    if ((unsigned)b >= (unsigned)n) continue;
    const f3 pb = xyz(snap[b]);
    const float def = (len(sub(p, pb)) - l0[s]) / l0[s];
    if (def > max_def && ks[s] != k_bend) {
      const float dif = (def - max_def) * l0[s];
      p = add(p, divs(mul(normalize(sub(pb, p)), dif), fixed[b] ? 1.f : 2.f));
    }
  }
  pos[id] = f4(p, p4.w);
  vel[id] = f4(divs(sub(p, xyz(prev[id])), dt), vel[id].w);
}

// AddSpring (Physics/CudaDeformableBody.cpp:288-311): the spring id1 -> id2 listed at
// id1, then id2 -> id1 listed at id2, both with the rest distance of the two
struct SpringNet {
  int cols;
  const float* xyz;
  std::vector<int> a, b;
  std::vector<float> ks, l0;
  std::vector<std::vector<int>> at;
  void add(int i1, int j1, int i2, int j2, float k) {
    const int id1 = i1 * cols + j1, id2 = i2 * cols + j2;
    const f3 p1 = mk3(xyz[3 * id1], xyz[3 * id1 + 1], xyz[3 * id1 + 2]);
    const f3 p2 = mk3(xyz[3 * id2], xyz[3 * id2 + 1], xyz[3 * id2 + 2]);
    const float d = len(sub(p2, p1));
    a.push_back(id1);
    b.push_back(id2);
    ks.push_back(k);
    l0.push_back(d);
    at[id1].push_back((int)a.size() - 1);
    a.push_back(id2);
    b.push_back(id1);
    ks.push_back(k);
    l0.push_back(d);
    at[id2].push_back((int)a.size() - 1);
  }
};

}  // namespace
}  // namespace cm

using namespace cm;

extern "C" int cm_ms_grid_springs(int32_t rows, int32_t cols, const float* xyz, float k_stretch,
                                  float k_shear, float k_bend, int32_t* a_ids, int32_t* b_ids,
                                  float* ks, float* l0s, int32_t* spring_info,
                                  int32_t* spring_ids, int32_t* n_springs) {
  if (rows < 1 || cols < 1 || (int64_t)rows * cols >= (1 << 28) || !xyz || !n_springs) {
    set_error("cm_ms_grid_springs: invalid argument");
    return CM_ERR_ARG;
  }
  SpringNet net;
  net.cols = cols;
  net.xyz = xyz;
  net.at.resize((size_t)rows * cols);
  // Physics/CudaDeformableBody.cpp:107-133, the square dims.x x dims.x loops taken
  // to rows x cols
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      if (i + 1 < rows) net.add(i, j, i + 1, j, k_stretch);
      if (j + 1 < cols) net.add(i, j, i, j + 1, k_stretch);
      if (i + 1 < rows && j + 1 < cols) {
        net.add(i, j, i + 1, j + 1, k_shear);
        net.add(i + 1, j, i, j + 1, k_shear);
      }
      if (i + 2 < rows) net.add(i, j, i + 2, j, k_bend);
      if (j + 2 < cols) net.add(i, j, i, j + 2, k_bend);
    }
  const int ns = (int)net.a.size();
  *n_springs = ns;
  for (int s = 0; s < ns; ++s) {
    if (a_ids) a_ids[s] = net.a[s];
    if (b_ids) b_ids[s] = net.b[s];
    if (ks) ks[s] = net.ks[s];
    if (l0s) l0s[s] = net.l0[s];
  }
  // springIDs: prefix of the per-vertex counts; linSpringInfo: the lists in vertex order
  int o = 0;
  for (size_t v = 0; v < net.at.size(); ++v) {
    if (spring_ids) spring_ids[v] = o;
    for (int s : net.at[v]) {
      if (spring_info) spring_info[o] = s;
      ++o;
    }
  }
  if (spring_ids) spring_ids[net.at.size()] = o;
  return CM_OK;
}

extern "C" int cm_ms_step(int32_t n, void* d_pos, void* d_prev, void* d_vel, const float* d_mass,
                          const uint8_t* d_fixed, const int32_t* d_a, const int32_t* d_b,
                          const float* d_ks, const float* d_l0, const int32_t* d_spring_info,
                          const int32_t* d_spring_ids, const float* gravity, float k_damp,
                          float dt, float k_bend, int32_t adjust_iters, float max_deformation,
                          void* d_temp, void* stream) {
  if (n < 1 || !d_pos || !d_prev || !d_vel || !d_mass || !d_fixed || !d_a || !d_b || !d_ks ||
      !d_l0 || !d_spring_info || !d_spring_ids || !gravity || !d_temp || adjust_iters < 0 ||
      !(dt > 0.f)) {
    set_error("cm_ms_step: invalid argument");
    return CM_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  float4* pos = (float4*)d_pos;
  float4* snap = (float4*)d_temp;
  const int g = nblocks(n);
  hipLaunchKernelGGL(k_ms_snapshot, dim3(g), dim3(kBlock), 0, s, n, pos, snap);
  hipLaunchKernelGGL(k_ms_engine, dim3(g), dim3(kBlock), 0, s, n, snap, pos, (float4*)d_prev,
                     (float4*)d_vel, d_mass, d_fixed, d_a, d_b, d_ks, d_l0, d_spring_info,
                     d_spring_ids, mk3(gravity[0], gravity[1], gravity[2]), k_damp, dt);
  for (int it = 0; it < adjust_iters; ++it) {
    hipLaunchKernelGGL(k_ms_snapshot, dim3(g), dim3(kBlock), 0, s, n, pos, snap);
    hipLaunchKernelGGL(k_ms_adjust, dim3(g), dim3(kBlock), 0, s, n, snap, pos,
                       (const float4*)d_prev, (float4*)d_vel, d_fixed, d_b, d_ks, d_l0,
                       d_spring_info, d_spring_ids, k_bend, max_deformation, dt);
  }
  CM_LAUNCH_CHECK();
  return CM_OK;
}
