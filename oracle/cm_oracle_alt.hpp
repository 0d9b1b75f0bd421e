// ORACLE — CPU restatement of two dead paths of the reference's GPU cloth, the
// mass-spring integrator and the impulse response. Test infrastructure like
// cm_oracle.hpp: only tests/ load it, as the checker of cm_ms_grid_springs,
// cm_ms_step and cm_impulses.
//
// Defined choices where the reference is racy (each pinned by
// tests/test_oracle_alt.py):
//  * ClothEngineStep and _AdjustSprings update positions in place while other
//    threads read them (Physics/CudaMassSpring.cu:56-86, 130-146); here a vertex
//    reads the other vertices' positions as they were when the pass started (the
//    pass works from a copy), and its own as its loop leaves it.
//  * vector length: sqrt(x*x+y*y+z*z) in place of norm3df, as in cm_oracle.hpp;
//  * springs with an end outside [0, n) and contacts with a vertex id outside it
//    are skipped (the reference assumes valid input).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "cm_oracle.hpp"

namespace orc {
namespace alt {

// Physics/CudaDeformableBody.cpp:288-311 (AddSpring) driven by the loops of :107-133
// over a rows x cols grid (the reference's loops are dims.x x dims.x).
struct Springs {
  std::vector<int> a, b, info, ids;
  std::vector<float> ks, l0;
};

inline Springs grid_springs(int rows, int cols, const float* xyz, float k_stretch, float k_shear,
                            float k_bend) {
  Springs S;
  std::vector<std::vector<int>> at((size_t)rows * cols);
  auto lin = [&](int i, int j) { return i * cols + j; };
  auto add = [&](int i1, int j1, int i2, int j2, float k) {
    const int id1 = lin(i1, j1), id2 = lin(i2, j2);
    const V3<float> p1 = mk(xyz[3 * id1], xyz[3 * id1 + 1], xyz[3 * id1 + 2]);
    const V3<float> p2 = mk(xyz[3 * id2], xyz[3 * id2 + 1], xyz[3 * id2 + 2]);
    const float l0 = len(p2 - p1);  // glm::distance(p1, p2)
    S.a.push_back(id1);
    S.b.push_back(id2);
    S.ks.push_back(k);
    S.l0.push_back(l0);
    at[id1].push_back((int)S.a.size() - 1);
    S.a.push_back(id2);
    S.b.push_back(id1);
    S.ks.push_back(k);
    S.l0.push_back(l0);
    at[id2].push_back((int)S.a.size() - 1);
  };
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      if (i + 1 < rows) add(i, j, i + 1, j, k_stretch);
      if (j + 1 < cols) add(i, j, i, j + 1, k_stretch);
      if (i + 1 < rows && j + 1 < cols) {
        add(i, j, i + 1, j + 1, k_shear);
        add(i + 1, j, i, j + 1, k_shear);
      }
      if (i + 2 < rows) add(i, j, i + 2, j, k_bend);
      if (j + 2 < cols) add(i, j, i, j + 2, k_bend);
    }
  S.ids.push_back(0);
  for (auto& l : at) {
    S.info.insert(S.info.end(), l.begin(), l.end());
    S.ids.push_back((int)S.info.size());
  }
  return S;
}

// ClothInternalDynamics (Physics/CudaDeformableBody.cpp:327-357) on 3N-float arrays.
template <class R>
void ms_step(int n, R* pos, R* prev, R* vel, const float* mass, const uint8_t* fixed,
             const Springs& S, const float* g, float k_damp, float dt, float k_bend,
             int adjust_iters, float max_def) {
  typedef V3<R> RV;
  auto get = [](const R* a, int i) { return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); };
  auto put = [](R* a, int i, RV v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
  };
  const RV grav = mk(R(g[0]), R(g[1]), R(g[2]));
  const R kd = R(k_damp), h = R(dt);
  // ClothEngineStep (Physics/CudaMassSpring.cu:22-88)
  std::vector<R> p0(pos, pos + 3 * (size_t)n);
  for (int id = 0; id < n; ++id) {
    if (fixed[id]) continue;
    const R m = R(mass[id]);
    const RV pi = get(p0.data(), id), qi = get(prev, id);
    RV force = mul(grav, m) - mul(pi - qi, kd);  // m g - kd (pos - prev)
    for (int k = S.ids[id]; k < S.ids[id + 1]; ++k) {
      const int s = S.info[k];
      if (S.a[s] < 0 || S.a[s] >= n || S.b[s] < 0 || S.b[s] >= n) continue;
      const RV dir = get(p0.data(), S.a[s]) - get(p0.data(), S.b[s]);
      const R l0 = R(S.l0[s]);
      if (len(dir) > l0) force = force - mul(dir - mul(normalize(dir), l0), R(S.ks[s]));
    }
    const RV accel = divs(force, m);
    const RV np = (mul(pi, R(2)) - qi) + mul(mul(accel, h), h);  // Verlet
    put(prev, id, pi);
    put(pos, id, np);
    put(vel, id, divs(np - pi, h));
  }
  // AdjustSprings (:95-118), _AdjustSprings (:121-148)
  for (int it = 0; it < adjust_iters; ++it) {
    std::vector<R> snap(pos, pos + 3 * (size_t)n);
    for (int id = 0; id < n; ++id) {
      if (fixed[id]) continue;
      RV p = get(snap.data(), id);
      for (int k = S.ids[id]; k < S.ids[id + 1]; ++k) {
        const int s = S.info[k];
        const int b = S.b[s];
        if (b < 0 || b >= n) continue;
        const RV pb = get(snap.data(), b);
        const R l0 = R(S.l0[s]);
        const R def = (len(p - pb) - l0) / l0;  // ComputeSpringDeformation (:90-93)
        if (def > R(max_def) && S.ks[s] != k_bend) {
          const R dif = (def - R(max_def)) * l0;
          p = p + divs(mul(normalize(pb - p), dif), fixed[b] ? R(1) : R(2));
        }
      }
      put(pos, id, p);
      put(vel, id, divs(p - get(prev, id), h));
    }
  }
}

// SolveCollisions (Physics/CudaDeformableBody.cpp:525-548): CreateImpulses
// (Collision/CudaCollisionResponse.cu:11-231) at x and v, the per-vertex mean in the
// order of a stable sort by vertex id (AccumulateImpulses :233-256), ApplyImpulses
// (:258-297) to v and p. Contacts with a vertex id outside [0, n) give no impulse.
template <class R>
void impulses(int n, const R* x, R* v, R* p, const uint8_t* fixed, const cm_contact* vf, int nvf,
              const cm_contact* ee, int nee, float stiffness, float vertex_mass, float dt,
              float thickness) {
  typedef V3<R> RV;
  auto get = [](const R* a, int i) { return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]); };
  const R m = R(vertex_mass), h = R(dt), th = R(thickness), kst = R(stiffness);
  std::vector<RV> imp(4 * (size_t)(nvf + nee));
  std::vector<int> key(imp.size());
  for (int c = 0; c < nvf + nee; ++c) {
    const bool isvf = c < nvf;
    const cm_contact& ct = isvf ? vf[c] : ee[c - nvf];
    bool valid = true;
    for (int k = 0; k < 4; ++k) valid = valid && ct.v[k] >= 0 && ct.v[k] < n;
    if (!valid) {
      for (int k = 0; k < 4; ++k) key[4 * c + k] = n;
      continue;
    }
    const R w1 = R(ct.w[0]), w2 = R(ct.w[1]), w3 = R(ct.w[2]), w4 = R(ct.w[3]);
    const RV u1 = get(v, ct.v[0]), u2 = get(v, ct.v[1]), u3 = get(v, ct.v[2]), u4 = get(v, ct.v[3]);
    RV cv1, cv2;
    if (isvf) {  // _CreateImpulseVF (:160-170)
      cv1 = mul(u1, w1);
      cv2 = (mul(u2, w2) + mul(u3, w3)) + mul(u4, w4);
    } else {  // _CreateImpulseEE (:172-182)
      cv1 = mul(u1, w1) + mul(u2, w2);
      cv2 = mul(u3, w3) + mul(u4, w4);
    }
    const RV vr = cv1 + cv2;
    const RV nn = mk(R(ct.n[0]), R(ct.n[1]), R(ct.n[2]));
    // _CreateImpulse (:184-231)
    const RV xs = ((mul(get(x, ct.v[0]), w1) + mul(get(x, ct.v[1]), w2)) + mul(get(x, ct.v[2]), w3)) +
                  mul(get(x, ct.v[3]), w4);
    const R d = th - dot(xs, nn);
    const R vn = dot(vr, nn);
    const R dts = (R(0.1f) * d) / h;
    R i = 0;
    if (vn < dts) i = -std::fmin(h * kst * d, m * (dts - vn));
    if (vn < 0) i += m * vn * R(0.5f);
    i = std::fmin(std::fmax(-m * R(0.1f) * th, i), m * R(0.1f) * th);
    const R ip = (2 * i) / (w1 * w1 + w2 * w2 + w3 * w3 + w4 * w4);
    const RV ipv = mul(nn, ip);
    const R w[4] = {w1, w2, w3, w4};
    for (int k = 0; k < 4; ++k) {
      key[4 * c + k] = ct.v[k];
      imp[4 * c + k] = mul(ipv, w[k]);
    }
  }
  // stable sort by vertex id = for each vertex, its impulses in index order
  std::vector<RV> acc(n, mk(R(0), R(0), R(0)));
  std::vector<int> cnt(n, 0);
  for (size_t r = 0; r < imp.size(); ++r) {
    if (key[r] >= n) continue;
    acc[key[r]] = acc[key[r]] + imp[r];
    ++cnt[key[r]];
  }
  for (int i = 0; i < n; ++i) {
    if (cnt[i]) acc[i] = divs(acc[i], R(cnt[i]));
    if (fixed[i]) continue;
    const RV nv = get(v, i) + divs(acc[i], m);
    const RV np = get(x, i) + mul(nv, h);
    v[3 * i] = nv.x;
    v[3 * i + 1] = nv.y;
    v[3 * i + 2] = nv.z;
    p[3 * i] = np.x;
    p[3 * i + 1] = np.y;
    p[3 * i + 2] = np.z;
  }
}

}  // namespace alt
}  // namespace orc
