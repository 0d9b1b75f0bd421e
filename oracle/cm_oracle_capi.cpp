// ORACLE — C entry points over cm_oracle.hpp for the Python test harness
// (tests/ and the bench's cpu_baseline leg load this through ctypes).
// Test infrastructure only: the product never links or calls it.
#include <chrono>
#include <memory>
#include <type_traits>

#include "cm_oracle.hpp"
#include "cm_oracle_alt.hpp"

using namespace orc;

namespace {
struct Handle {
  int precision = 32;
  std::unique_ptr<Sim<float>> f;
  std::unique_ptr<Sim<double>> d;
};

template <class R, class Fn>
auto with(Handle* h, Fn fn) {
  return h->precision == 64 ? fn(*h->d) : fn(*h->f);
}

template <class T>
int64_t copy_out(const std::vector<T>& src, T* out, int64_t cap) {
  int64_t n = (int64_t)src.size();
  if (out) std::memcpy(out, src.data(), sizeof(T) * (size_t)std::min(n, cap));
  return n;
}
}  // namespace

extern "C" {

void* orc_create(const cm_params* prm, const float* xyz, int n, const uint32_t* tri, int t,
                 const uint8_t* pinned, const float* inv_mass, const int32_t* dist_pairs,
                 int n_dist, const float* k_dist, const float* k_bend, int precision) {
  Handle* h = new Handle();
  h->precision = precision == 64 ? 64 : 32;
  if (h->precision == 64) {
    h->d.reset(new Sim<double>());
    h->d->create(*prm, xyz, n, tri, t, pinned, inv_mass, dist_pairs, n_dist, k_dist, k_bend);
  } else {
    h->f.reset(new Sim<float>());
    h->f->create(*prm, xyz, n, tri, t, pinned, inv_mass, dist_pairs, n_dist, k_dist, k_bend);
  }
  return h;
}

void orc_destroy(void* hp) { delete (Handle*)hp; }

int orc_step(void* hp, int substeps) {
  Handle* h = (Handle*)hp;
  for (int s = 0; s < substeps; ++s) with<float>(h, [](auto& sim) { sim.step(); return 0; });
  return 0;
}

// Times `substeps` substeps; returns seconds (CPU baseline leg).
double orc_time_steps(void* hp, int substeps) {
  auto t0 = std::chrono::steady_clock::now();
  orc_step(hp, substeps);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// OpenMP threads of the following calls on this thread (n < 1: the default);
// returns the count in effect.
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n >= 1) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int orc_set_sphere(void* hp, const float* c, float r) {
  Handle* h = (Handle*)hp;
  with<float>(h, [&](auto& sim) {
    sim.sphere_on = true;
    sim.sphere_c = mk(c[0], c[1], c[2]);
    sim.sphere_r = r;
    return 0;
  });
  return 0;
}

// x, p, v: 3N doubles each (p may be NULL: p = x); tri: 4T ints or NULL.
int orc_set_state(void* hp, const double* x, const double* p, const double* v, const int32_t* tri) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) {
    typedef typename std::remove_reference<decltype(sim.x[0].x)>::type R;
    for (int i = 0; i < sim.n; ++i) {
      sim.x[i] = mk<R>(R(x[3 * i]), R(x[3 * i + 1]), R(x[3 * i + 2]));
      const double* pp = p ? p : x;
      sim.p[i] = mk<R>(R(pp[3 * i]), R(pp[3 * i + 1]), R(pp[3 * i + 2]));
      sim.v[i] = mk<R>(R(v[3 * i]), R(v[3 * i + 1]), R(v[3 * i + 2]));
    }
    if (tri)
      for (int i = 0; i < sim.t; ++i) sim.tri[i] = Tri{tri[4 * i], tri[4 * i + 1], tri[4 * i + 2], tri[4 * i + 3]};
    return 0;
  });
}

// parents carry over between substeps (entries no node claims keep their value)
int orc_set_parents(void* hp, const int32_t* parents) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) {
    std::copy(parents, parents + sim.parents.size(), sim.parents.begin());
    return 0;
  });
}

int orc_get_state(void* hp, double* x, double* p, double* v) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) {
    for (int i = 0; i < sim.n; ++i) {
      if (x) { x[3 * i] = sim.x[i].x; x[3 * i + 1] = sim.x[i].y; x[3 * i + 2] = sim.x[i].z; }
      if (p) { p[3 * i] = sim.p[i].x; p[3 * i + 1] = sim.p[i].y; p[3 * i + 2] = sim.p[i].z; }
      if (v) { v[3 * i] = sim.v[i].x; v[3 * i + 1] = sim.v[i].y; v[3 * i + 2] = sim.v[i].z; }
    }
    return 0;
  });
}

// positions at the start of Jacobi iteration k (k = iterations: after the last) of
// recorded substep s; -1 when not recorded
int orc_tape_positions(void* hp, int s, int k, double* out) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) {
    if (s < 0 || s >= (int)sim.tape.size() || k < 0 || k >= (int)sim.tape[s].pk.size()) return -1;
    const auto& q = sim.tape[s].pk[k];
    for (int i = 0; i < sim.n; ++i) {
      out[3 * i] = q[i].x;
      out[3 * i + 1] = q[i].y;
      out[3 * i + 2] = q[i].z;
    }
    return 0;
  });
}

// vertex normals: with pos (3N) the normals of those positions, else the carried ones
int orc_get_normals(void* hp, const double* pos, double* out) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) {
    typedef typename std::decay<decltype(sim)>::type S;
    const std::vector<typename S::RV> carried = sim.vnrm;
    if (pos) {
      std::vector<typename S::RV> q(sim.n);
      for (int i = 0; i < sim.n; ++i)
        q[i] = mk<decltype(q[0].x)>(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
      sim.vertex_normals(q, false);
    }
    for (int i = 0; i < sim.n; ++i) {
      out[3 * i] = sim.vnrm[i].x;
      out[3 * i + 1] = sim.vnrm[i].y;
      out[3 * i + 2] = sim.vnrm[i].z;
    }
    sim.vnrm = carried;
    return 0;
  });
}

// counts: n, t, S, B, P, C_vf, C_ee, slots
void orc_counts(void* hp, int64_t* out) {
  Handle* h = (Handle*)hp;
  with<float>(h, [&](auto& sim) {
    out[0] = sim.n;
    out[1] = sim.t;
    out[2] = (int64_t)sim.tp.sid1.size();
    out[3] = (int64_t)sim.tp.bkp.size();
    out[4] = (int64_t)sim.pairs.size() / 2;
    out[5] = (int64_t)sim.vf.size();
    out[6] = (int64_t)sim.ee.size();
    out[7] = (int64_t)sim.tp.prefix.back();
    return 0;
  });
}

uint32_t orc_warnings(void* hp) {
  return with<float>((Handle*)hp, [](auto& sim) { return sim.warn; });
}

// which: 0 tri(4T) 1 lefts 2 rights 3 parents 4 rmll 5 rmlr 6 pairs(2P)
//        10 sid(2S) 11 bid(4B) 12 prefix(N+1) 13 soff(2S) 14 boff(4B) 15 init tri(4T)
int64_t orc_get_i32(void* hp, int which, int32_t* out, int64_t cap) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) -> int64_t {
    std::vector<int32_t> tmp;
    switch (which) {
      case 0:
      case 15: {
        const auto& tr = which == 0 ? sim.tri : sim.tp.tri;
        for (const Tri& q : tr) {
          tmp.push_back(q.v1);
          tmp.push_back(q.v2);
          tmp.push_back(q.v3);
          tmp.push_back(q.feat);
        }
        return copy_out(tmp, out, cap);
      }
      case 1: return copy_out(sim.lefts, out, cap);
      case 2: return copy_out(sim.rights, out, cap);
      case 3: return copy_out(sim.parents, out, cap);
      case 4: return copy_out(sim.rmll, out, cap);
      case 5: return copy_out(sim.rmlr, out, cap);
      case 6: return copy_out(sim.pairs, out, cap);
      case 10:
        for (size_t i = 0; i < sim.tp.sid1.size(); ++i) {
          tmp.push_back(sim.tp.sid1[i]);
          tmp.push_back(sim.tp.sid2[i]);
        }
        return copy_out(tmp, out, cap);
      case 11:
        for (size_t i = 0; i < sim.tp.bkp.size(); ++i)
          for (int r = 0; r < 4; ++r) tmp.push_back(sim.tp.bid[r][i]);
        return copy_out(tmp, out, cap);
      case 12: return copy_out(sim.tp.prefix, out, cap);
      case 13:
        for (size_t i = 0; i < sim.tp.sid1.size(); ++i) {
          tmp.push_back(sim.tp.soff[0][i]);
          tmp.push_back(sim.tp.soff[1][i]);
        }
        return copy_out(tmp, out, cap);
      case 14:
        for (size_t i = 0; i < sim.tp.bkp.size(); ++i)
          for (int r = 0; r < 4; ++r) tmp.push_back(sim.tp.boff[r][i]);
        return copy_out(tmp, out, cap);
      default: return -1;
    }
  });
}

// which: 0 sl0 1 skp 2 bkp 3 bphi0 4 lra_l0
int64_t orc_get_f32(void* hp, int which, float* out, int64_t cap) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) -> int64_t {
    switch (which) {
      case 0: return copy_out(sim.tp.sl0, out, cap);
      case 1: return copy_out(sim.tp.skp, out, cap);
      case 2: return copy_out(sim.tp.bkp, out, cap);
      case 3: return copy_out(sim.tp.bphi0, out, cap);
      case 4: return copy_out(sim.tp.lra_l0, out, cap);
      default: return -1;
    }
  });
}

int64_t orc_get_keys(void* hp, uint64_t* out, int64_t cap) {
  return with<float>((Handle*)hp, [&](auto& sim) { return copy_out(sim.keys, out, cap); });
}
int64_t orc_get_perm(void* hp, uint32_t* out, int64_t cap) {
  return with<float>((Handle*)hp, [&](auto& sim) { return copy_out(sim.perm, out, cap); });
}
// 3 floats per node, 2T-1 nodes
int64_t orc_get_aabb(void* hp, float* mins, float* maxs, int64_t cap) {
  return with<float>((Handle*)hp, [&](auto& sim) -> int64_t {
    int64_t n = (int64_t)sim.amin.size();
    for (int64_t i = 0; i < std::min(n, cap); ++i) {
      mins[3 * i] = sim.amin[i].x; mins[3 * i + 1] = sim.amin[i].y; mins[3 * i + 2] = sim.amin[i].z;
      maxs[3 * i] = sim.amax[i].x; maxs[3 * i + 1] = sim.amax[i].y; maxs[3 * i + 2] = sim.amax[i].z;
    }
    return n;
  });
}
// which: 0 VF contacts, 1 EE contacts, 2 VF candidates, 3 EE candidates (the CCD
// inputs: ownership gate and feature swept-AABB filter passed)
int64_t orc_get_contacts(void* hp, int which, cm_contact* out, int64_t cap) {
  return with<float>((Handle*)hp, [&](auto& sim) {
    const auto& src = which == 3 ? sim.cand_ee : which == 2 ? sim.cand_vf : which ? sim.ee : sim.vf;
    return copy_out(src, out, cap);
  });
}

// Contact activation decisions of the solver: mode 1 records them in the next
// substeps, 2 replays the ones set with orc_set_act, 0 evaluates them (default).
int orc_act_mode(void* hp, int mode) {
  return with<float>((Handle*)hp, [&](auto& sim) {
    sim.act_mode = mode;
    sim.act_mismatch = false;
    return 0;
  });
}
// bytes (iterations * contacts, VF first) of the last solver stage
int64_t orc_get_act(void* hp, uint8_t* out, int64_t cap) {
  return with<float>((Handle*)hp, [&](auto& sim) { return copy_out(sim.act, out, cap); });
}
int orc_set_act(void* hp, const uint8_t* in, int64_t n) {
  return with<float>((Handle*)hp, [&](auto& sim) {
    sim.act.assign(in, in + n);
    return 0;
  });
}
int orc_act_mismatch(void* hp) {
  return with<float>((Handle*)hp, [&](auto& sim) { return sim.act_mismatch ? 1 : 0; });
}

// ---- stage functions on plain arrays ------------------------------------
void orc_morton63(int n, const float* c3, uint64_t* out) {
  for (int i = 0; i < n; ++i) out[i] = morton63(mk(c3[3 * i], c3[3 * i + 1], c3[3 * i + 2]));
}
void orc_morton30(int n, const float* c3, const float* lo, const float* hi, uint64_t* out) {
  F3 l = mk(lo[0], lo[1], lo[2]), hh = mk(hi[0], hi[1], hi[2]);
  for (int i = 0; i < n; ++i) out[i] = morton30(mk(c3[3 * i], c3[3 * i + 1], c3[3 * i + 2]), l, hh);
}
// keys/vals sorted stably by key, in place
void orc_stable_sort(int n, uint64_t* keys, uint32_t* vals) {
  std::vector<uint32_t> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  std::vector<uint64_t> k(n);
  std::vector<uint32_t> v(n);
  for (int i = 0; i < n; ++i) {
    k[i] = keys[idx[i]];
    v[i] = vals[idx[i]];
  }
  std::memcpy(keys, k.data(), 8 * (size_t)n);
  std::memcpy(vals, v.data(), 4 * (size_t)n);
}
void orc_build(int t, const uint64_t* keys, int midpoint, int32_t* lefts, int32_t* rights,
               int32_t* parents, uint32_t* warn) {
  lbvh_build(t, keys, midpoint != 0, lefts, rights, parents, warn);
}
static void to_f3(const float* a, int64_t n, std::vector<F3>& out) {
  out.resize(n);
  for (int64_t i = 0; i < n; ++i) out[i] = mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
}
// amin3/amax3: (2T-1)*3 floats with leaves filled; internal nodes written.
void orc_refit(int t, const int32_t* lefts, const int32_t* rights, float* amin3, float* amax3,
               int32_t* rmll, int32_t* rmlr) {
  std::vector<F3> mn, mx;
  to_f3(amin3, 2 * (int64_t)t - 1, mn);
  to_f3(amax3, 2 * (int64_t)t - 1, mx);
  lbvh_refit(t, lefts, rights, mn.data(), mx.data(), rmll, rmlr);
  for (int64_t i = 0; i < 2 * (int64_t)t - 1; ++i) {
    amin3[3 * i] = mn[i].x; amin3[3 * i + 1] = mn[i].y; amin3[3 * i + 2] = mn[i].z;
    amax3[3 * i] = mx[i].x; amax3[3 * i + 1] = mx[i].y; amax3[3 * i + 2] = mx[i].z;
  }
}
// Returns the number of pairs written to out (2 ints each, capacity 2*T*cap).
int64_t orc_pairs(int t, const int32_t* lefts, const int32_t* rights, const float* amin3,
                  const float* amax3, const int32_t* rmll, const int32_t* rmlr, int cap,
                  int32_t* out, uint32_t* warn) {
  std::vector<F3> mn, mx;
  to_f3(amin3, 2 * (int64_t)t - 1, mn);
  to_f3(amax3, 2 * (int64_t)t - 1, mx);
  int64_t total = 0;
  std::vector<int32_t> buf(2 * (size_t)cap);
  for (int i = 0; i < t; ++i) {
    bool capped = false;
    int c = lbvh_leaf_pairs(t, i, lefts, rights, mn.data(), mx.data(), rmll, rmlr, cap, buf.data(),
                            warn, &capped);
    if (capped && warn) *warn |= CM_WARN_PAIR_CAP;
    std::memcpy(out + 2 * total, buf.data(), sizeof(int32_t) * 2 * (size_t)c);
    total += c;
  }
  return total;
}
int orc_coplanarity_times(const float* x12, const float* v12, float* times) {
  F3 x[4], v[4];
  for (int k = 0; k < 4; ++k) {
    x[k] = mk(x12[3 * k], x12[3 * k + 1], x12[3 * k + 2]);
    v[k] = mk(v12[3 * k], v12[3 * k + 1], v12[3 * k + 2]);
  }
  return coplanarity_times(x[0], x[1], x[2], x[3], v[0], v[1], v[2], v[3], times);
}
int orc_test_vf(const float* x12, float th, cm_contact* c) {
  F3 x[4];
  for (int k = 0; k < 4; ++k) x[k] = mk(x12[3 * k], x12[3 * k + 1], x12[3 * k + 2]);
  return test_vf(x[0], x[1], x[2], x[3], *c, th);
}
int orc_test_ee(const float* x12, float th, cm_contact* c) {
  F3 x[4];
  for (int k = 0; k < 4; ++k) x[k] = mk(x12[3 * k], x12[3 * k + 1], x12[3 * k + 2]);
  return test_ee(x[0], x[1], x[2], x[3], *c, th, 0.000000001f);
}

// ---- solver building blocks on plain arrays ----------------------------
// One contact projection (Sim::contact_vf_core / contact_ee_core) in FP32
// (precision 32) or FP64: p12 = the contact's four positions, w4 = their inverse
// masses, out12 = the four corrections; force < 0 evaluates the activation.
// Returns 1 when the constraint was active.
int orc_contact_project(int ee, int precision, const double* p12, const double* w4,
                        const cm_contact* k, double th, double* out12, int force) {
  auto run = [&](auto tag) {
    typedef decltype(tag) R;
    typedef V3<R> RV;
    RV p[4], c[4];
    R w[4];
    for (int r = 0; r < 4; ++r) {
      p[r] = mk<R>(R(p12[3 * r]), R(p12[3 * r + 1]), R(p12[3 * r + 2]));
      w[r] = R(w4[r]);
    }
    bool on = ee ? Sim<R>::contact_ee_core(*k, p, w, R(th), c, force)
                 : Sim<R>::contact_vf_core(*k, p, w, R(th), c, force);
    for (int r = 0; r < 4; ++r) {
      out12[3 * r] = c[r].x;
      out12[3 * r + 1] = c[r].y;
      out12[3 * r + 2] = c[r].z;
    }
    return on ? 1 : 0;
  };
  return precision == 64 ? run(double()) : run(float());
}

// Per-vertex sum of dynamic correction records (Sim::accumulate_records, FP64):
// record k belongs to vertex ids[k] with value vals3[3k..3k+2]; acc3 (3N) and cnt
// (N) are added to.
void orc_accumulate(int n, int64_t nrec, const uint32_t* ids, const double* vals3, double* acc3,
                    int32_t* cnt) {
  typedef V3<double> RV;
  std::vector<uint32_t> rid(ids, ids + nrec), order;
  std::vector<RV> rec(nrec), acc(n);
  std::vector<int> c(n);
  for (int64_t k = 0; k < nrec; ++k) rec[k] = mk(vals3[3 * k], vals3[3 * k + 1], vals3[3 * k + 2]);
  for (int i = 0; i < n; ++i) {
    acc[i] = mk(acc3[3 * i], acc3[3 * i + 1], acc3[3 * i + 2]);
    c[i] = cnt[i];
  }
  Sim<double>::record_order(rid, order);
  Sim<double>::accumulate_records(rid, order, rec, acc, c);
  for (int i = 0; i < n; ++i) {
    acc3[3 * i] = acc[i].x;
    acc3[3 * i + 1] = acc[i].y;
    acc3[3 * i + 2] = acc[i].z;
    cnt[i] = c[i];
  }
}

// Long-range attachment correction of one vertex at p3 (Sim::lra_vertex, FP64):
// pinned vertices fixed[j] at ps3 (positions of all vertices), rest lengths l0[j],
// flags[j]; a3 / *count receive the summed correction and its count.
void orc_lra_vertex(const double* p3, int F, const int32_t* fixed, const float* l0,
                    const uint8_t* flags, int n, const double* ps3, double* a3, int32_t* count) {
  typedef V3<double> RV;
  std::vector<RV> ps(n);
  for (int i = 0; i < n; ++i) ps[i] = mk(ps3[3 * i], ps3[3 * i + 1], ps3[3 * i + 2]);
  RV a;
  int c;
  Sim<double>::lra_vertex(mk(p3[0], p3[1], p3[2]), F, fixed, l0, flags, ps, a, c);
  a3[0] = a.x;
  a3[1] = a.y;
  a3[2] = a.z;
  *count = c;
}

// Global and rigid-mode velocity damping (Sim::damp, Physics/CudaPBD.cu:411-458) of
// vs3 in place at positions ps3 with per-vertex masses, FP64 (precision 64) or FP32.
void orc_damp(int precision, int n, const double* ps3, double* vs3, const double* mass, float kg,
              float kd) {
  auto run = [&](auto tag) {
    typedef decltype(tag) R;
    typedef V3<R> RV;
    Sim<R> sim;
    sim.n = n;
    std::memset(&sim.prm, 0, sizeof(sim.prm));
    sim.prm.global_vel_damp = kg;
    sim.prm.k_damp = kd;
    sim.mass.resize(n);
    std::vector<RV> ps(n), vs(n);
    for (int i = 0; i < n; ++i) {
      sim.mass[i] = R(mass[i]);
      ps[i] = mk<R>(R(ps3[3 * i]), R(ps3[3 * i + 1]), R(ps3[3 * i + 2]));
      vs[i] = mk<R>(R(vs3[3 * i]), R(vs3[3 * i + 1]), R(vs3[3 * i + 2]));
    }
    sim.template damp<R>(ps, vs);
    for (int i = 0; i < n; ++i) {
      vs3[3 * i] = vs[i].x;
      vs3[3 * i + 1] = vs[i].y;
      vs3[3 * i + 2] = vs[i].z;
    }
    return 0;
  };
  if (precision == 64)
    run(double());
  else
    run(float());
}

// ---- differentiable rollout (no reference counterpart) ----
// Recording on: later substeps keep what the adjoint needs (the tape is cleared).
int orc_record(void* hp, int on) {
  Handle* h = (Handle*)hp;
  with<float>(h, [&](auto& sim) {
    sim.record = on != 0;
    sim.tape.clear();
    return 0;
  });
  return 0;
}

// Replace the tape of dst (an FP64 handle) with the tape src (an FP32 handle)
// recorded: the same iterates, contact sets and activation decisions, widened to
// FP64, so dst's adjoint differentiates the FP32 forward at its own points.
int orc_copy_tape(void* dst_p, void* src_p) {
  Handle* dst = (Handle*)dst_p;
  Handle* src = (Handle*)src_p;
  if (dst->precision != 64 || src->precision != 32) return -1;
  auto& d = *dst->d;
  const auto& s = *src->f;
  if (d.n != s.n) return -1;
  auto widen = [](const std::vector<V3<float>>& a) {
    std::vector<V3<double>> out(a.size());
    for (size_t i = 0; i < a.size(); ++i) out[i] = cvt<double>(a[i]);
    return out;
  };
  d.tape.clear();
  for (const auto& t : s.tape) {
    d.tape.emplace_back();
    auto& u = d.tape.back();
    u.x = widen(t.x);
    u.v = widen(t.v);
    for (const auto& pk : t.pk) u.pk.push_back(widen(pk));
    u.vf = t.vf;
    u.ee = t.ee;
    u.act = t.act;
  }
  return 0;
}

// Adjoint of the recorded substeps. lx, lv: 3N doubles, dL/dx and dL/dv after the
// last substep in, dL/dx and dL/dv before the first out; gk_dist (S) and gk_bend (B):
// dL/dkp per distance / bend constraint, times dkp/dk = (1 - k)^(1/I - 1) / I when
// k_dist / k_bend (the stiffnesses) are given, else dL/dkp.
int orc_adjoint(void* hp, double* lx, double* lv, double* gk_dist, double* gk_bend,
                const float* k_dist, const float* k_bend) {
  Handle* h = (Handle*)hp;
  return with<float>(h, [&](auto& sim) {
    typedef typename std::remove_reference<decltype(sim)>::type S;
    typedef typename S::RV RV;
    typedef decltype(RV().x) R;
    std::vector<RV> x(sim.n), v(sim.n);
    for (int i = 0; i < sim.n; ++i) {
      x[i] = mk<R>(R(lx[3 * i]), R(lx[3 * i + 1]), R(lx[3 * i + 2]));
      v[i] = mk<R>(R(lv[3 * i]), R(lv[3 * i + 1]), R(lv[3 * i + 2]));
    }
    std::vector<R> gs, gb;
    sim.adjoint(x, v, gs, gb);
    for (int i = 0; i < sim.n; ++i) {
      lx[3 * i] = x[i].x; lx[3 * i + 1] = x[i].y; lx[3 * i + 2] = x[i].z;
      lv[3 * i] = v[i].x; lv[3 * i + 1] = v[i].y; lv[3 * i + 2] = v[i].z;
    }
    const double I = sim.prm.solver_iterations;
    for (size_t j = 0; j < gs.size(); ++j) {
      double f = 1.0;
      if (k_dist) f = std::pow(1.0 - (double)k_dist[j], 1.0 / I - 1.0) / I;
      if (gk_dist) gk_dist[j] = (double)gs[j] * f;
    }
    for (size_t j = 0; j < gb.size(); ++j) {
      double f = 1.0;
      if (k_bend) f = std::pow(1.0 - (double)k_bend[j], 1.0 / I - 1.0) / I;
      if (gk_bend) gk_bend[j] = (double)gb[j] * f;
    }
    return 0;
  });
}

// ---- dead paths (cm_oracle_alt.hpp), FP32 on 3N-float arrays ------------------
int orc_ms_grid_springs(int rows, int cols, const float* xyz, float k_stretch, float k_shear,
                        float k_bend, int32_t* a, int32_t* b, float* ks, float* l0,
                        int32_t* info, int32_t* ids, int32_t* n_springs) {
  const alt::Springs S = alt::grid_springs(rows, cols, xyz, k_stretch, k_shear, k_bend);
  *n_springs = (int32_t)S.a.size();
  for (size_t s = 0; s < S.a.size(); ++s) {
    if (a) a[s] = S.a[s];
    if (b) b[s] = S.b[s];
    if (ks) ks[s] = S.ks[s];
    if (l0) l0[s] = S.l0[s];
    if (info) info[s] = S.info[s];
  }
  for (size_t v = 0; v < S.ids.size(); ++v)
    if (ids) ids[v] = S.ids[v];
  return 0;
}

int orc_ms_step(int n, float* pos, float* prev, float* vel, const float* mass,
                const uint8_t* fixed, const int32_t* a, const int32_t* b, const float* ks,
                const float* l0, const int32_t* info, const int32_t* ids, int n_springs,
                const float* g, float k_damp, float dt, float k_bend, int adjust_iters,
                float max_def) {
  alt::Springs S;
  S.a.assign(a, a + n_springs);
  S.b.assign(b, b + n_springs);
  S.ks.assign(ks, ks + n_springs);
  S.l0.assign(l0, l0 + n_springs);
  S.ids.assign(ids, ids + n + 1);
  S.info.assign(info, info + ids[n]);
  alt::ms_step<float>(n, pos, prev, vel, mass, fixed, S, g, k_damp, dt, k_bend, adjust_iters,
                      max_def);
  return 0;
}

int orc_impulses(int n, const float* x, float* v, float* p, const uint8_t* fixed,
                 const cm_contact* vf, int nvf, const cm_contact* ee, int nee, float stiffness,
                 float vertex_mass, float dt, float thickness) {
  alt::impulses<float>(n, x, v, p, fixed, vf, nvf, ee, nee, stiffness, vertex_mass, dt,
                       thickness);
  return 0;
}

}  // extern "C"
