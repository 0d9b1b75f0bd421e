// cm_body: the simulation object behind the C ABI.
//
// Replaces Physics::CudaDeformableBody (LicentaV2/LicentaV2/Physics/CudaDeformableBody.cpp):
//   cm_create  <- constructor :30-285 with CudaPBD::Init (Physics/CudaPBD.cu:39-370) and
//                 FeatureList::Init/AssignFeatures (Collision/FeatureList.cu:12-132)
//   cm_step    <- FixedUpdate :550-627 (one call per substep, recorded once into a
//                 hipGraph and replayed; no host synchronisation inside a substep)
//   cm_read_*  <- Update / FinalVertUpdate :629-638, 473-495
//
// The constraint topology, slot layout and feature ownership are built on the
// host once; everything per substep runs on the body's stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "cm_internal.hpp"

namespace cm {

static_assert(sizeof(cm_params) == 144, "cm_params layout is part of the ABI");
static_assert(sizeof(cm_contact) == 48, "cm_contact layout is part of the ABI");

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

}  // namespace cm

using namespace cm;

namespace {

// Host-built topology; kept by the body so that stiffness changes can rebuild the
// tile entries.
struct HostTopo {
  std::vector<int2> sid;
  std::vector<float2> s_l0kp;
  std::vector<int4> bid;
  std::vector<float2> b_kpphi0;
  std::vector<int> prefix;
  std::vector<uint32_t> vslot;
  std::vector<int4> tri;
  std::vector<int> lra_fixed;
  std::vector<float> lra_l0;
  std::vector<int> tstart, eprefix;  // Jacobi tiles (empty: slot-parallel solve)
  std::vector<TileEntry> entries;
  std::vector<int> ecid;  // per entry: constraint index, bit 31 set for a bend
  std::vector<float> k_dist, k_bend;  // stiffness per constraint (before the kp mapping)
};

// Differentiable rollout: checkpoints of the state every `every` substeps, and the
// tape of one segment (each substep's Jacobi iterates and contact records).
struct AdjState {
  std::vector<void*> allocs;
  int nsub = 0, every = 0, nck = 0, iters = 0;
  bool ready = false;  // a forward run is on record
  float4 *ck_x = nullptr, *ck_v = nullptr, *ck_p = nullptr;  // (nck + 1) * n; last = final
  int4* ck_tri = nullptr;                                      // (nck + 1) * t
  float4* pk = nullptr;       // every * (iters + 1) * n
  int* dpre = nullptr;        // every * (n + 1)
  RecPayload* arena = nullptr;
  int64_t arena_cap = 0;
  int* ctl = nullptr;  // 2 + every
  AdjBufs ab;
  int* ecid = nullptr;
};

}  // namespace

struct cm_body {
  cm_params prm;
  int n = 0, t = 0, S = 0, B = 0, F = 0, nslots = 0, cap = 0;
  int key_bits = 30;  // bits of the Morton sort keys (a batch adds the cloth id)
  int dev = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // state
  float4 *x = nullptr, *p = nullptr, *p2 = nullptr, *v = nullptr;
  uint8_t* pinned = nullptr;
  float* mass = nullptr;
  // topology
  int2* sid = nullptr;
  float2* s_l0kp = nullptr;
  int4* bid = nullptr;
  float2* b_kpphi0 = nullptr;
  int* vprefix = nullptr;
  uint32_t* vslot = nullptr;
  int ntiles = 0;
  int* tstart = nullptr;
  int* eprefix = nullptr;
  TileEntry* entries = nullptr;
  int* lra_fixed = nullptr;
  float* lra_l0 = nullptr;
  uint8_t* lra_flag = nullptr;
  float4* ground = nullptr;
  float *damp_partials = nullptr, *damp_state = nullptr;
  // vertex normals: incident triangles per vertex (CSR, ascending triangle index,
  // vertex ids in the input order) and the normals
  int* vt_prefix = nullptr;
  int4* vt_tri = nullptr;
  float4* vnrm = nullptr;
  // collision
  int4 *tri = nullptr, *tri_alt = nullptr;
  uint64_t* keys = nullptr;
  uint32_t* perm = nullptr;
  void* sort_temp = nullptr;
  float4 *cen = nullptr, *box = nullptr, *tmp_min = nullptr, *tmp_max = nullptr;
  LbvhBufs lb;
  int2 *chunk = nullptr, *pairs = nullptr;
  int *leaf_cnt = nullptr, *leaf_off = nullptr;
  int* counters = nullptr;  // [0] npairs [1] ncand_vf [2] ncand_ee [3] nvf [4] nee ... [8] traversal hand-over count
  void* lb_temp = nullptr;  // LBVH build/refit scratch (carve_build)
  void* scan_temp = nullptr;
  NarrowBufs nb;
  uint32_t* warn = nullptr;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
  // launch shapes (cm_autotune): Jacobi tile vertices; traversal tiles per XCD chunk
  // and the traversal's visit budget live in lb
  int tile_v = kTileV;
  // host copies
  std::vector<float> inv_mass_h;
  std::vector<uint8_t> pinned_h;
  // counts of the host-built topology (for tests)
  std::vector<int> h_sid, h_bid, h_prefix;
  HostTopo topo;
  AdjState adj;
};

namespace {

template <class T>
int dalloc(cm_body* b, T** out, size_t count) {
  *out = nullptr;
  if (count == 0) count = 1;
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, sizeof(T) * count);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + std::to_string(sizeof(T) * count) +
              " bytes): " + hipGetErrorString(e));
    return CM_ERR_ALLOC;
  }
  b->allocs.push_back(ptr);
  *out = (T*)ptr;
  // CLOTHMI_LOG_ALLOC=1 lists the body's device allocations on stderr
  if (std::getenv("CLOTHMI_LOG_ALLOC"))
    std::fprintf(stderr, "clothmi alloc %zu bytes (%zu allocations)\n", sizeof(T) * count,
                 b->allocs.size());
  return CM_OK;
}

template <class T>
int upload(cm_body* b, T** out, const std::vector<T>& h) {
  int rc = dalloc(b, out, h.size());
  if (rc) return rc;
  if (!h.empty()) CM_HIP(hipMemcpy(*out, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return CM_OK;
}

float host_dist(const float* xyz, int a, int b) {
  float dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1],
        dz = xyz[3 * a + 2] - xyz[3 * b + 2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}
float kp_of(float k, int iters) { return 1.f - std::pow(1.f - k, 1.f / (float)iters); }

// the dihedral angle the reference computes at Physics/CudaPBD.cu:348-366
float dihedral_rest(const float* xyz, int p1, int p2, int p3, int p4) {
  auto d = [&](int a, int b) {
    return mk3(xyz[3 * a] - xyz[3 * b], xyz[3 * a + 1] - xyz[3 * b + 1],
               xyz[3 * a + 2] - xyz[3 * b + 2]);
  };
  f3 a = normalize(cross(d(p2, p1), d(p3, p1))), c = normalize(cross(d(p2, p1), d(p4, p1)));
  // the dot of two unit vectors can round past +-1
  return std::acos(fclamp(dot(a, c), -1.f, 1.f));
}

// Physics/CudaPBD.cu:289-346
void quad_of(const int t1[3], const int t2[3], int q[4]) {
  int p1 = -1, p2 = -1, p3 = -1, p4 = -1;
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
  for (int i = 0; i < 3 && p3 < 0; ++i)
    if (t1[i] != p1 && t1[i] != p2) p3 = t1[i];
  for (int i = 0; i < 3 && p4 < 0; ++i)
    if (t2[i] != p1 && t2[i] != p2) p4 = t2[i];
  q[0] = p1;
  q[1] = p2;
  q[2] = p3;
  q[3] = p4;
}

inline uint64_t edge_key(int a, int b) {
  uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
  return ((uint64_t)lo << 32) | hi;
}

// Cut the vertices into tiles of at most kTileV consecutive vertices and kTileSlots
// slots, and list for each tile every constraint that touches it once, with the
// tile-local slot of each of its roles. Leaves the tiles empty when one vertex has
// more slots than a tile holds.
void build_tiles(int n, HostTopo& h, int tile_v = kTileV) {
  const std::vector<int>& pre = h.prefix;
  std::vector<int> ts{0};
  for (int v = 0; v < n;) {
    const int start = v;
    while (v < n && v - start < tile_v && pre[v + 1] - pre[start] <= kTileSlots) ++v;
    if (v == start) return;  // a vertex with more than kTileSlots slots
    ts.push_back(v);
  }
  const int nt = (int)ts.size() - 1;
  std::vector<int> seen_s(h.sid.size(), -1), seen_b(h.bid.size(), -1);  // entry of this tile
  std::vector<int> ep{0};
  std::vector<TileEntry> ent;
  std::vector<int> cid;
  for (int tile = 0; tile < nt; ++tile) {
    const int s0 = pre[ts[tile]], s1 = pre[ts[tile + 1]];
    const int base = (int)ent.size();
    for (int k = s0; k < s1; ++k) {
      const uint32_t w = h.vslot[k];
      const int idx = (int)(w & 0x1fffffffu), role = (int)((w >> 29) & 3u);
      const bool bend = (w & 0x80000000u) != 0;
      int& seen = bend ? seen_b[idx] : seen_s[idx];
      if (seen < base) {  // first slot of this constraint in this tile
        seen = (int)ent.size();
        TileEntry e;
        if (bend) {
          e.ids = h.bid[idx];
          e.a = h.b_kpphi0[idx].x;
          e.b = h.b_kpphi0[idx].y;
        } else {
          e.ids = make_int4(h.sid[idx].x, h.sid[idx].y, -1, -1);
          e.a = h.s_l0kp[idx].x;
          e.b = h.s_l0kp[idx].y;
        }
        e.s01 = e.s23 = 0xffffffffu;
        ent.push_back(e);
        cid.push_back(bend ? (int)(0x80000000u | (uint32_t)idx) : idx);
      }
      TileEntry& e = ent[seen];
      const uint32_t off = (uint32_t)(k - s0);
      uint32_t& word = role < 2 ? e.s01 : e.s23;
      const int sh = (role & 1) * 16;
      word = (word & ~(0xffffu << sh)) | (off << sh);
    }
    // distance constraints first, then bends: a wave then mostly runs one path
    std::vector<int> order((size_t)ent.size() - base);
    for (size_t k = 0; k < order.size(); ++k) order[k] = base + (int)k;
    std::stable_partition(order.begin(), order.end(),
                          [&](int k) { return ent[k].ids.z < 0; });
    std::vector<TileEntry> te(order.size());
    std::vector<int> tc(order.size());
    for (size_t k = 0; k < order.size(); ++k) {
      te[k] = ent[order[k]];
      tc[k] = cid[order[k]];
    }
    std::copy(te.begin(), te.end(), ent.begin() + base);
    std::copy(tc.begin(), tc.end(), cid.begin() + base);
    ep.push_back((int)ent.size());
  }
  h.tstart = std::move(ts);
  h.eprefix = std::move(ep);
  h.entries = std::move(ent);
  h.ecid = std::move(cid);
}

int build_topology(const cm_params& P, const float* xyz, int n, const uint32_t* tri, int t,
                   const uint8_t* pinned, const int32_t* dist_pairs, int n_dist,
                   const float* k_dist, const float* k_bend, HostTopo& h) {
  const int I = P.solver_iterations;
  std::unordered_map<uint64_t, int> open;
  open.reserve((size_t)t * 2);
  std::vector<int2> s_edges;
  for (int i = 0; i < t; ++i) {
    const int tv[3] = {(int)tri[3 * i], (int)tri[3 * i + 1], (int)tri[3 * i + 2]};
    for (int j = 0; j < 3; ++j) {
      const int id1 = tv[j % 3], id2 = tv[(j + 1) % 3];
      const uint64_t k = edge_key(id1, id2);
      auto it = open.find(k);
      if (it != open.end()) {
        const int f = it->second;
        const int tf[3] = {(int)tri[3 * f], (int)tri[3 * f + 1], (int)tri[3 * f + 2]};
        int q[4];
        quad_of(tf, tv, q);
        const int bi = (int)h.bid.size();
        h.bid.push_back(make_int4(q[0], q[1], q[2], q[3]));
        const float kb = k_bend ? k_bend[bi] : P.k_bend;
        h.k_bend.push_back(kb);
        const float phi0 = (P.ref_compat & CM_RC_BEND_ABS) ? 3.14159265359f
                                                           : dihedral_rest(xyz, q[0], q[1], q[2], q[3]);
        h.b_kpphi0.push_back(make_float2(kp_of(kb, I), phi0));
        open.erase(it);
      } else {
        open[k] = i;
        s_edges.push_back(make_int2(id1, id2));
      }
    }
  }
  if (dist_pairs) {
    s_edges.resize(n_dist);
    for (int i = 0; i < n_dist; ++i) {
      if (dist_pairs[2 * i] < 0 || dist_pairs[2 * i] >= n || dist_pairs[2 * i + 1] < 0 ||
          dist_pairs[2 * i + 1] >= n) {
        set_error("cm_create: distance constraint vertex out of range");
        return CM_ERR_ARG;
      }
      s_edges[i] = make_int2(dist_pairs[2 * i], dist_pairs[2 * i + 1]);
    }
  }
  const int S = (int)s_edges.size(), B = (int)h.bid.size();
  if (S >= (1 << 29) || B >= (1 << 29)) {
    set_error("cm_create: too many constraints");
    return CM_ERR_ARG;
  }
  h.sid = s_edges;
  h.s_l0kp.resize(S);
  h.k_dist.resize(S);
  for (int i = 0; i < S; ++i) {
    h.k_dist[i] = k_dist ? k_dist[i] : P.k_stretch;
    h.s_l0kp[i] = make_float2(host_dist(xyz, s_edges[i].x, s_edges[i].y), kp_of(h.k_dist[i], I));
  }
  // slots: every stretch end in constraint order, then every bend vertex
  // (Physics/CudaPBD.cu:146-233)
  std::vector<int> cnt(n, 0);
  for (int i = 0; i < S; ++i) {
    cnt[s_edges[i].x]++;
    cnt[s_edges[i].y]++;
  }
  for (int i = 0; i < B; ++i) {
    cnt[h.bid[i].x]++;
    cnt[h.bid[i].y]++;
    cnt[h.bid[i].z]++;
    cnt[h.bid[i].w]++;
  }
  h.prefix.assign(n + 1, 0);
  for (int i = 1; i <= n; ++i) h.prefix[i] = h.prefix[i - 1] + cnt[i - 1];
  h.vslot.assign(h.prefix[n], 0);
  std::fill(cnt.begin(), cnt.end(), 0);
  for (int i = 0; i < S; ++i) {
    const int a = s_edges[i].x, b = s_edges[i].y;
    h.vslot[h.prefix[a] + cnt[a]++] = (uint32_t)i;
    h.vslot[h.prefix[b] + cnt[b]++] = (uint32_t)i | (1u << 29);
  }
  for (int i = 0; i < B; ++i) {
    const int q[4] = {h.bid[i].x, h.bid[i].y, h.bid[i].z, h.bid[i].w};
    for (int r = 0; r < 4; ++r)
      h.vslot[h.prefix[q[r]] + cnt[q[r]]++] = 0x80000000u | ((uint32_t)r << 29) | (uint32_t)i;
  }
  // feature ownership: the lowest triangle index claims each vertex / edge
  std::vector<uint8_t> vown(n, 0);
  std::unordered_map<uint64_t, int> eown;
  eown.reserve((size_t)t * 2);
  h.tri.resize(t);
  for (int i = 0; i < t; ++i) {
    const int a = (int)tri[3 * i], b = (int)tri[3 * i + 1], c = (int)tri[3 * i + 2];
    int f = 0;
    if (!vown[a]) { vown[a] = 1; f |= 1 << 5; }
    if (!vown[b]) { vown[b] = 1; f |= 1 << 4; }
    if (!vown[c]) { vown[c] = 1; f |= 1 << 3; }
    if (eown.emplace(edge_key(a, b), 1).second) f |= 1 << 2;
    if (eown.emplace(edge_key(b, c), 1).second) f |= 1 << 1;
    if (eown.emplace(edge_key(c, a), 1).second) f |= 1 << 0;
    h.tri[i] = make_int4(a, b, c, f);
  }
  // long-range attachments (Physics/CudaPBD.cu:60-81): N x F rest lengths, built only
  // when they are used (a batch of cloths with pinned edges would otherwise hold
  // every vertex's distance to every pinned vertex of every cloth)
  for (int i = 0; i < n && (P.flags & CM_F_LRA); ++i)
    if (pinned[i]) h.lra_fixed.push_back(i);
  const int F = (int)h.lra_fixed.size();
  h.lra_l0.assign((size_t)n * F, 0.f);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < F; ++j)
      if (!pinned[i])
        h.lra_l0[(size_t)i * F + j] = host_dist(xyz, i, h.lra_fixed[j]) * (1.f + P.strain_limit);
  build_tiles(n, h);
  return CM_OK;
}

PbdBufs pbd_bufs(cm_body* b) {
  PbdBufs pb;
  pb.n = b->n;
  pb.S = b->S;
  pb.B = b->B;
  pb.F = b->F;
  pb.x = b->x;
  pb.p = b->p;
  pb.p2 = b->p2;
  pb.v = b->v;
  pb.pinned = b->pinned;
  pb.sid = b->sid;
  pb.s_l0kp = b->s_l0kp;
  pb.bid = b->bid;
  pb.b_kpphi0 = b->b_kpphi0;
  pb.vprefix = b->vprefix;
  pb.vslot = b->vslot;
  pb.ntiles = b->ntiles;
  pb.tstart = b->tstart;
  pb.eprefix = b->eprefix;
  pb.entries = b->entries;
  pb.lra_fixed = b->lra_fixed;
  pb.lra_l0 = b->lra_l0;
  pb.lra_flag = b->lra_flag;
  pb.ground = b->ground;
  pb.mass = b->mass;
  pb.damp_partials = b->damp_partials;
  pb.damp_state = b->damp_state;
  return pb;
}

PbdParams pbd_params(const cm_body* b) {
  const cm_params& P = b->prm;
  PbdParams pp;
  pp.dt = P.timestep;
  pp.th = P.thickness;
  pp.gh = P.ground_height;
  pp.g = mk3(P.gravity[0], P.gravity[1], P.gravity[2]);
  pp.iters = P.solver_iterations;
  pp.flags = P.flags;
  pp.ref_compat = P.ref_compat;
  pp.wmin = mk3(P.world_min[0], P.world_min[1], P.world_min[2]);
  pp.wmax = mk3(P.world_max[0], P.world_max[1], P.world_max[2]);
  pp.sc = mk3(P.sphere_center[0], P.sphere_center[1], P.sphere_center[2]);
  pp.sr = P.sphere_radius;
  pp.kg = P.global_vel_damp;
  pp.kd = P.k_damp;
  pp.kf = P.k_friction;
  return pp;
}

// Where a recorded substep keeps its Jacobi iterates and contact records.
struct TapeSlot {
  float4* pk;
  int* dpre;
  int slot;
};

// One substep on b->stream (Physics/CudaDeformableBody.cpp:550-627 stage order).
// ev (optional): CM_NUM_STAGES + 1 events recorded at the stage boundaries.
// tape (optional): the same substep, keeping what its adjoint needs.
int record_substep(cm_body* b, hipEvent_t* ev = nullptr, const TapeSlot* tape = nullptr) {
  hipStream_t s = b->stream;
  const cm_params& P = b->prm;
  PbdBufs pb = pbd_bufs(b);
  PbdParams pp = pbd_params(b);
  auto mark = [&](int k) {
    if (ev) (void)hipEventRecord(ev[k], s);
  };
  mark(CM_STAGE_PREDICT);
  int rc = launch_predict(pb, pp, s);
  if (rc) return rc;
  // the carried vertex normals advance once per simulated substep: not again when
  // the adjoint recomputes a segment
  if ((P.flags & CM_F_VERTEX_NORMALS) && !tape) {
    rc = launch_vertex_normals(b->n, b->vt_prefix, b->vt_tri, b->p, b->vnrm, 1, 1, s);
    if (rc) return rc;
  }
  const bool coll = (P.flags & CM_F_SELF_COLLISION) && b->t >= 2;
  mark(CM_STAGE_TRIANGLES);
  if (coll) {
    if (b->lb.ncloth)
      rc = launch_tri_update_batch(b->lb, b->x, b->p, b->tri, P.thickness, b->keys, b->perm,
                                   b->cen, s);
    else
      rc = launch_tri_update(b->n, b->t, b->x, b->p, b->tri, P.thickness, P.morton_bits,
                           P.morton_normalize, b->lb.amin, b->lb.amax, b->keys, b->perm, b->cen,
                           b->box, s);
    if (rc) return rc;
  }
  mark(CM_STAGE_SORT);
  if (coll) {
    rc = radix_sort(b->keys, b->perm, b->t, b->key_bits, b->sort_temp, s);
    if (rc) return rc;
    rc = launch_gather_tris(b->t, b->perm, b->tri, b->tri_alt,
                            (P.ref_compat & CM_RC_STALE_LEAF_AABB) ? 0 : 1, b->lb.amin,
                            b->lb.amax, b->tmp_min, b->tmp_max, s);
    if (rc) return rc;
  }
  mark(CM_STAGE_TREE);
  if (coll) {
    rc = launch_build(b->lb, b->keys, (P.ref_compat & CM_RC_MIDPOINT_SPLIT) ? 1 : 0, s);
    if (rc) return rc;
    rc = launch_refit(b->lb, s);
    if (rc) return rc;
  }
  mark(CM_STAGE_PAIRS);
  if (coll) {
    rc = launch_pairs(b->lb, b->cap, b->chunk, b->leaf_cnt, b->leaf_off, b->pairs, b->counters,
                      b->scan_temp, s);
    if (rc) return rc;
  }
  mark(CM_STAGE_NARROW);
  if (coll) {
    rc = launch_narrow(b->nb, b->tri_alt, b->x, b->p, b->v, P.thickness, P.timestep,
                       (P.flags & CM_F_EE_CONTACTS) ? 1 : 0, (P.flags & CM_F_DCD) ? 1 : 0, s);
    if (rc) return rc;
    rc = launch_dynamic_csr(b->nb, s);
    if (rc) return rc;
    if (P.flags & CM_F_REVERT_TOC) {
      rc = launch_revert_toc(pb, b->nb, s);
      if (rc) return rc;
    }
    rc = launch_copy_i4(b->t, b->tri_alt, b->tri, s);
    if (rc) return rc;
    if (tape) {
      CM_HIP(hipMemcpyAsync(tape->dpre, b->nb.dprefix, sizeof(int) * ((size_t)b->n + 1),
                            hipMemcpyDeviceToDevice, s));
      rc = launch_tape_records(b->n, b->nb.dprefix, b->nb.rec, b->adj.arena, b->adj.arena_cap,
                               b->adj.ctl, tape->slot, s);
      if (rc) return rc;
    }
  }
  NarrowBufs nb = b->nb;
  if (!coll) nb.dprefix = nullptr;
  mark(CM_STAGE_SOLVER);
  rc = tape ? launch_solver_rec(pb, pp, nb, tape->pk, s) : launch_solver(pb, pp, nb, s);
  if (rc) return rc;
  mark(CM_STAGE_INTEGRATE);
  rc = tape ? launch_integrate_from(pb, pp, tape->pk + (int64_t)pp.iters * b->n, s)
            : launch_integrate(pb, pp, s);
  if (!rc && (P.flags & CM_F_FRICTION)) rc = launch_friction(pb, pp, nb, s);
  mark(CM_NUM_STAGES);
  return rc;
}

int ensure_graph(cm_body* b) {
  if (b->exec || !b->use_graph) return CM_OK;
  CM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = record_substep(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return CM_ERR_HIP;
  }
  b->graph = g;
  CM_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return CM_OK;
}

int body_check(cm_body* b) {
  if (!b) {
    set_error("null body");
    return CM_ERR_ARG;
  }
  CM_HIP(hipSetDevice(b->dev));
  return CM_OK;
}

int read_f4(cm_body* b, const float4* src, float* xyz) {
  std::vector<float4> h(b->n);
  CM_HIP(hipMemcpyAsync(h.data(), src, sizeof(float4) * b->n, hipMemcpyDeviceToHost, b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  for (int i = 0; i < b->n; ++i) {
    xyz[3 * i] = h[i].x;
    xyz[3 * i + 1] = h[i].y;
    xyz[3 * i + 2] = h[i].z;
  }
  return CM_OK;
}

}  // namespace

extern "C" {

const char* cm_last_error(void) { return g_err.c_str(); }

// The reference's ClothParams literal (Managers/MastersSimulationManager.cpp:33-62) and
// constants (Core/Utils.hpp:16-19, Physics/CudaDeformableUtils.cuh:16-17); wind off.
void cm_default_params(cm_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->timestep = 1.f / 60.f;
  p->gravity[1] = -0.0981f;
  p->k_stretch = 0.99f;
  p->k_bend = 0.1f;
  p->thickness = 0.f;
  p->solver_iterations = 6;
  p->ground_height = 0.f;
  p->world_max[0] = p->world_max[1] = p->world_max[2] = 1.f;
  p->global_vel_damp = 0.01f;
  p->k_damp = 0.05f;
  p->strain_limit = 0.15f;
  p->k_friction = 0.1f;
  p->pinned_mass = 1000.f;
  p->morton_bits = 63;
  p->morton_normalize = 0;
  p->pair_cap = 64;
  p->flags = CM_F_SELF_COLLISION | CM_F_GROUND | CM_F_WORLD_CLAMP | CM_F_DAMPING | CM_F_LRA |
             CM_F_EE_CONTACTS | CM_F_BEND_FP64 | CM_F_VERTEX_NORMALS;
  p->ref_compat = CM_RC_ALL;
}

}  // extern "C"

namespace {

// A batch of cloths packed into one body: cloth c owns vertices [vstart[c], vstart[c+1])
// and triangles [tstart[c], tstart[c+1]); optional per-cloth stiffness.
struct BatchSpec {
  int ncloth = 0;
  std::vector<int> vstart, tstart;
  const float* k_stretch = nullptr;
  const float* k_bend = nullptr;
};

int create_impl(const cm_params* params, const float* xyz, int32_t n, const uint32_t* tri,
                int32_t t, const uint8_t* pinned, const float* inv_mass, const int32_t* dist_pairs,
                int32_t n_dist, const float* k_dist, const float* k_bend, const BatchSpec* batch,
                cm_body** out) {
  if (out) *out = nullptr;
  // vertex ids share a word with 3 flag bits in the contact records
  if (!params || !xyz || n < 1 || n >= (1 << 29) || !tri || t < 1 || !inv_mass || !out ||
      params->solver_iterations < 1 || (params->morton_bits != 30 && params->morton_bits != 63) ||
      params->pair_cap < 1 || (dist_pairs && n_dist < 0) || params->timestep <= 0.f) {
    set_error("cm_create: invalid argument");
    return CM_ERR_ARG;
  }
  for (int64_t i = 0; i < 3 * (int64_t)t; ++i)
    if (tri[i] >= (uint32_t)n) {
      set_error("cm_create: triangle index out of range");
      return CM_ERR_ARG;
    }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("cm_create: no HIP device");
    return CM_ERR_NO_DEVICE;
  }
  cm_body* b = new cm_body();
  b->prm = *params;
  b->n = n;
  b->t = t;
  // CLOTHMI_GRAPH=0 launches every substep's kernels directly instead of replaying
  // the captured substep graph (same kernels, same order).
  if (const char* g = std::getenv("CLOTHMI_GRAPH")) b->use_graph = std::strcmp(g, "0") != 0;
  int rc = CM_OK;
  auto fail = [&](int code) {
    cm_destroy(b);
    return code;
  };
  if (hipGetDevice(&b->dev) != hipSuccess) return fail(CM_ERR_HIP);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(CM_ERR_HIP);
  }
  b->pinned_h.resize(n);
  b->inv_mass_h.assign(inv_mass, inv_mass + n);
  for (int i = 0; i < n; ++i) b->pinned_h[i] = pinned ? (pinned[i] ? 1 : 0) : (inv_mass[i] == 0.f);
  HostTopo h;
  rc = build_topology(b->prm, xyz, n, tri, t, b->pinned_h.data(), dist_pairs, n_dist, k_dist,
                      k_bend, h);
  if (rc) return fail(rc);
  if (batch && (batch->k_stretch || batch->k_bend)) {  // per-cloth stiffness
    std::vector<int> cloth_of(n);
    for (int c = 0; c < batch->ncloth; ++c)
      for (int i = batch->vstart[c]; i < batch->vstart[c + 1]; ++i) cloth_of[i] = c;
    const int I = b->prm.solver_iterations;
    if (batch->k_stretch)
      for (size_t i = 0; i < h.sid.size(); ++i) {
        h.k_dist[i] = batch->k_stretch[cloth_of[h.sid[i].x]];
        h.s_l0kp[i].y = kp_of(h.k_dist[i], I);
      }
    if (batch->k_bend)
      for (size_t i = 0; i < h.bid.size(); ++i) {
        h.k_bend[i] = batch->k_bend[cloth_of[h.bid[i].x]];
        h.b_kpphi0[i].x = kp_of(h.k_bend[i], I);
      }
    build_tiles(n, h);  // tile entries carry the stiffness
  }
  b->S = (int)h.sid.size();
  b->B = (int)h.bid.size();
  b->F = (int)h.lra_fixed.size();
  b->ntiles = h.tstart.empty() ? 0 : (int)h.tstart.size() - 1;
  b->nslots = h.prefix[n];
  for (auto& e : h.sid) {
    b->h_sid.push_back(e.x);
    b->h_sid.push_back(e.y);
  }
  for (auto& e : h.bid) {
    b->h_bid.push_back(e.x);
    b->h_bid.push_back(e.y);
    b->h_bid.push_back(e.z);
    b->h_bid.push_back(e.w);
  }
  b->h_prefix = h.prefix;
  // state
  std::vector<float4> x4(n), v4(n, make_float4(0, 0, 0, 0));
  std::vector<float> mass(n);
  for (int i = 0; i < n; ++i) {
    x4[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], inv_mass[i]);
    mass[i] = b->pinned_h[i] ? params->pinned_mass
                             : (inv_mass[i] > 0 ? 1.f / inv_mass[i] : params->pinned_mass);
  }
  if ((rc = upload(b, &b->x, x4)) || (rc = upload(b, &b->p, x4)) || (rc = dalloc(b, &b->p2, n)) ||
      (rc = upload(b, &b->v, v4)) || (rc = upload(b, &b->pinned, b->pinned_h)) ||
      (rc = upload(b, &b->mass, mass)) || (rc = upload(b, &b->sid, h.sid)) ||
      (rc = upload(b, &b->s_l0kp, h.s_l0kp)) || (rc = upload(b, &b->bid, h.bid)) ||
      (rc = upload(b, &b->b_kpphi0, h.b_kpphi0)) || (rc = upload(b, &b->vprefix, h.prefix)) ||
      (rc = upload(b, &b->vslot, h.vslot)) || (rc = upload(b, &b->lra_fixed, h.lra_fixed)) ||
      (!h.tstart.empty() && ((rc = upload(b, &b->tstart, h.tstart)) ||
                             (rc = upload(b, &b->eprefix, h.eprefix)) ||
                             (rc = upload(b, &b->entries, h.entries)) ||
                             (rc = upload(b, &b->adj.ecid, h.ecid)))) ||
      (rc = upload(b, &b->lra_l0, h.lra_l0)) ||
      (rc = dalloc(b, &b->lra_flag, (size_t)n * std::max(b->F, 1))) ||
      (rc = dalloc(b, &b->ground, n)) ||
      (rc = dalloc(b, &b->damp_partials, 9 * (size_t)nblocks(n))) ||
      (rc = dalloc(b, &b->damp_state, 16)) || (rc = upload(b, &b->tri, h.tri)) ||
      (rc = dalloc(b, &b->tri_alt, t)) || (rc = dalloc(b, &b->warn, 4)) ||
      (rc = dalloc(b, &b->counters, 16)))
    return fail(rc);
  CM_HIP(hipMemset(b->warn, 0, 16));
  CM_HIP(hipMemset(b->counters, 0, 64));
  CM_HIP(hipMemset(b->ground, 0, sizeof(float4) * n));
  {
    std::vector<int> vt_prefix(n + 1, 0);
    for (int64_t i = 0; i < 3 * (int64_t)t; ++i) ++vt_prefix[tri[i] + 1];
    for (int i = 0; i < n; ++i) vt_prefix[i + 1] += vt_prefix[i];
    std::vector<int> fill(vt_prefix.begin(), vt_prefix.end() - 1);
    std::vector<int4> vt_tri(3 * (size_t)t);
    for (int j = 0; j < t; ++j) {
      const int4 q = make_int4((int)tri[3 * j], (int)tri[3 * j + 1], (int)tri[3 * j + 2], j);
      for (int k = 0; k < 3; ++k) vt_tri[fill[tri[3 * j + k]]++] = q;
    }
    if ((rc = upload(b, &b->vt_prefix, vt_prefix)) || (rc = upload(b, &b->vt_tri, vt_tri)) ||
        (rc = dalloc(b, &b->vnrm, n)))
      return fail(rc);
    CM_HIP(hipMemset(b->vnrm, 0, sizeof(float4) * n));
    // the creation's normals: the raw (area-weighted) face normals of the input
    // positions (Physics/CudaDeformableBody.cpp:214, 269)
    if (params->flags & CM_F_VERTEX_NORMALS) {
      rc = launch_vertex_normals(n, b->vt_prefix, b->vt_tri, b->x, b->vnrm, 0, 0, b->stream);
      if (rc) return fail(rc);
      if (hipStreamSynchronize(b->stream) != hipSuccess) {
        set_error("cm_create: vertex normals failed");
        return fail(CM_ERR_HIP);
      }
    }
  }
  // collision buffers
  b->cap = std::min(params->pair_cap, t);
  const size_t m = 2 * (size_t)t - 1;
  const int64_t pcap = (int64_t)t * b->cap;
  b->lb.t = t;
  b->lb.warn = b->warn;
  b->key_bits = params->morton_bits == 30 ? 30 : 63;
  if (batch) {
    const int nc = batch->ncloth;
    std::vector<int> slot_cloth(t), ccap(nc);
    for (int c = 0; c < nc; ++c) {
      for (int i = batch->tstart[c]; i < batch->tstart[c + 1]; ++i) slot_cloth[i] = c;
      ccap[c] = std::min(params->pair_cap, batch->tstart[c + 1] - batch->tstart[c]);
    }
    int cbits = 0;
    while ((1 << cbits) < nc) ++cbits;
    b->key_bits = 30 + cbits;
    b->lb.ncloth = nc;
    int *d_tstart = nullptr, *d_slot = nullptr, *d_ccap = nullptr;
    if ((rc = upload(b, &d_tstart, batch->tstart)) || (rc = upload(b, &d_slot, slot_cloth)) ||
        (rc = upload(b, &d_ccap, ccap)) || (rc = dalloc(b, &b->lb.croot, (size_t)nc)) ||
        (rc = dalloc(b, &b->lb.cbox, 6 * (size_t)nc)))
      return fail(rc);
    b->lb.tstart = d_tstart;
    b->lb.slot_cloth = d_slot;
    b->lb.ccap = d_ccap;
  }
  if ((rc = dalloc(b, &b->keys, t)) || (rc = dalloc(b, &b->perm, t)) ||
      (rc = dalloc(b, (char**)&b->sort_temp, radix_sort_temp_bytes(t))) ||
      (rc = dalloc(b, &b->cen, t)) || (rc = dalloc(b, &b->box, 2 * (size_t)nblocks(t) + 2)) ||
      (rc = dalloc(b, &b->tmp_min, t)) || (rc = dalloc(b, &b->tmp_max, t)) ||
      (rc = dalloc(b, &b->lb.amin, m)) || (rc = dalloc(b, &b->lb.amax, m)) ||
      (rc = dalloc(b, &b->lb.lefts, t)) || (rc = dalloc(b, &b->lb.rights, t)) ||
      (rc = dalloc(b, &b->lb.parents, m)) || (rc = dalloc(b, &b->lb.rmll, t)) ||
      (rc = dalloc(b, &b->lb.rmlr, t)) ||
      (rc = dalloc(b, (char**)&b->lb_temp, lbvh_build_temp_bytes(t))) ||
      (rc = dalloc(b, &b->chunk, pcap)) || (rc = dalloc(b, &b->pairs, pcap)) ||
      (rc = dalloc(b, &b->leaf_cnt, t)) || (rc = dalloc(b, &b->leaf_off, t)))
    return fail(rc);
  // two-pass traversal: leaves still walking after the budget's visits are finished
  // by whole waves; up to a quarter of the leaves can be handed over, the rest walk on
  // in the first pass. Off for caps above 96 (the second pass keeps at most 256 pairs
  // per leaf). CLOTHMI_TRAV_BUDGET overrides the budget (0: one pass); the pairs are
  // the same with any budget.
  if ((rc = dalloc(b, &b->lb.spill, (size_t)t * kTravSpill))) return fail(rc);
  {
    int budget = params->pair_cap <= 96 ? 200 : 0;
    if (const char* e = std::getenv("CLOTHMI_TRAV_BUDGET")) budget = std::max(0, std::atoi(e));
    if (budget > 0 && t >= 2 && t < (1 << 26)) {
      b->lb.rcap = t / 4 + 1024;
      if ((rc = dalloc(b, &b->lb.trec, (size_t)b->lb.rcap * kTravRec))) return fail(rc);
      b->lb.hctl = b->counters + 8;  // zeroed above
      b->lb.trav_budget = budget;
    }
  }
  CM_HIP(hipMemset(b->sort_temp, 0, radix_sort_temp_bytes(t)));
  carve_build(t, b->lb_temp, &b->lb);
  CM_HIP(hipMemset(b->lb_temp, 0, lbvh_build_temp_bytes(t)));
  CM_HIP(hipMemset(b->lb.parents, 0, sizeof(int) * m));
  CM_HIP(hipMemset(b->lb.amin, 0, sizeof(float4) * m));
  CM_HIP(hipMemset(b->lb.amax, 0, sizeof(float4) * m));
  CM_HIP(hipMemset(b->lb.rmll, 0, sizeof(int) * t));
  CM_HIP(hipMemset(b->lb.rmlr, 0, sizeof(int) * t));
  CM_HIP(hipMemset(b->keys, 0, sizeof(uint64_t) * t));
  CM_HIP(hipMemset(b->perm, 0, sizeof(uint32_t) * t));
  // narrow phase capacities: candidates per triangle (reserved[0] / [1] override the
  // VF / EE budgets), CCD survivors per triangle (reserved[2])
  const int64_t vf_per = params->reserved[0] > 0 ? params->reserved[0] : 128;
  const int64_t ee_per = params->reserved[1] > 0 ? params->reserved[1] : 192;
  const int64_t out_per = params->reserved[2] > 0 ? params->reserved[2] : 16;
  NarrowBufs& nb = b->nb;
  nb.n = n;
  nb.t = t;
  nb.pair_cap = pcap;
  nb.vf_cap = std::min<int64_t>(vf_per * t + 1024, 6 * pcap);
  nb.ee_cap = std::min<int64_t>(ee_per * t + 1024, 9 * pcap);
  nb.vf_out_cap = std::min<int64_t>(out_per * t + 65536, nb.vf_cap);
  nb.ee_out_cap = std::min<int64_t>(out_per * t + 65536, nb.ee_cap);
  nb.pairs = b->pairs;
  nb.d_npairs = b->counters + 0;
  nb.d_ncand_vf = b->counters + 1;
  nb.d_ncand_ee = b->counters + 2;
  nb.d_nvf = b->counters + 3;
  nb.d_nee = b->counters + 4;
  nb.d_nrec = b->counters + 5;
  nb.d_ntotal = b->counters + 6;
  nb.d_nnear = b->counters + 7;
  nb.warn = b->warn;
  {
    // one scan temp for every scan of the substep, sized for the longest
    const int64_t smax = std::max<int64_t>({(int64_t)t, (int64_t)n, pcap, nb.vf_cap, nb.ee_cap});
    if ((rc = dalloc(b, (char**)&b->scan_temp, scan_temp_bytes((int)smax)))) return fail(rc);
    CM_HIP(hipMemset(b->scan_temp, 0, scan_temp_bytes((int)smax)));
  }
  nb.scan_temp = b->scan_temp;
  // tiles of 256 pairs or candidates (their counts, then the counts' prefixes)
  const int64_t ntile = (std::max<int64_t>(pcap, nb.vf_cap + nb.ee_cap) + 255) / 256;
  if ((rc = dalloc(b, &nb.mask, pcap)) || (rc = dalloc(b, &nb.tile_sum, ntile)) ||
      (rc = dalloc(b, &nb.tile_pre, 2 * ntile)) || (rc = dalloc(b, &nb.cand_vf, nb.vf_cap)) ||
      (rc = dalloc(b, &nb.cand_ee, nb.ee_cap)) ||
      (rc = dalloc(b, &nb.hit_flag, nb.vf_cap + nb.ee_cap)) ||
      (rc = dalloc(b, &nb.near_ids, nb.vf_cap + nb.ee_cap)) ||
      (rc = dalloc(b, &nb.cplan, nb.vf_cap + nb.ee_cap)) ||
      (rc = dalloc(b, &nb.ccp, nb.vf_cap + nb.ee_cap)) ||
      (rc = dalloc(b, &nb.croot, 3 * (nb.vf_cap + nb.ee_cap))) ||
      (rc = dalloc(b, &nb.vf_v, nb.vf_out_cap)) ||
      (rc = dalloc(b, &nb.vf_w, nb.vf_out_cap)) || (rc = dalloc(b, &nb.vf_tn, nb.vf_out_cap)) ||
      (rc = dalloc(b, &nb.ee_v, nb.ee_out_cap)) || (rc = dalloc(b, &nb.ee_w, nb.ee_out_cap)) ||
      (rc = dalloc(b, &nb.ee_tn, nb.ee_out_cap)) || (rc = dalloc(b, &nb.dcount, (size_t)n + 1)) ||
      (rc = dalloc(b, &nb.dprefix, (size_t)n + 1)) || (rc = dalloc(b, &nb.dcursor, n)) ||
      (rc = dalloc(b, &nb.drec, 4 * (size_t)(nb.vf_out_cap + nb.ee_out_cap))) ||
      (rc = dalloc(b, &nb.dtmp, 4 * (size_t)(nb.vf_out_cap + nb.ee_out_cap))) ||
      (rc = dalloc(b, &nb.rec, 4 * (size_t)(nb.vf_out_cap + nb.ee_out_cap))) ||
      (rc = dalloc(b, &nb.dlong, (size_t)n)))
    return fail(rc);
  // The contact-record CSR comes from one radix sort of the records by vertex id;
  // CLOTHMI_SORT_CSR=0 selects per-vertex counts with atomic placement instead (the
  // same lists; the sort needs record capacities below 2^31)
  {
    const int64_t rcap = 4 * (nb.vf_out_cap + nb.ee_out_cap);
    bool sorted = rcap < ((int64_t)1 << 31);
    if (const char* e = std::getenv("CLOTHMI_SORT_CSR")) sorted = sorted && std::atoi(e) != 0;
    if (sorted) {
      nb.sorted_csr = true;
      if ((rc = dalloc(b, &nb.skey, 2 * (size_t)rcap)) ||
          (rc = dalloc(b, (char**)&nb.sort_temp, radix_sort_dev_temp_bytes())))
        return fail(rc);
      CM_HIP(hipMemset(nb.sort_temp, 0, radix_sort_dev_temp_bytes()));
    }
  }
  CM_HIP(hipMemset(nb.dprefix, 0, sizeof(int) * ((size_t)n + 1)));
  CM_HIP(hipMemset(nb.dcount, 0, sizeof(int) * ((size_t)n + 1)));  // kept zero between substeps
  if (params->flags & CM_F_SPHERE) {
    b->prm.flags |= CM_F_SPHERE;
  }
  b->topo = std::move(h);
  CM_HIP(hipDeviceSynchronize());
  *out = b;
  return CM_OK;
}

}  // namespace

extern "C" {

int cm_create(const cm_params* params, const float* xyz, int32_t n, const uint32_t* tri,
              int32_t t, const uint8_t* pinned, const float* inv_mass, const int32_t* dist_pairs,
              int32_t n_dist, const float* k_dist, const float* k_bend, cm_body** out) {
  return create_impl(params, xyz, n, tri, t, pinned, inv_mass, dist_pairs, n_dist, k_dist, k_bend,
                     nullptr, out);
}

int cm_create_batch(const cm_params* params, int32_t ncloth, const int32_t* cloth_n,
                    const int32_t* cloth_t, const int32_t* cloth_ndist, const float* xyz,
                    const uint32_t* tri, const uint8_t* pinned, const float* inv_mass,
                    const int32_t* dist_pairs, const float* k_stretch, const float* k_bend,
                    cm_body** out) {
  if (out) *out = nullptr;
  if (!params || ncloth < 1 || !cloth_n || !cloth_t || !xyz || !tri || !inv_mass || !out ||
      (dist_pairs && !cloth_ndist)) {
    set_error("cm_create_batch: invalid argument");
    return CM_ERR_ARG;
  }
  if (params->morton_bits != 30 || !params->morton_normalize ||
      (params->flags & (CM_F_LRA | CM_F_DAMPING | CM_F_FRICTION)) || ncloth > (1 << 20)) {
    set_error("cm_create_batch: a batch needs 30-bit keys in each cloth's box, no LRA, no "
              "rigid-mode damping, no friction, and at most 2^20 cloths");
    return CM_ERR_ARG;
  }
  BatchSpec bs;
  bs.ncloth = ncloth;
  bs.k_stretch = k_stretch;
  bs.k_bend = k_bend;
  bs.vstart.assign(1, 0);
  bs.tstart.assign(1, 0);
  int64_t nd = 0;
  for (int c = 0; c < ncloth; ++c) {
    if (cloth_n[c] < 1 || cloth_t[c] < 1 || (cloth_ndist && cloth_ndist[c] < 0)) {
      set_error("cm_create_batch: empty cloth");
      return CM_ERR_ARG;
    }
    const int64_t nv = (int64_t)bs.vstart.back() + cloth_n[c];
    const int64_t nt = (int64_t)bs.tstart.back() + cloth_t[c];
    if (nv >= (1 << 29) || nt >= (1 << 30)) {
      set_error("cm_create_batch: batch too large");
      return CM_ERR_ARG;
    }
    bs.vstart.push_back((int)nv);
    bs.tstart.push_back((int)nt);
    if (cloth_ndist) nd += cloth_ndist[c];
  }
  const int n = bs.vstart.back(), t = bs.tstart.back();
  // cloth-local ids -> body ids
  std::vector<uint32_t> gtri(3 * (size_t)t);
  for (int c = 0; c < ncloth; ++c)
    for (int64_t i = 3 * (int64_t)bs.tstart[c]; i < 3 * (int64_t)bs.tstart[c + 1]; ++i) {
      if (tri[i] >= (uint32_t)cloth_n[c]) {
        set_error("cm_create_batch: triangle index out of range");
        return CM_ERR_ARG;
      }
      gtri[i] = tri[i] + (uint32_t)bs.vstart[c];
    }
  std::vector<int32_t> gdist;
  if (dist_pairs) {
    gdist.resize(2 * (size_t)nd);
    int64_t o = 0;
    for (int c = 0; c < ncloth; ++c)
      for (int64_t i = 0; i < 2 * (int64_t)cloth_ndist[c]; ++i, ++o) {
        if (dist_pairs[o] < 0 || dist_pairs[o] >= cloth_n[c]) {
          set_error("cm_create_batch: distance constraint vertex out of range");
          return CM_ERR_ARG;
        }
        gdist[o] = dist_pairs[o] + bs.vstart[c];
      }
  }
  return create_impl(params, xyz, n, gtri.data(), t, pinned, inv_mass,
                     dist_pairs ? gdist.data() : nullptr, (int32_t)nd, nullptr, nullptr, &bs, out);
}

int cm_destroy(cm_body* b) {
  if (!b) return CM_OK;
  hipSetDevice(b->dev);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (void* p : b->allocs) hipFree(p);
  for (void* p : b->adj.allocs) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return CM_OK;
}

int cm_step(cm_body* b, int32_t substeps) {
  int rc = body_check(b);
  if (rc) return rc;
  if (substeps < 0) {
    set_error("cm_step: negative substep count");
    return CM_ERR_ARG;
  }
  rc = ensure_graph(b);
  if (rc) return rc;
  for (int s = 0; s < substeps; ++s) {
    if (b->exec) {
      CM_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
      rc = record_substep(b);
      if (rc) return rc;
    }
  }
  return CM_OK;
}

int cm_sync(cm_body* b) {
  int rc = body_check(b);
  if (rc) return rc;
  CM_HIP(hipStreamSynchronize(b->stream));
  return CM_OK;
}

int cm_read_positions(cm_body* b, float* xyz) {
  int rc = body_check(b);
  if (rc) return rc;
  if (!xyz) {
    set_error("cm_read_positions: null output");
    return CM_ERR_ARG;
  }
  return read_f4(b, b->x, xyz);
}

int cm_read_velocities(cm_body* b, float* xyz) {
  int rc = body_check(b);
  if (rc) return rc;
  if (!xyz) {
    set_error("cm_read_velocities: null output");
    return CM_ERR_ARG;
  }
  return read_f4(b, b->v, xyz);
}

// With CM_F_VERTEX_NORMALS the normals carried through the substeps (computed from
// the predicted positions of each substep, as the reference's FixedUpdate does);
// without it, the normals of the current positions, computed now on the device.
int cm_read_normals(cm_body* b, float* out) {
  int rc = body_check(b);
  if (rc) return rc;
  if (!out) {
    set_error("cm_read_normals: null output");
    return CM_ERR_ARG;
  }
  if (!(b->prm.flags & CM_F_VERTEX_NORMALS)) {
    rc = launch_vertex_normals(b->n, b->vt_prefix, b->vt_tri, b->x, b->vnrm, 0, 1, b->stream);
    if (rc) return rc;
  }
  return read_f4(b, b->vnrm, out);
}

int cm_read_vertices(cm_body* b, float* verts) {
  int rc = body_check(b);
  if (rc) return rc;
  if (!verts) {
    set_error("cm_read_vertices: null output");
    return CM_ERR_ARG;
  }
  std::vector<float> xyz(3 * (size_t)b->n), nrm(3 * (size_t)b->n);
  if ((rc = cm_read_positions(b, xyz.data())) || (rc = cm_read_normals(b, nrm.data()))) return rc;
  for (int i = 0; i < b->n; ++i) {
    float* r = verts + (size_t)CM_VERTEX_FLOATS * i;
    for (int k = 0; k < 3; ++k) {
      r[CM_VERTEX_POSITION + k] = xyz[3 * (size_t)i + k];
      r[CM_VERTEX_NORMAL + k] = nrm[3 * (size_t)i + k];
    }
  }
  return CM_OK;
}

int cm_set_sphere(cm_body* b, const float center[3], float radius) {
  int rc = body_check(b);
  if (rc) return rc;
  CM_HIP(hipStreamSynchronize(b->stream));
  b->prm.sphere_center[0] = center[0];
  b->prm.sphere_center[1] = center[1];
  b->prm.sphere_center[2] = center[2];
  b->prm.sphere_radius = radius;
  b->prm.flags |= CM_F_SPHERE;
  // a recorded forward ran without this sphere: its adjoint would recompute the
  // segments with it, so the record is dropped
  b->adj.ready = false;
  // kernel arguments changed: re-record the substep graph
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  return CM_OK;
}

int cm_set_state(cm_body* b, const float* x, const float* p, const float* v,
                 const int32_t* tri) {
  int rc = body_check(b);
  if (rc) return rc;
  if (!x || !v) {
    set_error("cm_set_state: x and v are required");
    return CM_ERR_ARG;
  }
  std::vector<float4> hx(b->n), hp(b->n), hv(b->n);
  const float* pp = p ? p : x;
  for (int i = 0; i < b->n; ++i) {
    const float w = b->inv_mass_h[i];
    hx[i] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], w);
    hp[i] = make_float4(pp[3 * i], pp[3 * i + 1], pp[3 * i + 2], w);
    hv[i] = make_float4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.f);
  }
  CM_HIP(hipStreamSynchronize(b->stream));
  CM_HIP(hipMemcpy(b->x, hx.data(), sizeof(float4) * b->n, hipMemcpyHostToDevice));
  CM_HIP(hipMemcpy(b->p, hp.data(), sizeof(float4) * b->n, hipMemcpyHostToDevice));
  CM_HIP(hipMemcpy(b->v, hv.data(), sizeof(float4) * b->n, hipMemcpyHostToDevice));
  if (tri) {
    for (int64_t i = 0; i < 4 * (int64_t)b->t; i += 4)
      if (tri[i] < 0 || tri[i] >= b->n || tri[i + 1] < 0 || tri[i + 1] >= b->n || tri[i + 2] < 0 ||
          tri[i + 2] >= b->n) {
        set_error("cm_set_state: triangle index out of range");
        return CM_ERR_ARG;
      }
    CM_HIP(hipMemcpy(b->tri, tri, sizeof(int4) * b->t, hipMemcpyHostToDevice));
  }
  return CM_OK;
}

int cm_get_counts(cm_body* b, int64_t* counts) {
  int rc = body_check(b);
  if (rc) return rc;
  int c[8] = {0};
  CM_HIP(hipMemcpyAsync(c, b->counters, sizeof(int) * 8, hipMemcpyDeviceToHost, b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  counts[0] = b->n;
  counts[1] = b->t;
  counts[2] = b->S;
  counts[3] = b->B;
  counts[4] = c[0];
  counts[5] = c[3];
  counts[6] = c[4];
  counts[7] = c[1];
  counts[8] = c[2];
  return CM_OK;
}

int cm_get_warnings(cm_body* b, uint32_t* w) {
  int rc = body_check(b);
  if (rc) return rc;
  uint32_t h = 0;
  CM_HIP(hipMemcpyAsync(&h, b->warn, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  *w = h & 0xffffu;
  return CM_OK;
}

int cm_get_buffer(cm_body* b, int32_t which, void* host, int64_t capacity_bytes, int64_t* count) {
  int rc = body_check(b);
  if (rc) return rc;
  int c[8] = {0};
  CM_HIP(hipMemcpyAsync(c, b->counters, sizeof(int) * 8, hipMemcpyDeviceToHost, b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  const void* src = nullptr;
  size_t elem = 0;
  int64_t n = 0;
  const int t = b->t;
  switch (which) {
    case CM_BUF_X: src = b->x; elem = 16; n = b->n; break;
    case CM_BUF_P: src = b->p; elem = 16; n = b->n; break;
    case CM_BUF_V: src = b->v; elem = 16; n = b->n; break;
    case CM_BUF_TRI: src = b->tri; elem = 16; n = t; break;
    case CM_BUF_KEYS: src = b->keys; elem = 8; n = t; break;
    case CM_BUF_PERM: src = b->perm; elem = 4; n = t; break;
    case CM_BUF_LEFTS: src = b->lb.lefts; elem = 4; n = t - 1; break;
    case CM_BUF_RIGHTS: src = b->lb.rights; elem = 4; n = t - 1; break;
    case CM_BUF_PARENTS: src = b->lb.parents; elem = 4; n = 2 * (int64_t)t - 1; break;
    case CM_BUF_RMLL: src = b->lb.rmll; elem = 4; n = t - 1; break;
    case CM_BUF_RMLR: src = b->lb.rmlr; elem = 4; n = t - 1; break;
    case CM_BUF_AABB_MIN: src = b->lb.amin; elem = 16; n = 2 * (int64_t)t - 1; break;
    case CM_BUF_AABB_MAX: src = b->lb.amax; elem = 16; n = 2 * (int64_t)t - 1; break;
    case CM_BUF_PAIRS: src = b->pairs; elem = 8; n = std::min<int64_t>(c[0], b->nb.pair_cap); break;
    case CM_BUF_VF:
    case CM_BUF_EE: {
      const bool vf = which == CM_BUF_VF;
      const int64_t cnt = vf ? c[3] : c[4];
      if (count) *count = cnt;
      if (!host) return CM_OK;
      const int64_t k = std::min<int64_t>(cnt, capacity_bytes / (int64_t)sizeof(cm_contact));
      std::vector<int4> hv(k);
      std::vector<float4> hw(k), htn(k);
      if (k > 0) {
        CM_HIP(hipMemcpy(hv.data(), vf ? b->nb.vf_v : b->nb.ee_v, 16 * k, hipMemcpyDeviceToHost));
        CM_HIP(hipMemcpy(hw.data(), vf ? b->nb.vf_w : b->nb.ee_w, 16 * k, hipMemcpyDeviceToHost));
        CM_HIP(hipMemcpy(htn.data(), vf ? b->nb.vf_tn : b->nb.ee_tn, 16 * k, hipMemcpyDeviceToHost));
      }
      cm_contact* o = (cm_contact*)host;
      for (int64_t i = 0; i < k; ++i) {
        o[i].v[0] = hv[i].x; o[i].v[1] = hv[i].y; o[i].v[2] = hv[i].z; o[i].v[3] = hv[i].w;
        o[i].w[0] = hw[i].x; o[i].w[1] = hw[i].y; o[i].w[2] = hw[i].z; o[i].w[3] = hw[i].w;
        o[i].t = htn[i].x;
        o[i].n[0] = htn[i].y; o[i].n[1] = htn[i].z; o[i].n[2] = htn[i].w;
      }
      return CM_OK;
    }
    default:
      set_error("cm_get_buffer: unknown buffer");
      return CM_ERR_ARG;
  }
  if (n < 0) n = 0;
  if (count) *count = n;
  if (host && n > 0) {
    const int64_t k = std::min<int64_t>(n, capacity_bytes / (int64_t)elem);
    CM_HIP(hipMemcpy(host, src, elem * k, hipMemcpyDeviceToHost));
  }
  return CM_OK;
}

int cm_time_kernel(cm_body* b, int32_t which, int32_t reps, float* avg_ms) {
  int rc = body_check(b);
  if (rc) return rc;
  if (reps < 1 || !avg_ms ||
      (which != CM_KERNEL_SOLVE && which != CM_KERNEL_SUBSTEP && which != CM_KERNEL_TRAVERSE &&
       which != CM_KERNEL_CCD_PLAN)) {
    set_error("cm_time_kernel: invalid argument");
    return CM_ERR_ARG;
  }
  const bool coll = (b->prm.flags & CM_F_SELF_COLLISION) && b->t >= 2;
  if ((which == CM_KERNEL_TRAVERSE || which == CM_KERNEL_CCD_PLAN) && !coll) {
    set_error("cm_time_kernel: the body has no self-collision tree");
    return CM_ERR_STATE;
  }
  if (which == CM_KERNEL_SOLVE) reps += reps & 1;  // even: the result stays in p
  rc = ensure_graph(b);
  if (rc) return rc;
  hipEvent_t e0, e1;
  CM_HIP(hipEventCreate(&e0));
  CM_HIP(hipEventCreate(&e1));
  CM_HIP(hipEventRecord(e0, b->stream));
  if (which == CM_KERNEL_SOLVE) {
    NarrowBufs nb = b->nb;
    if (!coll) nb.dprefix = nullptr;
    rc = launch_solve_only(pbd_bufs(b), pbd_params(b), nb, reps, b->stream);
  } else if (which == CM_KERNEL_CCD_PLAN) {
    for (int r = 0; r < reps && !rc; ++r)
      rc = launch_ccd_plan_only(b->nb, b->x, b->v, b->prm.thickness, b->prm.timestep,
                                (b->prm.flags & CM_F_DCD) ? 1 : 0, b->stream);
  } else if (which == CM_KERNEL_TRAVERSE) {
    for (int r = 0; r < reps && !rc; ++r)
      rc = launch_traverse_only(b->lb, b->cap, b->chunk, b->leaf_cnt, b->leaf_off, b->stream);
  } else {
    for (int r = 0; r < reps && !rc; ++r) {
      if (b->exec) {
        hipError_t e = hipGraphLaunch(b->exec, b->stream);
        if (e != hipSuccess) {
          set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e));
          rc = CM_ERR_HIP;
        }
      } else {
        rc = record_substep(b);
      }
    }
  }
  CM_HIP(hipEventRecord(e1, b->stream));
  CM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  CM_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *avg_ms = ms / (float)reps;
  return CM_OK;
}

int cm_stage_times(cm_body* b, int32_t substeps, float* ms) {
  int rc = body_check(b);
  if (rc) return rc;
  if (substeps < 1 || !ms) {
    set_error("cm_stage_times: invalid argument");
    return CM_ERR_ARG;
  }
  hipEvent_t ev[CM_NUM_STAGES + 1];
  for (int k = 0; k <= CM_NUM_STAGES; ++k) CM_HIP(hipEventCreate(&ev[k]));
  for (int k = 0; k < CM_NUM_STAGES; ++k) ms[k] = 0.f;
  for (int s = 0; s < substeps && !rc; ++s) {
    rc = record_substep(b, ev);
    if (rc) break;
    CM_HIP(hipEventSynchronize(ev[CM_NUM_STAGES]));
    for (int k = 0; k < CM_NUM_STAGES; ++k) {
      float m = 0.f;
      CM_HIP(hipEventElapsedTime(&m, ev[k], ev[k + 1]));
      ms[k] += m;
    }
  }
  for (int k = 0; k <= CM_NUM_STAGES; ++k) (void)hipEventDestroy(ev[k]);
  return rc;
}

int cm_query_pairs(cm_body* b, int32_t* pairs, int32_t capacity, int32_t* n_pairs) {
  int64_t cnt = 0;
  int rc = cm_get_buffer(b, CM_BUF_PAIRS, pairs, (int64_t)capacity * 8, &cnt);
  if (n_pairs) *n_pairs = (int32_t)cnt;
  return rc;
}

int cm_query_contacts(cm_body* b, cm_contact* vf, int32_t cap_vf, int32_t* n_vf, cm_contact* ee,
                      int32_t cap_ee, int32_t* n_ee) {
  int64_t a = 0, c = 0;
  int rc = cm_get_buffer(b, CM_BUF_VF, vf, (int64_t)cap_vf * sizeof(cm_contact), &a);
  if (rc) return rc;
  rc = cm_get_buffer(b, CM_BUF_EE, ee, (int64_t)cap_ee * sizeof(cm_contact), &c);
  if (n_vf) *n_vf = (int32_t)a;
  if (n_ee) *n_ee = (int32_t)c;
  return rc;
}

}  // extern "C"

// ---- differentiable rollout (no reference counterpart) ----------------------------
namespace {

template <class T>
int adj_alloc(cm_body* b, T** out, size_t count) {
  *out = nullptr;
  if (count == 0) count = 1;
  void* ptr = nullptr;
  hipError_t e = hipMalloc(&ptr, sizeof(T) * count);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + std::to_string(sizeof(T) * count) +
              " bytes): " + hipGetErrorString(e));
    return CM_ERR_ALLOC;
  }
  b->adj.allocs.push_back(ptr);
  *out = (T*)ptr;
  return CM_OK;
}

void adj_release(cm_body* b) {
  AdjState& a = b->adj;
  for (void* ptr : a.allocs) hipFree(ptr);
  int* ecid = a.ecid;
  a = AdjState();
  a.ecid = ecid;
  a.ab.ecid = ecid;
}

int adj_check(cm_body* b, const char* who) {
  const uint32_t off =
      CM_F_GROUND | CM_F_LRA | CM_F_SPHERE | CM_F_DAMPING | CM_F_FRICTION | CM_F_REVERT_TOC;
  if (b->prm.flags & off) {
    set_error(std::string(who) + ": the differentiable substep has no ground, long-range "
                                 "attachments, sphere, damping, friction or contact-time rollback");
    return CM_ERR_ARG;
  }
  if (b->ntiles <= 0) {
    set_error(std::string(who) + ": a vertex has more constraint slots than a Jacobi tile holds");
    return CM_ERR_ARG;
  }
  return CM_OK;
}

// buffers for a forward of nsub substeps with a checkpoint every `every`
int adj_setup(cm_body* b, int nsub, int every) {
  AdjState& a = b->adj;
  const int I = b->prm.solver_iterations;
  const int nck = (nsub + every - 1) / every;
  if (a.pk && a.every == every && a.nck >= nck && a.iters == I) return CM_OK;
  adj_release(b);
  const size_t n = (size_t)b->n, t = (size_t)b->t, K = (size_t)nck + 1;
  int rc;
  if ((rc = adj_alloc(b, &a.ck_x, K * n)) || (rc = adj_alloc(b, &a.ck_v, K * n)) ||
      (rc = adj_alloc(b, &a.ck_p, K * n)) || (rc = adj_alloc(b, &a.ck_tri, K * t)) ||
      (rc = adj_alloc(b, &a.pk, (size_t)every * (I + 1) * n)) ||
      (rc = adj_alloc(b, &a.dpre, (size_t)every * (n + 1))) ||
      (rc = adj_alloc(b, &a.ctl, 2 + (size_t)every)) || (rc = adj_alloc(b, &a.ab.lx, n)) ||
      (rc = adj_alloc(b, &a.ab.lv, n)) || (rc = adj_alloc(b, &a.ab.lp, n)) ||
      (rc = adj_alloc(b, &a.ab.la, n)) || (rc = adj_alloc(b, &a.ab.av, n)) ||
      (rc = adj_alloc(b, &a.ab.dcnt, (size_t)I * n)) ||
      (rc = adj_alloc(b, &a.ab.gks, (size_t)b->S)) || (rc = adj_alloc(b, &a.ab.gkb, (size_t)b->B)))
    return rc;
  const int64_t per = b->prm.reserved[3] > 0 ? b->prm.reserved[3] : std::max<int64_t>(b->n, 65536);
  a.arena_cap = (int64_t)every * per;
  if ((rc = adj_alloc(b, &a.arena, (size_t)a.arena_cap))) return rc;
  a.every = every;
  a.nck = nck;
  a.iters = I;
  return CM_OK;
}

int adj_grow_arena(cm_body* b, int64_t need) {
  AdjState& a = b->adj;
  for (size_t k = 0; k < a.allocs.size(); ++k)
    if (a.allocs[k] == a.arena) {
      hipFree(a.arena);
      a.allocs.erase(a.allocs.begin() + k);
      break;
    }
  a.arena = nullptr;
  a.arena_cap = need + need / 4 + 1024;
  return adj_alloc(b, &a.arena, (size_t)a.arena_cap);
}

// checkpoint k <-> the body's substep state: positions, velocities, predicted
// positions and the triangle order (the sort's tie order); every other buffer is
// rewritten by the substep before it is read
int adj_ckpt(cm_body* b, int k, bool save) {
  AdjState& a = b->adj;
  const size_t n = (size_t)b->n, t = (size_t)b->t;
  auto cp = [&](void* ck, void* live, size_t bytes) -> hipError_t {
    return save ? hipMemcpyAsync(ck, live, bytes, hipMemcpyDeviceToDevice, b->stream)
                : hipMemcpyAsync(live, ck, bytes, hipMemcpyDeviceToDevice, b->stream);
  };
  CM_HIP(cp(a.ck_x + k * n, b->x, sizeof(float4) * n));
  CM_HIP(cp(a.ck_v + k * n, b->v, sizeof(float4) * n));
  CM_HIP(cp(a.ck_p + k * n, b->p, sizeof(float4) * n));
  CM_HIP(cp(a.ck_tri + k * t, b->tri, sizeof(int4) * t));
  return CM_OK;
}

int upload_xyz(cm_body* b, const float* src, float4* dst) {
  std::vector<float4> h(b->n, make_float4(0.f, 0.f, 0.f, 0.f));
  if (src)
    for (int i = 0; i < b->n; ++i) h[i] = make_float4(src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.f);
  CM_HIP(hipMemcpyAsync(dst, h.data(), sizeof(float4) * b->n, hipMemcpyHostToDevice, b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  return CM_OK;
}

// dkp/dk of kp = 1 - (1 - k)^(1 / I)
double dkp_dk(double k, int iters) {
  return std::pow(1.0 - k, 1.0 / iters - 1.0) / iters;
}

}  // namespace

extern "C" {

int cm_set_stiffness(cm_body* b, const float* k_dist, const float* k_bend) {
  int rc = body_check(b);
  if (rc) return rc;
  HostTopo& h = b->topo;
  const int I = b->prm.solver_iterations;
  for (int i = 0; k_dist && i < b->S; ++i)
    if (!(k_dist[i] >= 0.f && k_dist[i] <= 1.f)) {
      set_error("cm_set_stiffness: stiffness outside [0, 1]");
      return CM_ERR_ARG;
    }
  for (int i = 0; k_bend && i < b->B; ++i)
    if (!(k_bend[i] >= 0.f && k_bend[i] <= 1.f)) {
      set_error("cm_set_stiffness: stiffness outside [0, 1]");
      return CM_ERR_ARG;
    }
  for (int i = 0; k_dist && i < b->S; ++i) {
    h.k_dist[i] = k_dist[i];
    h.s_l0kp[i].y = kp_of(k_dist[i], I);
  }
  for (int i = 0; k_bend && i < b->B; ++i) {
    h.k_bend[i] = k_bend[i];
    h.b_kpphi0[i].x = kp_of(k_bend[i], I);
  }
  if (b->ntiles > 0) build_tiles(b->n, h, b->tile_v);
  CM_HIP(hipStreamSynchronize(b->stream));
  if (b->S)
    CM_HIP(hipMemcpy(b->s_l0kp, h.s_l0kp.data(), sizeof(float2) * b->S, hipMemcpyHostToDevice));
  if (b->B)
    CM_HIP(hipMemcpy(b->b_kpphi0, h.b_kpphi0.data(), sizeof(float2) * b->B,
                     hipMemcpyHostToDevice));
  if (b->ntiles > 0)
    CM_HIP(hipMemcpy(b->entries, h.entries.data(), sizeof(TileEntry) * h.entries.size(),
                     hipMemcpyHostToDevice));
  b->adj.ready = false;  // a recorded forward used the old stiffness
  return CM_OK;
}

int cm_forward_ckpt(cm_body* b, int32_t substeps, int32_t every) {
  int rc = body_check(b);
  if (rc) return rc;
  if (substeps < 1 || every < 1) {
    set_error("cm_forward_ckpt: substeps and every must be positive");
    return CM_ERR_ARG;
  }
  if ((rc = adj_check(b, "cm_forward_ckpt")) || (rc = adj_setup(b, substeps, every)) ||
      (rc = ensure_graph(b)))
    return rc;
  AdjState& a = b->adj;
  a.ready = false;
  for (int s = 0; s < substeps; ++s) {
    if (s % every == 0 && (rc = adj_ckpt(b, s / every, true))) return rc;
    if (b->exec) {
      CM_HIP(hipGraphLaunch(b->exec, b->stream));
    } else if ((rc = record_substep(b))) {
      return rc;
    }
  }
  if ((rc = adj_ckpt(b, a.nck, true))) return rc;
  a.nsub = substeps;
  a.ready = true;
  return CM_OK;
}

int cm_adjoint(cm_body* b, const float* dl_dx, const float* dl_dv, float* dl_dx0, float* dl_dv0,
               float* dl_dk_dist, float* dl_dk_bend) {
  int rc = body_check(b);
  if (rc) return rc;
  AdjState& a = b->adj;
  if (!a.ready) {
    set_error("cm_adjoint: no forward on record (cm_forward_ckpt first)");
    return CM_ERR_STATE;
  }
  // the recompute re-runs the substep with the body's current flags: they must
  // still describe the differentiable substep
  if ((rc = adj_check(b, "cm_adjoint"))) return rc;
  if (!dl_dx) {
    set_error("cm_adjoint: dl_dx is required");
    return CM_ERR_ARG;
  }
  const int n = b->n, I = a.iters;
  const bool coll = (b->prm.flags & CM_F_SELF_COLLISION) && b->t >= 2;
  if ((rc = upload_xyz(b, dl_dx, a.ab.lx)) || (rc = upload_xyz(b, dl_dv, a.ab.lv))) return rc;
  CM_HIP(hipMemsetAsync(a.ab.gks, 0, sizeof(double) * std::max(b->S, 1), b->stream));
  CM_HIP(hipMemsetAsync(a.ab.gkb, 0, sizeof(double) * std::max(b->B, 1), b->stream));
  const PbdBufs pb = pbd_bufs(b);
  const PbdParams pp = pbd_params(b);
  std::vector<int> ctl(2 + a.every);
  for (int seg = a.nck - 1; seg >= 0; --seg) {
    const int s0 = seg * a.every, len = std::min(a.every, a.nsub - s0);
    for (int attempt = 0;; ++attempt) {
      if ((rc = adj_ckpt(b, seg, false))) return rc;
      CM_HIP(hipMemsetAsync(a.ctl, 0, sizeof(int) * (2 + a.every), b->stream));
      for (int j = 0; j < len; ++j) {
        const TapeSlot ts{a.pk + (int64_t)j * (I + 1) * n, a.dpre + (int64_t)j * (n + 1), j};
        if ((rc = record_substep(b, nullptr, &ts))) return rc;
      }
      if (!coll) break;
      CM_HIP(hipMemcpyAsync(ctl.data(), a.ctl, sizeof(int) * (2 + len), hipMemcpyDeviceToHost,
                            b->stream));
      CM_HIP(hipStreamSynchronize(b->stream));
      if (!ctl[1]) break;
      if (attempt > 0) {
        set_error("cm_adjoint: contact tape overflow");
        return CM_ERR_ALLOC;
      }
      if ((rc = adj_grow_arena(b, ctl[0]))) return rc;
    }
    for (int j = len - 1; j >= 0; --j) {
      const float4* pk = a.pk + (int64_t)j * (I + 1) * n;
      rc = launch_adj_substep(pb, pp, a.ab, pk, coll ? a.dpre + (int64_t)j * (n + 1) : nullptr,
                              coll ? a.arena + ctl[2 + j] : nullptr, b->stream);
      if (rc) return rc;
    }
  }
  if ((rc = adj_ckpt(b, a.nck, false))) return rc;  // back to the state after the forward
  std::vector<float4> hx(n), hv(n);
  std::vector<double> gs(b->S), gb(b->B);
  CM_HIP(hipMemcpyAsync(hx.data(), a.ab.lx, sizeof(float4) * n, hipMemcpyDeviceToHost, b->stream));
  CM_HIP(hipMemcpyAsync(hv.data(), a.ab.lv, sizeof(float4) * n, hipMemcpyDeviceToHost, b->stream));
  if (b->S)
    CM_HIP(hipMemcpyAsync(gs.data(), a.ab.gks, sizeof(double) * b->S, hipMemcpyDeviceToHost,
                          b->stream));
  if (b->B)
    CM_HIP(hipMemcpyAsync(gb.data(), a.ab.gkb, sizeof(double) * b->B, hipMemcpyDeviceToHost,
                          b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  for (int i = 0; i < n; ++i) {
    if (dl_dx0) {
      dl_dx0[3 * i] = hx[i].x;
      dl_dx0[3 * i + 1] = hx[i].y;
      dl_dx0[3 * i + 2] = hx[i].z;
    }
    if (dl_dv0) {
      dl_dv0[3 * i] = hv[i].x;
      dl_dv0[3 * i + 1] = hv[i].y;
      dl_dv0[3 * i + 2] = hv[i].z;
    }
  }
  for (int i = 0; dl_dk_dist && i < b->S; ++i)
    dl_dk_dist[i] = (float)(gs[i] * dkp_dk(b->topo.k_dist[i], I));
  for (int i = 0; dl_dk_bend && i < b->B; ++i)
    dl_dk_bend[i] = (float)(gb[i] * dkp_dk(b->topo.k_bend[i], I));
  return CM_OK;
}

}  // extern "C"

// ---- launch-shape autotuner (no reference counterpart) ---------------------------
namespace {

// frees a buffer of the body's allocation list
void free_alloc(cm_body* b, void* ptr) {
  for (size_t k = 0; k < b->allocs.size(); ++k)
    if (b->allocs[k] == ptr) {
      hipFree(ptr);
      b->allocs.erase(b->allocs.begin() + k);
      return;
    }
}

void drop_graph(cm_body* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Jacobi tiles of tile_v vertices: rebuilt on the host and uploaded in place of the
// current ones (the substep graph is dropped: it holds the old buffers)
int set_tiles(cm_body* b, int tile_v) {
  HostTopo& h = b->topo;
  build_tiles(b->n, h, tile_v);
  if (h.tstart.empty()) {
    set_error("cm_autotune: the mesh has no tiled Jacobi layout");
    return CM_ERR_STATE;
  }
  CM_HIP(hipStreamSynchronize(b->stream));
  drop_graph(b);
  free_alloc(b, b->tstart);
  free_alloc(b, b->eprefix);
  free_alloc(b, b->entries);
  free_alloc(b, b->adj.ecid);
  // until the new tiles are in place the body uses the slot-parallel Jacobi form,
  // which needs no tiles (and stays so if an upload fails)
  b->ntiles = 0;
  b->tstart = b->eprefix = nullptr;
  b->entries = nullptr;
  b->adj.ecid = nullptr;
  b->adj.ab.ecid = nullptr;
  b->adj.ready = false;
  int rc;
  if ((rc = upload(b, &b->tstart, h.tstart)) || (rc = upload(b, &b->eprefix, h.eprefix)) ||
      (rc = upload(b, &b->entries, h.entries)) || (rc = upload(b, &b->adj.ecid, h.ecid)))
    return rc;
  b->adj.ab.ecid = b->adj.ecid;
  b->ntiles = (int)h.tstart.size() - 1;
  b->tile_v = tile_v;
  return CM_OK;
}

}  // namespace

extern "C" {

int cm_autotune(cm_body* b, int32_t substeps, int32_t* chosen) {
  int rc = body_check(b);
  if (rc) return rc;
  if (substeps < 1) {
    set_error("cm_autotune: substeps must be positive");
    return CM_ERR_ARG;
  }
  const size_t n = (size_t)b->n, t = (size_t)b->t;
  CM_HIP(hipStreamSynchronize(b->stream));
  drop_graph(b);
  // every trial starts from the current state
  float4 *sx = nullptr, *sv = nullptr, *sp = nullptr, *snrm = nullptr;
  int4* stri = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    hipFree(sx);
    hipFree(sv);
    hipFree(sp);
    hipFree(snrm);
    hipFree(stri);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
  };
  if (hipMalloc(&sx, sizeof(float4) * n) != hipSuccess ||
      hipMalloc(&sv, sizeof(float4) * n) != hipSuccess ||
      hipMalloc(&sp, sizeof(float4) * n) != hipSuccess ||
      hipMalloc(&snrm, sizeof(float4) * n) != hipSuccess ||
      hipMalloc(&stri, sizeof(int4) * t) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    set_error("cm_autotune: allocation failed");
    return CM_ERR_ALLOC;
  }
  auto copy_state = [&](bool save) -> hipError_t {
    hipError_t e = hipSuccess;
    auto cp = [&](void* keep, void* live, size_t bytes) {
      if (e == hipSuccess)
        e = save ? hipMemcpyAsync(keep, live, bytes, hipMemcpyDeviceToDevice, b->stream)
                 : hipMemcpyAsync(live, keep, bytes, hipMemcpyDeviceToDevice, b->stream);
    };
    cp(sx, b->x, sizeof(float4) * n);
    cp(sv, b->v, sizeof(float4) * n);
    cp(sp, b->p, sizeof(float4) * n);
    cp(snrm, b->vnrm, sizeof(float4) * n);  // carried vertex normals
    cp(stri, b->tri, sizeof(int4) * t);
    return e;
  };
  auto trial = [&](float* ms) -> int {
    if (copy_state(false) != hipSuccess) {
      set_error("cm_autotune: state copy failed");
      return CM_ERR_HIP;
    }
    CM_HIP(hipEventRecord(e0, b->stream));
    for (int s = 0; s < substeps; ++s) {
      const int r = record_substep(b);
      if (r) return r;
    }
    CM_HIP(hipEventRecord(e1, b->stream));
    CM_HIP(hipEventSynchronize(e1));
    CM_HIP(hipEventElapsedTime(ms, e0, e1));
    return CM_OK;
  };
  auto fail = [&](int code) {
    cleanup();
    return code;
  };
  if (copy_state(true) != hipSuccess) return fail(CM_ERR_HIP);
  // the trial substeps must not add to the body's accumulated warnings
  uint32_t warn_keep[4];
  CM_HIP(hipMemcpyAsync(warn_keep, b->warn, sizeof(warn_keep), hipMemcpyDeviceToHost, b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  if ((rc = trial(&ms))) return fail(rc);  // warm-up
  // one parameter at a time, each over its candidates, keeping the fastest
  if (b->ntiles > 0) {
    const int cand[3] = {64, 128, 256};
    int best = b->tile_v;
    float best_ms = 1e30f;
    for (int c : cand) {
      if ((rc = set_tiles(b, c)) || (rc = trial(&ms))) return fail(rc);
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
    if ((rc = set_tiles(b, best))) return fail(rc);
  }
  {
    const int cand[4] = {1, 2, 4, 8};
    int best = b->lb.trav_chunk;
    float best_ms = 1e30f;
    for (int c : cand) {
      b->lb.trav_chunk = c;
      if ((rc = trial(&ms))) return fail(rc);
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
    b->lb.trav_chunk = best;
  }
  if (b->lb.trav_budget > 0) {  // the two-pass traversal is on
    const int cand[4] = {128, 160, 200, 240};
    int best = b->lb.trav_budget;
    float best_ms = 1e30f;
    for (int c : cand) {
      b->lb.trav_budget = c;
      if ((rc = trial(&ms))) return fail(rc);
      if (ms < best_ms) { best_ms = ms; best = c; }
    }
    b->lb.trav_budget = best;
  }
  if (copy_state(false) != hipSuccess) return fail(CM_ERR_HIP);
  CM_HIP(hipMemcpyAsync(b->warn, warn_keep, sizeof(warn_keep), hipMemcpyHostToDevice, b->stream));
  CM_HIP(hipStreamSynchronize(b->stream));
  cleanup();
  if (chosen) {
    chosen[0] = b->tile_v;
    chosen[1] = b->lb.trav_chunk;
    chosen[2] = b->lb.trav_budget;
  }
  return CM_OK;
}

}  // extern "C"
