// Internal declarations of libclothmi: launch helpers shared between the
// kernel translation units and the body orchestration.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "../../include/clothmi.h"
#include "cm_math.hpp"

namespace cm {

void set_error(const std::string& msg);

#define CM_HIP(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::cm::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));              \
      return CM_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)

#define CM_LAUNCH_CHECK()                                                              \
  do {                                                                                 \
    hipError_t e_ = hipGetLastError();                                                 \
    if (e_ != hipSuccess) {                                                            \
      ::cm::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));         \
      return CM_ERR_HIP;                                                               \
    }                                                                                  \
  } while (0)

constexpr int kBlock = 256;  // 4 wave64s
constexpr int kXcds = 8;     // MI355X: workgroup i runs on XCD i mod 8

// Logical block for hardware block b of n so that each XCD gets a contiguous range
// of logical blocks (and so of data) instead of every eighth one. A bijection on
// [0, n).
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int q = n / kXcds, r = n % kXcds;  // XCD x holds q + (x < r) blocks
  const int x = b % kXcds, k = b / kXcds;
  return x * q + min(x, r) + k;
}
inline int nblocks(int64_t n, int block = kBlock) {
  int64_t b = (n + block - 1) / block;
  return (int)(b < 1 ? 1 : b);
}
inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// ---- primitives (k_prims.hip) --------------------------------------------
size_t scan_temp_bytes(int n_max);
// zero a scan temp before its first use (scans leave it zeroed for the next one)
int scan_temp_clear(void* temp, int n_max, hipStream_t s);
// exclusive prefix sum of in[0, n) into out[0, n), n = d_n ? min(*d_n, n_max) : n_max;
// *d_total (if non-null) receives the sum. out may alias in.
int scan_exclusive(const int* in, int* out, int n_max, const int* d_n, int* d_total, void* temp,
                   hipStream_t s);
// Exclusive scan of per-tile sums (tiles of per_tile items; ceil(*d_items / per_tile)
// of them, at most ntile_max; all ntile_max when d_items is null). packed: each sum
// holds two counts below 2^16, scanned into pre[2i] (low) and pre[2i+1] (high);
// else pre[i]. The totals (below 2^31), clamped to cap0 / cap1, go to *d_tot0 /
// *d_tot1. temp: a scan temp (scan_temp_bytes(n) with n >= ntile_max).
int scan_tiles(const int* d_items, int per_tile, int ntile_max, const int* sums, int* pre,
               int packed, int64_t cap0, int64_t cap1, int* d_tot0, int* d_tot1, void* temp,
               hipStream_t s);
size_t radix_sort_temp_bytes(int n);
int radix_sort(uint64_t* keys, uint32_t* vals, int n, int key_bits, void* temp, hipStream_t s);
int radix_sort_clear(void* temp, int n, hipStream_t s);  // before the first sort on a temp
// Stable LSD sort of the first min(*d_n, n_max) 32-bit keys (key_bits <= 32) with
// their values, with a launch shape independent of the count (graph-capturable). The
// result ends in keys_alt / vals_alt when *in_alt is set (an odd number of passes),
// else in keys / vals. temp: radix_sort_dev_temp_bytes(), cleared once before its
// first sort by radix_sort_dev_clear.
size_t radix_sort_dev_temp_bytes();
int radix_sort_dev_clear(void* temp, hipStream_t s);
int radix_sort_dev(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                   int n_max, const int* d_n, int key_bits, void* temp, int* in_alt,
                   hipStream_t s);
int fill_i32(int* p, int v, int64_t n, hipStream_t s);
int fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);

// ---- LBVH (k_lbvh.hip) -----------------------------------------------------
// ints per traversal hand-over record (leaf, pairs, depth, node, cap, visits, stack)
constexpr int kTravRec = 72;
// ints per leaf for the traversal stack entries below its LDS part
constexpr int kTravSpill = 47;

struct LbvhBufs {
  int t = 0;
  float4 *amin = nullptr, *amax = nullptr;  // 2T-1 (internal nodes first)
  int *lefts = nullptr, *rights = nullptr, *parents = nullptr, *rmll = nullptr, *rmlr = nullptr;
  int *claim_cnt = nullptr, *claim_hi = nullptr;  // 2T-1: claimants, highest claimant
  int2* ranges = nullptr;                           // T-1 key ranges
  // refit: per node 1 = finished by its chunk, 0 = deferred (top_index = list slot)
  int *state = nullptr, *top_list = nullptr, *top_index = nullptr, *ntop = nullptr;
  float4 *tmin = nullptr, *tmax = nullptr;  // deferred-node scratch (large trees)
  int *trm = nullptr, *tdone = nullptr;
  float4* nodes = nullptr;  // 4 float4 per internal node (children, rightmost leaves, boxes)
  int trav_chunk = 4;       // traversal: tiles per chunk dealt to one XCD (launch shape only)
  // traversal: visits a leaf makes in the first pass before a wave takes it over (0:
  // one pass); records of kTravRec ints for up to rcap such leaves; hctl[0] their
  // count, zero between substeps (the pair compaction clears it). The pair lists are
  // the same with any budget.
  int trav_budget = 0;
  int* trec = nullptr;
  int rcap = 0;
  int* hctl = nullptr;
  int* spill = nullptr;  // kTravSpill ints per leaf: deep traversal stack entries
  uint32_t* warn = nullptr;
  // batch of cloths (ncloth == 0: one cloth): cloth c owns triangle slots (and, after
  // the sort, leaves) [tstart[c], tstart[c + 1]); its Morton keys carry c above bit 30
  int ncloth = 0;
  const int* tstart = nullptr;      // ncloth + 1
  const int* slot_cloth = nullptr;  // T: cloth of each triangle slot / leaf
  int* croot = nullptr;             // ncloth: internal node spanning the cloth (-1: none)
  const int* ccap = nullptr;        // ncloth: per-leaf pair cap of the cloth
  uint32_t* cbox = nullptr;         // 6 * ncloth: centroid box, order-preserving encoding
};
size_t lbvh_build_temp_bytes(int t);
size_t carve_build(int t, void* d_temp, LbvhBufs* b);  // b == nullptr: size only
size_t lbvh_pairs_temp_bytes(int t, int cap);
int launch_tri_update(int n, int t, const float4* x, const float4* p, const int4* tri, float th,
                      int bits, int normalize, float4* amin, float4* amax, uint64_t* keys,
                      uint32_t* vals, float4* centroid_tmp, float4* box_tmp, hipStream_t s);
// batch: swept AABBs, per-cloth centroid boxes and keys (cloth << 30 | Morton30)
int launch_tri_update_batch(const LbvhBufs& b, const float4* x, const float4* p, const int4* tri,
                            float th, uint64_t* keys, uint32_t* vals, float4* centroid_tmp,
                            hipStream_t s);
int launch_build(const LbvhBufs& b, const uint64_t* keys, int midpoint, hipStream_t s);
int launch_refit(const LbvhBufs& b, hipStream_t s);
int launch_pack_nodes(const LbvhBufs& b, hipStream_t s);
// pairs: traversal into chunk (T*cap int2) + per-leaf counts, then compaction.
int launch_pairs(const LbvhBufs& b, int cap, int2* chunk, int* counts, int* offsets,
                 int2* pairs, int* d_npairs, void* scan_temp, hipStream_t s);
// the traversal kernel alone on the current tree (it rewrites the same pair chunk)
int launch_traverse_only(const LbvhBufs& b, int cap, int2* chunk, int* counts, int* offsets,
                         hipStream_t s);
int launch_gather_tris(int t, const uint32_t* perm, const int4* tri_in, int4* tri_out,
                       int fix_leaves, float4* amin, float4* amax, float4* tmp_min,
                       float4* tmp_max, hipStream_t s);
int launch_copy_i4(int64_t n, const int4* src, int4* dst, hipStream_t s);

// ---- narrow phase (k_narrow.hip) ------------------------------------------
// A contact record in vertex order, self-contained for the Jacobi kernel: the
// contact's four vertex ids (bits 29-30 of ids.x hold this record's role, bit 31 is
// set for a VF contact) and the contact weights.
struct alignas(16) RecPayload {
  int4 ids;
  float4 w;
};

struct NarrowBufs {
  int n = 0, t = 0;
  int64_t pair_cap = 0;       // capacity of pairs (T*cap)
  int64_t vf_cap = 0, ee_cap = 0;          // candidate capacities
  int64_t vf_out_cap = 0, ee_out_cap = 0;  // contact (CCD survivor) capacities
  const int2* pairs = nullptr;
  const int* d_npairs = nullptr;
  uint16_t* mask = nullptr;   // per pair
  int* tile_sum = nullptr;  // per tile of 256 pairs / candidates: packed counts
  int* tile_pre = nullptr;  // per tile: exclusive prefixes of the two counts
  int4 *cand_vf = nullptr, *cand_ee = nullptr;
  int *d_ncand_vf = nullptr, *d_ncand_ee = nullptr;
  uint8_t* hit_flag = nullptr;  // per candidate (VF then EE): near, then hit per near index
  int* near_ids = nullptr;       // candidates passing the box-gap rejection, in order
  int* d_nnear = nullptr;
  // CCD passes, per candidate (VF then EE, vf_cap + ee_cap), written for near
  // candidates only: plan word, critical points, and the secant root of each of its
  // (up to three) intervals
  int* cplan = nullptr;
  float2* ccp = nullptr;
  float* croot = nullptr;  // 3 * (vf_cap + ee_cap)
  int* d_ntotal = nullptr;
  int4 *vf_v = nullptr, *ee_v = nullptr;      // compacted contacts
  float4 *vf_w = nullptr, *vf_tn = nullptr, *ee_w = nullptr, *ee_tn = nullptr;
  int *d_nvf = nullptr, *d_nee = nullptr;
  // per-vertex dynamic record CSR
  int *dcount = nullptr, *dprefix = nullptr, *dcursor = nullptr;  // n + 1, n + 1, n
  uint32_t* drec = nullptr;  // 4 * (vf_out_cap + ee_out_cap)
  uint32_t* dtmp = nullptr;  // same size: staging for very long record lists
  // same size: record payloads in drec order; between the solver's last read and
  // k_rec_place it holds the CCD hits' weights, time and normal per near index
  RecPayload* rec = nullptr;
  int* dlong = nullptr;      // n: vertices whose record lists are ordered by a workgroup
  // sorted_csr: the CSR comes from one radix sort of the records by vertex id
  // (else from per-vertex counts and atomic placement). skey: record keys then
  // alternate keys (2 * 4 * (vf_out_cap + ee_out_cap)), sort_temp:
  // radix_sort_dev_temp_bytes()
  bool sorted_csr = false;
  uint32_t* skey = nullptr;
  void* sort_temp = nullptr;
  int* d_nrec = nullptr;     // records of the substep (sorted CSR)
  uint32_t* warn = nullptr;
  void* scan_temp = nullptr;
};
// dcd: one proximity test per candidate at x instead of the CCD (CM_F_DCD)
// the CCD plan pass alone over the current candidate lists (kernel timing)
int launch_ccd_plan_only(const NarrowBufs& nb, const float4* x, const float4* v, float th,
                         float dt, int dcd, hipStream_t s);
int launch_narrow(const NarrowBufs& nb, const int4* tri, const float4* x, const float4* p,
                  const float4* v, float th, float dt, int ee_on, int dcd, hipStream_t s);
int launch_dynamic_csr(const NarrowBufs& nb, hipStream_t s);

// ---- PBD (k_pbd.hip) -------------------------------------------------------
// Vertex tiles of the Jacobi kernel: at most kTileV consecutive vertices holding at
// most kTileSlots constraint slots. Each tile lists the constraints touching it once.
constexpr int kTileV = 128;
constexpr int kTileSlots = 3072;
struct alignas(16) TileEntry {
  int4 ids;            // vertex ids; ids.z < 0 marks a distance constraint (ids.x, ids.y)
  float a, b;          // distance: l0, kp; bend: kp, phi0
  uint32_t s01, s23;   // tile-local slot of roles 0..3, 16 bits each, 0xffff = not in tile
};
static_assert(sizeof(TileEntry) == 32, "TileEntry is two 16-byte loads");

struct PbdBufs {
  int n = 0, S = 0, B = 0, F = 0;
  float4 *x = nullptr, *p = nullptr, *p2 = nullptr, *v = nullptr;  // .w of x/p = inverse mass
  const uint8_t* pinned = nullptr;
  const int2* sid = nullptr;
  const float2* s_l0kp = nullptr;
  const int4* bid = nullptr;
  const float2* b_kpphi0 = nullptr;
  const int* vprefix = nullptr;
  const uint32_t* vslot = nullptr;
  int ntiles = 0;  // 0: slot-parallel solve (a vertex has more than kTileSlots slots)
  const int* tstart = nullptr;
  const int* eprefix = nullptr;
  const TileEntry* entries = nullptr;
  const int* lra_fixed = nullptr;
  const float* lra_l0 = nullptr;
  uint8_t* lra_flag = nullptr;
  float4* ground = nullptr;  // qc.xyz, flag
  const float* mass = nullptr;
  float* damp_partials = nullptr;  // per block sums
  float* damp_state = nullptr;     // omega, xcm, vcm
};
struct PbdParams {
  float dt, th, gh;
  f3 g;
  int iters;
  uint32_t flags, ref_compat;
  f3 wmin, wmax;
  f3 sc;
  float sr;
  float kg, kd, kf;
};
int launch_predict(const PbdBufs& b, const PbdParams& pp, hipStream_t s);
int launch_solver(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, hipStream_t s);
int launch_solve_only(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, int iters,
                      hipStream_t s);
int launch_integrate(const PbdBufs& b, const PbdParams& pp, hipStream_t s);
int launch_integrate_from(const PbdBufs& b, const PbdParams& pp, const float4* pfin,
                          hipStream_t s);
int launch_friction(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, hipStream_t s);
int launch_revert_toc(const PbdBufs& b, const NarrowBufs& nb, hipStream_t s);
int launch_vertex_normals(int n, const int* vt_prefix, const int4* vt_tri, const float4* pos,
                          float4* nrm, int carry, int unit, hipStream_t s);
// the solver stage with every iterate kept: pk[k * n, (k + 1) * n) = positions
// entering iteration k, pk[iters * n ...) the result
int launch_solver_rec(const PbdBufs& b, const PbdParams& pp, const NarrowBufs& nb, float4* pk,
                      hipStream_t s);

// ---- adjoint (k_pbd.hip) ---------------------------------------------------
struct AdjBufs {
  float4 *lx = nullptr, *lv = nullptr;  // dL/dx, dL/dv (n)
  float4 *lp = nullptr, *la = nullptr;  // dL/dp of the iterates (n each, alternating)
  float4* av = nullptr;                 // lp / count per vertex (n)
  int* dcnt = nullptr;                  // non-zero contact records per iteration and vertex (iters * n)
  double *gks = nullptr, *gkb = nullptr;  // dL/dkp per distance / bend constraint
  const int* ecid = nullptr;              // per tile entry: constraint index (bit 31: bend)
};
// The contact set of substep `slot`: the per-vertex records (dprefix[n] of them) are
// appended to arena; ctl[0] = records used, ctl[1] = overflow, ctl[2 + slot] = first.
int launch_tape_records(int n, const int* dprefix, const RecPayload* rec, RecPayload* arena,
                        int64_t cap, int* ctl, int slot, hipStream_t s);
// One substep backwards: (lx, lv) = dL/d(x, v) after the substep in, before it out;
// dL/dkp added to gks / gkb. pk: the substep's iterates (launch_solver_rec);
// dprefix / rec: its contact records (null: none).
int launch_adj_substep(const PbdBufs& b, const PbdParams& pp, const AdjBufs& ab, const float4* pk,
                       const int* dprefix, const RecPayload* rec, hipStream_t s);

}  // namespace cm
