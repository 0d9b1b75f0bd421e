/*
 * clothmi.h — C ABI of the MI355X-native PBD cloth substep with Morton/LBVH
 * self-collision.
 *
 * The entry points replace the reference's simulation-object interface and its
 * collision-method free functions (all paths relative to
 * LicentaV2/LicentaV2/ in Tureyz/Licenta):
 *
 *   cm_create          <- Physics::CudaDeformableBody::CudaDeformableBody
 *                         (Physics/CudaDeformableBody.h:23-24,
 *                          Physics/CudaDeformableBody.cpp:30-285)
 *   cm_step            <- Physics::CudaDeformableBody::FixedUpdate, one call per
 *                         substep (Physics/CudaDeformableBody.cpp:550-627)
 *   cm_read_positions  <- Physics::CudaDeformableBody::Update / FinalVertUpdate
 *                         (Physics/CudaDeformableBody.cpp:629-638, 473-495)
 *   cm_read_normals    <- DeformableUtils::UpdateVertexNormals + FinalVertUpdate
 *                         (Physics/CudaDeformableUtils.cu:190-249,317-344)
 *   cm_read_vertices   <- FinalVertUpdate's VertexFormat copy to the host
 *                         (Physics/CudaDeformableBody.cpp:473-495)
 *   cm_set_sphere      <- Physics::CudaDeformableBody::SetSpherePos
 *                         (Physics/CudaDeformableBody.cpp:640-643)
 *   cm_destroy         <- Physics::CudaDeformableBody::~CudaDeformableBody
 *                         (Physics/CudaDeformableBody.cpp:317-320)
 *   cm_query_pairs     <- CudaBVH::GetAABBCollisions + CubWrap::SelectFlagged
 *                         (Collision/CudaBVH.cu:9-113, Collision/CudaCollisionDetection.cu:17)
 *   cm_query_contacts  <- DeformableUtils::CCDTriangleTests output vfContacts/eeContacts
 *                         (Collision/CudaCollisionDetection.cu:223-283)
 *   cm_tri_update      <- DeformableUtils::_UpdateTrianglesContinuous
 *                         (Physics/CudaDeformableUtils.cu:120-151)
 *   cm_lbvh_sort       <- DeformableUtils::SortMortons -> CubWrap::SortByKey
 *                         (Physics/CudaDeformableUtils.cu:346-349, Core/CubWrappers.cuh:141-156)
 *   cm_lbvh_build      <- CudaBVH::GenerateBVH2 (Collision/CudaBVH.cu:152-242)
 *   cm_lbvh_refit      <- CudaBVH::ComputeTreeAABBs (Collision/CudaBVH.cu:115-149)
 *   cm_lbvh_pairs      <- CudaBVH::GetAABBCollisions (Collision/CudaBVH.cu:9-113)
 *
 * Conventions: every function returns 0 (CM_OK) or a negative CM_ERR_* code and
 * never exits or aborts; cm_last_error() gives the message of the last failure
 * on the calling thread. Host arrays are copied during the call and never
 * retained. Device pointers (d_*) belong to the caller and must be valid for
 * the given stream; the unit-level cm_tri_update/cm_lbvh_* calls are
 * stream-ordered and return without synchronising. A cm_body owns its device
 * memory and one HIP stream; one body must not be used from two threads at
 * once, separate bodies may.
 */
#ifndef CLOTHMI_H
#define CLOTHMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CM_OK 0
#define CM_ERR_ARG (-1)        /* invalid argument / size                    */
#define CM_ERR_HIP (-2)        /* HIP runtime error                          */
#define CM_ERR_ALLOC (-3)      /* device allocation failed                   */
#define CM_ERR_STATE (-4)      /* call not valid in the current state        */
#define CM_ERR_NO_DEVICE (-5)  /* no usable GPU                              */

/* warning bits accumulated on a body (cm_get_warnings) */
#define CM_WARN_STACK_OVERFLOW (1u << 0)  /* traversal stack (64) full: right subtree skipped */
#define CM_WARN_TRAVERSAL_LIMIT (1u << 1) /* traversal visit bound reached for some leaf     */
#define CM_WARN_BROKEN_TREE (1u << 2)     /* midpoint split on tied codes made a node shared */
#define CM_WARN_PAIR_CAP (1u << 3)        /* some leaf dropped pairs past the per-leaf cap    */
#define CM_WARN_CANDIDATE_CAP (1u << 4)   /* candidate/contact buffer full: list truncated    */
#define CM_WARN_REFIT_LIMIT (1u << 5)     /* refit climb bound reached                         */

/* ---- feature flags (cm_params.flags) ------------------------------------ */
#define CM_F_SELF_COLLISION (1u << 0) /* Morton/LBVH/CCD contacts every substep */
#define CM_F_GROUND (1u << 1)         /* ground plane at ground_height          */
#define CM_F_WORLD_CLAMP (1u << 2)    /* clamp predicted positions to the box   */
#define CM_F_DAMPING (1u << 3)        /* global + rigid-mode velocity damping   */
#define CM_F_LRA (1u << 4)            /* long-range attachments to pinned verts */
#define CM_F_SPHERE (1u << 5)         /* static/moving analytic sphere          */
#define CM_F_EE_CONTACTS (1u << 6)    /* edge-edge contacts in the solver       */
#define CM_F_FRICTION (1u << 7)       /* velocity friction after integration    */
/* (ground contacts, then the substep's VF / EE contacts; coefficient k_friction;
 * Physics/CudaPBD.cu:770-833). Not with cm_create_batch or cm_forward_ckpt. */
/* Evaluate the dihedral bend (Physics/CudaPBD.cu:1024-1147) in FP64 from the FP32
 * positions and round the corrections to FP32. On absolute positions the
 * reference formula cancels catastrophically in FP32 wherever two vertices are
 * nearly collinear with the origin; FP64 gives the formula's exact value there. */
#define CM_F_BEND_FP64 (1u << 8)
/* Vertex normals every substep, as the reference's FixedUpdate computes them
 * (DeformableUtils::UpdateVertexNormals, Physics/CudaDeformableUtils.cu:190-249):
 * the unit face normals of the predicted positions added to the previous
 * substep's normal, then normalised. Without it cm_read_normals computes the
 * normals of the current positions when called. */
#define CM_F_VERTEX_NORMALS (1u << 9)
/* Time-of-contact rollback before the solver: each vertex in contacts moves to
 * (1 - t) x + t p with t the smallest time of its contacts (CudaPBD::RevertToTOC,
 * Physics/CudaPBD.cu:727-768). The reference's call is commented out
 * (Physics/CudaDeformableBody.cpp:583), so it is off by default. Not with
 * cm_forward_ckpt. */
#define CM_F_REVERT_TOC (1u << 10)
/* Discrete instead of continuous narrow phase: each VF / EE candidate is tested
 * once at the substep's start positions x, with the same proximity tests and no
 * coplanarity roots (CudaPrimitiveTests::DCDTriangleTests, Collision/
 * CudaPrimitiveTests.cu:7-94, reached from HandleCollisionsDiscrete, whose call the
 * reference's FixedUpdate comments out: Physics/CudaDeformableBody.cpp:419-439,
 * 611). A contact's time is what the test leaves: 0, or -1 where the reference's
 * TestEE does not set it. */
#define CM_F_DCD (1u << 11)

/* ---- reference-compatibility bits (cm_params.ref_compat) ----------------- */
/* Leaf AABBs stay in pre-sort triangle slots after the Morton sort
 * (Physics/CudaDeformableUtils.cu:146-147 with Core/CubWrappers.cuh:141-156). */
#define CM_RC_STALE_LEAF_AABB (1u << 0)
/* Dihedral bend on absolute positions with rest angle pi
 * (Physics/CudaPBD.cu:1040-1046, 366-368); off = p1-relative Mueller bend with
 * the initial dihedral angle as rest angle. */
#define CM_RC_BEND_ABS (1u << 1)
/* FindSplit returns (first+last)>>1 when the end codes are equal
 * (Collision/CudaBVH.cu:219-220); off = index-augmented Karras split. */
#define CM_RC_MIDPOINT_SPLIT (1u << 2)
/* The ground step of the friction pass reaches only vertices below
 * 255 * ceil(N/256): the reference launches it with 255-thread blocks
 * (Physics/CudaPBD.cu:778); off = every vertex. */
#define CM_RC_RIGID_FRICTION_TAIL (1u << 3)
#define CM_RC_ALL \
  (CM_RC_STALE_LEAF_AABB | CM_RC_BEND_ABS | CM_RC_MIDPOINT_SPLIT | CM_RC_RIGID_FRICTION_TAIL)

/* POD replacement for Physics::ClothParams (Physics/Structs.h:23-61). */
typedef struct cm_params {
  float timestep;          /* substep dt in seconds                              */
  float gravity[3];        /* m/s^2                                              */
  float k_stretch;         /* distance stiffness when k_dist is NULL             */
  float k_bend;            /* bend stiffness when k_bend is NULL                 */
  float thickness;         /* collision thickness (no override, unlike :86)     */
  int32_t solver_iterations;
  float ground_height;     /* ground plane y (reference: worldMin.y)             */
  float world_min[3];      /* clamp box (reference: WORLD_MIN/WORLD_MAX)         */
  float world_max[3];
  float global_vel_damp;   /* kg: v *= (1-kg)                                    */
  float k_damp;            /* rigid-mode damping coefficient                     */
  float strain_limit;      /* LRA rest-length slack                              */
  float k_friction;
  float pinned_mass;       /* mass used for pinned vertices in damping (ref 1000)*/
  float sphere_center[3];
  float sphere_radius;
  int32_t morton_bits;     /* 63 (reference) or 30                               */
  int32_t morton_normalize;/* 0: raw coords clamped to [0,1]; 1: centroid box    */
  int32_t pair_cap;        /* per-leaf triangle-pair cap (reference 64)          */
  uint32_t flags;          /* CM_F_*                                             */
  uint32_t ref_compat;     /* CM_RC_*                                            */
  /* capacity overrides, 0 = default: [0] / [1] VF / EE candidates per triangle
   * (128 / 192), [2] CCD survivors per triangle (16), [3] contact records per
   * substep that cm_adjoint's tape holds at first (max(N, 65536); it grows when a
   * segment needs more); [4..6] unused */
  int32_t reserved[7];
} cm_params;

/* Same fields as Physics::PrimitiveContact (Physics/Structs.h:106-129), 48 B. */
typedef struct cm_contact {
  int32_t v[4];
  float w[4];
  float t;
  float n[3];
} cm_contact;

typedef struct cm_body cm_body;

/* Internal buffers readable with cm_get_buffer (element type in brackets). */
enum cm_buffer {
  CM_BUF_X = 0,          /* [float4 N] carried positions (m_dPrevPositions)      */
  CM_BUF_P = 1,          /* [float4 N] predicted positions (m_dPositions)        */
  CM_BUF_V = 2,          /* [float4 N] velocities                                */
  CM_BUF_TRI = 3,        /* [int4 T] triangles in current (sorted) order, .w = feature bits */
  CM_BUF_KEYS = 4,       /* [uint64 T] sorted Morton keys                        */
  CM_BUF_PERM = 5,       /* [uint32 T] sort permutation (pre-sort slot of sorted i) */
  CM_BUF_LEFTS = 6,      /* [int32 T-1]                                          */
  CM_BUF_RIGHTS = 7,     /* [int32 T-1]                                          */
  CM_BUF_PARENTS = 8,    /* [int32 2T-1]                                         */
  CM_BUF_RMLL = 9,       /* [int32 T-1]                                          */
  CM_BUF_RMLR = 10,      /* [int32 T-1]                                          */
  CM_BUF_AABB_MIN = 11,  /* [float4 2T-1] internal nodes then leaves             */
  CM_BUF_AABB_MAX = 12,  /* [float4 2T-1]                                        */
  CM_BUF_PAIRS = 13,     /* [int2 P] compacted triangle pairs                    */
  CM_BUF_VF = 14,        /* [cm_contact C_vf]                                    */
  CM_BUF_EE = 15,        /* [cm_contact C_ee]                                    */
  CM_BUF_UNUSED_KEYS_PRESORT = 16
};

void cm_default_params(cm_params* params);
const char* cm_last_error(void);

/* Body lifecycle. xyz: 3N floats; tri: 3T vertex ids; pinned: N bytes or NULL
 * (NULL: pinned where inv_mass == 0); inv_mass: N floats; dist_pairs: 2*n_dist
 * vertex ids or NULL (NULL: one distance constraint per unique mesh edge, as
 * Physics/CudaPBD.cu:111-142); k_dist: per distance constraint or NULL;
 * k_bend: per bend constraint or NULL. */
int cm_create(const cm_params* params, const float* xyz, int32_t n, const uint32_t* tri,
              int32_t t, const uint8_t* pinned, const float* inv_mass, const int32_t* dist_pairs,
              int32_t n_dist, const float* k_dist, const float* k_bend, cm_body** out);
int cm_destroy(cm_body* body);

/* A batch of independent cloths stepped as one body (no counterpart in the
 * reference, whose bodies are stepped one at a time; each cloth follows
 * cm_create's semantics on its own, Physics/CudaDeformableBody.cpp:550-627).
 * Cloth c has cloth_n[c] vertices and cloth_t[c] triangles; xyz, pinned and
 * inv_mass hold all vertices cloth after cloth, tri all triangles with cloth-local
 * vertex ids; dist_pairs (NULL: mesh edges) holds cloth_ndist[c] cloth-local pairs
 * per cloth; k_stretch / k_bend give one stiffness per cloth (NULL: params). Each
 * cloth has its own LBVH (30-bit codes in the cloth's centroid box; cloths never
 * collide with each other). Positions and buffers are read back for all vertices,
 * cloth after cloth. Requires morton_bits 30 with morton_normalize, and no LRA,
 * rigid-mode damping or friction. */
int cm_create_batch(const cm_params* params, int32_t ncloth, const int32_t* cloth_n,
                    const int32_t* cloth_t, const int32_t* cloth_ndist, const float* xyz,
                    const uint32_t* tri, const uint8_t* pinned, const float* inv_mass,
                    const int32_t* dist_pairs, const float* k_stretch, const float* k_bend,
                    cm_body** out);

/* Runs `substeps` substeps on the body's stream (graph-replayed); does not sync. */
int cm_step(cm_body* body, int32_t substeps);
int cm_sync(cm_body* body);

int cm_read_positions(cm_body* body, float* xyz);   /* 3N, syncs */
int cm_read_velocities(cm_body* body, float* xyz);  /* 3N, syncs */
int cm_read_normals(cm_body* body, float* xyz);     /* 3N, syncs */
/* Rendering::VertexFormat readback (Physics/CudaDeformableBody.cpp:473-495,
 * Physics/CudaDeformableUtils.cu:317-344): verts holds N records of
 * CM_VERTEX_FLOATS floats laid out as Rendering/VertexFormat.h:24-54 (position 3,
 * color 4, normal 3, uv 2; 48 bytes). Writes each record's position (the carried
 * positions) and normal (as cm_read_normals) and leaves color and uv. Syncs. */
#define CM_VERTEX_FLOATS 12
#define CM_VERTEX_POSITION 0
#define CM_VERTEX_NORMAL 7
int cm_read_vertices(cm_body* body, float* verts);
/* Also drops a forward recorded by cm_forward_ckpt (its cm_adjoint then returns
 * CM_ERR_STATE). */
int cm_set_sphere(cm_body* body, const float center[3], float radius);
/* Replace the state. x, p, v: 3N floats each (p may be NULL: p = x);
 * tri: 4T ints (current triangle order with feature bits) or NULL. */
int cm_set_state(cm_body* body, const float* x, const float* p, const float* v,
                 const int32_t* tri);
/* counts[9] = n, t, s, bd, pairs, vf contacts, ee contacts, vf candidates,
 * ee candidates (the last substep's; candidates are the CCD inputs) */
int cm_get_counts(cm_body* body, int64_t* counts);
int cm_get_buffer(cm_body* body, int32_t which, void* host, int64_t capacity_bytes,
                  int64_t* count);
int cm_get_warnings(cm_body* body, uint32_t* warnings);
int cm_query_pairs(cm_body* body, int32_t* pairs, int32_t capacity, int32_t* n_pairs);
int cm_query_contacts(cm_body* body, cm_contact* vf, int32_t cap_vf, int32_t* n_vf,
                      cm_contact* ee, int32_t cap_ee, int32_t* n_ee);

/* Kernel timing for roofline reporting, in the spirit of the reference's stage
 * timers (Simulation/GPUBenchmark.cu:14-77): launches kernel `which`
 * (CM_KERNEL_*) `reps` times on the body's stream between two HIP events and
 * returns the average duration in milliseconds. CM_KERNEL_SOLVE re-runs Jacobi
 * iterations on the current predicted positions (it changes them). */
#define CM_KERNEL_SOLVE 0
#define CM_KERNEL_SUBSTEP 1
/* the BVH traversal of the current tree (rewrites the same per-leaf pair chunk;
 * the compacted pair list is unchanged) */
#define CM_KERNEL_TRAVERSE 2
/* the CCD plan pass (cubic, secant roots, near test) over the last substep's
 * candidates, with the current positions and velocities (rewrites its own plan
 * words and near flags; the contact lists are unchanged) */
#define CM_KERNEL_CCD_PLAN 3
int cm_time_kernel(cm_body* body, int32_t which, int32_t reps, float* avg_ms);

/* Per-stage timing of `substeps` substeps run eagerly on the body's stream with a
 * HIP event between stages, as the reference's GPUBenchmark stage timers
 * (Simulation/GPUBenchmark.cu:14-77, stage names at Physics/CudaDeformableBody.cpp:556-594).
 * ms[CM_NUM_STAGES] receives the summed milliseconds per stage. */
#define CM_STAGE_PREDICT 0    /* external forces, damping, Euler prediction   */
#define CM_STAGE_TRIANGLES 1  /* swept AABBs and Morton keys                  */
#define CM_STAGE_SORT 2       /* radix sort and triangle gather               */
#define CM_STAGE_TREE 3       /* Karras build and refit                       */
#define CM_STAGE_PAIRS 4      /* traversal and pair compaction                */
#define CM_STAGE_NARROW 5     /* candidates, CCD, contact compaction, records */
#define CM_STAGE_SOLVER 6     /* constraint set-up and Jacobi iterations      */
#define CM_STAGE_INTEGRATE 7  /* velocity update and world clamp              */
#define CM_NUM_STAGES 8
int cm_stage_times(cm_body* body, int32_t substeps, float* ms);

/* ---- differentiable rollout ------------------------------------------------
 * No counterpart in the reference (its substep is not differentiated); for callers
 * that fit per-constraint stiffness and initial velocities by gradient descent
 * (e.g. a torch.autograd.Function over this library). The differentiable substep
 * is prediction, the Jacobi iterations over the distance and bend constraints and
 * the substep's VF / EE contacts (frozen: contact sets and Jacobi counts do not
 * carry gradient), and the velocity update; the Morton / sort / LBVH / CCD stages
 * carry no gradient. Bodies with ground, long-range attachments, a sphere,
 * damping, friction or contact-time rollback are rejected (CM_ERR_ARG). */

/* Replace the stiffness of every distance (k_dist, S floats) and bend (k_bend,
 * B floats) constraint, in [0, 1]; either may be NULL (kept). Per-iteration
 * stiffness is kp = 1 - (1 - k)^(1/iterations), as at create time
 * (Physics/CudaPBD.cu:990, 1029). Syncs; invalidates a recorded forward. */
int cm_set_stiffness(cm_body* body, const float* k_dist, const float* k_bend);
/* `substeps` substeps from the current state, as cm_step, keeping the state every
 * `every` substeps (and after the last) for cm_adjoint. Does not sync. */
int cm_forward_ckpt(cm_body* body, int32_t substeps, int32_t every);
/* Gradients of a loss L over the last cm_forward_ckpt: in, dl_dx / dl_dv = dL/dx,
 * dL/dv of the final positions / velocities (3N floats; dl_dv may be NULL = 0);
 * out (each may be NULL), dL/dx and dL/dv of the state before the first substep
 * (3N), and dL/dk per distance (S) and bend (B) constraint through kp(k) (infinite
 * at k = 1). Each segment of `every` substeps is recomputed from its checkpoint,
 * keeping its Jacobi iterates and contact records, then swept backwards. Leaves
 * the body in its state after the forward. Syncs. CM_ERR_STATE without a forward. */
int cm_adjoint(cm_body* body, const float* dl_dx, const float* dl_dv, float* dl_dx0,
               float* dl_dv0, float* dl_dk_dist, float* dl_dk_bend);

/* ---- launch-shape autotuner ------------------------------------------------
 * No counterpart in the reference (its launch shapes are fixed). Times `substeps`
 * substeps from the current state for each candidate of, one at a time: the Jacobi
 * tile size (64, 128 or 256 vertices), the traversal's tiles per XCD chunk (1, 2, 4,
 * 8) and its first pass's visit budget (128, 160, 200, 240; kept as it is when the
 * two-pass traversal is off), keeps the fastest of each,
 * and restores the state (the simulation does not advance). None of them changes a
 * result. chosen (3 ints, may be NULL) receives the picks in that order. Syncs; the
 * substep graph is captured again at the next cm_step. The body's warnings
 * (cm_get_warnings) keep their value from before the call; the counts and the
 * pair / contact query buffers describe the last trial substep until the next
 * cm_step. */
int cm_autotune(cm_body* body, int32_t substeps, int32_t* chosen);

/* ---- unit-level LBVH entry points on caller device buffers -------------- */
/* Swept leaf AABBs (written to leaf slots [T-1, 2T-1) of d_aabb_min/max) and
 * Morton keys for triangles d_tri in their current order. The values that
 * cm_lbvh_sort permutes with the keys (the triangle slots 0..T-1) are the
 * caller's to fill. With morton_normalize it allocates a centroid scratch of
 * 16 (T + 2 ceil(T/256) + 2) bytes on the stream (hipMallocAsync) and frees it. */
int cm_tri_update(int32_t n, int32_t t, const void* d_x /*float4 N*/, const void* d_p /*float4 N*/,
                  const void* d_tri /*int4 T*/, float thickness, int32_t morton_bits,
                  int32_t morton_normalize, void* d_aabb_min /*float4 2T-1*/,
                  void* d_aabb_max, uint64_t* d_keys, void* stream);
/* cm_tri_update with no device allocation. No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The normalised 30-bit path keeps its
 * centroid scratch in a static device array (4,227,104 bytes on each GPU the
 * library runs on, held until the process ends), so t is at most 262,144
 * (CM_ERR_ARG beyond). The stream must belong to the current device. Calls that
 * overlap on one GPU share that array and race; results are those of
 * cm_tri_update. */
int cm_tri_update_noalloc(int32_t n, int32_t t, const void* d_x, const void* d_p,
                          const void* d_tri, float thickness, int32_t morton_bits,
                          int32_t morton_normalize, void* d_aabb_min, void* d_aabb_max,
                          uint64_t* d_keys, void* stream);
size_t cm_lbvh_sort_temp_bytes(int32_t t);
/* Stable LSD radix sort of (d_keys, d_vals) by the low `key_bits` bits, in place. */
int cm_lbvh_sort(int32_t t, int32_t key_bits, uint64_t* d_keys, uint32_t* d_vals, void* d_temp,
                 size_t temp_bytes, void* stream);
size_t cm_lbvh_build_temp_bytes(int32_t t);
int cm_lbvh_build(int32_t t, const uint64_t* d_keys, int32_t midpoint_split, int32_t* d_lefts,
                  int32_t* d_rights, int32_t* d_parents, uint32_t* d_warn, void* d_temp,
                  size_t temp_bytes, void* stream);
int cm_lbvh_refit(int32_t t, const int32_t* d_lefts, const int32_t* d_rights,
                  const int32_t* d_parents, void* d_aabb_min, void* d_aabb_max, int32_t* d_rmll,
                  int32_t* d_rmlr, uint32_t* d_warn, void* d_temp, size_t temp_bytes,
                  void* stream);
size_t cm_lbvh_pairs_temp_bytes(int32_t t, int32_t cap);
/* Pairs (i, j), j > i, per leaf in depth-first order, at most `cap` per leaf,
 * compacted in leaf order into d_pairs (capacity T*cap int2); *d_n_pairs set. */
int cm_lbvh_pairs(int32_t t, const int32_t* d_lefts, const int32_t* d_rights,
                  const void* d_aabb_min, const void* d_aabb_max, const int32_t* d_rmll,
                  const int32_t* d_rmlr, int32_t cap, int32_t* d_pairs, int32_t* d_n_pairs,
                  uint32_t* d_warn, void* d_temp, size_t temp_bytes, void* stream);

/* ---- unit-level dead paths of the reference's GPU cloth ------------------ */
/* Neither is reached by any reference configuration: FixedUpdate's calls of
 * ClothInternalDynamics and HandleCollisionsContinuous are commented out
 * (Physics/CudaDeformableBody.cpp:605-613). */

/* Spring network of a rows x cols grid of vertices (vertex i * cols + j), as the
 * commented-out Physics/CudaDeformableBody.cpp:107-152 builds it with AddSpring
 * (:288-311), its square dims.x x dims.x loops taken to rows x cols: for each
 * vertex (i, j), i-major, stretch springs to (i+1, j) and (i, j+1), shear springs
 * (i, j)-(i+1, j+1) and (i+1, j)-(i, j+1), bend springs to (i+2, j) and (i, j+2);
 * each spring twice, a -> b listed at a and b -> a listed at b, with the rest
 * distance of xyz (3 floats per vertex). *n_springs receives the number of directed
 * springs S; a_ids, b_ids, ks, l0s (S each), spring_info (S: the springs listed at
 * each vertex, vertex after vertex) and spring_ids (rows*cols + 1 offsets into
 * spring_info) are written when non-null. Host memory; no device work. */
int cm_ms_grid_springs(int32_t rows, int32_t cols, const float* xyz, float k_stretch,
                       float k_shear, float k_bend, int32_t* a_ids, int32_t* b_ids, float* ks,
                       float* l0s, int32_t* spring_info, int32_t* spring_ids,
                       int32_t* n_springs);
/* One step of the mass-spring cloth, ClothInternalDynamics
 * (Physics/CudaDeformableBody.cpp:327-357): ClothEngineStep
 * (Physics/CudaMassSpring.cu:22-88: force m g - k_damp (pos - prev) minus
 * ks (d - l0 d/|d|) for every stretched spring listed at the vertex, Verlet step,
 * prev = old pos, vel = (pos - prev) / dt) then adjust_iters passes of _AdjustSprings
 * (:95-148; the reference runs 5 with max_deformation 0.1: a vertex whose spring
 * that is not a bend spring (ks != k_bend) is stretched by more than
 * max_deformation moves towards the other end by the excess, half of it when the
 * other end is free; vel = (pos - prev) / dt). Fixed vertices (d_fixed != 0) keep
 * position, previous position and velocity. Device buffers: d_pos, d_prev, d_vel
 * float4 N (.w kept), d_mass float N, d_fixed N bytes, the spring arrays of
 * cm_ms_grid_springs; gravity: 3 floats on the host; d_temp: N float4 of scratch.
 * Springs with an end outside [0, n) are skipped. The reference updates positions
 * in place while other threads read them; here every read of another vertex's
 * position sees the value it had when its pass started. */
int cm_ms_step(int32_t n, void* d_pos, void* d_prev, void* d_vel, const float* d_mass,
               const uint8_t* d_fixed, const int32_t* d_a, const int32_t* d_b, const float* d_ks,
               const float* d_l0, const int32_t* d_spring_info, const int32_t* d_spring_ids,
               const float* gravity, float k_damp, float dt, float k_bend, int32_t adjust_iters,
               float max_deformation, void* d_temp, void* stream);

size_t cm_impulses_temp_bytes(int32_t n, int32_t n_contacts);
/* Impulse response to a substep's contacts, SolveCollisions
 * (Physics/CudaDeformableBody.cpp:525-548): CreateImpulses
 * (Collision/CudaCollisionResponse.cu:11-231) at the start positions d_x and the
 * velocities d_v (float4 N), for d_vf (n_vf) then d_ee (n_ee) device contacts, one
 * impulse per contact vertex; each vertex's impulse is the mean of its impulses in
 * the order of a stable sort by vertex id (AccumulateImpulses :233-256); then
 * ApplyImpulses (:258-297): v += impulse / vertex_mass and p = x + v dt for every
 * vertex with d_fixed == 0 (d_p float4 N; .w of v and p kept). stiffness is the
 * reference's kStretch argument. A contact with a vertex id outside [0, n) gives
 * no impulse. d_temp: cm_impulses_temp_bytes(n, n_vf + n_ee) bytes. */
int cm_impulses(int32_t n, const void* d_x, void* d_v, void* d_p, const uint8_t* d_fixed,
                const cm_contact* d_vf, int32_t n_vf, const cm_contact* d_ee, int32_t n_ee,
                float stiffness, float vertex_mass, float dt, float thickness, void* d_temp,
                size_t temp_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CLOTHMI_H */
