"""Host-side mirror of the reference's simulation object.

``ClothBody`` keeps the shape of Physics::CudaDeformableBody
(LicentaV2/LicentaV2/Physics/CudaDeformableBody.h:18-28):

    constructor(verts, indices, params, fixed_verts)  -> cm_create
    fixed_update()        one substep                -> cm_step      (FixedUpdate, :550-627)
    update()              read positions + normals   -> cm_read_*    (Update, :629-638)
    set_sphere_pos(p, r)                              -> cm_set_sphere (:640-643)

and adds the contact-pair query of the collision pipeline (cm_query_pairs /
cm_query_contacts) plus access to every internal buffer for parity tests.
All compute runs in libclothmi.so on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import CONTACT_DTYPE, VERTEX_DTYPE, check, lib, make_params, ptr

BUF = dict(x=0, p=1, v=2, tri=3, keys=4, perm=5, lefts=6, rights=7, parents=8, rmll=9, rmlr=10,
           aabb_min=11, aabb_max=12, pairs=13, vf=14, ee=15)
_BUF_DTYPE = dict(x=(np.float32, 4), p=(np.float32, 4), v=(np.float32, 4), tri=(np.int32, 4),
                  keys=(np.uint64, 1), perm=(np.uint32, 1), lefts=(np.int32, 1),
                  rights=(np.int32, 1), parents=(np.int32, 1), rmll=(np.int32, 1),
                  rmlr=(np.int32, 1), aabb_min=(np.float32, 4), aabb_max=(np.float32, 4),
                  pairs=(np.int32, 2))

WARN_NAMES = {1: "stack_overflow", 2: "traversal_limit", 4: "broken_tree", 8: "pair_cap",
              16: "candidate_cap", 32: "refit_limit"}


def warning_names(w: int) -> list:
    """The CM_WARN_* bits of a warning word, by name."""
    return [name for bit, name in sorted(WARN_NAMES.items()) if w & bit]


def leaf_reach(lefts, rights) -> dict:
    """Coverage of a Karras tree over T leaves: internal nodes 0..T-2 with children
    lefts/rights, a child >= T-1 being leaf child - (T-1), the root node 0. With the
    midpoint split on tied Morton codes (CM_RC_MIDPOINT_SPLIT) two internal nodes can
    take the same child, and the leaves of the subtrees no path from the root enters
    are never tested for collisions. Returns the leaf count, the leaves reached from
    the root, the leaves not reached, and the nodes listed as a child more than once."""
    lefts = np.asarray(lefts, np.int64)
    rights = np.asarray(rights, np.int64)
    nin = len(lefts)
    t = nin + 1
    if nin == 0:
        return dict(leaves=t, reached=t, unreachable_leaves=0, shared_nodes=0)
    ch_all = np.concatenate([lefts, rights])
    shared = int((np.bincount(ch_all, minlength=2 * t - 1) > 1).sum())
    seen_int = np.zeros(nin, bool)
    seen_leaf = np.zeros(t, bool)
    seen_int[0] = True
    frontier = np.array([0], np.int64)
    while frontier.size:
        ch = np.concatenate([lefts[frontier], rights[frontier]])
        seen_leaf[ch[ch >= nin] - nin] = True
        inner = np.unique(ch[ch < nin])
        frontier = inner[~seen_int[inner]]
        seen_int[frontier] = True
    reached = int(seen_leaf.sum())
    return dict(leaves=t, reached=reached, unreachable_leaves=t - reached, shared_nodes=shared)


class ClothBody:
    def __init__(self, xyz, tri, params: dict, fixed_verts=None, inv_mass=None, dist_pairs=None,
                 k_dist=None, k_bend=None):
        L = lib()
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        self.tri = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3)
        self.n = len(self.xyz)
        self.t = len(self.tri)
        if inv_mass is None:
            inv_mass = np.ones(self.n, dtype=np.float32)
        self.inv_mass = np.ascontiguousarray(inv_mass, dtype=np.float32)
        self.pinned = None if fixed_verts is None else np.ascontiguousarray(fixed_verts, dtype=np.uint8)
        self.dist_pairs = None if dist_pairs is None else np.ascontiguousarray(dist_pairs, dtype=np.int32)
        self.k_dist = None if k_dist is None else np.ascontiguousarray(k_dist, dtype=np.float32)
        self.k_bend = None if k_bend is None else np.ascontiguousarray(k_bend, dtype=np.float32)
        self.params = dict(params)
        self._p = make_params(self.params)
        h = C.c_void_p()
        nd = 0 if self.dist_pairs is None else len(self.dist_pairs)
        check(L.cm_create(C.byref(self._p), ptr(self.xyz), self.n, ptr(self.tri), self.t,
                          ptr(self.pinned), ptr(self.inv_mass), ptr(self.dist_pairs), nd,
                          ptr(self.k_dist), ptr(self.k_bend), C.byref(h)))
        self.h = h

    @classmethod
    def batch(cls, scenes: list, params: dict | None = None) -> "ClothBody":
        """Independent cloths stepped together as one body (cm_create_batch). Every
        cloth keeps its own LBVH and its k_stretch / k_bend from its scene; the other
        parameters come from ``params`` (default: the first scene's). Vertices of
        cloth c follow those of cloth c - 1 in every read-back."""
        L = lib()
        self = cls.__new__(cls)
        self.params = dict(params or scenes[0]["params"])
        self._p = make_params(self.params)
        self.cloth_n = np.array([len(s["xyz"]) for s in scenes], dtype=np.int32)
        self.cloth_t = np.array([len(s["tri"]) for s in scenes], dtype=np.int32)
        self.xyz = np.ascontiguousarray(np.concatenate([s["xyz"] for s in scenes]), dtype=np.float32)
        self.tri = np.ascontiguousarray(np.concatenate([s["tri"] for s in scenes]), dtype=np.uint32)
        self.n, self.t = len(self.xyz), len(self.tri)
        self.inv_mass = np.ascontiguousarray(np.concatenate([s["inv_mass"] for s in scenes]),
                                             dtype=np.float32)
        self.pinned = np.ascontiguousarray(np.concatenate(
            [s["pinned"] if s.get("pinned") is not None else (s["inv_mass"] == 0) for s in scenes]),
            dtype=np.uint8)
        if all(s.get("dist_pairs") is not None for s in scenes):
            self.cloth_nd = np.array([len(s["dist_pairs"]) for s in scenes], dtype=np.int32)
            self.dist_pairs = np.ascontiguousarray(np.concatenate([s["dist_pairs"] for s in scenes]),
                                                   dtype=np.int32)
        else:
            self.cloth_nd, self.dist_pairs = None, None
        ks = np.array([s["params"].get("k_stretch", self.params["k_stretch"]) for s in scenes],
                      dtype=np.float32)
        kb = np.array([s["params"].get("k_bend", self.params["k_bend"]) for s in scenes],
                      dtype=np.float32)
        self.k_dist = self.k_bend = None
        h = C.c_void_p()
        check(L.cm_create_batch(C.byref(self._p), len(scenes), ptr(self.cloth_n), ptr(self.cloth_t),
                                ptr(self.cloth_nd), ptr(self.xyz), ptr(self.tri), ptr(self.pinned),
                                ptr(self.inv_mass), ptr(self.dist_pairs), ptr(ks), ptr(kb),
                                C.byref(h)))
        self.h = h
        return self

    @classmethod
    def from_scene(cls, scene: dict) -> "ClothBody":
        return cls(scene["xyz"], scene["tri"], scene["params"], scene.get("pinned"),
                   scene["inv_mass"], scene.get("dist_pairs"), scene.get("k_dist"),
                   scene.get("k_bend"))

    def close(self):
        if getattr(self, "h", None):
            lib().cm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- reference-shaped interface ------------------------------------------
    def fixed_update(self, substeps: int = 1) -> None:
        check(lib().cm_step(self.h, substeps))

    def update(self):
        """Readback of positions and vertex normals (Update / FinalVertUpdate)."""
        return self.positions(), self.normals()

    def set_sphere_pos(self, pos, radius: float) -> None:
        c = np.ascontiguousarray(pos, dtype=np.float32)
        check(lib().cm_set_sphere(self.h, ptr(c), float(radius)))

    # ---- state -------------------------------------------------------------------
    def sync(self):
        check(lib().cm_sync(self.h))

    def positions(self) -> np.ndarray:
        out = np.zeros((self.n, 3), dtype=np.float32)
        check(lib().cm_read_positions(self.h, ptr(out)))
        return out

    def velocities(self) -> np.ndarray:
        out = np.zeros((self.n, 3), dtype=np.float32)
        check(lib().cm_read_velocities(self.h, ptr(out)))
        return out

    def normals(self) -> np.ndarray:
        out = np.zeros((self.n, 3), dtype=np.float32)
        check(lib().cm_read_normals(self.h, ptr(out)))
        return out

    def read_vertices(self, verts: np.ndarray) -> None:
        """Fill the position and normal fields of a VERTEX_DTYPE array in place (the
        reference's Update readback); color and uv are left as they are."""
        if verts.dtype != VERTEX_DTYPE or verts.shape != (self.n,) or not verts.flags.c_contiguous:
            raise ValueError("read_vertices needs a contiguous (n,) VERTEX_DTYPE array")
        check(lib().cm_read_vertices(self.h, ptr(verts)))

    def set_state(self, x, v, p=None, tri=None) -> None:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float32).reshape(-1, 3)
        p = None if p is None else np.ascontiguousarray(p, dtype=np.float32).reshape(-1, 3)
        tri = None if tri is None else np.ascontiguousarray(tri, dtype=np.int32).reshape(-1, 4)
        check(lib().cm_set_state(self.h, ptr(x), ptr(p), ptr(v), ptr(tri)))

    def counts(self) -> dict:
        out = np.zeros(9, dtype=np.int64)
        check(lib().cm_get_counts(self.h, ptr(out)))
        return dict(zip(("n", "t", "s", "b", "pairs", "vf", "ee", "vf_cand", "ee_cand"),
                        out.tolist()))

    def warnings(self) -> int:
        w = C.c_uint32()
        check(lib().cm_get_warnings(self.h, C.byref(w)))
        return int(w.value)

    def tree_coverage(self) -> dict:
        """leaf_reach over the tree of the last substep's build."""
        return leaf_reach(self.buffer("lefts"), self.buffer("rights"))

    def buffer(self, name: str) -> np.ndarray:
        L = lib()
        cnt = C.c_int64()
        if name in ("vf", "ee"):
            check(L.cm_get_buffer(self.h, BUF[name], None, 0, C.byref(cnt)))
            out = np.zeros(cnt.value, dtype=CONTACT_DTYPE)
            check(L.cm_get_buffer(self.h, BUF[name], ptr(out), out.nbytes, C.byref(cnt)))
            return out
        dt, width = _BUF_DTYPE[name]
        check(L.cm_get_buffer(self.h, BUF[name], None, 0, C.byref(cnt)))
        out = np.zeros((cnt.value, width) if width > 1 else cnt.value, dtype=dt)
        check(L.cm_get_buffer(self.h, BUF[name], ptr(out), out.nbytes, C.byref(cnt)))
        return out

    def pairs(self) -> np.ndarray:
        return self.buffer("pairs")

    def autotune(self, substeps: int = 3) -> dict:
        """Pick the fastest launch shapes on the current state (cm_autotune); the
        state is restored and no result changes."""
        out = np.zeros(3, dtype=np.int32)
        check(lib().cm_autotune(self.h, int(substeps), ptr(out)))
        return dict(tile_vertices=int(out[0]), traversal_chunk=int(out[1]),
                    traversal_budget=int(out[2]))

    # ---- differentiable rollout ------------------------------------------------------
    def set_stiffness(self, k_dist=None, k_bend=None) -> None:
        """Per-constraint distance / bend stiffness in [0, 1] (None: kept)."""
        kd = None if k_dist is None else np.ascontiguousarray(k_dist, dtype=np.float32)
        kb = None if k_bend is None else np.ascontiguousarray(k_bend, dtype=np.float32)
        c = self.counts()
        if kd is not None and kd.shape != (c["s"],):
            raise ValueError(f"k_dist needs {c['s']} values")
        if kb is not None and kb.shape != (c["b"],):
            raise ValueError(f"k_bend needs {c['b']} values")
        check(lib().cm_set_stiffness(self.h, ptr(kd), ptr(kb)))

    def forward_ckpt(self, substeps: int, every: int) -> None:
        """`substeps` substeps keeping a checkpoint every `every` for adjoint()."""
        check(lib().cm_forward_ckpt(self.h, int(substeps), int(every)))

    def adjoint(self, dl_dx, dl_dv=None):
        """Gradients of a loss over the last forward_ckpt, given dL/dx (and dL/dv) of
        the final state: (dL/dx0, dL/dv0, dL/dk_dist, dL/dk_bend)."""
        c = self.counts()
        lx = np.ascontiguousarray(dl_dx, dtype=np.float32).reshape(self.n, 3)
        lv = None if dl_dv is None else np.ascontiguousarray(dl_dv, dtype=np.float32).reshape(self.n, 3)
        gx = np.zeros((self.n, 3), dtype=np.float32)
        gv = np.zeros((self.n, 3), dtype=np.float32)
        gkd = np.zeros(c["s"], dtype=np.float32)
        gkb = np.zeros(c["b"], dtype=np.float32)
        check(lib().cm_adjoint(self.h, ptr(lx), ptr(lv), ptr(gx), ptr(gv), ptr(gkd), ptr(gkb)))
        return gx, gv, gkd, gkb

    def contacts(self):
        return self.buffer("vf"), self.buffer("ee")
