"""Pins the oracle's LBVH stages (Collision/CudaBVH.cu:9-242) without retyping them:

* unique keys: lefts/rights equal a top-down radix tree built here by scanning for
  the highest differing bit (a different algorithm), refit equals the union of the
  leaves of each node's key range, rightmost-leaf ids are closed forms, and the
  pair list equals brute-force O(T^2) inclusive-overlap enumeration;
* tied keys: a 5-leaf example derived by hand for the reference's midpoint split
  (a leaf left unclaimed, another claimed twice, duplicate pairs) and for the
  index-augmented split (valid tree, all 10 pairs);
* BVHTestUnit invariants (Collision/CudaBVH.cu:244-376).
"""
import numpy as np
import pytest


def topdown(keys):
    """Karras node numbering from a recursive split at the highest differing bit."""
    t = len(keys)
    nin = t - 1
    lefts = np.full(nin, -1, np.int64)
    rights = np.full(nin, -1, np.int64)
    rng = {}
    split_of = {}

    def rec(node, f, l):
        rng[node] = (f, l)
        kf, kl = int(keys[f]), int(keys[l])
        hb = (kf ^ kl).bit_length() - 1
        g = f
        while g + 1 <= l and ((int(keys[g + 1]) >> hb) & 1) == ((kf >> hb) & 1):
            g += 1
        split_of[node] = g
        left = g if g != f else nin + g
        right = g + 1 if g + 1 != l else nin + g + 1
        lefts[node], rights[node] = left, right
        if left < nin:
            rec(left, f, g)
        if right < nin:
            rec(right, g + 1, l)

    rec(0, 0, t - 1)
    return lefts, rights, rng, split_of


def unique_keys(rng, t):
    k = np.unique(rng.integers(0, 1 << 62, size=3 * t, dtype=np.int64))[:t]
    return np.sort(k).astype(np.uint64)


@pytest.mark.parametrize("t", [2, 3, 7, 64, 1000])
def test_build_unique_matches_topdown(oracle_mod, t):
    keys = unique_keys(np.random.default_rng(t), t)
    for midpoint in (True, False):
        lefts, rights, parents, warn = oracle_mod.build(keys, midpoint)
        tl, tr, _, _ = topdown(keys)
        assert np.array_equal(lefts, tl)
        assert np.array_equal(rights, tr)
        assert warn == 0
        # BVHTestUnit: every node but the root has exactly one parent, parents[0] stays 0
        nin = t - 1
        cnt = np.zeros(2 * t - 1, int)
        for i in range(nin):
            cnt[lefts[i]] += 1
            cnt[rights[i]] += 1
            assert parents[lefts[i]] == i and parents[rights[i]] == i
        assert cnt[0] == 0 and (cnt[1:] == 1).all()
        assert parents[0] == 0


def test_refit_and_pairs_unique(oracle_mod):
    rng = np.random.default_rng(5)
    t = 300
    keys = unique_keys(rng, t)
    lefts, rights, parents, _ = oracle_mod.build(keys, True)
    lo = rng.uniform(0, 1, size=(t, 3)).astype(np.float32)
    hi = lo + rng.uniform(0.0, 0.15, size=(t, 3)).astype(np.float32)
    amin = np.zeros((2 * t - 1, 3), np.float32)
    amax = np.zeros((2 * t - 1, 3), np.float32)
    amin[t - 1:] = lo
    amax[t - 1:] = hi
    amin, amax, rmll, rmlr = oracle_mod.refit(lefts, rights, amin, amax)
    _, _, rngs, split_of = topdown(keys)
    nin = t - 1
    for node, (f, l) in rngs.items():
        assert np.array_equal(amin[node], lo[f:l + 1].min(0))
        assert np.array_equal(amax[node], hi[f:l + 1].max(0))
        assert rmll[node] == nin + split_of[node]
        assert rmlr[node] == nin + l
    pairs, warn = oracle_mod.pairs(lefts, rights, amin, amax, rmll, rmlr, cap=t)
    ov = ((lo[:, None, :] <= hi[None, :, :]) & (hi[:, None, :] >= lo[None, :, :])).all(-1)
    brute = {(i, j) for i in range(t) for j in range(i + 1, t) if ov[i, j]}
    got = [tuple(p) for p in pairs.tolist()]
    assert len(got) == len(set(got))
    assert set(got) == brute
    # leaf order preserved by the compaction
    assert all(got[k][0] <= got[k + 1][0] for k in range(len(got) - 1))
    assert warn == 0


def test_pair_cap_keeps_first_in_dfs_order(oracle_mod):
    rng = np.random.default_rng(6)
    t = 40
    keys = unique_keys(rng, t)
    lefts, rights, _, _ = oracle_mod.build(keys, True)
    amin = np.zeros((2 * t - 1, 3), np.float32)
    amax = np.ones((2 * t - 1, 3), np.float32)  # everything overlaps
    amin, amax, rmll, rmlr = oracle_mod.refit(lefts, rights, amin, amax)
    full, _ = oracle_mod.pairs(lefts, rights, amin, amax, rmll, rmlr, cap=t)
    capped, warn = oracle_mod.pairs(lefts, rights, amin, amax, rmll, rmlr, cap=5)
    assert warn & 8  # CM_WARN_PAIR_CAP
    for i in range(t):
        fi = full[full[:, 0] == i]
        ci = capped[capped[:, 0] == i]
        assert np.array_equal(ci, fi[:5])


# keys 1, 4, 4, 4, 8 (leaves 0..4, nodes 4..8): derived by hand in the module docstring's terms
TIE_KEYS = np.array([1, 4, 4, 4, 8], dtype=np.uint64)


def test_midpoint_split_on_ties_by_hand(oracle_mod):
    parents_in = np.full(9, 77, np.int32)
    lefts, rights, parents, warn = oracle_mod.build(TIE_KEYS, True, parents_in)
    assert lefts.tolist() == [3, 2, 6, 4]
    assert rights.tolist() == [8, 7, 7, 1]
    assert warn & 4  # CM_WARN_BROKEN_TREE: leaf 7 claimed by nodes 1 and 2
    # highest claimant wins; the unclaimed leaf 5 and the root keep their old value
    assert parents.tolist() == [77, 3, 1, 0, 3, 77, 2, 2, 0]
    amin = np.zeros((9, 3), np.float32)
    amax = np.zeros((9, 3), np.float32)
    for leaf in range(5):
        amin[4 + leaf] = [leaf, 0, 0]
        amax[4 + leaf] = [leaf + 0.5, 1, 1]
    amin, amax, rmll, rmlr = oracle_mod.refit(lefts, rights, amin, amax)
    # node 2 = L2|L3, node 1 = node2|L3, node 3 = L0|node1, node 0 = node3|L4: leaf 1 missing
    assert amin[3, 0] == 0 and amax[3, 0] == 3.5
    assert amin[1, 0] == 2 and amin[2, 0] == 2
    assert rmll.tolist() == [7, 7, 6, 4] and rmlr.tolist() == [8, 7, 7, 7]
    big_min = np.full((9, 3), -1, np.float32)
    big_max = np.full((9, 3), 10, np.float32)
    pairs, _ = oracle_mod.pairs(lefts, rights, big_min, big_max, rmll, rmlr, cap=16)
    assert [tuple(p) for p in pairs.tolist()] == [
        (0, 4), (0, 3), (0, 2), (0, 3), (1, 4), (1, 3), (1, 2), (1, 3), (2, 4), (2, 3), (2, 3),
        (3, 4)]


def test_augmented_split_on_ties_by_hand(oracle_mod):
    lefts, rights, parents, warn = oracle_mod.build(TIE_KEYS, False)
    assert lefts.tolist() == [3, 5, 6, 4]
    assert rights.tolist() == [8, 2, 7, 1]
    assert warn == 0
    amin = np.zeros((9, 3), np.float32)
    amax = np.ones((9, 3), np.float32)
    amin, amax, rmll, rmlr = oracle_mod.refit(lefts, rights, amin, amax)
    assert rmll.tolist() == [7, 5, 6, 4] and rmlr.tolist() == [8, 7, 7, 7]
    pairs, _ = oracle_mod.pairs(lefts, rights, amin, amax, rmll, rmlr, cap=16)
    assert [tuple(p) for p in pairs.tolist()] == [
        (0, 4), (0, 1), (0, 2), (0, 3), (1, 4), (1, 2), (1, 3), (2, 4), (2, 3), (3, 4)]


def test_all_equal_keys_midpoint(oracle_mod):
    # every key equal: each node's range comes from the index tie-break; the
    # midpoint split then reuses nodes (BVHTestUnit would flag it); the oracle's
    # refit still terminates and each node is the union of its two children
    t = 9
    keys = np.full(t, 12345, np.uint64)
    lefts, rights, parents, warn = oracle_mod.build(keys, True)
    nin = t - 1
    assert ((lefts > 0) & (lefts < 2 * t - 1)).all() and ((rights > 0) & (rights < 2 * t - 1)).all()
    amin = np.zeros((2 * t - 1, 3), np.float32)
    amax = np.zeros((2 * t - 1, 3), np.float32)
    amin[nin:, 0] = np.arange(t)
    amax[nin:, 0] = np.arange(t) + 1
    amin2, amax2, rmll, rmlr = oracle_mod.refit(lefts, rights, amin, amax)
    for q in range(nin):
        l, r = lefts[q], rights[q]
        assert amin2[q, 0] == min(amin2[l, 0], amin2[r, 0])
        assert amax2[q, 0] == max(amax2[l, 0], amax2[r, 0])
    lefts2, rights2, _, warn2 = oracle_mod.build(keys, False)
    assert warn2 == 0
    tl, tr, _, _ = topdown(np.arange(t, dtype=np.uint64))  # index bits decide
    assert np.array_equal(lefts2, tl) and np.array_equal(rights2, tr)


def _walk_leaves(lefts, rights):
    """Leaves reached from the root by a recursive walk over the children lists."""
    nin = len(lefts)
    seen, leaves = set(), set()

    def visit(q):
        if q >= nin:
            leaves.add(q - nin)
            return
        if q in seen:
            return
        seen.add(q)
        visit(int(lefts[q]))
        visit(int(rights[q]))

    visit(0)
    return leaves


def test_leaf_reach_on_the_tie_trees(oracle_mod):
    """The bench's tree-coverage count (licenta_amd.body.leaf_reach) on the
    hand-derived 5-leaf tie trees: the midpoint split loses leaf 1 and gives leaf 3
    two parents; the index-augmented split reaches all five leaves."""
    from licenta_amd.body import leaf_reach
    lefts, rights, _, _ = oracle_mod.build(TIE_KEYS, True)
    assert leaf_reach(lefts, rights) == dict(leaves=5, reached=4, unreachable_leaves=1,
                                             shared_nodes=1)
    lefts, rights, _, _ = oracle_mod.build(TIE_KEYS, False)
    assert leaf_reach(lefts, rights) == dict(leaves=5, reached=5, unreachable_leaves=0,
                                             shared_nodes=0)
    assert leaf_reach(np.zeros(0, np.int32), np.zeros(0, np.int32))["unreachable_leaves"] == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_leaf_reach_matches_a_recursive_walk(oracle_mod, seed):
    """Keys drawn from a few values (long runs of ties): leaf_reach's count equals a
    recursive walk's over the oracle's midpoint-split tree, and the augmented split's
    tree reaches every leaf."""
    from licenta_amd.body import leaf_reach
    rng = np.random.default_rng(seed)
    keys = np.sort(rng.integers(0, 12, 400)).astype(np.uint64)
    lefts, rights, _, warn = oracle_mod.build(keys, True)
    walked = _walk_leaves(lefts, rights)
    r = leaf_reach(lefts, rights)
    assert r["reached"] == len(walked) and r["unreachable_leaves"] == 400 - len(walked)
    assert r["unreachable_leaves"] > 0  # these runs of ties do lose leaves
    assert (r["shared_nodes"] > 0) == bool(warn & 4)
    lefts, rights, _, warn = oracle_mod.build(keys, False)
    assert warn == 0 and leaf_reach(lefts, rights)["unreachable_leaves"] == 0
    assert len(_walk_leaves(lefts, rights)) == 400
