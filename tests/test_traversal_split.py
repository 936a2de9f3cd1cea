"""The full traversal hands walks longer than a cap to a split kernel, which
walks the rest of the walk's pre-order range [node, end) as sub-ranges in
parallel (collision_amd/csrc/cs_bvh.hip, k_traverse / k_traverse_split).  The
split is exact because a rope walk started at any node x visits x's subtree
and then the following subtrees, i.e. exactly the nodes from x on.

This checks that argument on the CPU, on trees laid out like the device's: a
binary tree over sorted leaves in depth-first pre-order, an internal node's
first child the next record, every node's rope the first node after its
subtree (-1 at the end).  For every leaf, the pairs found by its walk equal
the union of the pairs found by walks over a split of the same range -- at
arbitrary split points, at several depths of splitting, and with the split
taken after an arbitrary number of steps of the original walk.
"""
from __future__ import annotations

import numpy as np
import pytest


def build_tree(lo, hi):
    """Pre-order tree over the leaves (already in order): lists of node
    boxes, left (leaf id or -1), rope."""
    n = len(lo)
    blo, bhi, left, rope = [], [], [], []

    def rec(a, b):  # leaves [a, b)
        me = len(left)
        blo.append(None)
        bhi.append(None)
        left.append(-1)
        rope.append(-1)
        if b - a == 1:
            blo[me], bhi[me], left[me] = lo[a], hi[a], a
        else:
            m = (a + b) // 2 if (b - a) % 3 else a + 1  # uneven splits too
            c1 = rec(a, m)
            c2 = rec(m, b)
            blo[me] = np.minimum(blo[c1], blo[c2])
            bhi[me] = np.maximum(bhi[c1], bhi[c2])
        return me

    rec(0, n)
    # ropes: the next node after the subtree = the node after the last
    # descendant in pre-order
    nn = len(left)
    end = [0] * nn

    def subtree_end(x):
        if left[x] >= 0:
            end[x] = x + 1
            return end[x]
        e1 = subtree_end(x + 1)
        end[x] = subtree_end(e1)
        return end[x]

    subtree_end(0)
    for x in range(nn):
        rope[x] = end[x] if end[x] < nn else -1
    return np.array(blo), np.array(bhi), np.array(left), np.array(rope)


def overlap(alo, ahi, blo, bhi):
    return bool(np.all(alo <= bhi) and np.all(ahi >= blo))


def walk(T, qlo, qhi, start, stop, q, max_steps=None):
    """Rope walk from `start` over nodes < stop; returns (pairs, next node or
    None when the walk ended, steps)."""
    blo, bhi, left, rope = T
    nd, steps, found = start, 0, set()
    while nd >= 0 and nd < stop:
        if max_steps is not None and steps == max_steps:
            return found, nd, steps
        steps += 1
        if not overlap(blo[nd], bhi[nd], qlo, qhi):
            nd = rope[nd]
        elif left[nd] < 0:
            nd += 1
        else:
            found.add((q, int(left[nd])))
            nd = rope[nd]
    return found, None, steps


def split_walk(T, qlo, qhi, lo, hi, q, parts, rng, depth):
    """Walk [lo, hi) as `parts` sub-ranges (random cut points); sub-ranges
    that run longer than a few steps are split again."""
    cuts = sorted(set([lo, hi] + list(rng.integers(lo, hi, size=parts - 1))))
    found = set()
    for a, b in zip(cuts[:-1], cuts[1:]):
        got, rest, _ = walk(T, qlo, qhi, a, b, q, max_steps=3 if depth > 0 else None)
        found |= got
        if rest is not None:
            found |= split_walk(T, qlo, qhi, rest, b, q, parts, rng, depth - 1)
    return found


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_split_walks_find_the_same_pairs(seed):
    rng = np.random.default_rng(seed)
    n = 300
    c = rng.random((n, 3))
    c = c[np.lexsort((c[:, 2], c[:, 1], c[:, 0]))]  # some spatial order
    ext = rng.random((n, 1)) * 0.05
    ext[rng.integers(0, n, size=5)] = 0.6  # a few large boxes (long walks)
    lo, hi = c - ext, c + ext
    T = build_tree(lo, hi)
    nn = len(T[2])
    leaf_node = {int(T[2][x]): x for x in range(nn) if T[2][x] >= 0}
    total_full = total_split = 0
    for q in range(n):
        x = leaf_node[q]
        start = T[3][x]
        if start < 0:
            continue
        full, _, _ = walk(T, lo[q], hi[q], start, nn, q)
        # the device's scheme: walk a few steps, then hand [nd, end) over
        first, rest, _ = walk(T, lo[q], hi[q], start, nn, q, max_steps=int(rng.integers(1, 8)))
        split = set(first)
        if rest is not None:
            split |= split_walk(T, lo[q], hi[q], rest, nn, q, parts=int(rng.integers(2, 9)), rng=rng,
                                depth=2)
        assert split == full, q
        total_full += len(full)
        total_split += len(split)
    assert total_full == total_split > 0


def test_every_pair_once_over_all_leaves():
    """A leaf's walk sees exactly the leaves after it in pre-order whose boxes
    overlap; over all leaves each overlapping pair is found once."""
    rng = np.random.default_rng(7)
    n = 200
    c = rng.random((n, 3))
    lo, hi = c - 0.08, c + 0.08
    T = build_tree(lo, hi)
    nn = len(T[2])
    pairs = []
    for x in range(nn):
        q = int(T[2][x])
        if q < 0 or T[3][x] < 0:
            continue
        got, _, _ = walk(T, lo[q], hi[q], T[3][x], nn, q)
        pairs += [tuple(sorted(p)) for p in got]
    brute = {(i, j) for i in range(n) for j in range(i + 1, n) if overlap(lo[i], hi[i], lo[j], hi[j])}
    assert len(pairs) == len(set(pairs))
    assert set(pairs) == brute
