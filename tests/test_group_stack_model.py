"""Host model of the occlusion kernel with several lanes per ray
(k_occlude_group in pbrt_rust_amd/csrc/trace.hip): its group stack of GSTK
entries never overflows and every walk ends.  A step takes
k = min(G, top, room) entries, room = (lim - top) // 3 when top + 6 <= lim
and 1 otherwise, with lim = GSTK - 3 * depth (depth: record levels); each
taken record pushes up to four children, a leaf nothing.  The model walks
trees of every shape the bound has to cover -- chains with three leaf
siblings a level (the deepest stack a depth-first walk builds), bushy
trees whose whole frontier passes, and random mixes -- with every child
passing (the most pushes a step can make), and checks the largest stack
against GSTK and the entries popped against the tree's size."""
import random

import pytest

GSTK = 128


def walk(tree, depth, G):
    """Model of one ray's walk; tree[i] = list of children (ints: records,
    None: leaves).  Returns (largest top, entries popped)."""
    lim = GSTK - 3 * depth
    assert 3 * depth + 8 <= GSTK  # the launcher's condition for the group kernel
    stack = [0]
    top_max, popped = 1, 0
    while stack:
        top = len(stack)
        k = min(G, top, (lim - top) // 3 if top + 6 <= lim else 1)
        assert k >= 1
        taken = [stack.pop() for _ in range(k)]
        popped += k
        pushes = []
        for e in taken:
            if e is not None:
                pushes.extend(tree[e])
        assert len(stack) + len(pushes) <= GSTK, (len(stack), len(pushes))
        stack.extend(pushes)
        top_max = max(top_max, len(stack))
    return top_max, popped


def chain(depth):
    # each record: three leaves and the next record (the last: four leaves)
    tree = []
    for d in range(depth):
        tree.append([None, None, None] + ([d + 1] if d + 1 < depth else [None]))
    return tree


def bushy(depth, fan, rng, budget=20000):
    # records with `fan` record children down to `depth` levels (as far as the
    # budget of records allows), the rest leaves
    tree, frontier = [[]], [(0, 1)]
    while frontier:
        i, d = frontier.pop(0)
        kids = []
        for _ in range(4):
            if d < depth and len(tree) < budget and len([c for c in kids if c is not None]) < fan:
                tree.append([])
                kids.append(len(tree) - 1)
                frontier.append((len(tree) - 1, d + 1))
            else:
                kids.append(None)
        tree[i] = kids
    return tree


@pytest.mark.parametrize("G", [4, 8, 16])
@pytest.mark.parametrize("depth", [1, 2, 10, 20, 30, 36, 40])
def test_chain_stays_within_the_group_stack(G, depth):
    tree = chain(depth)
    top_max, popped = walk(tree, depth, G)
    assert top_max <= GSTK
    assert popped == 1 + 4 * depth  # the root, then every child once


@pytest.mark.parametrize("G", [4, 8, 16])
@pytest.mark.parametrize("depth,fan", [(6, 4), (9, 3), (14, 2), (30, 2), (40, 1)])
def test_bushy_frontiers_stay_within_the_group_stack(G, depth, fan):
    tree = bushy(depth, fan, random.Random(depth * 7 + fan))
    lvl = {0: 1}
    for i, kids in enumerate(tree):
        for c in kids:
            if c is not None:
                lvl[c] = lvl[i] + 1
    d = max(lvl.values())
    top_max, popped = walk(tree, d, G)
    assert top_max <= GSTK
    assert popped == 1 + sum(len(k) for k in tree)


@pytest.mark.parametrize("seed", range(12))
def test_random_trees_stay_within_the_group_stack(seed):
    rng = random.Random(seed)
    G = rng.choice([4, 8, 16])
    depth = rng.randint(3, 40)
    # random records: each level's records get 1..4 record children with
    # random probabilities, capped in total
    tree, frontier = [[]], [(0, 1)]
    while frontier and len(tree) < 30000:
        i, d = frontier.pop(rng.randrange(len(frontier)))
        kids = []
        for _ in range(4):
            if d < depth and rng.random() < 0.45:
                tree.append([])
                kids.append(len(tree) - 1)
                frontier.append((len(tree) - 1, d + 1))
            else:
                kids.append(None)
        tree[i] = kids
    for i, d in frontier:
        tree[i] = [None] * 4
    lvl = {0: 1}
    for i, kids in enumerate(tree):
        for c in kids:
            if c is not None:
                lvl[c] = lvl[i] + 1
    d = max(lvl.values())
    top_max, popped = walk(tree, d, G)
    assert top_max <= GSTK
    assert popped == 1 + sum(len(k) for k in tree)
