"""Pins the oracle's octree and interaction lists (bbfmm.cpp:8-280) by the
invariants the construction must satisfy, plus the tie-breaking rule."""
import numpy as np
import pytest

import geometry as G

SETS = {"shell": lambda: G.shell(3000, seed=1)[0], "clusters": lambda: G.clusters(2500, seed=2)[0],
        "cube": lambda: G.cube(2000, seed=3)[0]}


@pytest.fixture(scope="module", params=sorted(SETS))
def tree(request, oracle):
    pts = SETS[request.param]()
    return pts, oracle.Tree(pts, max_point=24, max_level=8)


def test_root_is_source_bounding_box(tree):
    pts, t = tree
    np.testing.assert_array_equal(t.centre[0], (pts.max(0) + pts.min(0)) / 2)
    np.testing.assert_array_equal(t.radius[0], (pts.max(0) - pts.min(0)) / 2)


def test_children_geometry(tree):
    pts, t = tree
    for i in range(1, t.nnodes):
        p = t.parent[i]
        s = t.slot[i]
        bits = np.array([s & 1, (s >> 1) & 1, s >> 2])
        np.testing.assert_array_equal(t.radius[i], t.radius[p] * 0.5)
        np.testing.assert_array_equal(t.centre[i], t.centre[p] + (bits - 0.5) * t.radius[p])
        assert t.level[i] == t.level[p] + 1


def test_leaves_partition_points_and_contain_them(tree):
    pts, t = tree
    seen = np.zeros(len(pts), int)
    for i in range(t.nnodes):
        if not t.leaf[i]:
            continue
        idx = t.points(i)
        seen[idx] += 1
        assert np.all(np.diff(idx) > 0), "ascending original order"
        if len(idx) == 0:
            assert t.empty[i]
            continue
        assert len(idx) <= 24 or t.level[i] == 8
        # every point satisfies the descent comparisons of all its ancestors
        node = i
        while t.parent[node] >= 0:
            p = t.parent[node]
            s = t.slot[node]
            bits = np.array([s & 1, (s >> 1) & 1, s >> 2])
            upper = pts[idx] >= t.centre[p]
            assert np.all(upper == bits.astype(bool)), "point on the wrong side of a centre plane"
            node = p
    assert np.all(seen == 1)


def test_split_rule(tree):
    pts, t = tree
    for i in range(t.nnodes):
        if t.leaf[i]:
            assert t.nsrc[i] <= 24 or t.level[i] == 8 or t.nsrc[i] == 0
        else:
            assert t.nsrc[i] > 24 and t.level[i] < 8


def test_list_symmetries(tree):
    pts, t = tree
    U = {i: set(t.list(i, "U")) for i in range(t.nnodes) if t.leaf[i]}
    V = {i: set(t.list(i, "V")) for i in range(t.nnodes)}
    W = {i: set(t.list(i, "W")) for i in range(t.nnodes) if t.leaf[i]}
    X = {i: set(t.list(i, "X")) for i in range(t.nnodes)}
    for a, us in U.items():
        assert a in us
        for b in us:
            assert a in U[b], "U is symmetric between leaves"
    for a, vs in V.items():
        for b in vs:
            assert a in V[b] and t.level[a] == t.level[b]
    for a, ws in W.items():
        for b in ws:
            assert a in X[b], "W(a) contains b exactly when X(b) contains a"
    for b, xs in X.items():
        for a in xs:
            assert t.leaf[a] and b in W[a]


def test_ties_go_to_upper_child(oracle):
    pts = np.array([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0], [0.0, -0.5, 0.5],
                    [-0.5, 0.0, 0.2], [0.3, 0.3, 0.0]])
    t = oracle.Tree(pts, max_point=1, max_level=1)
    slot_of = {}
    for i in range(1, t.nnodes):
        for j in t.points(i):
            slot_of[int(j)] = int(t.slot[i])
    # the origin is the root centre: x, y, z all >= centre -> slot 7
    assert slot_of[2] == 7
    assert slot_of[3] == 1 + 4      # x on the plane -> 1; y below; z above
    assert slot_of[4] == 2 + 4      # y on the plane -> upper
    assert slot_of[5] == 1 + 2 + 4  # z on the plane -> upper
