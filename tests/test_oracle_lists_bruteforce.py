"""The oracle's U/V/W/X lists (its restatement of bbfmm.cpp:158-280) against
the textbook definitions of the adaptive FMM lists, evaluated by brute force
over all pairs of boxes of the oracle's own tree.  Two boxes touch when they
share a boundary point without overlapping: every centre distance is at most
the sum of the half-widths and one of them equals it (within 1e-12, the
reference's EPS).  For a non-root box B:

  V(B) = same-level boxes not touching B whose parent touches or is B's parent
  U(B) = the leaves touching B, and B            (leaves B only)
  W(B) = finer boxes not touching B whose parent touches B   (leaves B only)
  X(B) = the leaves A with B in W(A)

A box filed under the wrong list (a V box taken as U, a W box as X) changes
these sets even when every pair is still counted once, which the symmetry and
duality checks of test_oracle_tree.py cannot see."""
import numpy as np
import pytest

import geometry as G

EPS = 1e-12


def lattice(n=9, jitter=0.0, seed=0):
    g = np.linspace(-1, 1, n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    if jitter:
        pts = pts + np.random.default_rng(seed).uniform(-jitter, jitter, pts.shape)
    return pts


SETS = {
    "shell": lambda: (G.shell(2500, seed=1)[0], 32),
    "clusters": lambda: (G.clusters(2000, seed=2)[0], 32),
    "cube": lambda: (G.cube(1500, seed=3)[0], 32),
    # points on box faces, edges and corners, and within 1e-13 of them
    "lattice": lambda: (lattice(), 4),
    "lattice_jitter": lambda: (lattice(jitter=1e-13, seed=4), 4),
}


def textbook_lists(t):
    c, r, lev, par, leaf = t.centre, t.radius, t.level, t.parent, t.leaf.astype(bool)
    n = len(lev)

    def touching(b):
        d = np.abs(c - c[b])
        s = r + r[b]
        return np.all(s >= d - EPS, axis=1) & np.any(np.abs(d - s) < EPS, axis=1)

    touch = np.array([touching(b) for b in range(n)])  # symmetric
    out = {k: [set() for _ in range(n)] for k in "UVWX"}
    for b in range(1, n):
        pb = par[b]
        colleagues_of_parent = np.flatnonzero((lev == lev[pb]) & (touch[pb] | (np.arange(n) == pb)))
        v = (lev == lev[b]) & ~touch[b] & np.isin(par, colleagues_of_parent)
        v[b] = False
        out["V"][b] = set(np.flatnonzero(v).tolist())
        if leaf[b]:
            out["U"][b] = set(np.flatnonzero(leaf & touch[b]).tolist()) | {b}
            has_par = par >= 0
            w = (lev > lev[b]) & ~touch[b] & has_par & touch[b][np.where(has_par, par, 0)]
            out["W"][b] = set(np.flatnonzero(w).tolist())
    for a in range(n):
        for b in out["W"][a]:
            out["X"][b].add(a)
    return out


@pytest.mark.parametrize("name", sorted(SETS))
def test_oracle_lists_are_the_textbook_lists(oracle, name):
    pts, leaf_cap = SETS[name]()
    t = oracle.Tree(pts, max_point=leaf_cap, max_level=10)
    try:
        ref = textbook_lists(t)
        checked = 0
        for b in range(1, t.nnodes):
            for kind in "VX" + ("UW" if t.leaf[b] else ""):
                got = set(t.list(b, kind).tolist())
                assert got == ref[kind][b], (name, b, kind, sorted(got ^ ref[kind][b])[:8])
                checked += 1
        # the sets exercise every list
        assert all(sum(len(s) for s in ref[k]) > 0 for k in "UVWX"), name
        assert checked >= 2 * (t.nnodes - 1)
    finally:
        t.close()
