"""Parity of the HIP FMM path with the CPU oracle (the reference algorithm
restated in oracle/ibim_oracle.c), called through the C ABI.

Bars: tree boxes, point sets and interaction lists bit-exact; FP64 FMM and
matvec results within 1e-10 relative L2 of the oracle (the parity tolerance
for the FP64 path; the two differ only in summation order, typically
1e-14..1e-12); FMM-exact polynomial kernels within 1e-10 of the direct
O(N^2) sum."""
import numpy as np
import pytest

import geometry as G

pytestmark = pytest.mark.gpu

PB = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0, vacant_radius=0.05)
TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def node_keys(parent, slot):
    keys = [()] * len(parent)
    for i in range(len(parent)):
        if parent[i] >= 0:
            keys[i] = keys[parent[i]] + (int(slot[i]),)
    return keys


SETS = {
    "shell": lambda: G.shell(2500, seed=1),
    "clusters": lambda: G.clusters(2000, seed=2),
    "cube": lambda: G.cube(1500, seed=3),
}


@pytest.fixture(scope="module", params=sorted(SETS))
def case(request, ibim, oracle, gpu):
    pts, nrm, w = SETS[request.param]()
    print(f"[case] {request.param} n={len(pts)}", flush=True)
    plan = ibim.Plan(pts, nrm, w, p=4, max_points=32, max_level=10, **PB)
    otree = oracle.Tree(pts, max_point=32, max_level=10)
    yield request.param, pts, nrm, w, plan, otree
    plan.close()
    otree.close()


def test_tree_boxes_bit_exact(case):
    name, pts, nrm, w, plan, ot = case
    gt = plan.tree()
    gk = node_keys(gt["parent"], gt["slot"])
    ok = node_keys(ot.parent, ot.slot)
    gmap = {k: i for i, k in enumerate(gk)}
    omap = {k: i for i, k in enumerate(ok)}
    assert set(gmap) == set(omap), "different box sets"
    for k, oi in omap.items():
        gi = gmap[k]
        assert gt["level"][gi] == ot.level[oi]
        assert gt["leaf"][gi] == ot.leaf[oi] and gt["empty"][gi] == ot.empty[oi]
        assert gt["nsrc"][gi] == ot.nsrc[oi]
        assert np.array_equal(gt["centre"][gi], ot.centre[oi]), "centre not bit-exact"
        assert np.array_equal(gt["radius"][gi], ot.radius[oi]), "radius not bit-exact"
    # point membership of every leaf, bit-exact (ascending original index)
    for k, oi in omap.items():
        if ot.leaf[oi] and ot.nsrc[oi] > 0:
            assert np.array_equal(plan.points(gmap[k]), ot.points(oi)), f"leaf {k} points differ"


def test_lists_match(case):
    name, pts, nrm, w, plan, ot = case
    gt = plan.tree()
    gk = node_keys(gt["parent"], gt["slot"])
    ok = node_keys(ot.parent, ot.slot)
    omap = {k: i for i, k in enumerate(ok)}
    for kind in "UVWX":
        off, items = plan.lists(kind)
        for gi, k in enumerate(gk):
            oi = omap[k]
            glist = {gk[j] for j in items[off[gi]:off[gi + 1]]}
            if kind == "U" and not gt["leaf"][gi]:
                # only leaves use U; the reference also records coarser
                # adjacent leaves for non-leaf boxes, which are never read
                continue
            if kind == "W" and not gt["leaf"][gi]:
                continue
            olist = {ok[j] for j in ot.list(oi, kind)}
            assert glist == olist, f"{kind} list of {k} differs"


@pytest.mark.parametrize("kernel", ["K1x", "K2", "K3xx", "K3xy", "K3zx", "K3yz", "K4y"])
def test_kernel_fmm_matches_oracle(case, kernel):
    name, pts, nrm, w, plan, ot = case
    q = np.random.default_rng(7).normal(size=len(pts))
    g = plan.kernel_fmm(kernel, q)
    o = ot.fmm(4, kernel, q, **PB)
    assert rel(g, o) < TOL, (kernel, rel(g, o))


def test_all_16_kernels_shell(ibim, oracle, gpu):
    pts, nrm, w = G.shell(1200, seed=5)
    q = np.random.default_rng(8).normal(size=len(pts))
    with ibim.Plan(pts, nrm, w, p=5, max_points=24, max_level=10, **PB) as plan:
        ot = oracle.Tree(pts, max_point=24, max_level=10)
        for k in ibim.KERNELS:
            g = plan.kernel_fmm(k, q)
            o = ot.fmm(5, k, q, **PB)
            assert rel(g, o) < TOL, (k, rel(g, o))


@pytest.mark.parametrize("p", [3, 4, 6, 8])
def test_polynomial_kernels_exact(ibim, oracle, gpu, p):
    """With polynomial kernels the Chebyshev FMM is exact, so the GPU FMM must
    equal the direct sum over all pairs: pins tree, lists and all operators."""
    pts, nrm, w = G.clusters(1500, seed=11)
    q = np.random.default_rng(3).normal(size=len(pts))
    with ibim.Plan(pts, nrm, w, p=p, max_points=20, max_level=10, kernel_mode=ibim.KERNEL_POLY,
                   vacant_radius=0.0) as plan:
        for k in ibim.KERNELS:
            g = plan.kernel_fmm(k, q)
            d = oracle.direct(k, pts, pts, q, mode=1)
            assert rel(g, d) < TOL, (k, rel(g, d))


def test_matvec_matches_oracle(case, oracle):
    name, pts, nrm, w, plan, ot = case
    x = np.random.default_rng(4).uniform(-1, 1, size=2 * len(pts))
    g = plan.matvec(x)
    o = oracle.dmapping(pts, nrm, w, x, p=4, max_point=32, max_level=10, **PB)
    assert rel(g, o) < TOL, rel(g, o)


def test_matvec_laplace_kappa0(ibim, oracle, gpu):
    pts, nrm, w = G.sphere_tube(h=0.25, lo=-4.0, npts=32)
    prm = dict(kappa=0.0, eps_in=1.0, eps_out=80.0, vacant_radius=0.25)
    x = np.random.default_rng(1).uniform(-1, 1, size=2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=4, max_points=64, max_level=10, **prm) as plan:
        g = plan.matvec(x)
    o = oracle.dmapping(pts, nrm, w, x, p=4, max_point=64, max_level=10, **prm)
    assert rel(g, o) < TOL


def test_identity_operator(ibim, gpu):
    """eps_in = eps_out and kappa = 0 make every kernel vanish: A = I."""
    pts, nrm, w = G.shell(800, seed=9)
    x = np.random.default_rng(2).normal(size=2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=4, max_points=32, max_level=10, kappa=0.0, eps_in=3.0, eps_out=3.0,
                   vacant_radius=0.05) as plan:
        y = plan.matvec(x)
    assert np.array_equal(y, x)


def test_rhs_matches_oracle(ibim, oracle, gpu):
    pts, nrm, w = G.shell(3000, seed=4)
    rng = np.random.default_rng(5)
    atoms = rng.uniform(-1, 1, size=(37, 3))
    q = rng.uniform(-0.5, 0.5, size=37)
    with ibim.Plan(pts, nrm, w, p=4, max_points=32, max_level=10, **PB) as plan:
        g = plan.rhs(atoms, q)
    o = oracle.rhs(pts, nrm, atoms, q, PB["eps_in"], PB["eps_out"])
    assert rel(g, o) < TOL


def test_polarize_matches_oracle(ibim, oracle, gpu):
    pts, nrm, w = G.shell(2500, seed=6)
    rng = np.random.default_rng(6)
    # atoms inside the shell, one on a box centre plane, two outside the root box
    atoms = np.concatenate([rng.uniform(-1.2, 1.2, size=(40, 3)), [[0.0, 0.0, 0.0]],
                            [[3.0, 0.1, 0.2], [-0.3, -2.6, 0.4]]])
    x = rng.uniform(-1, 1, size=2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=4, max_points=32, max_level=10, **PB) as plan:
        g = plan.polarize(atoms, x)
    o = oracle.polarize(pts, nrm, w, atoms, x, p=4, max_point=32, max_level=10, **PB)
    assert rel(g, o) < TOL, rel(g, o)


@pytest.mark.parametrize("tol", [1e-10, 1e-13])
def test_gmres_matches_oracle(ibim, oracle, gpu, tol):
    """GMRES(8) with restarts on the plan's operator against the oracle's
    GMRES on its own: the same iteration count (the two operators differ by
    ~1e-15, so a residual estimate may cross tol one step apart: +-1), and
    solutions within the stopping tolerance's reach; at tol 1e-13 they agree
    to the FP64 bar, 1e-10."""
    pts, nrm, w = G.shell(700, seed=12)
    rng = np.random.default_rng(13)
    atoms = rng.uniform(-0.8, 0.8, size=(5, 3))
    q = rng.uniform(-0.5, 0.5, size=5)
    b = oracle.rhs(pts, nrm, atoms, q, PB["eps_in"], PB["eps_out"])
    with ibim.Plan(pts, nrm, w, p=4, max_points=32, max_level=10, **PB) as plan:
        xg, rc, info = plan.gmres(b, restart=8, maxit=200, tol=tol)
    xo, orc, oit, ores, onmv = oracle.solve(pts, nrm, w, b, p=4, max_point=32, max_level=10, restart=8,
                                            maxit=200, tol=tol, **PB)
    print(f"tol {tol:g}: iterations {info['iterations']} / {oit}, x rel {rel(xg, xo):.1e}")
    assert rc == 0 and orc == 0
    assert abs(info["iterations"] - oit) <= 1, (info, oit)
    assert rel(xg, xo) < (1e-8 if tol > 1e-12 else 1e-10)


def test_small_and_degenerate_trees(ibim, oracle, gpu):
    rng = np.random.default_rng(21)
    for n, mp, ml in [(1, 4, 10), (7, 32, 10), (300, 16, 0), (300, 16, 1), (400, 1, 3)]:
        pts = rng.uniform(-1, 1, size=(n, 3))
        if n == 1:
            pts = np.concatenate([pts, pts + 0.3])  # a single point has a zero-size box
        nrm = rng.normal(size=pts.shape)
        w = np.full(len(pts), 0.02)
        x = rng.uniform(-1, 1, size=2 * len(pts))
        with ibim.Plan(pts, nrm, w, p=4, max_points=mp, max_level=ml, **PB) as plan:
            g = plan.matvec(x)
        o = oracle.dmapping(pts, nrm, w, x, p=4, max_point=mp, max_level=ml, **PB)
        assert rel(g, o) < TOL, (n, mp, ml, rel(g, o))


def test_ties_on_box_centres(ibim, oracle, gpu):
    """Points exactly on a box's centre planes go to the upper child (`<`)."""
    g = np.linspace(-1, 1, 9)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    with ibim.Plan(pts, p=3, max_points=4, max_level=4, kernel_mode=ibim.KERNEL_POLY) as plan:
        gt = plan.tree()
        ot = oracle.Tree(pts, max_point=4, max_level=4)
        gk, ok = node_keys(gt["parent"], gt["slot"]), node_keys(ot.parent, ot.slot)
        omap = {k: i for i, k in enumerate(ok)}
        for gi, k in enumerate(gk):
            oi = omap[k]
            if gt["leaf"][gi] and gt["nsrc"][gi]:
                assert np.array_equal(plan.points(gi), ot.points(oi))


def test_bad_arguments_fail_loudly(ibim, gpu):
    pts, nrm, w = G.shell(100)
    with pytest.raises(ibim.IbimError):
        ibim.Plan(pts, nrm, w, p=9)
    with pytest.raises(ibim.IbimError):
        ibim.Plan(pts, nrm, w, p=4, vacant_radius=0.0)  # PB kernels need tau*h > 0
    with ibim.Plan(pts, nrm, w, p=4, vacant_radius=0.1) as plan:
        with pytest.raises(ibim.IbimError):
            plan.kernel_fmm(16, np.zeros(len(pts)))


def test_near_field_split_into_units(ibim, oracle, gpu):
    """Leaves holding more than 64 targets and more than 1024 near-field
    sources (max_level caps the tree at 64 leaves of ~125 points): each
    leaf's near field is cut into several work units whose partial fields are
    summed, in FP64 and in the mixed path."""
    rng = np.random.default_rng(31)
    n = 8000
    pts = rng.uniform(-1, 1, size=(n, 3))
    nrm = rng.normal(size=pts.shape)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    w = np.full(n, 1e-3)
    x = rng.uniform(-1, 1, size=2 * n)
    with ibim.Plan(pts, nrm, w, p=4, max_points=64, max_level=2, **PB) as plan:
        t = plan.tree()
        leaves = t["leaf"].astype(bool)
        assert t["nsrc"][leaves].max() > 64
        g = plan.matvec(x)
        plan.set_precision(ibim.PREC_MIXED)
        gm = plan.matvec(x)
    o = oracle.dmapping(pts, nrm, w, x, p=4, max_point=64, max_level=2, **PB)
    assert rel(g, o) < TOL, rel(g, o)
    assert rel(gm, o) < 1e-5, rel(gm, o)
