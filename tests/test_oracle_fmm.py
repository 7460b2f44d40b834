"""Pins the oracle's FMM engine and matvec (bbfmm.h:500-623,
electric.cpp:294-413, 442-485):

* with the polynomial surrogate kernels (degree <= 2 per coordinate) the
  Chebyshev FMM is exact, so it must equal the direct O(N^2) sum to rounding:
  a dropped or doubly counted interaction, a wrong transfer index or a sign
  error in any of the 16 closures fails this;
* with the PB kernels it converges to the direct sum as p grows;
* eps_in = eps_out and kappa = 0 give A = I exactly;
* the RHS and the reaction potential match direct formulas.
"""
import numpy as np
import pytest

import geometry as G

PB = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0, vacant_radius=0.05)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("maker", [lambda: G.shell(1500, seed=3), lambda: G.clusters(1200, seed=4)])
def test_polynomial_kernels_exact(oracle, maker):
    pts, nrm, w = maker()
    q = np.random.default_rng(0).normal(size=len(pts))
    t = oracle.Tree(pts, max_point=16, max_level=10)
    assert (t.nx > 0).any() and (t.nw > 0).any(), "the point set exercises the W and X lists"
    for k in oracle.KERNELS:
        f = t.fmm(3, k, q, mode=1)
        d = oracle.direct(k, pts, pts, q, mode=1)
        assert rel(f, d) < 1e-12, k


def test_polynomial_exact_with_external_targets(oracle):
    """Targets outside the source box are extrapolated (bbfmm.cpp:43-65); exact
    for polynomials, which pins the target assignment and L2P."""
    pts, nrm, w = G.shell(1000, seed=5)
    tgt = np.concatenate([np.random.default_rng(1).uniform(-1, 1, (50, 3)), [[3.0, 0, 0], [0, -3.5, 1.0]]])
    q = np.random.default_rng(2).normal(size=len(pts))
    t = oracle.Tree(pts, tgt, max_point=16, max_level=10)
    for k in ["K1y", "K2", "K3zz", "K3xz", "K4x"]:
        assert rel(t.fmm(4, k, q, mode=1), oracle.direct(k, pts, tgt, q, mode=1)) < 1e-12


def test_pb_kernels_converge_with_p(oracle):
    pts, nrm, w = G.shell(1500, seed=6)
    q = np.random.default_rng(3).normal(size=len(pts))
    t = oracle.Tree(pts, max_point=24, max_level=10)
    for k in ["K1x", "K2", "K3xy", "K4z"]:
        d = oracle.direct(k, pts, pts, q, **PB)
        errs = [rel(t.fmm(p, k, q, **PB), d) for p in (3, 5, 7)]
        assert errs[2] < errs[1] < errs[0], (k, errs)
        assert errs[2] < 5e-4, (k, errs)


def test_dmapping_matches_dense_operator(oracle):
    pts, nrm, w = G.shell(900, seed=7)
    x = np.random.default_rng(4).uniform(-1, 1, size=2 * len(pts))
    yd = oracle.dmapping_direct(pts, nrm, w, x, **PB)
    y = oracle.dmapping(pts, nrm, w, x, p=7, max_point=24, max_level=10, **PB)
    assert rel(y, yd) < 1e-4
    # one tree per kernel (as the reference) or shared: identical results
    y16 = oracle.dmapping(pts, nrm, w, x, p=4, max_point=24, max_level=10, share_tree=False, **PB)
    y1 = oracle.dmapping(pts, nrm, w, x, p=4, max_point=24, max_level=10, share_tree=True, **PB)
    np.testing.assert_array_equal(y16, y1)


def test_dmapping_identity(oracle):
    pts, nrm, w = G.shell(600, seed=8)
    x = np.random.default_rng(5).normal(size=2 * len(pts))
    y = oracle.dmapping(pts, nrm, w, x, p=4, max_point=24, max_level=10, kappa=0.0, eps_in=5.0, eps_out=5.0,
                        vacant_radius=0.05)
    np.testing.assert_array_equal(y, x)


def test_dmapping_combination_of_16_fmms(oracle):
    """y = x + [(sum K1a[qa] - K2[qpsi]) / c1 ; (sum_a n_a (sum_b K3ab[qb] - K4a[qpsi])) / c2]
    assembled here from 16 separate direct sums (electric.cpp:369-384)."""
    pts, nrm, w = G.shell(500, seed=9)
    n = len(pts)
    x = np.random.default_rng(6).normal(size=2 * n)
    q = {"x": x[:n] * nrm[:, 0] * w, "y": x[:n] * nrm[:, 1] * w, "z": x[:n] * nrm[:, 2] * w, "p": x[n:] * w}
    K = lambda k, c: oracle.direct(k, pts, pts, q[c], **PB)
    f0 = K("K1x", "x") + K("K1y", "y") + K("K1z", "z") - K("K2", "p")
    H = {"x": K("K3xx", "x") + K("K3xy", "y") + K("K3zx", "z"),
         "y": K("K3xy", "x") + K("K3yy", "y") + K("K3yz", "z"),
         "z": K("K3zx", "x") + K("K3yz", "y") + K("K3zz", "z")}
    f1 = sum(nrm[:, i] * (H[a] - K("K4" + a, "p")) for i, a in enumerate("xyz"))
    c1 = 0.5 * (1 + PB["eps_out"] / PB["eps_in"])
    c2 = 0.5 * (1 + PB["eps_in"] / PB["eps_out"])
    expect = np.concatenate([x[:n] + f0 / c1, x[n:] + f1 / c2])
    got = oracle.dmapping_direct(pts, nrm, w, x, **PB)
    assert rel(got, expect) < 1e-13


def test_rhs_formula(oracle):
    pts, nrm, w = G.shell(400, seed=10)
    rng = np.random.default_rng(7)
    atoms = rng.uniform(-1, 1, (6, 3))
    qa = rng.uniform(-1, 1, 6)
    b = oracle.rhs(pts, nrm, atoms, qa, 2.0, 80.0)
    n = len(pts)
    d = pts[:, None, :] - atoms[None, :, :]
    r = np.linalg.norm(d, axis=2)
    c1, c2 = 0.5 * (1 + 80.0 / 2.0), 0.5 * (1 + 2.0 / 80.0)
    b1 = (qa / (4 * np.pi * 2.0 * r)).sum(1) / c1
    b2 = -(qa * np.einsum("nad,nd->na", d, nrm) / (4 * np.pi * 2.0 * r ** 3)).sum(1) / c2
    assert rel(b, np.concatenate([b1, b2])) < 1e-14


def test_polarize_matches_direct_sum(oracle):
    pts, nrm, w = G.shell(1200, seed=11)
    rng = np.random.default_rng(8)
    atoms = rng.uniform(-1, 1, (20, 3))
    x = rng.uniform(-1, 1, 2 * len(pts))
    n = len(pts)
    q = {"x": x[:n] * nrm[:, 0] * w, "y": x[:n] * nrm[:, 1] * w, "z": x[:n] * nrm[:, 2] * w, "p": x[n:] * w}
    direct = oracle.direct("K2", pts, atoms, q["p"], **PB) - sum(
        oracle.direct("K1" + a, pts, atoms, q[a], **PB) for a in "xyz")
    got = oracle.polarize(pts, nrm, w, atoms, x, p=7, max_point=24, max_level=10, **PB)
    assert rel(got, direct) < 1e-4
