"""Pins the oracle's Chebyshev operators (bbfmm.h:273-433) against closed
forms and polynomial reproduction, independent of how the oracle computes
them."""
import numpy as np
import pytest


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6, 8])
def test_nodes_closed_form(oracle, p):
    x = oracle.cheb_nodes(p)
    theta = (np.arange(p) + 0.5) * np.pi / p
    np.testing.assert_allclose(x, -np.cos(theta), rtol=0, atol=1e-15)
    assert np.all(np.diff(x) > 0), "nodes ascend from -1 to 1"


@pytest.mark.parametrize("p", [3, 5, 8])
def test_chebyshev_polynomials_are_cos_k_theta(oracle, p):
    theta = np.linspace(0.01, np.pi - 0.01, 37)
    T = oracle.cheb_poly(p, np.cos(theta))
    for k in range(p):
        np.testing.assert_allclose(T[:, k], np.cos(k * theta), atol=1e-13)


@pytest.mark.parametrize("p", [3, 4, 6])
def test_sp_is_cardinal_at_nodes(oracle, p):
    x = oracle.cheb_nodes(p)
    S = oracle.sp(p, x)
    np.testing.assert_allclose(S, np.eye(p), atol=1e-14)


@pytest.mark.parametrize("p", [3, 4, 6])
def test_sp_reproduces_polynomials(oracle, p):
    """sum_m S_p(x, xbar_m) f(xbar_m) = f(x) for deg f < p, also outside [-1,1]
    (the extrapolation used for targets outside the root box)."""
    rng = np.random.default_rng(p)
    xb = oracle.cheb_nodes(p)
    x = np.linspace(-1.6, 1.6, 29)
    S = oracle.sp(p, x)
    for deg in range(p):
        c = rng.normal(size=deg + 1)
        f = np.polynomial.polynomial.polyval
        np.testing.assert_allclose(S @ f(xb, c), f(x, c), atol=1e-12 * (1 + np.abs(f(x, c)).max()))
    # a degree-p polynomial is NOT reproduced: the check can fail
    assert np.abs(S @ xb ** p - x ** p).max() > 1e-3


@pytest.mark.parametrize("p", [3, 4])
def test_transfer_R_interpolates_parent_to_child(oracle, p):
    """R[id] maps values at the parent's tensor grid to the child's grid for
    any polynomial of degree < p per axis; child id bits (x,y,z) = (0,1,2),
    index = z*p^2 + y*p + x (bbfmm.h:328-367, 456-479)."""
    rng = np.random.default_rng(1)
    xb = oracle.cheb_nodes(p)
    R = oracle.transfer_R(p)
    coef = rng.normal(size=(p, p, p))  # f(x,y,z) = sum c_abc x^a y^b z^c

    def f(x, y, z):
        return sum(coef[a, b, c] * x ** a * y ** b * z ** c for a in range(p) for b in range(p) for c in range(p))

    Zp, Yp, Xp = np.meshgrid(xb, xb, xb, indexing="ij")  # flat index z*p^2 + y*p + x
    parent_vals = f(Xp, Yp, Zp).ravel()
    for cid in range(8):
        bx, by, bz = cid & 1, (cid >> 1) & 1, (cid >> 2) & 1
        off = lambda b: 0.5 if b else -0.5
        cx, cy, cz = 0.5 * xb + off(bx), 0.5 * xb + off(by), 0.5 * xb + off(bz)
        Zc, Yc, Xc = np.meshgrid(cz, cy, cx, indexing="ij")
        child_vals = f(Xc, Yc, Zc).ravel()
        np.testing.assert_allclose(R[cid] @ parent_vals, child_vals, atol=1e-11 * np.abs(child_vals).max())
