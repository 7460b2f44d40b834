"""Pins the oracle's 16 kernel closures (electric.cpp:67-291) against
derivatives of the Laplace and Yukawa Green's functions, taken by central
finite differences of closed forms that the oracle does not use:

  G0 = 1/(4 pi r),  Gk = exp(-kappa r)/(4 pi r),  f = (1 - exp(-kappa r))/r
  K1a = d/da_a (G0 - eps Gk)        K4a = -d/da_a (G0 - Gk/eps)
  K2  = G0 - Gk                     K3ab = d^2 f / da_a db_b   (no 1/(4 pi))
with a = source, b = target, eps = eps_out/eps_in.
"""
import numpy as np
import pytest
from scipy import integrate

KAP, EI, EO = 0.1257, 2.0, 80.0
EPS = EO / EI


def G0(a, b):
    return 1.0 / (4 * np.pi * np.linalg.norm(a - b))


def Gk(a, b, kap=KAP):
    r = np.linalg.norm(a - b)
    return np.exp(-kap * r) / (4 * np.pi * r)


def f(a, b, kap=KAP):
    r = np.linalg.norm(a - b)
    return (1 - np.exp(-kap * r)) / r


def grad_a(fun, a, b, h=1e-5):
    g = np.zeros(3)
    for i in range(3):
        e = np.zeros(3)
        e[i] = h
        g[i] = (fun(a + e, b) - fun(a - e, b)) / (2 * h)
    return g


def hess_ab(fun, a, b, h=1e-4):
    H = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            ei = np.zeros(3)
            ej = np.zeros(3)
            ei[i] = h
            ej[j] = h
            H[i, j] = (fun(a + ei, b + ej) - fun(a + ei, b - ej) - fun(a - ei, b + ej) + fun(a - ei, b - ej)) / (4 * h * h)
    return H


POINTS = [(np.array([0.3, -0.2, 0.5]), np.array([-0.4, 0.6, -0.1])),
          (np.array([1.5, 0.1, -2.0]), np.array([0.2, 0.3, 0.4])),
          (np.array([0.01, 0.02, 0.9]), np.array([0.0, 0.0, 0.0]))]
KW = dict(kappa=KAP, eps_in=EI, eps_out=EO, vacant_radius=1e-3)


@pytest.mark.parametrize("a,b", POINTS)
def test_k1_k4_are_gradients(oracle, a, b):
    g1 = grad_a(lambda x, y: G0(x, y) - EPS * Gk(x, y), a, b)
    g4 = -grad_a(lambda x, y: G0(x, y) - Gk(x, y) / EPS, a, b)
    for i, ax in enumerate("xyz"):
        assert oracle.kernel_eval("K1" + ax, a, b, **KW) == pytest.approx(g1[i], rel=1e-7, abs=1e-12)
        assert oracle.kernel_eval("K4" + ax, a, b, **KW) == pytest.approx(g4[i], rel=1e-7, abs=1e-12)


@pytest.mark.parametrize("a,b", POINTS)
def test_k2_and_k3(oracle, a, b):
    assert oracle.kernel_eval("K2", a, b, **KW) == pytest.approx(G0(a, b) - Gk(a, b), rel=1e-13)
    H = hess_ab(f, a, b)
    names = {(0, 0): ["K3xx"], (1, 1): ["K3yy"], (2, 2): ["K3zz"], (0, 1): ["K3xy", "K3yx"],
             (1, 2): ["K3yz", "K3zy"], (0, 2): ["K3zx", "K3xz"]}
    for (i, j), ks in names.items():
        for k in ks:
            assert abs(oracle.kernel_eval(k, a, b, **KW) - H[i, j]) <= 1e-6 * np.abs(H).max(), k


def test_special_cases(oracle):
    a, b = POINTS[0]
    # kappa = 0: K2 = K3 = 0
    kw0 = dict(kappa=0.0, eps_in=EI, eps_out=EO, vacant_radius=1e-3)
    assert oracle.kernel_eval("K2", a, b, **kw0) == 0.0
    for k in ["K3xx", "K3xy", "K3zz"]:
        assert abs(oracle.kernel_eval(k, a, b, **kw0)) < 1e-15
    # eps_in = eps_out and kappa = 0: every kernel vanishes (A = I)
    kwi = dict(kappa=0.0, eps_in=3.0, eps_out=3.0, vacant_radius=1e-3)
    for k in oracle.KERNELS:
        assert abs(oracle.kernel_eval(k, a, b, **kwi)) < 1e-15


def test_regularisation_below_vacant_radius(oracle):
    """Below tau*h: K1 = K3 = K4 = 0 and K2 is the average of K2 over a disk
    of radius tau*h (electric.cpp:91-101), checked by quadrature."""
    vr = 0.3
    a = np.zeros(3)
    b = np.array([0.1, 0.05, 0.0])
    kw = dict(kappa=KAP, eps_in=EI, eps_out=EO, vacant_radius=vr)
    for k in oracle.KERNELS:
        if k != "K2":
            assert oracle.kernel_eval(k, a, b, **kw) == 0.0
    avg = integrate.quad(lambda r: (1 - np.exp(-KAP * r)) / (4 * np.pi * r) * 2 * np.pi * r, 0, vr)[0] / (np.pi * vr * vr)
    assert oracle.kernel_eval("K2", a, b, **kw) == pytest.approx(avg, rel=1e-12)
    assert oracle.kernel_eval("K2", a, a, **kw) == pytest.approx(avg, rel=1e-12)  # the self pair
    assert oracle.kernel_eval("K2", a, b, kappa=0.0, eps_in=EI, eps_out=EO, vacant_radius=vr) == 0.0
