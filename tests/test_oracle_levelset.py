"""Pins the oracle's level-set pre-step (levelset.cpp:91-586, main.cpp:43-62,
electric.cpp:17-41) against what the mathematics fixes:

* setGradient (WENO5) is exact on quadratics, whose smoothness indicators
  are all equal, and fifth-order on a smooth function: a wrong stencil
  coefficient or a sign on a correction term fails one of the two;
* getNorm is |grad u| for a linear field and upwinds at a kink: the
  minimum of |x| has norm 0 with (Dun, Dup) and 1 with (Dup, Dun);
* evolve moves a distance function by vel * t exactly, and solves
  u_t + |u_x| = 0 on a quadratic, whose solution is quadratic in t (a
  first-order time step fails it);
* reinitialize leaves a distance function untouched (zero steps) and turns a
  distorted level set with the same zero set into the distance function;
* Surface: nodes of a plane lie on it with the exact normal and a weight sum
  of exactly 1 per grid column (eight samples of the raised cosine span its
  period); a sphere's weights sum to 4 pi (R^2 + eps^2 (1/3 - 2/pi^2)), the
  kernel-weighted area of the parallel spheres; the scale converts to the
  atoms' unit; a cavity behind a shell thicker than the tube is pushed out
  of the tube, point for point; the fill starts at grid point 0 only if
  that point lies above -tube;
* the whole pipeline on one atom gives the sphere of the atom's radius, and
  on two atoms farther apart than the probe's diameter two such spheres.
"""
import math

import numpy as np
import pytest

C_EPS2 = 1.0 / 3.0 - 2.0 / math.pi ** 2  # <d^2> / eps^2 under the raised-cosine kernel


def sphere_area(R, eps):
    return 4 * math.pi * (R * R + C_EPS2 * eps * eps)


def test_weno_exact_on_quadratics(oracle):
    rng = np.random.default_rng(0)
    lp = oracle.ls_params((13, 12, 11), (-1.0, 0.5, -0.3), 0.17)
    P = oracle.grid_points(lp)
    A = rng.normal(size=(3, 3))
    A = A + A.T
    b = rng.normal(size=3)
    g = 0.5 * np.einsum("...i,ij,...j->...", P, A, P) + P @ b + 0.3
    for (i, j, k) in [(3, 3, 3), (6, 5, 4), (9, 8, 7), (5, 4, 6)]:
        up, un = oracle.ls_gradient(lp, g, i, j, k)
        exact = A @ P[i, j, k] + b
        assert np.allclose(up, exact, rtol=0, atol=1e-11), (up, exact)
        assert np.allclose(un, exact, rtol=0, atol=1e-11), (un, exact)


def test_weno_fifth_order(oracle):
    errs = []
    for h in (0.2, 0.1, 0.05):
        n = int(round(2 / h)) + 1
        lp = oracle.ls_params((n, 7, 7), (-1.0, 0, 0), h)
        P = oracle.grid_points(lp)
        g = np.sin(1.3 * P[..., 0] + 0.4)
        i = n // 2
        up, un = oracle.ls_gradient(lp, g, i, 3, 3)
        exact = 1.3 * math.cos(1.3 * P[i, 3, 3, 0] + 0.4)
        errs.append(max(abs(up[0] - exact), abs(un[0] - exact)))
        assert abs(up[1]) < 1e-14 and abs(un[2]) < 1e-14
    for e0, e1 in zip(errs, errs[1:]):
        assert e0 / e1 > 2 ** 4.3, errs


def test_godunov_norm(oracle):
    rng = np.random.default_rng(1)
    for _ in range(5):
        v = rng.normal(size=3)
        assert oracle.ls_norm(v, v) == pytest.approx(np.linalg.norm(v), rel=1e-15)
    # the kink of |x - x0|: backward derivative -1, forward +1
    lp = oracle.ls_params((9, 7, 7), (0, 0, 0), 0.5)
    P = oracle.grid_points(lp)
    g = np.abs(P[..., 0] - 2.0)
    up, un = oracle.ls_gradient(lp, g, 4, 3, 3)
    assert up[0] == pytest.approx(1.0) and un[0] == pytest.approx(-1.0)
    assert oracle.ls_norm(un, up) == 0.0
    assert oracle.ls_norm(up, un) == pytest.approx(1.0)


def test_evolve_moves_a_distance_function(oracle):
    lp = oracle.ls_params((41, 41, 41), (-5, -5, -5), 0.25)
    P = oracle.grid_points(lp)
    r = np.linalg.norm(P, axis=-1)
    g, steps = oracle.ls_evolve(lp, 4.0 - r, 1.0, vel=1.0, cfl=0.2)
    assert steps == int(1.0 / (0.2 * 0.25)) + 1
    band = (r >= 1.0) & (r <= 3.0)
    assert np.abs(g - (3.0 - r))[band].max() < 5e-4
    # the same distance at half the speed over twice the time
    g2, _ = oracle.ls_evolve(lp, 4.0 - r, 2.0, vel=0.5, cfl=0.2)
    assert np.abs(g2 - (3.0 - r))[band].max() < 5e-4


def test_evolve_quadratic_time_order(oracle):
    """u0 = (x + 1)^2 / 2 under u_t + |u_x| = 0 is u = (x + 1 - t)^2 / 2 while
    u_x > 0; WENO is exact on it, so only the time integration errs."""
    h = 0.1
    lp = oracle.ls_params((101, 7, 7), (0.0, 0, 0), h)
    P = oracle.grid_points(lp)
    x = P[..., 0]
    g, steps = oracle.ls_evolve(lp, 0.5 * (x + 1) ** 2, 0.5, vel=1.0, cfl=0.2)
    mid = (x > 3.0) & (x < 6.0)
    err = np.abs(g - 0.5 * (x + 1 - 0.5) ** 2)[mid].max()
    # one forward-Euler step per dt would leave 0.5 * t * dt = 2.5e-3
    assert err < 1e-9, err


def test_reinit_keeps_distance_and_repairs_distortion(oracle):
    lp = oracle.ls_params((41, 41, 41), (-5, -5, -5), 0.25, thickness=4, thres=0.05)
    P = oracle.grid_points(lp)
    r = np.linalg.norm(P, axis=-1)
    d = r - 3.0
    # a distance function: no tube point is bad, so no step is taken
    g, steps, q = oracle.ls_reinitialize(lp, d, 2.0)
    assert steps == 0 and np.array_equal(g, d) and q[1] == 0.0
    # the same zero set with a gradient between 0.6 and 1.4
    phi0 = d * (1 + 0.4 * P[..., 0] / 5)
    g, steps, q = oracle.ls_reinitialize(lp, phi0, 2.0)
    assert 0 < steps < int(2.0 / (0.5 * 0.25)) + 1 and q[1] == 0.0
    tube = np.abs(g) < 1.0
    assert np.abs(g - d)[tube].max() < 0.02
    assert np.abs(g - d)[np.abs(d) < 0.3].max() < 0.006


def test_surface_plane(oracle):
    lp = oracle.ls_params((24, 24, 32), (0, 0, -4), 0.25, thickness=4)
    P = oracle.grid_points(lp)
    xyz, nrm, w, inc = oracle.surface_from_grid(lp, 0.1 - P[..., 2])
    assert inc == 0 and len(w) == 24 * 24 * 8
    assert np.abs(xyz[:, 2] - 0.1).max() < 1e-14
    assert np.abs(nrm - [0, 0, -1]).max() < 1e-14
    # per column: sum_k h K(d_k) = 1 exactly; times the column's area h^2
    assert w.sum() == pytest.approx(24 * 24 * 0.25 ** 2, rel=1e-13)


def test_surface_sphere_and_scale(oracle):
    R, h, T = 4.0, 0.25, 4
    lp = oracle.ls_params((48, 48, 48), (-6, -6, -6), h, thickness=T)
    P = oracle.grid_points(lp)
    r = np.linalg.norm(P, axis=-1)
    xyz, nrm, w, inc = oracle.surface_from_grid(lp, r - R)
    assert inc == 0 and len(w) == int((np.abs(r - R) < T * h).sum())
    assert np.abs(np.linalg.norm(xyz, axis=1) - R).max() < 1e-5
    assert np.abs(nrm - xyz / np.linalg.norm(xyz, axis=1)[:, None]).max() < 1e-4
    assert w.sum() == pytest.approx(sphere_area(R, T * h), rel=2e-5)
    # the same sphere on a grid in units scaled by s = 2
    s = 2.0
    lp2 = oracle.ls_params((49, 49, 49), (-12, -12, -12), 2 * h, thickness=T, scale=s)
    r2 = np.linalg.norm(oracle.grid_points(lp2), axis=-1)
    xyz2, _, w2, _ = oracle.surface_from_grid(lp2, r2 - s * R)
    assert np.abs(np.linalg.norm(xyz2, axis=1) - R).max() < 1e-5
    assert w2.sum() == pytest.approx(sphere_area(R, T * h), rel=2e-5)


def test_surface_cavity_pushed_out(oracle):
    lp = oracle.ls_params((49, 49, 49), (-6, -6, -6), 0.25, thickness=4)
    r = np.linalg.norm(oracle.grid_points(lp), axis=-1)
    phi = np.maximum(r - 5.0, 2.0 - r)  # material between radii 2 and 5, negative inside it
    xyz, nrm, w, inc = oracle.surface_from_grid(lp, phi)
    tube = np.abs(phi) < 1.0
    assert inc == int((tube & (r < 3.5)).sum()) > 0
    assert len(w) == int((tube & (r > 3.5)).sum())
    assert np.abs(np.linalg.norm(xyz, axis=1) - 5.0).max() < 1e-5


def test_flood_starts_from_point_zero(oracle):
    """The fill spreads from a reached point with phi > -tube (levelset.cpp:446-524):
    when grid point 0 itself lies deeper than the tube nothing is reached and
    every tube point is pushed out, even though its neighbours are outside."""
    h, T = 0.25, 4
    lp = oracle.ls_params((12, 12, 12), (0, 0, 0), h, thickness=T)
    P = oracle.grid_points(lp)
    s = P.sum(axis=-1)
    for c, reached in [(T * h + 0.5 * h, False), (T * h - 0.5 * h, True)]:
        phi = s - c
        tube = int((np.abs(phi) < T * h).sum())
        xyz, nrm, w, inc = oracle.surface_from_grid(lp, phi)
        assert (len(w), inc) == ((tube, 0) if reached else (0, tube))


def test_pipeline_one_atom(oracle):
    h, T, rad = 0.25, 4, 2.0
    lp = oracle.ls_params((48, 48, 48), (-6, -6, -6), h, probe=1.4, band=8, thickness=T, thres=0.05,
                          reinit_time=2.0)
    c = np.array([0.01, 0.02, -0.03])
    xyz, nrm, w, info, g = oracle.surface_from_atoms(lp, [c], [rad], -2.0 * 48)
    assert info.evolve_steps == int(1.4 / (0.2 * h)) + 1
    assert info.inclusions == 0 and info.quality[1] == 0.0
    # nodes at the tube's edge sit off by up to (|grad phi| - 1) phi <= thres * tube
    assert np.abs(np.linalg.norm(xyz - c, axis=1) - rad).max() < 0.05 * T * h + 0.02
    assert w.sum() == pytest.approx(sphere_area(rad, T * h), rel=1e-3)


def test_pipeline_two_separate_atoms(oracle):
    """Atoms 8 apart with radius 2: the probe (1.4) passes between them, so the
    solvent-excluded surface is the two spheres."""
    h, T, rad = 0.25, 4, 2.0
    lp = oracle.ls_params((65, 41, 41), (-8, -5, -5), h, probe=1.4, band=8, thickness=T, thres=0.05,
                          reinit_time=2.0)
    c = np.array([[-4.0, 0.02, 0.01], [4.0, -0.01, 0.03]])
    xyz, nrm, w, info, g = oracle.surface_from_atoms(lp, c, [rad, rad], -2.0 * 65)
    d = np.min(np.linalg.norm(xyz[:, None, :] - c[None], axis=2), axis=1)
    assert np.abs(d - rad).max() < 0.05 * T * h + 0.02
    assert w.sum() == pytest.approx(2 * sphere_area(rad, T * h), rel=1e-3)
