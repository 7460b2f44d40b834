"""Seeded point sets for tests: surface-like shells, boxes, clusters, and the
tube of a sphere on a grid (the IBIM quadrature layout of levelset.cpp:437-586
with the analytic distance function)."""
import numpy as np


def shell(n, radius=2.0, thickness=0.2, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    rad = radius + thickness * (rng.random((n, 1)) - 0.5)
    pts = v * rad
    w = 0.01 + 0.01 * rng.random(n)
    return pts, v.copy(), w


def cube(n, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    pts = scale * (rng.random((n, 3)) - 0.5)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    w = 0.01 + 0.01 * rng.random(n)
    return pts, nrm, w


def clusters(n, seed=0):
    """Strongly non-uniform: a few tight blobs plus a sparse background, so
    the tree has deep and shallow leaves side by side (W and X lists)."""
    rng = np.random.default_rng(seed)
    centres = rng.random((4, 3)) * 4 - 2
    k = n // 5
    parts = [c + 0.05 * rng.normal(size=(k, 3)) for c in centres]
    parts.append(rng.random((n - 4 * k, 3)) * 4 - 2)
    pts = np.concatenate(parts)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    w = 0.01 + 0.01 * rng.random(n)
    return pts, nrm, w


def sphere_tube(a=2.0, h=0.125, lo=-4.0, npts=64, thickness=2):
    """Tube nodes of the sphere |x| = a on the grid x_i = lo + i h: projection
    P - n phi, unit normal, cosine weight times h^3 (levelset.cpp:552-576,
    electric.cpp:33)."""
    g = lo + h * np.arange(npts)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    r = np.linalg.norm(P, axis=1)
    phi = r - a
    tube = thickness * h
    m = np.abs(phi) < tube
    P, r, phi = P[m], r[m], phi[m]
    n = P / r[:, None]
    nodes = P - n * phi[:, None]
    w = 0.5 * (1.0 + np.cos(np.pi * phi / tube)) / tube * h ** 3
    return nodes, n, w
