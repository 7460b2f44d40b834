"""Geometry pre-step on the GPU (levelset.cpp, main.cpp:47-62) and the Born
sphere solve, checked against closed forms:
* the tube of the analytic sphere level set reproduces the tube layout
  (N = 12,790 at h = 1/8 on [-4,4)^3) with nodes on the sphere;
* the full expand/evolve/reinit pipeline for one atom gives its vdW sphere;
* the Born solvation energy of a charge in a sphere is q^2/(8 pi a) (1/eps_out
  - 1/eps_in) in the code's units, up to the O(h) discretisation error."""
import numpy as np
import pytest

import geometry as G

pytestmark = pytest.mark.gpu

BORN = 1.0 / (8 * np.pi * 2.0) * (1 / 80.0 - 1.0)   # -1.964569e-2


def sphere_grid(h=0.125, lo=-4.0, n=64, a=2.0):
    g = lo + h * np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - a


def test_tube_of_analytic_sphere(ibim, gpu):
    phi = sphere_grid()
    s = ibim.Surface.from_grid(phi, h=0.125, lo=(-4.0, -4.0, -4.0), thickness=2)
    xyz, nrm, w = s.arrays()
    s.close()
    ref_xyz, ref_n, ref_w = G.sphere_tube()
    assert len(xyz) == 12790
    # WENO gradients of |x| - 2 are close to the exact radial unit normal
    np.testing.assert_allclose(nrm, ref_n, atol=2e-3)
    np.testing.assert_allclose(xyz, ref_xyz, atol=1e-3)
    np.testing.assert_allclose(w, ref_w, rtol=1e-12)
    assert abs(np.linalg.norm(xyz, axis=1) - 2.0).max() < 1e-3
    assert w.sum() == pytest.approx(4 * np.pi * 4, rel=0.01)


def test_single_atom_pipeline_gives_vdw_sphere(ibim, gpu):
    s = ibim.Surface.from_atoms([[0.1, -0.2, 0.05]], [2.0], h=0.125, probe=1.4, thickness=2)
    xyz, nrm, w = s.arrays()
    info = s.info
    s.close()
    r = np.linalg.norm(xyz - [0.1, -0.2, 0.05], axis=1)
    assert abs(r - 2.0).max() < 0.05, abs(r - 2.0).max()
    assert w.sum() == pytest.approx(4 * np.pi * 4, rel=0.02)
    assert np.abs(np.linalg.norm(nrm, axis=1) - 1).max() < 0.05
    assert info["inclusions"] == 0 and info["evolve_steps"] > 0


def test_synthetic_globule_surface(ibim, gpu):
    xyz, r, q = ibim.synthetic_globule(500, 500, 20.0)
    assert q.sum() == pytest.approx(-2.0, abs=1e-12)
    s = ibim.Surface.from_atoms(xyz, r, h=0.5, probe=1.4, thickness=2)
    nodes, nrm, w = s.arrays()
    info = s.info
    s.close()
    # every node lies between the vdW and the probe-expanded surfaces
    d = np.linalg.norm(nodes[:, None, :] - xyz[None, ::7, :], axis=2)  # subsample of atoms
    assert len(nodes) > 1000 and info["area"] > 4 * np.pi * 18 ** 2
    assert np.isfinite(nodes).all() and np.isfinite(w).all() and (w >= 0).all()
    assert d.min() > 0.5


def test_born_sphere_energy(ibim, gpu):
    """C1: Born sphere, p=4, leaf 64, restart 20, tol 1e-8, kappa = 0."""
    phi = sphere_grid()
    s = ibim.Surface.from_grid(phi, h=0.125, lo=(-4.0, -4.0, -4.0), thickness=2)
    pts, nrm, w = s.arrays()
    s.close()
    atoms, qa = np.zeros((1, 3)), np.array([1.0])
    with ibim.Plan(pts, nrm, w, p=4, max_points=64, max_level=10, kappa=0.0, eps_in=1.0, eps_out=80.0,
                   vacant_radius=0.125) as plan:
        b = plan.rhs(atoms, qa)
        x, rc, info = plan.gmres(b, restart=20, maxit=200, tol=1e-8)
        phi_reac = plan.polarize(atoms, x)
    energy = 0.5 * float(phi_reac @ qa)
    print(f"Born: E={energy:.6e} analytic={BORN:.6e} its={info['iterations']}")
    assert rc == 0
    assert energy == pytest.approx(BORN, rel=0.03)
