"""The benchmark configurations on the GPU, through the C ABI.

* Born ion (the analytic 64^3 sphere tube, N = 12,790, p = 4, leaf 64,
  kappa = 0): the FP64 matvec against the oracle run here, and the solve's
  energy against the oracle's solve at tol 1e-12 (tests/golden/born_ion.json,
  written by tests/golden/make_born_ion.py); bar 1e-10 relative.
* 500-atom globule (h = 0.25 A, 847k tube points, p = 5, leaf 192): the FP64
  matvec on one octant of its surface (~101k points) against the oracle at
  1e-10; at full size, FP64 and FP16-mixed solves converge, their energies
  agree within 1e-6, the FP16 matvec repeats bit for bit, and the operator is
  exactly the identity when eps_in = eps_out and kappa = 0.
* 5000-atom globule (45 A sphere, h = 0.25 A, 8.3M tube points, p = 6, leaf
  256): FP64 and FP16-mixed solves converge and agree within 1e-6.

The reference path these stand in for is dmapping, load, GMRES and
polarizeMap (electric.cpp:294-505)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import geometry as G

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BORN = dict(kappa=0.0, eps_in=1.0, eps_out=80.0, vacant_radius=0.125)
GLOBULE = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def born_golden():
    with open(os.path.join(HERE, "golden", "born_ion.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def born_case():
    pts, nrm, w = G.sphere_tube(a=2.0, h=0.125, lo=-4.0, npts=64)
    assert len(pts) == 12790
    return pts, nrm, w


def test_born_ion_matvec_matches_oracle(ibim, oracle, gpu, born_case, born_golden):
    pts, nrm, w = born_case
    x = np.random.default_rng(born_golden["matvec_seed"]).uniform(-1, 1, 2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=4, max_points=64, max_level=10, **BORN) as plan:
        y = plan.matvec(x)
    yo = oracle.dmapping(pts, nrm, w, x, p=4, max_point=64, max_level=10, **BORN)
    print(f"born ion matvec: rel err vs oracle {rel(y, yo):.2e}")
    assert rel(y, yo) < 1e-10
    # the oracle here is the one that wrote the fixture
    assert abs(np.linalg.norm(yo) - born_golden["matvec_norm"]) <= 1e-12 * born_golden["matvec_norm"]
    assert rel(y[:8], born_golden["matvec_first"]) < 1e-10


def test_born_ion_solve_energy_matches_oracle(ibim, gpu, born_case, born_golden):
    pts, nrm, w = born_case
    atom, q = np.zeros((1, 3)), np.ones(1)
    with ibim.Plan(pts, nrm, w, p=4, max_points=64, max_level=10, **BORN) as plan:
        b = plan.rhs(atom, q)
        x, rc, info = plan.gmres(b, restart=20, maxit=400, tol=1e-12)
        e = 0.5 * float(plan.polarize(atom, x) @ q)
    g = born_golden
    born = 1.0 / (8 * np.pi * 2.0) * (1 / 80.0 - 1.0)
    print(f"born ion: E {e:.12e} oracle {g['energy']:.12e} ({abs(e - g['energy']) / abs(g['energy']):.1e}), "
          f"its {info['iterations']} / {g['iterations']}, analytic {born:.6e}")
    assert rc == 0 and info["converged"] == 1
    assert abs(np.linalg.norm(b) - g["b_norm"]) <= 1e-12 * g["b_norm"]
    assert abs(info["iterations"] - g["iterations"]) <= 1
    assert abs(e - g["energy"]) <= 1e-10 * abs(g["energy"])
    assert abs(np.linalg.norm(x) - g["x_norm"]) <= 1e-10 * g["x_norm"]
    assert rel(x[:8], g["x_first"]) < 1e-10
    assert abs(e - born) < 0.03 * abs(born)  # the discretisation error, O(h)


def globule_surface(ibim, atoms, seed, confine, h):
    xyz, rad, q = ibim.synthetic_globule(atoms, seed, confine)
    return xyz, q, ibim.Surface.from_atoms(xyz, rad, h=h, probe=1.4, thickness=2)


@pytest.fixture(scope="module")
def globule500(ibim, gpu):
    xyz, q, surf = globule_surface(ibim, 500, 500, 20.0, 0.25)
    print(f"[case] globule500: {surf.n} tube points", flush=True)
    yield xyz, q, surf
    surf.close()


def test_globule500_octant_matvec_matches_oracle(ibim, oracle, globule500):
    xyz, q, surf = globule500
    pts, nrm, w = surf.arrays()
    c = np.median(pts, axis=0)
    sel = np.all(pts > c, axis=1)
    pts, nrm, w = pts[sel], nrm[sel], w[sel]
    prm = dict(GLOBULE, vacant_radius=0.25)
    x = np.random.default_rng(1).uniform(-1, 1, 2 * len(pts))
    print(f"[case] globule500 octant: {len(pts)} points", flush=True)
    with ibim.Plan(pts, nrm, w, p=5, max_points=192, max_level=10, **prm) as plan:
        y = plan.matvec(x)
        st = plan.stats()
    yo = oracle.dmapping(pts, nrm, w, x, p=5, max_point=192, max_level=10, **prm)
    print(f"globule500 octant: rel err vs oracle {rel(y, yo):.2e} ({st['v_pairs']} V, {st['x_pairs']} X)")
    assert len(pts) > 90000 and st["v_pairs"] > 0 and st["x_pairs"] > 0
    assert rel(y, yo) < 1e-10


def device_solve(ibim, torch, plan, xyz, q, restart=40, tol=1e-8):
    """RHS, GMRES and the solvation energy with device-resident vectors."""
    dev = torch.device("cuda:0")
    n, m = plan.n, len(q)
    d_at = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev)
    d_q = torch.from_numpy(np.ascontiguousarray(q)).to(dev)
    d_b = torch.empty(2 * n, dtype=torch.float64, device=dev)
    d_x = torch.zeros(2 * n, dtype=torch.float64, device=dev)
    d_phi = torch.empty(m, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    P = lambda t: C.c_void_p(t.data_ptr())
    plan.rhs_device(m, P(d_at), P(d_q), P(d_b))
    rc, info = plan.gmres_device(P(d_b), P(d_x), restart=restart, maxit=2000, tol=tol)
    plan.polarize_device(m, P(d_at), P(d_x), P(d_phi))
    plan.sync()
    return rc, info, 0.5 * float(torch.dot(d_phi, d_q).item()), d_b, d_x


def true_residual(torch, plan, ibim, d_b, d_x):
    """|b - A x| / |b| with the FP64 operator."""
    P = lambda t: C.c_void_p(t.data_ptr())
    y = torch.empty_like(d_b)
    plan.set_precision(ibim.PREC_FP64)
    plan.matvec_device(P(d_x), P(y))
    plan.sync()
    return float(torch.linalg.norm(d_b - y) / torch.linalg.norm(d_b))


def test_globule500_full_size(ibim, globule500):
    import torch
    xyz, q, surf = globule500
    ptrs = surf.device_ptrs()
    kw = dict(p=5, max_points=192, max_level=10, vacant_radius=0.25, **GLOBULE)
    out = {}
    for name, prec in (("fp64", ibim.PREC_FP64), ("mixed16", ibim.PREC_MIXED_F16)):
        with ibim.Plan(device_ptrs=ptrs, precision=prec, **kw) as plan:
            rc, info, e, d_b, d_x = device_solve(ibim, torch, plan, xyz, q)
            res = true_residual(torch, plan, ibim, d_b, d_x)
        out[name] = (rc, info, e, res)
        print(f"globule500 {name}: E {e:.12e}, its {info['iterations']}, outer {info['outer_steps']}, "
              f"FP64 residual {res:.1e}")
    (rc64, i64, e64, r64), (rc16, i16, e16, r16) = out["fp64"], out["mixed16"]
    assert rc64 == 0 and rc16 == 0 and i64["converged"] == 1 and i16["converged"] == 1
    assert r64 < 1e-8 and r16 < 1e-8
    assert abs(e16 - e64) <= 1e-6 * abs(e64)
    # the FP16 matvec repeats bit for bit
    P = lambda t: C.c_void_p(t.data_ptr())
    dev = torch.device("cuda:0")
    x = torch.from_numpy(np.random.default_rng(2).uniform(-1, 1, 2 * surf.n)).to(dev)
    y1, y2 = torch.empty_like(x), torch.empty_like(x)
    torch.cuda.synchronize()
    with ibim.Plan(device_ptrs=ptrs, precision=ibim.PREC_MIXED_F16, **kw) as plan:
        plan.matvec_device(P(x), P(y1))
        plan.matvec_device(P(x), P(y2))
        plan.sync()
    assert torch.equal(y1, y2)
    # eps_in = eps_out and kappa = 0: every kernel vanishes and A = I exactly
    with ibim.Plan(device_ptrs=ptrs, p=5, max_points=192, max_level=10, kappa=0.0, eps_in=2.0, eps_out=2.0,
                   vacant_radius=0.25) as plan:
        plan.matvec_device(P(x), P(y1))
        plan.sync()
    assert torch.equal(y1, x)


def test_globule5000_full_size(ibim, gpu):
    import torch
    xyz, q, surf = globule_surface(ibim, 5000, 5000, 45.0, 0.25)
    try:
        print(f"[case] globule5000: {surf.n} tube points", flush=True)
        ptrs = surf.device_ptrs()
        kw = dict(p=6, max_points=256, max_level=10, vacant_radius=0.25, **GLOBULE)
        out = {}
        for name, prec in (("mixed16", ibim.PREC_MIXED_F16), ("fp64", ibim.PREC_FP64)):
            with ibim.Plan(device_ptrs=ptrs, precision=prec, **kw) as plan:
                rc, info, e, d_b, d_x = device_solve(ibim, torch, plan, xyz, q)
                res = true_residual(torch, plan, ibim, d_b, d_x)
            del d_b, d_x
            out[name] = (rc, info, e, res)
            print(f"globule5000 {name}: E {e:.12e}, its {info['iterations']}, residual {res:.1e}", flush=True)
        (rc64, i64, e64, r64), (rc16, i16, e16, r16) = out["fp64"], out["mixed16"]
        assert surf.n > 5_000_000
        assert rc64 == 0 and rc16 == 0 and r64 < 1e-8 and r16 < 1e-8
        assert abs(e16 - e64) <= 1e-6 * abs(e64)
    finally:
        surf.close()
