"""Mixed refinement whose inner solves use a lower Chebyshev order
(`inner_order`): the FP64 residual keeps the plan's order p, so the solve
ends at the order-p FP64 system's answer.  Bars: converged, the FP64
residual (measured again here with the order-p FP64 operator) below the
1e-8 tolerance, the solution within 1e-6 and the solvation energy within
1e-7 relative of the FP64 solve's (both solves stop at a residual below
1e-8; the observed differences are 1e-8 and below); the inner plan's stages
are in the plan's stage report; a precision switch reaches the companion
plan; served solves with a lower inner order give the FP64 energies within
1e-7, fused small solves too."""
import numpy as np
import pytest

import geometry as G

pytestmark = pytest.mark.gpu

PB = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0)


@pytest.fixture(scope="module")
def sphere_case(ibim, gpu):
    return G.sphere_tube(a=2.0, h=0.125, lo=-4.0, npts=64)


def solve(ibim, case, precision, inner_order=0, profile=False):
    pts, nrm, w = case
    atoms = np.array([[0.0, 0.0, 0.0], [0.4, -0.3, 0.2]])
    qa = np.array([1.0, -0.5])
    with ibim.Plan(pts, nrm, w, p=5, max_points=64, vacant_radius=0.125, precision=precision,
                   inner_order=inner_order, **PB) as plan:
        if profile:
            plan.profile(True)
        b = plan.rhs(atoms, qa)
        x, rc, info = plan.gmres(b, restart=20, maxit=300, tol=1e-8)
        rep = plan.stage_report() if profile else None
        fp = plan.footprint()
        plan.set_precision(ibim.PREC_FP64)
        res = np.linalg.norm(b - plan.matvec(x)) / np.linalg.norm(b)
        e = 0.5 * float(plan.polarize(atoms, x) @ qa)
    return x, e, rc, info, res, rep, fp


@pytest.mark.parametrize("mode,order", [("mixed16", 4), ("mixed16", 3), ("mixed", 4)])
def test_lower_inner_order_solve_matches_fp64(ibim, sphere_case, mode, order):
    low = {"mixed": ibim.PREC_MIXED, "mixed16": ibim.PREC_MIXED_F16}[mode]
    x64, e64, rc64, i64, r64, _, _ = solve(ibim, sphere_case, ibim.PREC_FP64)
    x, e, rc, info, res, rep, fp = solve(ibim, sphere_case, low, order, profile=True)
    print(f"{mode} inner order {order}: E64={e64:.12e} E={e:.12e} its {info['iterations']} outer "
          f"{info['outer_steps']} res {res:.2e} footprint {fp / 2**20:.1f} MiB")
    assert rc64 == 0 and rc == 0, info
    assert r64 < 1e-8 and res < 1e-8
    assert info["outer_steps"] >= 2
    assert abs(e - e64) <= 1e-7 * abs(e64)
    assert np.linalg.norm(x - x64) <= 1e-6 * np.linalg.norm(x64)
    # the inner plan's mixed matvecs are in the plan's report
    assert rep["MATVEC_MIXED"]["launches"] >= info["iterations"], rep["MATVEC_MIXED"]
    assert rep["MATVEC"]["launches"] >= info["outer_steps"]


def test_plan_with_inner_order_builds_its_own_mixed_matvec(ibim, sphere_case):
    """A plan whose inner solves use a lower order skips its own mixed
    operators until a mixed matvec asks for them; that matvec then equals an
    order-p FP16 plan's bit for bit, before and after a solve."""
    pts, nrm, w = sphere_case
    kw = dict(p=5, max_points=64, vacant_radius=0.125, precision=ibim.PREC_MIXED_F16, **PB)
    x = np.random.default_rng(5).uniform(-1, 1, 2 * len(pts))
    with ibim.Plan(pts, nrm, w, **kw) as ref:
        y_ref = ref.matvec(x)
    with ibim.Plan(pts, nrm, w, inner_order=4, **kw) as plan:
        fp0 = plan.footprint()
        y0 = plan.matvec(x)
        fp1 = plan.footprint()
        b = plan.rhs(np.array([[0.0, 0.0, 0.0]]), np.array([1.0]))
        _, rc, info = plan.gmres(b, restart=20, maxit=300, tol=1e-8)
        fp2 = plan.footprint()
        y1 = plan.matvec(x)
    print(f"footprint {fp0 / 2**20:.1f} -> {fp1 / 2**20:.1f} (own mixed operators) -> {fp2 / 2**20:.1f} MiB "
          f"(inner plan); {info}")
    assert rc == 0
    assert fp0 < fp1 < fp2
    assert np.array_equal(y0, y_ref) and np.array_equal(y1, y_ref)


def test_precision_switch_carries_to_the_inner_plan(ibim, sphere_case):
    """Switching a plan with a lower inner order from the FP16 to the FP32
    mixed form switches the companion plan built by its first solve: that
    plan builds its FP32 operators at once (the footprint grows), and the
    next solve still ends at the FP64 energy within 1e-7."""
    pts, nrm, w = sphere_case
    atoms = np.array([[0.0, 0.0, 0.0], [0.4, -0.3, 0.2]])
    qa = np.array([1.0, -0.5])
    _, e64, rc64, _, _, _, _ = solve(ibim, sphere_case, ibim.PREC_FP64)
    with ibim.Plan(pts, nrm, w, p=5, max_points=64, vacant_radius=0.125, precision=ibim.PREC_MIXED_F16,
                   inner_order=4, **PB) as plan:
        b = plan.rhs(atoms, qa)
        _, rc1, _ = plan.gmres(b, restart=20, maxit=300, tol=1e-8)
        fp1 = plan.footprint()
        plan.set_precision(ibim.PREC_MIXED)
        fp2 = plan.footprint()
        x, rc2, info = plan.gmres(b, restart=20, maxit=300, tol=1e-8)
        e = 0.5 * float(plan.polarize(atoms, x) @ qa)
    print(f"footprint {fp1 / 2**20:.1f} -> {fp2 / 2**20:.1f} MiB; {info}; E64={e64:.12e} E={e:.12e}")
    assert rc64 == 0 and rc1 == 0 and rc2 == 0, info
    assert fp2 > fp1
    assert info["outer_steps"] >= 2
    assert abs(e - e64) <= 1e-7 * abs(e64)


def test_inner_order_rejected_out_of_range(ibim, sphere_case):
    pts, nrm, w = sphere_case
    for order in (1, 5, 6):
        with pytest.raises(ibim.IbimError, match="inner_order"):
            ibim.Plan(pts, nrm, w, p=5, max_points=64, vacant_radius=0.125, precision=ibim.PREC_MIXED_F16,
                      inner_order=order, **PB)


@pytest.mark.parametrize("fuse", [False, True])
def test_served_lower_inner_order_matches_fp64(ibim, gpu, fuse):
    mols = [ibim.synthetic_globule(m, 3100 + m, R) for m, R in [(40, 8.0), (25, 7.0), (60, 9.0)]]
    fmm = dict(p=5, max_points=64, max_level=10, vacant_radius=0.5, **PB)
    with ibim.Server(slots=1) as srv:
        ref = [srv.wait(srv.submit(a, r, q, h=0.5, tol=1e-8, precision=ibim.PREC_FP64, **fmm)) for a, r, q in mols]
    with ibim.Server(slots=3, fuse_points=10 ** 7 if fuse else 0, fuse_wait_us=20000) as srv:
        ids = [srv.submit(a, r, q, h=0.5, tol=1e-8, precision=ibim.PREC_MIXED_F16, inner_order=4, **fmm)
               for a, r, q in mols]
        got = [srv.wait(i) for i in ids]
        groups, _ = srv.fusion_stats()
    print(f"fuse={fuse}: {groups} fused matvecs; outer steps {[g['info']['outer_steps'] for g in got]}")
    if fuse:
        assert groups > 0  # fused solves keep their own order-p operator in the inner solves
    for a, b in zip(ref, got):
        assert a["status"] == 0 and b["status"] == 0
        assert abs(a["energy"] - b["energy"]) <= 1e-7 * abs(a["energy"]), (a["energy"], b["energy"])
