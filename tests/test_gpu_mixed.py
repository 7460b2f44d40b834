"""Mixed precision (FP32 near and far field inside FP64 refinement) against
the FP64 path on the same plan geometry.  Bars: the mixed matvec within 1e-5
relative of FP64; the refined solve meets the FP64 residual tolerance with
the FP64 operator, and its solvation energy is within 1e-6 relative of the
FP64 solve's."""
import numpy as np
import pytest

import geometry as G

pytestmark = pytest.mark.gpu

PB = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def sphere_case(ibim, gpu):
    pts, nrm, w = G.sphere_tube(a=2.0, h=0.125, lo=-4.0, npts=64)
    return pts, nrm, w


# matvec bars against FP64: FP32 everywhere 1e-5; FP16 V-list M2L (10-bit
# mantissa after power-of-two scaling; observed 3-5e-6) 2e-5
MATVEC_BAR = {"mixed": 1e-5, "mixed16": 2e-5}


# p = 5 runs the 32x32 FP16 GEMM, p = 6 the 16x16 one, p = 7 and 8 the
# tiled FP16 GEMM that adds its results into L, p = 3 the smallest tiles
@pytest.mark.parametrize("p", [3, 5, 6, 7, 8])
@pytest.mark.parametrize("mode", ["mixed", "mixed16"])
def test_mixed_matvec_close_to_fp64(ibim, sphere_case, mode, p):
    pts, nrm, w = sphere_case
    prec = {"mixed": ibim.PREC_MIXED, "mixed16": ibim.PREC_MIXED_F16}[mode]
    x = np.random.default_rng(0).uniform(-1, 1, 2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=p, max_points=64, vacant_radius=0.125, **PB) as plan:
        y64 = plan.matvec(x)
        plan.set_precision(prec)
        y32 = plan.matvec(x)
        far = plan.stats()["v_pairs"]
    err = rel(y32, y64)
    print(f"{mode} p={p} matvec rel err {err:.2e} ({far} V interactions)")
    assert far > 0
    assert err < MATVEC_BAR[mode]
    assert err > 0  # the mixed path really ran


@pytest.mark.parametrize("mode,p", [("mixed", 5), ("mixed16", 5), ("mixed16", 8)])
@pytest.mark.parametrize("kappa", [0.0, 0.1257])
def test_mixed_solve_energy_matches_fp64(ibim, sphere_case, kappa, mode, p):
    pts, nrm, w = sphere_case
    atoms = np.array([[0.0, 0.0, 0.0], [0.4, -0.3, 0.2]])
    qa = np.array([1.0, -0.5])
    low = {"mixed": ibim.PREC_MIXED, "mixed16": ibim.PREC_MIXED_F16}[mode]
    out = {}
    for prec in (ibim.PREC_FP64, low):
        with ibim.Plan(pts, nrm, w, p=p, max_points=64, vacant_radius=0.125, precision=prec, kappa=kappa,
                       eps_in=1.0, eps_out=80.0) as plan:
            b = plan.rhs(atoms, qa)
            x, rc, info = plan.gmres(b, restart=20, maxit=300, tol=1e-8)
            assert rc == 0, info
            # the residual is measured with the FP64 operator in both cases
            plan.set_precision(ibim.PREC_FP64)
            res = np.linalg.norm(b - plan.matvec(x)) / np.linalg.norm(b)
            phi = plan.polarize(atoms, x)
        out[prec] = (x, 0.5 * float(phi @ qa), info, res)
    (x64, e64, i64, r64), (x32, e32, i32, r32) = out[ibim.PREC_FP64], out[low]
    print(f"{mode} p={p} kappa={kappa}: E64={e64:.10e} Emix={e32:.10e} its {i64['iterations']} / {i32['iterations']} "
          f"outer {i32['outer_steps']} res {r64:.1e} / {r32:.1e}")
    assert r64 < 1e-8 and r32 < 1e-8
    assert i32["outer_steps"] >= 1
    assert abs(e32 - e64) / abs(e64) < 1e-6
    assert rel(x32, x64) < 1e-6


@pytest.mark.parametrize("p", [7, 8])
def test_mixed16_large_p_with_x_list(ibim, sphere_case, p):
    """p >= 7: the FP16 stage adds both lists into L, the X list through
    reflected multipoles and reflected result rows; a wrong reflection or
    sign would put the X list's share of the field (percent level) in the
    error."""
    pts, nrm, w = sphere_case
    x = np.random.default_rng(4).uniform(-1, 1, 2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=p, max_points=32, vacant_radius=0.125, **PB) as plan:
        y64 = plan.matvec(x)
        plan.set_precision(ibim.PREC_MIXED_F16)
        a = plan.matvec(x)
        b = plan.matvec(x)
        st = plan.stats()
    print(f"p={p}: {st['v_pairs']} V and {st['x_pairs']} X interactions, rel err vs FP64 {rel(a, y64):.2e}, "
          f"repeat {rel(a, b):.1e}")
    assert st["x_pairs"] > 0
    assert rel(a, y64) < MATVEC_BAR["mixed16"]
    assert rel(a, b) < 1e-14


@pytest.mark.parametrize("p", [5, 6])
def test_mixed16_matvec_repeatable(ibim, sphere_case, p):
    """The FP16 form's matvec runs its near field on a second stream beside
    the M2L reduction and reflects the X list's multipoles and results; two
    matvecs in a row, and one after an FP64 matvec on the same plan, give the
    same bits (every sum has a fixed order), and the X list is exercised."""
    pts, nrm, w = sphere_case
    x = np.random.default_rng(3).uniform(-1, 1, 2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=p, max_points=32, vacant_radius=0.125, precision=ibim.PREC_MIXED_F16,
                   **PB) as plan:
        a = plan.matvec(x)
        b = plan.matvec(x)
        plan.set_precision(ibim.PREC_FP64)
        y64 = plan.matvec(x)
        plan.set_precision(ibim.PREC_MIXED_F16)
        c = plan.matvec(x)
        st = plan.stats()
    print(f"p={p}: {st['v_pairs']} V and {st['x_pairs']} X interactions, rel err vs FP64 {rel(a, y64):.2e}")
    assert st["x_pairs"] > 0
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert rel(a, y64) < MATVEC_BAR["mixed16"]


def test_stage_timers(ibim, sphere_case):
    """Stage timers (HIP events on the plan's streams): with every stage
    recorded, each FP16 matvec adds one launch to its matvec and M2L stages;
    with only the M2L stage selected, no other stage records; switched off,
    nothing records; and recording never changes the result's bits."""
    pts, nrm, w = sphere_case
    x = np.random.default_rng(4).uniform(-1, 1, 2 * len(pts))
    with ibim.Plan(pts, nrm, w, p=5, max_points=32, vacant_radius=0.125, precision=ibim.PREC_MIXED_F16,
                   **PB) as plan:
        ref = plan.matvec(x)
        plan.profile(True)
        plan.stage_times(reset=True)
        a = plan.matvec(x)
        b = plan.matvec(x)
        st_all = plan.stage_times(reset=True)
        plan.profile(True, stages=["M2L_MIXED"])
        c = plan.matvec(x)
        st_m2l = plan.stage_times(reset=True)
        plan.profile(False)
        d = plan.matvec(x)
        st_off = plan.stage_times(reset=True)
    assert st_all["MATVEC_MIXED"][1] == 2 and st_all["M2L_MIXED"][1] == 2
    assert st_all["MATVEC_MIXED"][0] > 0 and st_all["M2L_MIXED"][0] > 0
    assert st_all["MATVEC_MIXED"][0] >= st_all["M2L_MIXED"][0]
    assert st_m2l["M2L_MIXED"][1] == 1 and st_m2l["M2L_MIXED"][0] > 0
    assert all(cnt == 0 for name, (ms, cnt) in st_m2l.items() if name != "M2L_MIXED")
    assert all(cnt == 0 for ms, cnt in st_off.values())
    for y in (a, b, c, d):
        assert np.array_equal(y, ref)


def test_stage_report_counts_algorithmic_work(ibim, sphere_case):
    """The stage report's work counts follow the matvec's roofline model: a
    whole FP16 mixed matvec carries 60 flops per near-field pair and 32 p^6
    flops per V or X interaction (the tree statistics count both
    independently, on the host); the near field's two launches on the second
    stream carry all of its pairs between them, and the FP16 M2L stage all the
    interactions."""
    pts, nrm, w = sphere_case
    x = np.random.default_rng(5).uniform(-1, 1, 2 * len(pts))
    p = 5
    with ibim.Plan(pts, nrm, w, p=p, max_points=32, vacant_radius=0.125, precision=ibim.PREC_MIXED_F16,
                   **PB) as plan:
        st = plan.stats()
        plan.profile(True, stages=["MATVEC_MIXED", "P2P_MIXED", "M2L_MIXED"])
        plan.matvec(x)
        plan.matvec(x)
        rep = plan.stage_report()
    r = p ** 3
    nf, m2l = 60.0 * st["p2p_pairs"], 32.0 * r * r * (st["v_pairs"] + st["x_pairs"])
    mv = rep["MATVEC_MIXED"]
    assert mv["launches"] == 2
    assert mv["valu_flops"] == 2 * nf and mv["mfma_flops"] == 2 * m2l
    assert rep["P2P_MIXED"]["launches"] == 4 and rep["P2P_MIXED"]["valu_flops"] == 2 * nf
    assert rep["M2L_MIXED"]["launches"] == 2 and rep["M2L_MIXED"]["mfma_flops"] == 2 * m2l
    nb = st["n_nodes"] - st["n_empty_leaves"]  # boxes with sources (empty boxes are leaves)
    assert mv["bytes"] == 2 * 8.0 * (14 * len(pts) + 8 * r * nb)
