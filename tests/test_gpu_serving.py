"""Serving: solves of different molecules admitted while others run.  Each
result must be the same alone and in the mix (FP64: within 1e-10 relative,
only the order of floating-point atomics differs), and requests must really
overlap in time when slots allow it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FMM = dict(p=4, max_points=64, max_level=10, kappa=0.1257, eps_in=2.0, eps_out=80.0, vacant_radius=0.5)


def molecules(ibim):
    out = []
    for i, (m, R) in enumerate([(60, 9.0), (150, 12.0), (30, 7.0), (100, 11.0), (45, 8.0)]):
        out.append(ibim.synthetic_globule(m, 1000 + i, R))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed16"])
def test_same_alone_and_in_the_mix(ibim, gpu, precision):
    """Bar: energy within 1e-10 relative alone vs in the mix, in FP64 and in
    the FP16-M2L mixed form (whose FP64 outer residual uses the FP64 GEMM's
    atomics, so its iterates may differ at rounding level: iteration counts
    within one)."""
    fmm = dict(FMM, precision={"fp64": ibim.PREC_FP64, "mixed16": ibim.PREC_MIXED_F16}[precision])
    mols = molecules(ibim)
    alone = []
    with ibim.Server(slots=1) as srv:
        for xyz, r, q in mols:
            res = srv.wait(srv.submit(xyz, r, q, h=0.5, tol=1e-8, **fmm))
            assert res["status"] == 0, res
            alone.append(res)
    with ibim.Server(slots=4) as srv:
        ids = [srv.submit(xyz, r, q, h=0.5, tol=1e-8, **fmm) for xyz, r, q in mols]
        mixed = [srv.wait(i) for i in ids]
    for a, b in zip(alone, mixed):
        assert b["status"] == 0
        assert a["n_points"] == b["n_points"]
        assert abs(a["energy"] - b["energy"]) <= 1e-10 * abs(a["energy"]), (a["energy"], b["energy"])
        assert abs(a["info"]["iterations"] - b["info"]["iterations"]) <= (0 if precision == "fp64" else 1)
    # with 4 slots some requests ran at the same time
    spans = sorted((r["t_start"], r["t_end"]) for r in mixed)
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))


def test_admission_by_memory_budget(ibim, gpu):
    """With a budget smaller than two requests, they run one at a time."""
    mols = molecules(ibim)[:3]
    with ibim.Server(slots=3, mem_budget=1) as srv:
        ids = [srv.submit(xyz, r, q, h=0.5, tol=1e-6, **FMM) for xyz, r, q in mols]
        res = [srv.wait(i) for i in ids]
    spans = sorted((r["t_start"], r["t_end"]) for r in res)
    for k in range(len(spans) - 1):
        assert spans[k + 1][0] >= spans[k][1] - 1e-6
    assert all(r["status"] == 0 for r in res)


def test_small_request_passes_a_blocked_large_one(ibim, gpu):
    """A request that does not fit beside the running one waits, and a later
    small request that fits starts before it (admission in submission order
    among the requests that fit); every result is still correct."""
    mols = molecules(ibim)
    big, mid, small = mols[1], mols[3], mols[2]  # 150, 100 and 30 atoms
    fmm = dict(FMM, precision=ibim.PREC_FP64)
    with ibim.Server(slots=2) as srv:  # footprint estimates (budget = 90% of free memory)
        est = [srv.wait(srv.submit(a, r, q, h=0.5, tol=1e-6, **fmm)) for a, r, q in (mid, big, small)]
    e_mid, e_big, e_small = (r["footprint_est"] for r in est)
    assert e_small < e_big
    budget = e_mid + (e_small + e_big) // 2  # mid + small fit together, mid + big do not
    assert e_mid + e_small <= budget < e_mid + e_big and e_big <= budget
    with ibim.Server(slots=2, mem_budget=budget) as srv:
        ids = [srv.submit(a, r, q, h=0.5, tol=1e-6, **fmm) for a, r, q in (mid, big, small)]
        res = [srv.wait(i) for i in ids]
    r_mid, r_big, r_small = res
    assert all(r["status"] == 0 for r in res)
    assert r_small["t_start"] < r_big["t_start"], "the small request should pass the blocked large one"
    assert r_big["t_start"] >= min(r_mid["t_end"], r_small["t_end"]) - 1e-6
    for a, b in zip(est, res):
        assert abs(a["energy"] - b["energy"]) <= 1e-10 * abs(a["energy"])
