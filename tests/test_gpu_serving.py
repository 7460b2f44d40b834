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
