"""The largest single solve at full size: a 50,000-atom synthetic globule in a
100 A sphere, h = 0.3 A (58.5M tube points), p = 8, leaf 512,
kappa = 0.1257 / A, eps 2 / 80.

* The whole FMM at full size: with the polynomial surrogate kernels the
  Chebyshev FMM is exact, and their exact sums over all 58.5M sources come
  from source moments (tests/poly_moments.py), so every tree box, list entry,
  M2L tile and near-field unit is checked without a direct sum.
* A = I exactly at eps_in = eps_out, kappa = 0.
* The FP64 solve and the FP16-mixed solve (FP16 V-list M2L with per-box
  scaling, FP32 near field, FP64 refinement) both reach an FP64 relative
  residual below 1e-8, and their solvation energies agree within 1e-6.  (At
  p = 8 the FP16 stage adds its results into L by atomics, so two mixed
  matvecs agree to rounding, not bit for bit.)

The two solves take about 10 minutes of GPU time, so they run only when
IBIM_GLOBULE50000=1 (GMRES prints its iteration table; run pytest with -s to
see it).  The quick check always runs (about a minute): one kernel's FMM at
full size against the exact polynomial sums, and the FP16-mixed matvec
against the FP64 one on the same plan.
"""
import ctypes as C
import os
import time

import numpy as np
import pytest

from poly_moments import poly_kernel_sum

pytestmark = pytest.mark.gpu
long_run = pytest.mark.skipif(os.environ.get("IBIM_GLOBULE50000") != "1",
                              reason="about 10 minutes of GPU time: set IBIM_GLOBULE50000=1 to run")

H, P, LEAF = 0.3, 8, 512
PB = dict(p=P, max_points=LEAF, max_level=10, kappa=0.1257, eps_in=2.0, eps_out=80.0, vacant_radius=H)


def say(msg):
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


@pytest.fixture(scope="module")
def globule(ibim, gpu):
    xyz, rad, q = ibim.synthetic_globule(50000, 50000, 100.0)
    t = time.time()
    s = ibim.Surface.from_atoms(xyz, rad, h=H, probe=1.4, thickness=2)
    say(f"surface: {s.n} tube points, grid {s.info['dims']}, {time.time() - t:.1f} s")
    yield xyz, q, s
    s.close()


def test_quick_full_size_fmm_and_mixed_matvec(ibim, globule, gpu):
    """K2 with the polynomial surrogate kernel exact at full size (every box,
    list entry, M2L tile and near-field unit of the p = 8 tree), and the
    FP16-mixed matvec within 2e-5 of the FP64 one on the PB plan (the bar of
    the smaller mixed-precision tests), repeating to rounding."""
    import torch
    _, _, s = globule
    pts, _, _ = s.arrays()
    q = np.random.default_rng(50001).normal(size=s.n)
    with ibim.Plan(device_ptrs=s.device_ptrs(), p=P, max_points=LEAF, max_level=10,
                   kernel_mode=ibim.KERNEL_POLY) as plan:
        t = time.time()
        g = plan.kernel_fmm("K2", q)
        ref = poly_kernel_sum("K2", pts, q)
        err = np.abs(g - ref).max() / np.abs(ref).max()
        say(f"K2 (polynomial) at full size: {time.time() - t:.1f} s, max error {err:.2e} of max |sum|")
        assert err < 1e-9
    x = torch.from_numpy(np.random.default_rng(50002).uniform(-1, 1, 2 * s.n)).to(gpu)
    y64, y16, y16b = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
    torch.cuda.synchronize()
    Pt = lambda v: C.c_void_p(v.data_ptr())
    with ibim.Plan(device_ptrs=s.device_ptrs(), precision=ibim.PREC_FP64, **PB) as plan:
        t = time.time()
        plan.matvec_device(Pt(x), Pt(y64))
        plan.sync()
        say(f"FP64 matvec {time.time() - t:.1f} s")
        plan.set_precision(ibim.PREC_MIXED_F16)
        t = time.time()
        plan.matvec_device(Pt(x), Pt(y16))
        plan.matvec_device(Pt(x), Pt(y16b))
        plan.sync()
        say(f"FP16 operators and two mixed matvecs {time.time() - t:.1f} s")
    err = float(torch.linalg.norm(y16 - y64) / torch.linalg.norm(y64))
    rep = float(torch.linalg.norm(y16b - y16) / torch.linalg.norm(y16))
    say(f"mixed vs FP64 matvec: {err:.2e}; two mixed matvecs differ by {rep:.1e}")
    assert 0 < err < 2e-5
    assert rep < 1e-12


@long_run
def test_fmm_exact_for_polynomial_kernels(ibim, globule):
    _, _, s = globule
    pts, _, _ = s.arrays()
    q = np.random.default_rng(50000).normal(size=s.n)
    with ibim.Plan(device_ptrs=s.device_ptrs(), p=P, max_points=LEAF, max_level=10,
                   kernel_mode=ibim.KERNEL_POLY) as plan:
        st = plan.stats()
        say(f"tree: {st['n_nodes']} boxes, {st['v_pairs']} V and {st['x_pairs']} X interactions")
        for k in ("K2", "K3xy", "K1z"):
            t = time.time()
            g = plan.kernel_fmm(k, q)
            ref = poly_kernel_sum(k, pts, q)
            err = np.abs(g - ref).max() / np.abs(ref).max()
            say(f"{k}: FMM {time.time() - t:.1f} s, max error {err:.2e} of max |sum|")
            assert err < 1e-9, k


@long_run
def test_identity_when_media_match(ibim, globule):
    _, _, s = globule
    x = np.random.default_rng(1).normal(size=2 * s.n)
    with ibim.Plan(device_ptrs=s.device_ptrs(), **dict(PB, kappa=0.0, eps_in=1.0, eps_out=1.0)) as plan:
        y = plan.matvec(x)
    assert np.array_equal(y, x)


@long_run
def test_fp64_and_mixed_solves_agree(ibim, globule, gpu):
    import torch
    xyz, qa, s = globule
    m = len(qa)
    d_atoms = torch.from_numpy(np.ascontiguousarray(xyz)).to(gpu)
    d_q = torch.from_numpy(np.ascontiguousarray(qa)).to(gpu)
    d_b = torch.empty(2 * s.n, dtype=torch.float64, device=gpu)
    d_x = torch.empty_like(d_b)
    d_phi = torch.empty(m, dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    Pt = lambda v: C.c_void_p(v.data_ptr())
    out = {}
    for name, prec in (("fp64", ibim.PREC_FP64), ("mixed16", ibim.PREC_MIXED_F16)):
        with ibim.Plan(device_ptrs=s.device_ptrs(), precision=prec, **PB) as plan:
            plan.set_trace(True)
            plan.rhs_device(m, Pt(d_atoms), Pt(d_q), Pt(d_b))
            d_x.zero_()
            torch.cuda.synchronize()
            t = time.time()
            rc, info = plan.gmres_device(Pt(d_b), Pt(d_x), restart=40, maxit=2000, tol=1e-8)
            plan.polarize_device(m, Pt(d_atoms), Pt(d_x), Pt(d_phi))
            plan.sync()
            e = 0.5 * float(torch.dot(d_phi, d_q).item())
            say(f"{name}: rc {rc}, {info}, energy {e:.12e}, {time.time() - t:.1f} s, "
                f"footprint {plan.footprint() / 2**30:.1f} GiB")
            assert rc == 0 and info["rel_residual"] < 1e-8
            out[name] = e
            if name == "mixed16":
                y1 = torch.empty_like(d_b)
                y2 = torch.empty_like(d_b)
                plan.matvec_device(Pt(d_x), Pt(y1))
                plan.matvec_device(Pt(d_x), Pt(y2))
                plan.sync()
                diff = float(torch.linalg.norm(y1 - y2) / torch.linalg.norm(y1))
                say(f"two mixed matvecs differ by {diff:.1e} (atomic summation order)")
                assert diff < 1e-12
    rel = abs(out["mixed16"] - out["fp64"]) / abs(out["fp64"])
    say(f"energy: FP64 {out['fp64']:.12e}, mixed {out['mixed16']:.12e}, relative difference {rel:.2e}")
    assert rel < 1e-6
