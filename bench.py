#!/usr/bin/env python3
"""Benchmark: PB solves/s on one MI355X.

Workload (default): the synthetic 500-atom globule -- seeded self-avoiding
walk of 3.8 A steps in a 20 A sphere, radii U[1.2,2.0] A, charges U[-0.5,0.5]
shifted to net -2 e -- meshed at h = 0.25 A (tube width 2h, probe 1.4 A),
eps_in = 2, eps_out = 80, kappa = 0.1257 /A, Chebyshev order p = 5, leaf
capacity 64, max level 10, tau = 1, GMRES(40) to 1e-8, FP64 throughout.

One step = one complete PB solve with every input already in HBM (surface
nodes, normals, weights, atoms): octree + interaction lists, right-hand side,
GMRES with the FMM matvec, reaction potential at the atoms and the
solvation energy.  The level-set meshing runs once before the timed region.

Prints one JSON line.  The roofline entry is computed live from HIP events
recorded on the library's stream around its dominant stage; the cpu_baseline
entry times the CPU oracle (a restatement of the reference algorithm) on a
bounded sample of the same surface.
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0

CONFIGS = {
    # name: molecule (atoms, seed, confinement), grid h, p, leaf, kappa, eps_in, eps_out
    "globule500": dict(atoms=500, seed=500, confine=20.0, h=0.25, p=5, leaf=64, max_level=10, kappa=0.1257,
                       eps_in=2.0, eps_out=80.0, tau=1.0, restart=40, tol=1e-8),
}


def load_ibim():
    import importlib.util
    if "ibim_amd" in sys.modules:
        return sys.modules["ibim_amd"]
    spec = importlib.util.spec_from_file_location("ibim_amd", os.path.join(ROOT, "ibim-levelset_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ibim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_cpu_baseline(cfg, pts, nrm, w, matvecs_per_solve, budget_s, threads):
    """Time the CPU oracle on a bounded sample: one of the 16 kernel FMMs (K2,
    tree build included, as the reference does per kernel) on the surface
    points of one octant of the molecule, then scale to the full solve:
    16 kernels per matvec, points/sample points (the FMM is O(N)), and the
    solve's matvec count plus 4 FMMs for the energy."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    c = np.median(pts, axis=0)
    sel = np.all(pts > c, axis=1)
    sample = pts[sel]
    q = np.random.default_rng(0).normal(size=len(sample))
    env_threads = os.environ.get("OMP_NUM_THREADS")
    t0 = time.perf_counter()
    tree = O.Tree(sample, max_point=cfg["leaf"], max_level=cfg["max_level"])
    tree.fmm(cfg["p"], "K2", q, kappa=cfg["kappa"], eps_in=cfg["eps_in"], eps_out=cfg["eps_out"],
             vacant_radius=cfg["tau"] * cfg["h"])
    t_fmm = time.perf_counter() - t0
    tree.close()
    scale = len(pts) / len(sample)
    t_matvec = 16 * t_fmm * scale
    t_solve = (matvecs_per_solve + 0.25) * t_matvec
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": (f"oracle K2 FMM (tree build + up/down pass) on the {len(sample)} surface points of one octant "
                       f"({t_fmm:.2f} s on {threads} threads, {cpu_model()}); scaled x16 kernels x{scale:.2f} points "
                       f"x({matvecs_per_solve}+0.25) matvecs per solve"),
            "matvec_s_est": t_matvec, "omp_threads_env": env_threads}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="globule500", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision solves")
    ap.add_argument("--leaf", type=int, default=None, help="override the leaf capacity")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (solves of different molecules "
              f"are independent: N GPUs run N request streams with no exchange)", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    ibim = load_ibim()
    cfg = dict(CONFIGS[args.config])
    if args.leaf:
        cfg["leaf"] = args.leaf

    # ---- setup (untimed): molecule, level set, tube nodes -- all in HBM
    xyz, rad, qa = ibim.synthetic_globule(cfg["atoms"], cfg["seed"], cfg["confine"])
    t0 = time.perf_counter()
    surf = ibim.Surface.from_atoms(xyz, rad, h=cfg["h"], probe=1.4, thickness=2)
    t_mesh = time.perf_counter() - t0
    n = surf.n
    dev = torch.device("cuda:0")
    d_atoms = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev)
    d_q = torch.from_numpy(np.ascontiguousarray(qa)).to(dev)
    d_b = torch.empty(2 * n, dtype=torch.float64, device=dev)
    d_x = torch.empty(2 * n, dtype=torch.float64, device=dev)
    d_phi = torch.empty(cfg["atoms"], dtype=torch.float64, device=dev)
    ptrs = surf.device_ptrs()
    torch.cuda.synchronize()
    import ctypes as C
    P = lambda t: C.c_void_p(t.data_ptr())
    m = cfg["atoms"]

    def solve(profile=False, precision=ibim.PREC_FP64):
        plan = ibim.Plan(device_ptrs=ptrs, p=cfg["p"], max_points=cfg["leaf"], max_level=cfg["max_level"],
                         kappa=cfg["kappa"], eps_in=cfg["eps_in"], eps_out=cfg["eps_out"],
                         vacant_radius=cfg["tau"] * cfg["h"], precision=precision)
        if profile:
            plan.profile(True)
        plan.rhs_device(m, P(d_atoms), P(d_q), P(d_b))
        d_x.zero_()
        torch.cuda.synchronize()
        rc, info = plan.gmres_device(P(d_b), P(d_x), restart=cfg["restart"], maxit=2000, tol=cfg["tol"])
        plan.polarize_device(m, P(d_atoms), P(d_x), P(d_phi))
        plan.sync()
        energy = 0.5 * float(torch.dot(d_phi, d_q).item())
        stages = plan.stage_times() if profile else None
        stats = plan.stats() if profile else None
        foot = plan.footprint()
        plan.close()
        return rc, info, energy, stages, stats, foot

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    lat = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        rc, info, energy, _, _, foot = solve()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start

    # the mixed-precision form of the same solve (FP32 far and near field
    # inside FP64 iterative refinement), timed the same way
    mixed = None
    if not args.no_mixed:
        for _ in range(args.warmup):
            solve(precision=ibim.PREC_MIXED)
        torch.cuda.synchronize()
        lat_m = []
        tm0 = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            rc_m, info_m, energy_m, _, _, _ = solve(precision=ibim.PREC_MIXED)
            lat_m.append(time.perf_counter() - ts)
        total_m = time.perf_counter() - tm0
        lat_m = sorted(1e3 * x for x in lat_m)
        mixed = {"solves_per_s": args.steps / total_m, "speedup_vs_fp64": (args.steps / total_m) / (args.steps / total),
                 "energy": energy_m, "rel_energy_err_vs_fp64": abs(energy_m - energy) / abs(energy),
                 "latency_ms": {"p50": float(np.percentile(lat_m, 50)), "p99": float(np.percentile(lat_m, 99))},
                 "solve": info_m}

    # one more solve with stage timers on (outside the timed region)
    rc_p, info_p, energy_p, stages, stats, _ = solve(profile=True)
    r = cfg["p"] ** 3
    nmv = info_p["matvecs"]
    p2p_flops = 60.0 * stats["p2p_pairs"]
    # V-list GEMM: 32 r^2 flops per V interaction; the X list runs in its own launch
    v_flops = 32.0 * r * r * stats["v_pairs"]
    per_launch = {"M2L": (stages["M2L"][0] / max(stages["M2L"][1], 1), v_flops),
                  "P2P": (stages["P2P"][0] / max(stages["P2P"][1], 1), p2p_flops)}
    dom = max(per_launch, key=lambda k: per_launch[k][0])
    ms_launch, flops = per_launch[dom]
    achieved = flops / (ms_launch * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "kernel": dom,
                "ms_per_launch": ms_launch, "algorithmic_flops_per_launch": flops,
                "note": "FP64; M2L counted as 32 r^2 flops per V/X interaction, P2P as 60 flops per pair"}
    stage_ms = {k: v[0] for k, v in stages.items()}
    lat_ms = sorted(1e3 * x for x in lat)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        pts, nrm, w = surf.arrays()
        try:
            cpu = run_cpu_baseline(cfg, pts, nrm, w, nmv, args.cpu_budget, threads)
        except Exception as e:  # report, do not hide
            cpu = {"error": repr(e)}
    surf.close()

    value = args.steps / total
    out = {
        "metric": "PB solves/s (FP64 GPU; p50/p99 latency reported)",
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: synthetic 500-atom globule, h=0.25 A, p=5, leaf {cfg['leaf']}, "
                               f"kappa=0.1257, eps 2/80, GMRES(40) tol 1e-8, FP64",
                   "tube_points": n, "atoms": m, "parallelism": "single GPU"},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                       "all": lat_ms},
        "solve": {"iterations": info["iterations"], "matvecs": info["matvecs"], "rel_residual": info["rel_residual"],
                  "converged": info["converged"], "energy": energy, "plan_bytes": foot},
        "matvecs_per_s": info["matvecs"] * value, "tube_points_per_s": n * info["matvecs"] * value,
        "stage_ms_per_solve": stage_ms, "tree": stats, "mesh_s": t_mesh,
        "mixed": mixed, "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
