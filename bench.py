#!/usr/bin/env python3
"""Benchmark: PB solves/s and latency on one MI355X, mixed precision vs FP64.

Workload (default): the synthetic 500-atom globule -- seeded self-avoiding
walk of 3.8 A steps in a 20 A sphere, radii U[1.2,2.0] A, charges U[-0.5,0.5]
shifted to net -2 e -- meshed at h = 0.25 A (tube width 2h, probe 1.4 A),
eps_in = 2, eps_out = 80, kappa = 0.1257 /A, Chebyshev order p = 5, leaf
capacity 192 (the fastest FP64 matvec of a 128..384 sweep), max level 10,
tau = 1, GMRES(40) to 1e-8.

One step = one complete PB solve with every input already in HBM (surface
nodes, normals, weights, atoms): octree + interaction lists + M2L operators,
right-hand side, GMRES with the FMM matvec, reaction potential at the atoms
and the solvation energy.  The level-set meshing runs once before the timed
region.  The solve is timed in FP64 throughout and in two mixed forms
inside FP64 iterative refinement (FP32 near field and X list, with the
V-list M2L on FP32 MFMA or on FP16 MFMA with power-of-two scaling); `value`
is the faster mixed form's solves/s among those whose energy is within 1e-6
of the FP64 solve's (`energy_within_1e-6`).

A served leg then pushes a burst of seeded globules of the same size through
the continuous-batching server (geometry included per request) and reports
its solves/s and p50/p99 latency.

Prints one JSON line.  The roofline entry is computed from HIP events that
the library records on its own stream around each FMM stage during the timed
mixed solves (the V-list M2L stage is the dominant one); its traffic
figure comes from the committed rocprofv3 counter summary when present.  The
cpu_baseline entry times the CPU oracle (a restatement of the reference
algorithm) on a bounded sample of the same surface.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "f16": 2500.0}   # MI355X dense MFMA
METRIC = "PB solves/s + p50/p99 latency, mixed vs FP64 GPU at <=1e-6 rel energy error"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "m2l_gemm_globule500_traffic.json")

CONFIGS = {
    # name: molecule (atoms, seed, confinement), grid h, p, leaf, kappa, eps_in, eps_out
    "globule500": dict(atoms=500, seed=500, confine=20.0, h=0.25, p=5, leaf=192, max_level=10, kappa=0.1257,
                       eps_in=2.0, eps_out=80.0, tau=1.0, restart=40, tol=1e-8),
    # the 5000-atom globule in a 45 A sphere at p = 6 (not the default workload)
    "globule5000": dict(atoms=5000, seed=5000, confine=45.0, h=0.25, p=6, leaf=256, max_level=10, kappa=0.1257,
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


def run_cpu_baseline(cfg, pts, nrm, w, matvecs_per_solve, threads):
    """Time the CPU oracle on a bounded sample: one of the 16 kernel FMMs (K2,
    tree build included, as the reference does per kernel) on the surface
    points of one octant of the molecule, then scale to the full solve:
    16 kernels per matvec, points/sample points (the FMM is O(N)), and the
    FP64 solve's matvec count plus 4 FMMs for the energy."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    c = np.median(pts, axis=0)
    sel = np.all(pts > c, axis=1)
    sample = pts[sel]
    q = np.random.default_rng(0).normal(size=len(sample))
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
            "sample": (f"oracle K2 FMM (tree build + up/down pass, OpenMP) on the {len(sample)} surface points of "
                       f"one octant ({t_fmm:.2f} s on {threads} threads, {cpu_model()}); scaled x16 kernels "
                       f"x{scale:.2f} points x({matvecs_per_solve}+0.25) matvecs per FP64 solve"),
            "matvec_s_est": t_matvec}


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def serve_stream(ibim, n, slots, precision, cfg):
    """A request stream like the serving configuration, scaled down: n
    globules with atom counts exp(U[ln 100, ln 2000]) (seed 4), molecule i from
    seed 10000 + i in a sphere of radius 2.6 m^(1/3) A, h = 0.35 A, the
    workload's FMM settings.  A burst measures the saturated rate mu; the
    same molecules then arrive as a Poisson process at 0.7 mu."""
    import threading
    rng = np.random.default_rng(4)
    sizes = np.exp(rng.uniform(np.log(100), np.log(2000), n)).astype(int)
    mols = [ibim.synthetic_globule(int(m), 10000 + i, 2.6 * m ** (1 / 3)) for i, m in enumerate(sizes)]
    fmm = dict(p=cfg["p"], max_points=cfg["leaf"], max_level=cfg["max_level"], kappa=cfg["kappa"],
               eps_in=cfg["eps_in"], eps_out=cfg["eps_out"], vacant_radius=cfg["tau"] * 0.35, precision=precision)
    out = {"requests": n, "slots": slots, "atoms": [int(m) for m in sizes], "h": 0.35}
    with ibim.Server(slots=slots) as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(a, r, q, h=0.35, restart=cfg["restart"], tol=cfg["tol"], **fmm) for a, r, q in mols]
        res = [srv.wait(i) for i in ids]
        mu = n / (time.perf_counter() - t0)
        out["burst"] = {"solves_per_s": mu, "all_ok": all(r["status"] == 0 for r in res),
                        "tube_points": [int(r["n_points"]) for r in res]}
        lam = 0.7 * mu
        gaps = np.random.default_rng(5).exponential(1.0 / lam, n)
        ids = []

        def producer():
            for (a, r, q), g in zip(mols, gaps):
                time.sleep(g)
                ids.append(srv.submit(a, r, q, h=0.35, restart=cfg["restart"], tol=cfg["tol"], **fmm))

        t1 = time.perf_counter()
        th = threading.Thread(target=producer)
        th.start()
        th.join()
        res = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t1
    lat = [1e3 * (r["t_end"] - r["t_submit"]) for r in res]
    out["poisson"] = {"rate_per_s": lam, "solves_per_s": n / wall, "latency_ms": {"p50": pct(lat, 50),
                                                                                  "p99": pct(lat, 99)},
                      "all_ok": all(r["status"] == 0 and r["info"]["converged"] == 1 for r in res)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="globule500", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--serve", type=int, default=8, help="requests in the served burst (0 = skip)")
    ap.add_argument("--slots", type=int, default=2, help="server slots of the served burst")
    ap.add_argument("--stream", type=int, default=0, metavar="N",
                    help="also serve a stream of N seeded globules of 100..2000 atoms (h = 0.35 A): a burst for the "
                         "saturated rate, then Poisson arrivals at 70%% of it")
    ap.add_argument("--leaf", type=int, default=None, help="override the leaf capacity")
    ap.add_argument("--matvecs-only", type=int, default=0, metavar="K",
                    help="counter-collection mode: one 256 MiB device copy (calibration), then K matvecs in each "
                         "precision (FP16-M2L mixed, FP32 mixed, FP64) on the workload's plan; prints a short JSON "
                         "line, not the benchmark line")
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
    PRECS = {"mixed_f32": ibim.PREC_MIXED, "mixed_f16": ibim.PREC_MIXED_F16}
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
    stage_acc = {}   # timed mixed solves: the roofline stage only
    stage_full = {}  # one untimed solve per mixed form: every stage

    if args.matvecs_only:
        cal = torch.ones(32 << 20, dtype=torch.float64, device=dev)  # 256 MiB
        cal2 = torch.empty_like(cal)
        cal2.copy_(cal)
        plan = ibim.Plan(device_ptrs=ptrs, p=cfg["p"], max_points=cfg["leaf"], max_level=cfg["max_level"],
                         kappa=cfg["kappa"], eps_in=cfg["eps_in"], eps_out=cfg["eps_out"],
                         vacant_radius=cfg["tau"] * cfg["h"], precision=ibim.PREC_MIXED)
        d_x.uniform_(-1, 1)
        torch.cuda.synchronize()
        for prec in (ibim.PREC_MIXED_F16, ibim.PREC_MIXED, ibim.PREC_FP64):
            plan.set_precision(prec)
            for _ in range(args.matvecs_only):
                plan.matvec_device(P(d_x), P(d_b))
        plan.sync()
        print(json.dumps({"mode": "matvecs-only", "copy_bytes": 2 * cal.numel() * 8, "matvecs": args.matvecs_only,
                          "tree": plan.stats()}))
        plan.close()
        surf.close()
        return

    def solve(precision, profile=None, query=True, acc=None):
        plan = ibim.Plan(device_ptrs=ptrs, p=cfg["p"], max_points=cfg["leaf"], max_level=cfg["max_level"],
                         kappa=cfg["kappa"], eps_in=cfg["eps_in"], eps_out=cfg["eps_out"],
                         vacant_radius=cfg["tau"] * cfg["h"], precision=precision)
        if profile == "all":
            plan.profile(True)
        elif profile == "m2l":  # only the stage the roofline is quoted on (two events per matvec)
            plan.profile(True, stages=["M2L_MIXED"])
        plan.rhs_device(m, P(d_atoms), P(d_q), P(d_b))
        d_x.zero_()
        torch.cuda.synchronize()  # the library's stream does not wait for torch's
        rc, info = plan.gmres_device(P(d_b), P(d_x), restart=cfg["restart"], maxit=2000, tol=cfg["tol"])
        plan.polarize_device(m, P(d_atoms), P(d_x), P(d_phi))
        plan.sync()
        energy = 0.5 * float(torch.dot(d_phi, d_q).item())
        if profile:
            for k, (ms, cnt) in plan.stage_times().items():
                a = acc.setdefault(k, [0.0, 0])
                a[0] += ms
                a[1] += cnt
        # tree statistics and footprint are the same for every solve of one
        # geometry: the timed solves skip these queries (a few ms of small
        # device-to-host copies) and report those of the warmup solve
        stats = plan.stats() if query else None
        foot = plan.footprint() if query else None
        plan.close()
        return rc, info, energy, stats, foot

    def leg(precision, profile=False):
        stats = foot = None
        for _ in range(args.warmup):
            _, _, _, stats, foot = solve(precision)
        torch.cuda.synchronize()
        lat = []
        t_start = time.perf_counter()
        for i in range(args.steps):
            ts = time.perf_counter()
            rc, info, energy, st_, ft_ = solve(precision, "m2l" if profile else None,
                                               query=stats is None and i == args.steps - 1, acc=stage_acc)
            lat.append(1e3 * (time.perf_counter() - ts))
            if st_ is not None:
                stats, foot = st_, ft_
        torch.cuda.synchronize()
        total = time.perf_counter() - t_start
        if profile:  # the per-stage breakdown: one more solve, untimed, with every stage's events
            solve(precision, "all", query=False, acc=stage_full)
        return {"solves_per_s": args.steps / total, "ms_per_solve": 1e3 * total / args.steps,
                "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "all": lat}, "energy": energy,
                "solve": dict(info, rc=rc), "plan_bytes": foot}, stats

    def progress(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    progress(f"{args.config}: {surf.n} tube points")
    fp64, stats = leg(ibim.PREC_FP64)
    progress(f"fp64: {fp64['solves_per_s']:.3f} solves/s")
    forms = {}
    for name, prec in (("mixed_f32", ibim.PREC_MIXED), ("mixed_f16", ibim.PREC_MIXED_F16)):
        stage_acc.clear()
        stage_full.clear()
        res, _ = leg(prec, profile=True)
        res["rel_energy_err_vs_fp64"] = abs(res["energy"] - fp64["energy"]) / abs(fp64["energy"])
        res["energy_within_1e-6"] = bool(res["rel_energy_err_vs_fp64"] <= 1e-6 and res["solve"]["converged"] == 1
                                         and fp64["solve"]["converged"] == 1)
        res["stages"] = ({k: list(v) for k, v in stage_acc.items()}, {k: list(v) for k, v in stage_full.items()})
        forms[name] = res
        progress(f"{name}: {res['solves_per_s']:.3f} solves/s, energy error {res['rel_energy_err_vs_fp64']:.1e}")
    ok = [k for k in forms if forms[k]["energy_within_1e-6"]]
    best = max(ok or forms, key=lambda k: forms[k]["solves_per_s"])
    mixed = forms[best]
    rel_err = mixed["rel_energy_err_vs_fp64"]
    energy_ok = mixed["energy_within_1e-6"]
    timed_st, full_st = mixed.pop("stages")
    stage_acc = {k: tuple(v) for k, v in timed_st.items()}
    stage_full = {k: tuple(v) for k, v in full_st.items()}
    for f in forms.values():
        f.pop("stages", None)

    # ---- served burst: seeded globules of the same size through the server
    serving = None
    if args.serve > 0:
        mols = [ibim.synthetic_globule(cfg["atoms"], cfg["seed"] + 1 + i, cfg["confine"]) for i in range(args.serve)]
        fmm = dict(p=cfg["p"], max_points=cfg["leaf"], max_level=cfg["max_level"], kappa=cfg["kappa"],
                   eps_in=cfg["eps_in"], eps_out=cfg["eps_out"], vacant_radius=cfg["tau"] * cfg["h"],
                   precision=PRECS[best])
        with ibim.Server(slots=args.slots) as srv:
            t0 = time.perf_counter()
            ids = [srv.submit(a, r, q, h=cfg["h"], restart=cfg["restart"], tol=cfg["tol"], **fmm) for a, r, q in mols]
            res = [srv.wait(i) for i in ids]
            wall = time.perf_counter() - t0
        lat = [1e3 * (r["t_end"] - r["t_submit"]) for r in res]
        serving = {"requests": args.serve, "slots": args.slots, "precision": best,
                   "solves_per_s": args.serve / wall, "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
                   "all_ok": all(r["status"] == 0 and r["info"]["converged"] == 1 for r in res),
                   "note": "burst submitted at t=0; latency = completion - submission, geometry (level set, tube) "
                           "and plan build included per request"}

    stream = serve_stream(ibim, args.stream, args.slots, PRECS[best], cfg) if args.stream > 0 else None

    # ---- roofline of the dominant kernel in the timed mixed solves
    r = cfg["p"] ** 3
    totals = {k: v[0] for k, v in stage_full.items() if v[1] > 0 and k not in ("MATVEC", "MATVEC_MIXED")}
    # M2L_MIXED / P2P_MIXED: the inner (mixed) matvecs; M2L / P2P: the FP64 outer residuals
    dom = max(totals, key=totals.get)  # the stage with the most time per solve
    ms_gemm = stage_acc["M2L_MIXED"][0] / stage_acc["M2L_MIXED"][1]
    p = cfg["p"]
    # the FP16 stage also evaluates the X list (as GEMMs on the same kernel)
    x_in_stage = best == "mixed_f16" and p <= 6 and stats["x_pairs"] > 0
    pairs = stats["v_pairs"] + (stats["x_pairs"] if x_in_stage else 0)
    flops = 32.0 * r * r * pairs  # 2 x (4r) x (4r) per V or X interaction
    achieved = flops / (ms_gemm * 1e-3) / 1e12
    gemm16 = (f"k_m2l_gemm_h32<{p},8> (V) + k_m2l_gemm_h32<{p},4> (X)" if p <= 5 else
              f"k_m2l_gemm_hu<{p},16>" if p == 6 else f"k_m2l_gemm<{p},float>")
    kern, peak = {"mixed_f32": (f"k_m2l_gemm<{p},float> (FP32 MFMA)", PEAK_TFLOPS["f32"]),
                  "mixed_f16": (f"k_m_to_half + {gemm16} (FP16 MFMA)" if p <= 6 else
                                f"{gemm16} (FP32 MFMA; the FP16 tiles stop at p = 6)",
                                PEAK_TFLOPS["f16"] if p <= 6 else PEAK_TFLOPS["f32"])}[best]
    traffic = None
    if args.config == "globule500" and cfg["leaf"] == 192 and best == "mixed_f16" and os.path.exists(TRAFFIC_FILE):
        try:
            traffic = json.load(open(TRAFFIC_FILE)).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic, "kernel": kern,
                "dominant_stage": dom, "ms_per_launch": ms_gemm, "launches": stage_acc["M2L_MIXED"][1],
                "algorithmic_flops_per_launch": flops,
                "interactions_per_launch": pairs,
                "note": "M2L stage of the inner (mixed) matvecs: 32 r^2 flops per interaction (r = p^3), V pairs "
                        + ("and X pairs " if x_in_stage else "") + "per launch; time from HIP events on the plan "
                        "stream around the stage over the timed mixed solves (the only stage recorded there): the "
                        "multipole conversion and the GEMMs (the per-target reduction of their results runs beside "
                        "the near field and is not in it); dominant_stage from one extra untimed solve with every "
                        "stage recorded"}
    stage_ms = {k: v[0] for k, v in stage_full.items()}  # one solve

    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", min(16, os.cpu_count() or 1)))
        pts, nrm, w = surf.arrays()
        try:
            cpu = run_cpu_baseline(cfg, pts, nrm, w, fp64["solve"]["matvecs"], threads)
        except Exception as e:  # report, do not hide
            cpu = {"error": repr(e)}
    surf.close()

    value = mixed["solves_per_s"]
    out = {
        "metric": METRIC,
        "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_solve"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64", "data": "synthetic",
        "config": {"workload": f"{args.config}: synthetic {cfg['atoms']}-atom globule, h={cfg['h']} A, p={cfg['p']}, "
                               f"leaf {cfg['leaf']}, "
                               f"kappa=0.1257, eps 2/80, GMRES(40) tol 1e-8; value = the faster mixed-precision "
                               f"form meeting 1e-6 energy error ({best}: FP32 near field and X list, "
                               f"{'FP16' if best == 'mixed_f16' else 'FP32'} V-list M2L, inside FP64 iterative "
                               f"refinement), FP64 solve timed beside it",
                   "mixed_form": best,
                   "tube_points": n, "atoms": m, "parallelism": "single GPU"},
        "latency_ms": {"p50": mixed["latency_ms"]["p50"], "p99": mixed["latency_ms"]["p99"]},
        "mixed": mixed, "fp64": fp64, "mixed_forms": {k: {"solves_per_s": v["solves_per_s"],
                                                          "rel_energy_err_vs_fp64": v["rel_energy_err_vs_fp64"],
                                                          "energy_within_1e-6": v["energy_within_1e-6"],
                                                          "solve": v["solve"]} for k, v in forms.items()},
        "mixed_speedup_vs_fp64": mixed["solves_per_s"] / fp64["solves_per_s"],
        "rel_energy_err_vs_fp64": rel_err, "energy_within_1e-6": energy_ok,
        "fp64_matvecs_per_s": fp64["solve"]["matvecs"] * fp64["solves_per_s"],
        "fp64_tube_points_per_s": n * fp64["solve"]["matvecs"] * fp64["solves_per_s"],
        "serving": serving, "stream": stream, "stage_ms_per_mixed_solve": stage_ms, "tree": stats, "mesh_s": t_mesh,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
