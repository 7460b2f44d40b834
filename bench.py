#!/usr/bin/env python3
"""Benchmark: served PB solves/s and p50/p99 latency on one MI355X, mixed
precision vs FP64, on the serving stream of 256 synthetic globules.

Workload (default, `--config stream256`): 256 seeded synthetic globules whose
atom counts are exp(U[ln 100, ln 20000]) (seed 4); molecule i is the
self-avoiding walk of seed 10000 + i (3.8 A steps, 3.0 A exclusion) in a
sphere of radius 2.6 m^(1/3) A, radii U[1.2, 2.0] A, charges U[-0.5, 0.5]
shifted to net -2 e.  Each request is meshed at h = 0.35 A (tube width 2h,
probe 1.4 A) and solved with eps_in = 2, eps_out = 80, kappa = 0.1257 /A,
tau = 1, Chebyshev order p = 5, leaf capacity 192, max level 10, GMRES(40)
to 1e-8, then its solvation energy is formed from the reaction potential at
the atoms.  A request is the whole chain: level set, tube nodes, plan (tree,
lists, operators), right-hand side, GMRES, reaction potential, energy.  The
mixed form refines in FP64 at order p; its inner solves apply the FP16-mixed
operator of order p - 1 (`--inner-order`, 0 for order p).

One step = one served request.  The whole stream holds about 9.2e8 tube
points (the walk-in-a-sphere globules are porous: ~800 tube points per atom
at h = 0.35 A), about 9 minutes of mixed-precision GPU time and half an hour
in FP64, so the timed runs serve stated subsamples of it, every k-th request
in stream order, which keeps the log-uniform size mix.  The requests go
through the continuous-batching server (`slots` requests at once, admission
by device footprint):

1. Like-for-like capacity on the FP64 subsample (every `--fp64-every`-th
   request, 8 by default): a burst in the mixed form and one in FP64.  Their
   ratio is the served speed-up of the mixed form; the energies of the two
   runs give the mixed form's energy error.
2. The mixed stream (the headline) on every `--every`-th request (32 by
   default), arriving as a Poisson process at `--load` times the mixed
   capacity (scaled from step 1 by mean atom count).  value = completed
   solves / wall time from the first arrival to the last completion, with
   p50/p99 latency (completion - arrival).  Stage timers in every request's
   plan give the roofline of the dominant kernel (the near field), of the M2L
   stage and of the whole matvec over this timed region.
3. The FP64 stream: the FP64 subsample as a Poisson process at `--load` times
   its FP64 capacity: served FP64 solves/s and p50/p99.

4. Fusion (small molecules): 32 globules of 10..40 atoms (11k..32k tube
   points, where one solve's matvec leaves most of the GPU idle) served as a
   burst on 16 slots with fused stage launches across the concurrent solves
   (ibim_server_set_fusion) and without; and one matvec of each of 8 of them,
   issued on their own streams against one fused group matvec.  The
   headline server fuses solves of at most --fuse-points tube points (the
   stream's smallest requests, about 100k points, fill the GPU alone: fusion
   measured neutral there).

--steps does not change the stream: its subsample fixes the timed requests.
--warmup W serves W small molecules in each precision before timing.

The cpu_baseline entry times the CPU oracle (a restatement of the reference
algorithm) on this machine's cores: one full matvec (the 16 kernel FMMs,
each with its own tree, as the reference runs them) on a bounded sample of
the stream's first molecule, scaled to the stream; and one complete Born
ion solve (the 64^3 sphere, N = 12,790) end to end, with its energy against
the GPU's.

`--config globule500` / `globule5000` / `globule50000` time single solves
of one molecule instead (h = 0.25, 0.25 and 0.3 A; p = 5, 6 and 8), FP64 and
both mixed forms.  `--matvecs-only K` runs K matvecs in each form (the
counter-collection mode used for the HBM traffic figures, and the kernel
profile of the 50,000-atom globule).
"""
import argparse
import json
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

# MI355X published peaks: vector FP32 / FP64 and dense MFMA, TFLOP/s; HBM TB/s
PEAK = {"valu_f32": 157.3, "valu_f64": 78.6, "mfma_f16": 2500.0, "mfma_f32": 157.3, "mfma_f64": 78.6, "hbm": 8.0}
METRIC = "PB solves/s + p50/p99 latency, mixed vs FP64 GPU at <=1e-6 rel energy error"
PREC_FP64 = 0  # (IBIM_PREC_FP64 of include/ibim.h)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "p2p_mixed_stream256_traffic.json")

FMM = dict(p=5, leaf=192, max_level=10, kappa=0.1257, eps_in=2.0, eps_out=80.0, tau=1.0, restart=40, tol=1e-8)
CONFIGS = {
    # (the mixed inner solves at order p - 1; the FP64 residual stays at order p)
    "stream256": dict(FMM, requests=256, h=0.35, atoms_lo=100, atoms_hi=20000, size_seed=4, mol_seed=10000,
                      confine=2.6, inner_order=4),
    # single molecules (not the default workload)
    "globule500": dict(FMM, atoms=500, seed=500, confine=20.0, h=0.25),
    "globule5000": dict(FMM, atoms=5000, seed=5000, confine=45.0, h=0.25, p=6, leaf=256, inner_order=5),
    "globule50000": dict(FMM, atoms=50000, seed=50000, confine=100.0, h=0.3, p=8, leaf=512, inner_order=7),
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
    """The host CPU's model name, with its family/model/stepping numbers (some
    hosts report only a short model name)."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping", "vendor_id") and k not in info:
                info[k] = v.strip()
    except OSError:
        return platform.processor()
    return (f"{info.get('model name', platform.processor())} ({info.get('vendor_id', '?')} family "
            f"{info.get('cpu family', '?')} model {info.get('model', '?')} stepping {info.get('stepping', '?')}, "
            f"{os.cpu_count()} logical CPUs visible)")


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def fmm_kw(cfg, h, precision):
    """Plan parameters; the mixed forms' inner solves run at cfg's inner_order
    (0: the plan's order p)."""
    return dict(p=cfg["p"], max_points=cfg["leaf"], max_level=cfg["max_level"], kappa=cfg["kappa"],
                eps_in=cfg["eps_in"], eps_out=cfg["eps_out"], vacant_radius=cfg["tau"] * h, precision=precision,
                inner_order=cfg.get("inner_order", 0) if precision != PREC_FP64 else 0)


# ------------------------------------------------------------ the stream --
def stream_molecules(ibim, cfg):
    rng = np.random.default_rng(cfg["size_seed"])
    sizes = np.exp(rng.uniform(np.log(cfg["atoms_lo"]), np.log(cfg["atoms_hi"]), cfg["requests"])).astype(int)
    mols = [ibim.synthetic_globule(int(m), cfg["mol_seed"] + i, cfg["confine"] * m ** (1 / 3))
            for i, m in enumerate(sizes)]
    return sizes, mols


def serve(ibim, mols, cfg, precision, slots, rate=None, profile=None, arrival_seed=5, label=None, fuse=0):
    """Push molecules through a server: all at t = 0 (rate None) or as a
    Poisson process of the given rate.  Returns the results (in molecule
    order), the wall time from the first arrival to the last completion and
    the server's stage report."""
    h = cfg["h"]
    kw = fmm_kw(cfg, h, precision)
    with ibim.Server(slots=slots, fuse_points=fuse) as srv:
        if profile:
            srv.profile(profile)
        ids = [None] * len(mols)
        gaps = (np.zeros(len(mols)) if rate is None else
                np.random.default_rng(arrival_seed).exponential(1.0 / rate, len(mols)))

        def producer():
            t_next = time.perf_counter()
            for i, (a, r, q) in enumerate(mols):
                t_next += gaps[i]
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids[i] = srv.submit(a, r, q, h=h, restart=cfg["restart"], tol=cfg["tol"], **kw)

        t0 = time.perf_counter()
        win = [time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.clock_gettime_ns(time.CLOCK_BOOTTIME)]
        th = threading.Thread(target=producer)
        th.start()
        res = []
        for k in range(len(mols)):
            while ids[k] is None:  # not submitted yet
                time.sleep(0.01)
            res.append(srv.wait(ids[k]))
            if label:
                progress(f"{label}: request {k + 1}/{len(mols)} done ({res[-1]['n_points']} points, "
                         f"{res[-1]['t_end'] - res[-1]['t_submit']:.2f} s)")
        th.join()
        wall = time.perf_counter() - t0
        win += [time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.clock_gettime_ns(time.CLOCK_BOOTTIME)]
        rep = srv.stage_report() if profile else None
        fused = srv.fusion_stats()
    if profile is not None:
        rep = dict(rep or {}, _window_ns={"monotonic": [win[0], win[2]], "boottime": [win[1], win[3]]},
                   _fused=fused)
    return res, wall, rep


def summarize(res, wall):
    lat = [1e3 * (r["t_end"] - r["t_submit"]) for r in res]
    mv = sum(r["info"]["matvecs"] for r in res)
    return {"requests": len(res), "solves_per_s": len(res) / wall, "wall_s": wall,
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat)},
            "tube_points": int(sum(r["n_points"] for r in res)),
            "matvecs_per_s": mv / wall,
            "tube_point_matvecs_per_s": sum(r["n_points"] * r["info"]["matvecs"] for r in res) / wall,
            "all_converged": all(r["status"] == 0 and r["info"]["converged"] == 1 for r in res)}


def roofline_entries(rep, p):
    """Rooflines over the timed mixed stream from the stage report: the near
    field (vector FP32), the FP16 M2L stage (matrix cores) and the whole
    mixed matvec (vector + matrix + HBM terms of the matvec's roofline)."""
    out = {}
    nf = rep.get("P2P_MIXED")
    if nf and nf["ms"] > 0:
        ach = nf["valu_flops"] / (nf["ms"] * 1e-3) / 1e12
        per = nf["valu_flops"] / nf["launches"]
        out["near_field"] = {"bound": "valu", "achieved": ach, "peak": PEAK["valu_f32"], "unit": "TFLOP/s",
                             "frac": ach / PEAK["valu_f32"], "kernel": "k_p2p_unit<true,false> (FP32 near field)",
                             "launches": nf["launches"], "ms_per_launch": nf["ms"] / nf["launches"],
                             "algorithmic_flops_per_launch": per, "pairs_per_launch": per / 60.0}
    m2l = rep.get("M2L_MIXED")
    if m2l and m2l["ms"] > 0:
        f16 = p <= 6
        peak = PEAK["mfma_f16"] if f16 else PEAK["mfma_f32"]
        ach = m2l["mfma_flops"] / (m2l["ms"] * 1e-3) / 1e12
        out["m2l"] = {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                      "kernel": (f"k_m_to_half + k_m2l_gemm_h32<{p},8> (V) + k_m2l_gemm_h32<{p},4> (X), FP16 MFMA"
                                 if p <= 5 else f"k_m2l_gemm_hu<{p},16>" if f16 else f"k_m2l_gemm<{p},float>"),
                      "launches": m2l["launches"], "ms_per_launch": m2l["ms"] / m2l["launches"],
                      "algorithmic_flops_per_launch": m2l["mfma_flops"] / m2l["launches"]}
    mv = rep.get("MATVEC_MIXED")
    if mv and mv["ms"] > 0:
        t_roof = (mv["valu_flops"] / (PEAK["valu_f32"] * 1e12) +
                  mv["mfma_flops"] / ((PEAK["mfma_f16"] if p <= 6 else PEAK["mfma_f32"]) * 1e12) +
                  mv["bytes"] / (PEAK["hbm"] * 1e12))
        out["matvec"] = {"frac": t_roof / (mv["ms"] * 1e-3), "launches": mv["launches"],
                         "ms_per_launch": mv["ms"] / mv["launches"], "roofline_ms_per_launch":
                         1e3 * t_roof / mv["launches"],
                         "note": "sum of the near field at the FP32 vector peak, the V+X interactions at the FP16 "
                                 "MFMA peak and the point and expansion bytes at 8 TB/s, over the measured time "
                                 "of the mixed matvecs"}
    return out


def run_stream(ibim, args, cfg):
    import torch  # noqa: F401  (device initialised for the library's streams)
    slots = args.slots
    t0 = time.perf_counter()
    sizes, mols = stream_molecules(ibim, cfg)
    progress(f"stream: {len(mols)} molecules of {sizes.min()}..{sizes.max()} atoms (mean {sizes.mean():.0f}) "
             f"generated in {time.perf_counter() - t0:.1f} s")
    sub = list(range(0, len(mols), args.fp64_every))
    sub_mols = [mols[i] for i in sub]
    head = list(range(0, len(mols), args.every))
    head_mols = [mols[i] for i in head]
    # warm-up: W small molecules per precision (code objects, block cache)
    warm = [ibim.synthetic_globule(120 + 20 * i, 30000 + i, cfg["confine"] * (120 + 20 * i) ** (1 / 3))
            for i in range(args.warmup)]
    for prec in (ibim.PREC_MIXED_F16, ibim.PREC_FP64):
        if warm:
            serve(ibim, warm, cfg, prec, slots)
    progress(f"warm-up: {args.warmup} requests per precision")

    # 1. capacity on the FP64 subsample, both precisions (bursts)
    res_m_sub, wall_m_sub, _ = serve(ibim, sub_mols, cfg, ibim.PREC_MIXED_F16, slots, label="burst mixed")
    mu_m_sub = len(sub) / wall_m_sub
    progress(f"burst mixed: {len(sub)} requests, {mu_m_sub:.3f} solves/s")
    res_d_sub, wall_d_sub, _ = serve(ibim, sub_mols, cfg, ibim.PREC_FP64, slots, label="burst fp64")
    mu_d_sub = len(sub) / wall_d_sub
    progress(f"burst fp64: {mu_d_sub:.3f} solves/s")
    errs = [abs(a["energy"] - b["energy"]) / abs(b["energy"]) for a, b in zip(res_m_sub, res_d_sub)]

    # 2. the mixed stream (headline subsample), Poisson arrivals
    scale = float(np.mean([len(m[0]) for m in sub_mols]) / np.mean([len(m[0]) for m in head_mols]))
    lam_m = args.load * mu_m_sub * scale
    stages = ["P2P_MIXED", "M2L_MIXED", "MATVEC_MIXED", "P2P", "M2L", "MATVEC"]
    res_m, wall_m, rep = serve(ibim, head_mols, cfg, ibim.PREC_MIXED_F16, slots, rate=lam_m, profile=stages,
                               label="mixed stream", fuse=args.fuse_points)
    mixed = summarize(res_m, wall_m)
    mixed["arrival_rate_per_s"] = lam_m
    mixed["subsample"] = f"every {args.every}th request of the {len(mols)} ({len(head)} requests)"
    progress(f"mixed stream: {mixed['solves_per_s']:.3f} solves/s, p50 {mixed['latency_ms']['p50']:.0f} ms, "
             f"p99 {mixed['latency_ms']['p99']:.0f} ms")
    # requests of both subsamples: the mixed stream's energies against FP64
    pos = {i: k for k, i in enumerate(head)}
    errs_in_stream = [abs(res_m[pos[i]]["energy"] - b["energy"]) / abs(b["energy"])
                      for i, b in zip(sub, res_d_sub) if i in pos]

    # 3. the FP64 stream on the FP64 subsample
    lam_d = args.load * mu_d_sub
    res_d, wall_d, _ = serve(ibim, sub_mols, cfg, ibim.PREC_FP64, slots, rate=lam_d, arrival_seed=6,
                             label="fp64 stream", fuse=args.fuse_points)
    fp64 = summarize(res_d, wall_d)
    fp64["arrival_rate_per_s"] = lam_d
    fp64["subsample"] = f"every {args.fp64_every}th request of the {len(mols)} ({len(sub)} requests)"
    progress(f"fp64 stream: {fp64['solves_per_s']:.3f} solves/s, p99 {fp64['latency_ms']['p99']:.0f} ms")

    max_err = max(errs + errs_in_stream)
    ok = max_err <= 1e-6 and mixed["all_converged"] and fp64["all_converged"]
    window = rep.pop("_window_ns")
    mixed["fused_group_matvecs"], mixed["fused_member_matvecs"] = rep.pop("_fused")
    roof = roofline_entries(rep, cfg.get("inner_order") or cfg["p"])  # (the order of the mixed matvecs)
    stage_tot = {k: {"ms": v["ms"], "launches": v["launches"]} for k, v in rep.items()}
    return {
        "sizes": sizes, "mols": head_mols, "res_m": res_m, "res_d_sub": res_d_sub,
        "sub_in_head": [pos[i] for i in sub if i in pos],
        "mixed": mixed, "fp64": fp64,
        "capacity": {"subsample": f"every {args.fp64_every}th request ({len(sub)} requests), bursts",
                     "mixed_solves_per_s": mu_m_sub, "fp64_solves_per_s": mu_d_sub,
                     "mixed_speedup_vs_fp64": mu_m_sub / mu_d_sub},
        "energy": {"max_rel_err_mixed_vs_fp64": max_err, "within_1e-6": bool(max_err <= 1e-6),
                   "compared": f"the {len(sub)} FP64-subsample requests: mixed burst and mixed stream against the "
                               f"FP64 burst"},
        "ok": ok, "roof": roof, "stages": stage_tot, "window": window,
    }


# ----------------------------------------------------------------- fusion --
def run_fusion(ibim, cfg):
    """Small molecules: served bursts with and without fused stage launches
    across the concurrent solves, and a fused group matvec against the same
    matvecs issued one per plan on their own streams."""
    import ctypes as C
    import torch
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(9)
    atoms = np.exp(rng.uniform(np.log(10), np.log(40), 32)).astype(int)
    mols = [ibim.synthetic_globule(int(m), 50000 + i, cfg["confine"] * m ** (1 / 3)) for i, m in enumerate(atoms)]
    kw = dict(fmm_kw(cfg, cfg["h"], ibim.PREC_MIXED_F16), inner_order=0)  # (fused solves use order p)
    out = {"workload": f"32 synthetic globules of {atoms.min()}..{atoms.max()} atoms (log-uniform, seed 9), h="
                       f"{cfg['h']} A, p={cfg['p']}, leaf {cfg['leaf']}, FP16 mixed form, served as one burst on "
                       f"16 slots"}

    def burst(fuse):
        with ibim.Server(slots=16, fuse_points=fuse, fuse_wait_us=2000) as srv:
            t0 = time.perf_counter()
            ids = [srv.submit(a, r, q, h=cfg["h"], restart=cfg["restart"], tol=cfg["tol"], **kw) for a, r, q in mols]
            res = [srv.wait(i) for i in ids]
            wall = time.perf_counter() - t0
            return res, wall, srv.fusion_stats()

    burst(0)  # warm-up
    res0, w0, _ = burst(0)
    res1, w1, st = burst(10 ** 7)
    out["tube_points"] = [int(r["n_points"]) for r in res0]
    out["unfused_solves_per_s"] = len(res0) / w0
    out["fused_solves_per_s"] = len(res1) / w1
    out["fused_speedup"] = w0 / w1
    out["fused_group_matvecs"], out["fused_member_matvecs"] = st
    out["max_rel_energy_diff_fused_vs_unfused"] = max(abs(a["energy"] - b["energy"]) / abs(a["energy"])
                                                      for a, b in zip(res0, res1))
    out["all_converged"] = all(r["status"] == 0 for r in res0 + res1)
    # one matvec per plan on its own stream vs one fused group matvec
    surfs = [ibim.Surface.from_atoms(a, r, h=cfg["h"], probe=1.4, thickness=2) for a, r, _ in mols[:8]]
    plans = [ibim.Plan(device_ptrs=s.device_ptrs(), **kw) for s in surfs]
    xs = [torch.rand(2 * pl.n, dtype=torch.float64, device=dev) for pl in plans]
    ys = [torch.empty_like(x) for x in xs]
    P = lambda t: C.c_void_p(t.data_ptr())
    torch.cuda.synchronize()
    with ibim.Group(plans, mixed=True) as grp:
        def singles():
            for pl, x, y in zip(plans, xs, ys):
                pl.matvec_device(P(x), P(y))
            for pl in plans:
                pl.sync()

        def fused():
            grp.matvec_device([P(x) for x in xs], [P(y) for y in ys])
            for pl in plans:
                pl.sync()

        t = {}
        for name, f in (("single", singles), ("group", fused)):
            f()
            t0 = time.perf_counter()
            for _ in range(40):
                f()
            t[name] = 1e3 * (time.perf_counter() - t0) / 40
    for pl in plans:
        pl.close()
    for s_ in surfs:
        s_.close()
    out["matvec_round_8_plans_ms"] = {"one_per_plan_on_own_streams": t["single"], "fused_group": t["group"],
                                      "points": [int(pl_n) for pl_n in [len(x) // 2 for x in xs]]}
    return out


# ---------------------------------------------------------- CPU baseline --
def born_sphere_tube(a=2.0, h=0.125, lo=-4.0, npts=64, thickness=2):
    """Tube nodes of the sphere |x| = a on the grid lo + i h (analytic level
    set; the node layout of levelset.cpp:437-586 and electric.cpp:33)."""
    g = lo + h * np.arange(npts)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    r = np.linalg.norm(P, axis=1)
    phi = r - a
    tube = thickness * h
    m = np.abs(phi) < tube
    P, r, phi = P[m], r[m], phi[m]
    n = P / r[:, None]
    return P - n * phi[:, None], n, 0.5 * (1.0 + np.cos(np.pi * phi / tube)) / tube * h ** 3


def cpu_baseline(ibim, cfg, stream, threads):
    """The CPU oracle on this machine's cores (OpenMP, `threads` threads)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    out = {"unit": "solves/s", "cores": threads, "kind": "port", "cpu": cpu_model()}
    # (a) one full matvec (16 kernel FMMs, 16 trees) on a bounded sample of
    # the stream's first molecule: its tube points in a box around the centre
    xyz, rad, _ = stream["mols"][0]
    surf = ibim.Surface.from_atoms(xyz, rad, h=cfg["h"], probe=1.4, thickness=2)
    pts, nrm, w = surf.arrays()
    surf.close()
    c = np.median(pts, axis=0)
    d = np.max(np.abs(pts - c), axis=1)
    n_sample = min(len(pts), 20000)
    sel = np.argsort(d, kind="stable")[:n_sample]
    x = np.random.default_rng(1).uniform(-1, 1, 2 * n_sample)
    t0 = time.perf_counter()
    O.dmapping(pts[sel], nrm[sel], w[sel], x, p=cfg["p"], max_point=cfg["leaf"], max_level=cfg["max_level"],
               share_tree=False, kappa=cfg["kappa"], eps_in=cfg["eps_in"], eps_out=cfg["eps_out"],
               vacant_radius=cfg["tau"] * cfg["h"])
    t_mv = time.perf_counter() - t0
    # scaled to the stream: O(N) per matvec, FP64 matvecs per solve from the
    # GPU FP64 runs of the subsample (their mean where a request has none)
    mv64 = {k: r["info"]["matvecs"] for k, r in zip(stream["sub_in_head"], stream["res_d_sub"])}
    mean_mv = float(np.mean([r["info"]["matvecs"] for r in stream["res_d_sub"]]))
    t_stream = sum(mv64.get(k, mean_mv) * t_mv * r["n_points"] / n_sample for k, r in enumerate(stream["res_m"]))
    out["value"] = len(stream["res_m"]) / t_stream
    out["sample"] = (f"oracle dmapping (16 kernel FMMs, one tree each, OpenMP) on {n_sample} tube points of the "
                     f"stream's first molecule ({t_mv:.2f} s on {threads} threads), scaled by tube points (O(N)) "
                     f"and by FP64 matvecs per solve to the {len(stream['res_m'])} requests of the timed mixed "
                     f"stream")
    out["matvec_s"] = t_mv
    out["matvec_points"] = n_sample
    # (b) the Born ion solved end to end on the CPU: RHS, GMRES(20), reaction
    # potential, energy; and on the GPU for comparison
    pts, nrm, w = born_sphere_tube()
    prm = dict(kappa=0.0, eps_in=1.0, eps_out=80.0, vacant_radius=0.125)
    atom, q = np.zeros((1, 3)), np.ones(1)
    t0 = time.perf_counter()
    b = O.rhs(pts, nrm, atom, q, prm["eps_in"], prm["eps_out"])
    xo, rc, it, res, nmv = O.solve(pts, nrm, w, b, p=4, max_point=64, max_level=10, restart=20, maxit=200,
                                   tol=1e-8, share_tree=False, **prm)
    e_cpu = 0.5 * float(O.polarize(pts, nrm, w, atom, xo, p=4, max_point=64, max_level=10, **prm) @ q)
    t_c1 = time.perf_counter() - t0
    with ibim.Plan(pts, nrm, w, p=4, max_points=64, max_level=10, **prm) as plan:
        xg, _, info = plan.gmres(plan.rhs(atom, q), restart=20, maxit=200, tol=1e-8)
        e_gpu = 0.5 * float(plan.polarize(atom, xg) @ q)
    born = 1.0 / (8 * np.pi * 2.0) * (1 / 80.0 - 1.0)
    out["born_ion_solve"] = {"points": len(pts), "cpu_s": t_c1, "cpu_matvecs": nmv, "cpu_energy": e_cpu,
                             "gpu_energy": e_gpu, "gpu_matvecs": info["matvecs"],
                             "rel_diff_gpu_vs_cpu": abs(e_gpu - e_cpu) / abs(e_cpu),
                             "analytic_born": born, "rel_err_vs_born": abs(e_cpu - born) / abs(born)}
    return out


# ------------------------------------------------------- single molecule --
def run_single(ibim, args, cfg):
    """FP64 and both mixed forms of one molecule's solve, `steps` times each."""
    import ctypes as C
    import torch
    dev = torch.device("cuda:0")
    xyz, rad, qa = ibim.synthetic_globule(cfg["atoms"], cfg["seed"], cfg["confine"])
    surf = ibim.Surface.from_atoms(xyz, rad, h=cfg["h"], probe=1.4, thickness=2)
    n, m = surf.n, cfg["atoms"]
    d_atoms = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev)
    d_q = torch.from_numpy(np.ascontiguousarray(qa)).to(dev)
    d_b = torch.empty(2 * n, dtype=torch.float64, device=dev)
    d_x = torch.empty(2 * n, dtype=torch.float64, device=dev)
    d_phi = torch.empty(m, dtype=torch.float64, device=dev)
    ptrs = surf.device_ptrs()
    torch.cuda.synchronize()
    P = lambda t: C.c_void_p(t.data_ptr())

    def plan_for(prec):
        return ibim.Plan(device_ptrs=ptrs, **fmm_kw(cfg, cfg["h"], prec))

    if args.matvecs_only:
        cal = torch.ones(32 << 20, dtype=torch.float64, device=dev)  # 256 MiB
        cal2 = torch.empty_like(cal)
        cal2.copy_(cal)
        plan = plan_for(ibim.PREC_MIXED)
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
        return None

    def solve(prec, rep_acc=None):
        plan = plan_for(prec)
        if rep_acc is not None:
            plan.profile(True, stages=["P2P_MIXED", "M2L_MIXED", "MATVEC_MIXED"])
        plan.rhs_device(m, P(d_atoms), P(d_q), P(d_b))
        d_x.zero_()
        torch.cuda.synchronize()  # the library's stream does not wait for torch's
        rc, info = plan.gmres_device(P(d_b), P(d_x), restart=cfg["restart"], maxit=2000, tol=cfg["tol"])
        plan.polarize_device(m, P(d_atoms), P(d_x), P(d_phi))
        plan.sync()
        energy = 0.5 * float(torch.dot(d_phi, d_q).item())
        if rep_acc is not None:
            for k, v in plan.stage_report().items():
                a = rep_acc.setdefault(k, dict.fromkeys(v, 0))
                for f in v:
                    a[f] += v[f]
        plan.close()
        return rc, info, energy

    legs = {}
    acc = {}
    forms = (("fp64", ibim.PREC_FP64), ("mixed_f32", ibim.PREC_MIXED), ("mixed_f16", ibim.PREC_MIXED_F16))
    for name, prec in forms:
        if args.forms and name not in args.forms.split(","):
            continue
        for _ in range(args.warmup):
            solve(prec)
        torch.cuda.synchronize()
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ts = time.perf_counter()
            rc, info, e = solve(prec, acc if name == "mixed_f16" else None)
            lat.append(1e3 * (time.perf_counter() - ts))
        wall = time.perf_counter() - t0
        legs[name] = {"solves_per_s": args.steps / wall, "ms_per_solve": 1e3 * wall / args.steps,
                      "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)}, "energy": e, "solve": info}
        progress(f"{name}: {legs[name]['solves_per_s']:.3f} solves/s")
    for k in ("mixed_f32", "mixed_f16"):
        if k in legs and "fp64" in legs:
            legs[k]["rel_energy_err_vs_fp64"] = abs(legs[k]["energy"] - legs["fp64"]["energy"]) / abs(
                legs["fp64"]["energy"])
    with plan_for(ibim.PREC_FP64) as pl:
        stats = pl.stats()
    surf.close()
    return {"legs": legs, "tube_points": n, "atoms": m, "tree": stats,
            "roof": roofline_entries(acc, cfg.get("inner_order") or cfg["p"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="single-molecule configs: timed solves per form")
    ap.add_argument("--warmup", type=int, default=2, help="warm-up requests (stream) or solves (single molecule)")
    ap.add_argument("--config", default="stream256", choices=sorted(CONFIGS))
    ap.add_argument("--slots", type=int, default=2, help="server slots (concurrent requests)")
    ap.add_argument("--load", type=float, default=0.7, help="Poisson arrival rate as a fraction of the saturated rate")
    ap.add_argument("--every", type=int, default=8, help="the timed mixed stream serves every k-th request")
    ap.add_argument("--fp64-every", type=int, default=32,
                    help="the capacity bursts and the FP64 stream serve every k-th request")
    ap.add_argument("--requests", type=int, default=None, help="stream length (default: the workload's 256)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fusion", action="store_true", help="skip the small-molecule fusion leg")
    ap.add_argument("--fusion-only", action="store_true",
                    help="run only the small-molecule fusion leg and print its JSON (kernel-trace mode)")
    ap.add_argument("--fuse-points", type=int, default=150000,
                    help="the headline server fuses the matvecs of concurrent solves of at most this many tube "
                         "points (0: off)")
    ap.add_argument("--leaf", type=int, default=None, help="override the leaf capacity")
    ap.add_argument("--inner-order", type=int, default=None,
                    help="Chebyshev order of the mixed forms' inner solves (0: the plan's p; default: the "
                         "workload's, p - 1 for the serving stream)")
    ap.add_argument("--molecule", type=int, default=None, metavar="I",
                    help="stream256: time single solves of the stream's molecule I instead of serving")
    ap.add_argument("--forms", default=None, help="single-solve forms to time (comma list of fp64, mixed_f32, "
                                                   "mixed_f16; default all)")
    ap.add_argument("--matvecs-only", type=int, default=0, metavar="K",
                    help="single-molecule counter mode: a 256 MiB device copy (calibration), then K matvecs in "
                         "each precision on the molecule's plan; prints a short JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the stream's solves are "
              f"independent: N GPUs would serve N request streams with no exchange)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    ibim = load_ibim()
    cfg = dict(CONFIGS[args.config])
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    if args.leaf:
        cfg["leaf"] = args.leaf
    if args.inner_order is not None:
        cfg["inner_order"] = args.inner_order
    if args.config == "stream256" and (args.matvecs_only or args.molecule is not None):
        # one molecule of the stream (counter mode: the median-size one)
        sizes, _ = stream_molecules(ibim, dict(cfg, requests=cfg["requests"]))
        i = args.molecule if args.molecule is not None else int(np.argsort(sizes)[len(sizes) // 2])
        m = int(sizes[i])
        cfg = dict(cfg, atoms=m, seed=cfg["mol_seed"] + i, confine=cfg["confine"] * m ** (1 / 3))
        if args.matvecs_only:
            run_single(ibim, args, cfg)
            return
    if args.config != "stream256" or args.molecule is not None:
        single = run_single(ibim, args, cfg)
        if single is None:
            return
        legs = single["legs"]
        best = legs.get("mixed_f16") or next(iter(legs.values()))
        out = {"metric": METRIC, "value": best["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup, "ms_per_step": best["ms_per_solve"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f16/f32/f64",
               "data": "synthetic",
               "config": {"workload": f"{args.config}: single solves of one synthetic {cfg['atoms']}-atom globule, "
                                      f"h={cfg['h']} A, p={cfg['p']}, leaf {cfg['leaf']} (not the metric's "
                                      f"serving workload)", "parallelism": "single GPU"},
               "latency_ms": best["latency_ms"], "legs": legs,
               "mixed_speedup_vs_fp64": (best["solves_per_s"] / legs["fp64"]["solves_per_s"]) if "fp64" in legs
               else None,
               "tube_points": single["tube_points"], "tree": single["tree"],
               "roofline": single["roof"].get("near_field"), "roofline_m2l": single["roof"].get("m2l"),
               "roofline_matvec": single["roof"].get("matvec"), "cpu_baseline": None}
        print(json.dumps(out))
        return

    if args.requests:
        cfg["requests"] = args.requests
    if args.fusion_only:
        print(json.dumps({"mode": "fusion-only", "fusion_small_molecules": run_fusion(ibim, cfg)}))
        return
    st = run_stream(ibim, args, cfg)
    fusion = None
    if not args.no_fusion:
        fusion = run_fusion(ibim, cfg)
        progress(f"fusion (small molecules): {fusion['fused_solves_per_s']:.1f} vs {fusion['unfused_solves_per_s']:.1f} "
                 f"solves/s, matvec round {fusion['matvec_round_8_plans_ms']['fused_group']:.2f} vs "
                 f"{fusion['matvec_round_8_plans_ms']['one_per_plan_on_own_streams']:.2f} ms")
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(ibim, cfg, st, threads)
            progress(f"cpu baseline: {cpu['value']:.3g} solves/s")
        except Exception as e:  # report, do not hide
            cpu = {"error": repr(e)}
    mixed, fp64 = st["mixed"], st["fp64"]
    roof = st["roof"].get("near_field")
    if roof is not None:
        traffic = None
        if os.path.exists(TRAFFIC_FILE):  # HBM bytes per near-field pair from the counter runs
            try:
                traffic = json.load(open(TRAFFIC_FILE))["bytes_per_pair"] * roof["pairs_per_launch"]
            except (OSError, ValueError, KeyError):
                traffic = None
        roof = dict(roof, traffic=traffic,
                    note="the dominant kernel of the mixed solves: the FP32 near field (60 flops per pair), two "
                         "launches per mixed matvec on the second stream; time from HIP events on that stream "
                         "around each launch, summed over every mixed matvec of the timed stream; traffic: HBM "
                         "bytes per pair of the same kernel from rocprofv3 FETCH_SIZE + WRITE_SIZE on the stream's "
                         "median molecule (profiles/p2p_mixed_stream256_traffic.json) times this run's pairs per "
                         "launch")
    out = {
        "metric": METRIC,
        "value": mixed["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": mixed["requests"], "warmup": args.warmup, "ms_per_step": 1e3 * mixed["wall_s"] / mixed["requests"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16/f32/f64 (FP16 M2L on MFMA with FP32 accumulation, FP32 near field, FP64 translations, "
                 "Krylov and refinement" + (f"; inner solves at Chebyshev order {cfg['inner_order']}, the FP64 "
                                            f"residual at order {cfg['p']}" if cfg.get("inner_order") else "") + ")",
        "data": "synthetic",
        "config": {"workload": f"stream256: serving stream of {cfg['requests']} synthetic globules, "
                               f"{cfg['atoms_lo']}..{cfg['atoms_hi']} atoms log-uniform (seed {cfg['size_seed']}), "
                               f"h={cfg['h']} A, p={cfg['p']}"
                               + (f" (mixed inner solves at order {cfg['inner_order']})" if cfg.get("inner_order")
                                  else "") + f", leaf {cfg['leaf']}, kappa={cfg['kappa']}, eps "
                               f"{cfg['eps_in']:g}/{cfg['eps_out']:g}, GMRES({cfg['restart']}) tol {cfg['tol']}; "
                               f"Poisson arrivals at {args.load:g} of the saturated rate; geometry, plan, solve and "
                               f"energy per request; {args.slots} server slots, fused matvecs of concurrent solves "
                               f"of at most {args.fuse_points} tube points; timed on every {args.every}th "
                               f"request (mixed) and every {args.fp64_every}th (FP64)",
                   "steps": "one step = one served request of the timed mixed stream",
                   "requests": cfg["requests"], "timed_requests": mixed["requests"],
                   "slots": args.slots, "tube_points": mixed["tube_points"], "parallelism": "single GPU"},
        "latency_ms": {"p50": mixed["latency_ms"]["p50"], "p99": mixed["latency_ms"]["p99"]},
        "mixed": mixed, "fp64": fp64,
        "capacity": st["capacity"],
        "mixed_speedup_vs_fp64": st["capacity"]["mixed_speedup_vs_fp64"],
        "target_speedup": 3.0,
        "energy": st["energy"],
        "roofline": roof, "roofline_m2l": st["roof"].get("m2l"), "roofline_matvec": st["roof"].get("matvec"),
        "stage_totals_mixed_stream": dict(st["stages"], note="HIP-event spans summed over the stream's requests; "
                                          "spans on the two streams of a matvec, and of two concurrent requests, "
                                          "overlap in time, so they do not add up to the wall time"),
        "timed_window_ns": st["window"],
        "fusion_small_molecules": fusion,
        "cpu_baseline": cpu,
        "mixed_vs_cpu": (mixed["solves_per_s"] / cpu["value"]) if cpu and "value" in cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
