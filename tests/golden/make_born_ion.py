"""Writes tests/golden/born_ion.json: the oracle's solve of the Born ion on
its analytic tube (one +1 e atom at the origin, radius 2 A, eps 1/80,
kappa 0; the sphere level set on the 64^3 grid x_i = -4 + 0.125 i, tube
|phi| < 2h: N = 12,790 nodes), p = 4, leaf 64, GMRES(20) to 1e-12.

Run from the repository root:  python tests/golden/make_born_ion.py
(about a minute on 8 cores).  The oracle (oracle/ibim_oracle.c) restates the
reference's path; the fixture is the data the GPU solve is compared with."""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))
import geometry as G  # noqa: E402
import oracle as O  # noqa: E402

PRM = dict(kappa=0.0, eps_in=1.0, eps_out=80.0, vacant_radius=0.125)
FMM = dict(p=4, max_point=64, max_level=10)


def main():
    pts, nrm, w = G.sphere_tube(a=2.0, h=0.125, lo=-4.0, npts=64)
    atom, q = np.zeros((1, 3)), np.ones(1)
    t0 = time.perf_counter()
    b = O.rhs(pts, nrm, atom, q, PRM["eps_in"], PRM["eps_out"])
    x, rc, it, res, nmv = O.solve(pts, nrm, w, b, restart=20, maxit=400, tol=1e-12, **FMM, **PRM)
    phi = O.polarize(pts, nrm, w, atom, x, **FMM, **PRM)
    # a seeded matvec as well (the operator itself, not only the solve)
    xr = np.random.default_rng(7).uniform(-1, 1, 2 * len(pts))
    y = O.dmapping(pts, nrm, w, xr, **FMM, **PRM)
    out = {
        "what": "Born ion, analytic sphere tube on the 64^3 grid of [-4, 4]^3 A, oracle FP64",
        "n": len(pts), "fmm": FMM, "params": PRM, "restart": 20, "tol": 1e-12,
        "rc": rc, "iterations": it, "matvecs": nmv, "rel_residual": res,
        "energy": 0.5 * float(phi @ q),
        "b_norm": float(np.linalg.norm(b)),
        "x_norm": float(np.linalg.norm(x)), "x_sum": float(x.sum()),
        "x_first": x[:8].tolist(), "x_mid": x[len(pts):len(pts) + 8].tolist(),
        "matvec_seed": 7, "matvec_norm": float(np.linalg.norm(y)), "matvec_sum": float(y.sum()),
        "matvec_first": y[:8].tolist(),
        "seconds": time.perf_counter() - t0,
    }
    with open(os.path.join(HERE, "born_ion.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("n", "iterations", "matvecs", "energy", "seconds")}))


if __name__ == "__main__":
    main()
