"""Exact sums of the polynomial surrogate kernels over all sources.

The surrogates (IBIM_KERNEL_POLY; the oracle's eval_poly) are quadratics in
d = source - target:
    K1a = d_a,  K2 = 0.5 + 0.25 |d|^2,  K3ab = delta_ab + 0.5 d_a d_b,  K4a = -0.75 d_a,
so sum_s K(x_s - x_t) q_s expands into moments of the sources (sum q,
sum x q, sum x_a x_b q) and powers of x_t: O(N) at any size, with no tree.
Test helper: checks the whole FMM (tree, lists, every M2L tile, near field)
on point sets far beyond what a direct sum reaches."""
import numpy as np

KERNELS = ["K1x", "K1y", "K1z", "K2", "K3xx", "K3yy", "K3zz", "K3xy", "K3yz", "K3zx",
           "K3yx", "K3zy", "K3xz", "K4x", "K4y", "K4z"]
_AX = {"x": 0, "y": 1, "z": 2}


def poly_kernel_sum(kernel, pts, q):
    """out[t] = sum_s K(pts[s] - pts[t]) q[s] for one surrogate kernel."""
    pts = np.asarray(pts, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    Q0 = q.sum()
    S1 = pts.T @ q                        # sum_s x_s q_s, (3,)
    t = pts.T                              # targets' coordinates, (3, n)
    name = kernel[:2]
    if name in ("K1", "K4"):
        a = _AX[kernel[2]]
        v = S1[a] - t[a] * Q0
        return v if name == "K1" else -0.75 * v
    if name == "K2":
        S2 = (pts * pts).sum(axis=1) @ q
        return 0.5 * Q0 + 0.25 * (S2 - 2.0 * (S1 @ t) + (t * t).sum(axis=0) * Q0)
    a, b = _AX[kernel[2]], _AX[kernel[3]]
    Sab = (pts[:, a] * pts[:, b]) @ q
    v = 0.5 * (Sab - t[b] * S1[a] - t[a] * S1[b] + t[a] * t[b] * Q0)
    return v + Q0 if a == b else v
