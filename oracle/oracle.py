"""ctypes binding of the CPU oracle (oracle/ibim_oracle.c, oracle/ibim_levelset_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product.  The functions restate
the reference's FMM matvec path (see ibim_oracle.h for the file:line map).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libibim_oracle.so")
_lib = None

KERNELS = ["K1x", "K1y", "K1z", "K2", "K3xx", "K3yy", "K3zz", "K3xy", "K3yz", "K3zx",
           "K3yx", "K3zy", "K3xz", "K4x", "K4y", "K4z"]
KID = {k: i for i, k in enumerate(KERNELS)}


class KParams(C.Structure):
    _fields_ = [("kappa", C.c_double), ("eps_in", C.c_double), ("eps_out", C.c_double),
                ("vacant_radius", C.c_double), ("mode", C.c_int)]


class FmmParams(C.Structure):
    _fields_ = [("p", C.c_int), ("max_point", C.c_int), ("max_level", C.c_int), ("k", KParams)]


class LsParams(C.Structure):
    _fields_ = [("dims", C.c_int * 3), ("lo", C.c_double * 3), ("h", C.c_double), ("probe", C.c_double),
                ("band", C.c_int), ("thickness", C.c_int), ("thres", C.c_double), ("reinit_time", C.c_double),
                ("scale", C.c_double)]


class LsInfo(C.Structure):
    _fields_ = [("evolve_steps", C.c_int), ("reinit_steps", C.c_int * 2), ("quality", C.c_double * 2),
                ("inclusions", C.c_int64), ("n_nodes", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        d = C.POINTER(C.c_double)
        i = C.POINTER(C.c_int)
        vp = C.c_void_p
        _lib.orc_tree_build.argtypes = [vp, d, C.c_int, d, C.c_int, C.c_int, C.c_int]
        _lib.orc_tree_free.argtypes = [vp]
        _lib.orc_tree_nnodes.argtypes = [vp]
        _lib.orc_tree_export.argtypes = [vp, i, i, i, i, i, i, i, d, d, i, i, i, i]
        _lib.orc_tree_list.argtypes = [vp, C.c_int, C.c_int, i, C.c_int]
        _lib.orc_tree_points.argtypes = [vp, C.c_int, C.c_int, i, C.c_int]
        _lib.orc_cheb_nodes.argtypes = [C.c_int, d]
        _lib.orc_cheb_poly.argtypes = [C.c_int, C.c_int, d, d]
        _lib.orc_sp.argtypes = [C.c_int, C.c_int, d, d]
        _lib.orc_transfer_R.argtypes = [C.c_int, d]
        _lib.orc_kernel_eval.argtypes = [C.c_int, C.POINTER(KParams), d, d]
        _lib.orc_kernel_eval.restype = C.c_double
        _lib.orc_fmm_run.argtypes = [vp, C.c_int, C.c_int, C.POINTER(KParams), d, d]
        _lib.orc_direct.argtypes = [C.c_int, C.POINTER(KParams), d, C.c_int, d, C.c_int, d, d]
        _lib.orc_dmapping.argtypes = [C.c_int, d, d, d, C.POINTER(FmmParams), d, d, C.c_int]
        _lib.orc_dmapping_direct.argtypes = [C.c_int, d, d, d, C.POINTER(KParams), d, d]
        _lib.orc_rhs.argtypes = [C.c_int, d, d, C.c_int, d, d, C.c_double, C.c_double, d]
        _lib.orc_polarize.argtypes = [C.c_int, d, d, d, C.c_int, d, C.POINTER(FmmParams), d, d]
        _lib.orc_gmres_dense.argtypes = [C.c_int, d, d, d, C.c_int, C.c_int, C.c_double, i, d]
        _lib.orc_solve.argtypes = [C.c_int, d, d, d, C.POINTER(FmmParams), d, d, C.c_int, C.c_int,
                                   C.c_double, C.c_int, i, d, i]
        i64 = C.POINTER(C.c_int64)
        lp = C.POINTER(LsParams)
        _lib.orc_ls_gradient.argtypes = [lp, d, C.c_int, C.c_int, C.c_int, d, d]
        _lib.orc_ls_gradient.restype = None
        _lib.orc_ls_norm.argtypes = [d, d]
        _lib.orc_ls_norm.restype = C.c_double
        _lib.orc_ls_run.argtypes = [lp, C.c_int, d, d, C.c_double, C.c_double, C.c_double, d]
        _lib.orc_surface_from_grid.argtypes = [lp, d, d, d, d, C.c_int64, i64]
        _lib.orc_surface_from_grid.restype = C.c_int64
        _lib.orc_surface_from_atoms.argtypes = [lp, C.c_int, d, d, C.c_double, d, d, d, C.c_int64,
                                                C.POINTER(LsInfo), d]
        _lib.orc_surface_from_atoms.restype = C.c_int64
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def kparams(kappa=0.0, eps_in=1.0, eps_out=80.0, vacant_radius=0.0, mode=0):
    return KParams(kappa, eps_in, eps_out, vacant_radius, mode)


def fmm_params(p=4, max_point=64, max_level=10, **kw):
    return FmmParams(p, max_point, max_level, kparams(**kw))


# ----------------------------------------------------------------- tree --
class _TreeBuf(C.Structure):
    # orc_tree is opaque to Python; reserve generous storage for it
    _fields_ = [("raw", C.c_char * 256)]


class Tree:
    """The reference octree with U/V/W/X lists (bbfmm.cpp:8-280)."""

    def __init__(self, src, tgt=None, max_point=64, max_level=10):
        self._src = _f64(src, (-1, 3))
        self._tgt = self._src if tgt is None else _f64(tgt, (-1, 3))
        self._buf = _TreeBuf()
        rc = lib().orc_tree_build(C.byref(self._buf), _d(self._src), len(self._src), _d(self._tgt),
                                  len(self._tgt), max_point, max_level)
        if rc != 0:
            raise ValueError("orc_tree_build failed")
        n = lib().orc_tree_nnodes(C.byref(self._buf))
        self.nnodes = n
        ints = {k: np.zeros(n, np.int32) for k in
                ["level", "parent", "slot", "leaf", "empty", "nsrc", "ntgt", "nu", "nv", "nw", "nx"]}
        self.centre = np.zeros((n, 3))
        self.radius = np.zeros((n, 3))
        lib().orc_tree_export(C.byref(self._buf), _i(ints["level"]), _i(ints["parent"]), _i(ints["slot"]),
                              _i(ints["leaf"]), _i(ints["empty"]), _i(ints["nsrc"]), _i(ints["ntgt"]),
                              _d(self.centre), _d(self.radius), _i(ints["nu"]), _i(ints["nv"]),
                              _i(ints["nw"]), _i(ints["nx"]))
        for k, v in ints.items():
            setattr(self, k, v)

    def list(self, node, kind):
        kinds = {"U": 0, "V": 1, "W": 2, "X": 3}
        cap = int({"U": self.nu, "V": self.nv, "W": self.nw, "X": self.nx}[kind][node])
        out = np.zeros(max(cap, 1), np.int32)
        lib().orc_tree_list(C.byref(self._buf), node, kinds[kind], _i(out), cap)
        return out[:cap].copy()

    def points(self, node, which="src"):
        cap = int(self.nsrc[node] if which == "src" else self.ntgt[node])
        out = np.zeros(max(cap, 1), np.int32)
        lib().orc_tree_points(C.byref(self._buf), node, 0 if which == "src" else 1, _i(out), cap)
        return out[:cap].copy()

    def fmm(self, p, kernel, charge, **kw):
        kid = KID[kernel] if isinstance(kernel, str) else int(kernel)
        q = _f64(charge)
        out = np.zeros(len(self._tgt))
        kp = kparams(**kw)
        rc = lib().orc_fmm_run(C.byref(self._buf), p, kid, C.byref(kp), _d(q), _d(out))
        if rc != 0:
            raise RuntimeError("orc_fmm_run failed")
        return out

    def close(self):
        if self._buf is not None:
            lib().orc_tree_free(C.byref(self._buf))
            self._buf = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ chebyshev --
def cheb_nodes(p):
    x = np.zeros(p)
    lib().orc_cheb_nodes(p, _d(x))
    return x


def cheb_poly(p, x):
    x = _f64(x)
    T = np.zeros(p * len(x))
    lib().orc_cheb_poly(p, len(x), _d(x), _d(T))
    return T.reshape(p, len(x)).T  # (n, p)


def sp(p, x):
    x = _f64(x)
    S = np.zeros(p * len(x))
    lib().orc_sp(p, len(x), _d(x), _d(S))
    return S.reshape(p, len(x)).T  # (n, p)


def transfer_R(p):
    r = p ** 3
    R = np.zeros(8 * r * r)
    lib().orc_transfer_R(p, _d(R))
    return R.reshape(8, r, r).transpose(0, 2, 1)  # R[id][row, col]


# -------------------------------------------------------------- kernels --
def kernel_eval(kernel, a, b, **kw):
    kid = KID[kernel] if isinstance(kernel, str) else int(kernel)
    kp = kparams(**kw)
    return lib().orc_kernel_eval(kid, C.byref(kp), _d(_f64(a)), _d(_f64(b)))


def direct(kernel, src, tgt, charge, **kw):
    kid = KID[kernel] if isinstance(kernel, str) else int(kernel)
    src = _f64(src, (-1, 3))
    tgt = _f64(tgt, (-1, 3))
    out = np.zeros(len(tgt))
    kp = kparams(**kw)
    lib().orc_direct(kid, C.byref(kp), _d(src), len(src), _d(tgt), len(tgt), _d(_f64(charge)), _d(out))
    return out


# --------------------------------------------------------------- matvec --
def dmapping(xyz, nrm, w, x, p=4, max_point=64, max_level=10, share_tree=True, **kw):
    xyz = _f64(xyz, (-1, 3))
    n = len(xyz)
    nrm = _f64(nrm, (-1, 3))
    w = _f64(w)
    x = _f64(x)
    y = np.zeros(2 * n)
    fp = fmm_params(p, max_point, max_level, **kw)
    rc = lib().orc_dmapping(n, _d(xyz), _d(nrm), _d(w), C.byref(fp), _d(x), _d(y), int(share_tree))
    if rc != 0:
        raise RuntimeError("orc_dmapping failed")
    return y


def dmapping_direct(xyz, nrm, w, x, **kw):
    xyz = _f64(xyz, (-1, 3))
    n = len(xyz)
    y = np.zeros(2 * n)
    kp = kparams(**kw)
    lib().orc_dmapping_direct(n, _d(xyz), _d(_f64(nrm, (-1, 3))), _d(_f64(w)), C.byref(kp), _d(_f64(x)), _d(y))
    return y


def rhs(xyz, nrm, atom_xyz, atom_q, eps_in, eps_out):
    xyz = _f64(xyz, (-1, 3))
    atom_xyz = _f64(atom_xyz, (-1, 3))
    n, m = len(xyz), len(atom_xyz)
    b = np.zeros(2 * n)
    lib().orc_rhs(n, _d(xyz), _d(_f64(nrm, (-1, 3))), m, _d(atom_xyz), _d(_f64(atom_q)), eps_in, eps_out, _d(b))
    return b


def polarize(xyz, nrm, w, atom_xyz, x, p=4, max_point=64, max_level=10, **kw):
    xyz = _f64(xyz, (-1, 3))
    atom_xyz = _f64(atom_xyz, (-1, 3))
    n, m = len(xyz), len(atom_xyz)
    out = np.zeros(m)
    fp = fmm_params(p, max_point, max_level, **kw)
    rc = lib().orc_polarize(n, _d(xyz), _d(_f64(nrm, (-1, 3))), _d(_f64(w)), m, _d(atom_xyz), C.byref(fp),
                            _d(_f64(x)), _d(out))
    if rc != 0:
        raise RuntimeError("orc_polarize failed")
    return out


def gmres_dense(A, b, restart=20, maxit=100, tol=1e-10, x0=None):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    it = np.zeros(1, np.int32)
    res = np.zeros(1)
    rc = lib().orc_gmres_dense(n, A.ctypes.data_as(C.POINTER(C.c_double)), _d(_f64(b)), _d(x), restart, maxit,
                               tol, _i(it), _d(res))
    return x, int(rc), int(it[0]), float(res[0])


def solve(xyz, nrm, w, b, p=4, max_point=64, max_level=10, restart=20, maxit=100, tol=1e-8,
          share_tree=True, **kw):
    xyz = _f64(xyz, (-1, 3))
    n = len(xyz)
    x = np.zeros(2 * n)
    it = np.zeros(1, np.int32)
    nmv = np.zeros(1, np.int32)
    res = np.zeros(1)
    fp = fmm_params(p, max_point, max_level, **kw)
    rc = lib().orc_solve(n, _d(xyz), _d(_f64(nrm, (-1, 3))), _d(_f64(w)), C.byref(fp), _d(_f64(b)), _d(x),
                         restart, maxit, tol, int(share_tree), _i(it), _d(res), _i(nmv))
    return x, int(rc), int(it[0]), float(res[0]), int(nmv[0])


# ------------------------------------------------------------ level set --
def ls_params(dims, lo, h, probe=1.4, band=8, thickness=4, thres=0.05, reinit_time=2.0, scale=1.0):
    """Grid and pipeline parameters of the level-set pre-step (levelset.cpp)."""
    return LsParams((C.c_int * 3)(*[int(v) for v in dims]), (C.c_double * 3)(*[float(v) for v in lo]), float(h),
                    float(probe), int(band), int(thickness), float(thres), float(reinit_time), float(scale))


def grid_points(lp):
    """Coordinates (nx, ny, nz, 3) of the grid points lo + (i, j, k) h."""
    ax = [lp.lo[q] + lp.h * np.arange(lp.dims[q]) for q in range(3)]
    return np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1)


def ls_gradient(lp, g, i, j, k):
    """WENO5 one-sided derivatives (Dup, Dun) at grid point (i, j, k) (levelset.cpp:330-396)."""
    g = _f64(g).ravel()
    up, un = np.zeros(3), np.zeros(3)
    lib().orc_ls_gradient(C.byref(lp), _d(g), i, j, k, _d(up), _d(un))
    return up, un


def ls_norm(a, b):
    """getNorm(a, b) (levelset.cpp:294-301)."""
    return float(lib().orc_ls_norm(_d(_f64(a)), _d(_f64(b))))


def ls_evolve(lp, g, final_t, vel=1.0, cfl=0.2):
    """evolve (levelset.cpp:111-187) of a copy of g; returns (g, steps)."""
    g = _f64(g).copy()
    steps = lib().orc_ls_run(C.byref(lp), 0, _d(g.ravel()), None, final_t, vel, cfl, None)
    return g, int(steps)


def ls_reinitialize(lp, g, final_t, vel=1.0, cfl=0.5, phi0=None):
    """reinitialize (levelset.cpp:191-287) of a copy of g against phi0 (default g);
    returns (g, steps, (mean error, bad fraction))."""
    g = _f64(g).copy()
    p0 = g.copy() if phi0 is None else _f64(phi0).copy()
    q = np.zeros(2)
    steps = lib().orc_ls_run(C.byref(lp), 1, _d(g.ravel()), _d(p0.ravel()), final_t, vel, cfl, _d(q))
    return g, int(steps), (float(q[0]), float(q[1]))


def surface_from_grid(lp, phi):
    """Surface (levelset.cpp:437-586, electric.cpp:17-41) of the level set phi:
    (xyz, normals, weights, inclusions), nodes in grid order."""
    phi = _f64(phi).ravel()
    inc = np.zeros(1, np.int64)
    n = lib().orc_surface_from_grid(C.byref(lp), _d(phi), None, None, None, 0, inc.ctypes.data_as(C.POINTER(C.c_int64)))
    xyz, nrm, w = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    m = lib().orc_surface_from_grid(C.byref(lp), _d(phi), _d(xyz), _d(nrm), _d(w), n,
                                    inc.ctypes.data_as(C.POINTER(C.c_int64)))
    assert m == n
    return xyz, nrm, w, int(inc[0])


def surface_from_atoms(lp, atom_xyz, radii, init):
    """The whole pre-step (main.cpp:43-62) on the grid of lp with the grid
    starting at init: (xyz, normals, weights, info, level set before Surface)."""
    axyz = _f64(atom_xyz, (-1, 3))
    rad = _f64(radii)
    g = np.zeros(tuple(lp.dims))
    info = LsInfo()
    n = lib().orc_surface_from_atoms(C.byref(lp), len(axyz), _d(axyz), _d(rad), init, None, None, None, 0,
                                     C.byref(info), _d(g.ravel()))
    # Surface only reads g, so the node pass can start from the returned level set
    xyz, nrm, w, inc = surface_from_grid(lp, g)
    assert len(w) == n and inc == info.inclusions
    return xyz, nrm, w, info, g
