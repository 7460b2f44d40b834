"""Python view of the MI355X IBIM-PB FMM library (ctypes over include/ibim.h).

The product is the C ABI in ``_build/libibim_amd.so`` (HIP, gfx950) and the
C++ host code beside it; this module only binds that ABI for tests and the
benchmark.  There is no CPU fallback: if the library is missing, or no GPU is
visible, calls fail with an error.

The names follow the reference's interfaces (GaZ3ll3/ibim-levelset):
``Plan.kernel_fmm`` is ``bbfmm::kernel::run`` (bbfmm.h:244-270),
``Plan.matvec`` is the ``dmapping`` lambda (electric.cpp:294-388),
``Plan.rhs`` the ``load`` loop (electric.cpp:395-413),
``Plan.polarize`` is ``polarizeMap`` (electric.cpp:442-485) and
``Plan.gmres`` is ``GMRES`` (gmres.cpp:50-164) on the plan's matvec.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libibim_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ibim.h")

KERNELS = ["K1x", "K1y", "K1z", "K2", "K3xx", "K3yy", "K3zz", "K3xy", "K3yz", "K3zx",
           "K3yx", "K3zy", "K3xz", "K4x", "K4y", "K4z"]
KID = {k: i for i, k in enumerate(KERNELS)}

PREC_FP64, PREC_MIXED, PREC_MIXED_F16 = 0, 1, 2
KERNEL_PB, KERNEL_POLY = 0, 1
IBIM_OK, IBIM_NOT_CONVERGED = 0, 1
STAGES = ["P2M", "M2M", "M2L", "L2L", "P2P", "MATVEC", "ORTH", "RHS", "POLARIZE", "M2L_X", "MATVEC_MIXED", "M2L_MIXED",
          "P2P_MIXED"]


class IbimError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [("p", C.c_int), ("max_points", C.c_int), ("max_level", C.c_int),
                ("kappa", C.c_double), ("eps_in", C.c_double), ("eps_out", C.c_double),
                ("vacant_radius", C.c_double), ("precision", C.c_int), ("kernel_mode", C.c_int),
                ("inner_order", C.c_int)]


class TreeStats(C.Structure):
    _fields_ = [("n_points", C.c_int64), ("n_nodes", C.c_int64), ("n_leaves", C.c_int64),
                ("n_empty_leaves", C.c_int64), ("depth", C.c_int),
                ("n_u", C.c_int64), ("n_v", C.c_int64), ("n_w", C.c_int64), ("n_x", C.c_int64),
                ("p2p_pairs", C.c_int64), ("m2l_pairs", C.c_int64), ("v_pairs", C.c_int64),
                ("x_pairs", C.c_int64), ("p2p_pairs_u", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("matvecs", C.c_int), ("rel_residual", C.c_double),
                ("converged", C.c_int), ("outer_steps", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StageReport(C.Structure):
    _fields_ = [("ms", C.c_double * 16), ("launches", C.c_int64 * 16), ("valu_flops", C.c_double * 16),
                ("mfma_flops", C.c_double * 16), ("bytes", C.c_double * 16)]

    def as_dict(self):
        """{stage: {ms, launches, valu_flops, mfma_flops, bytes}} for the stages that recorded."""
        out = {}
        for i, name in enumerate(STAGES):
            if self.launches[i]:
                out[name] = {"ms": self.ms[i], "launches": int(self.launches[i]), "valu_flops": self.valu_flops[i],
                             "mfma_flops": self.mfma_flops[i], "bytes": self.bytes[i]}
        return out


def stage_mask(stages):
    mask = 0
    for name in stages:
        mask |= 1 << STAGES.index(name)
    return mask


class GeomParams(C.Structure):
    _fields_ = [("h", C.c_double), ("lo", C.c_double * 3), ("dims", C.c_int * 3), ("pad", C.c_double),
                ("probe", C.c_double), ("band", C.c_int), ("thickness", C.c_int), ("thres", C.c_double),
                ("reinit_time", C.c_double), ("scale", C.c_double), ("init", C.c_double)]


class SurfaceInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("dims", C.c_int * 3), ("lo", C.c_double * 3), ("h", C.c_double),
                ("inclusions", C.c_int64), ("evolve_steps", C.c_int), ("reinit_steps", C.c_int * 2),
                ("quality", C.c_double * 2), ("area", C.c_double)]

    def as_dict(self):
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = list(v) if hasattr(v, "__len__") else v
        return out


class Request(C.Structure):
    _fields_ = [("n_atoms", C.c_int64), ("atom_xyz", C.POINTER(C.c_double)), ("radii", C.POINTER(C.c_double)),
                ("charges", C.POINTER(C.c_double)), ("geom", GeomParams), ("fmm", Params), ("restart", C.c_int),
                ("maxit", C.c_int), ("tol", C.c_double)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("energy", C.c_double), ("n_points", C.c_int64), ("info", SolveInfo),
                ("t_submit", C.c_double), ("t_start", C.c_double), ("t_end", C.c_double),
                ("footprint_est", C.c_int64), ("footprint", C.c_int64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "info"}
        d["info"] = self.info.as_dict()
        return d


def build(jobs=8):
    """Compile the HIP library (hipcc cross-compiles gfx950 without a GPU)."""
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", HERE], check=True)


_lib = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IbimError(f"{LIB_PATH} is missing: build it with `make -C {HERE}`")
    L = C.CDLL(LIB_PATH)
    d = C.POINTER(C.c_double)
    i = C.POINTER(C.c_int)
    vp = C.c_void_p
    pp = C.POINTER(C.c_void_p)
    i64 = C.c_int64
    L.ibim_plan_create.argtypes = [C.POINTER(Params), i64, d, d, d, pp]
    L.ibim_plan_create_device.argtypes = [C.POINTER(Params), i64, vp, vp, vp, pp]
    L.ibim_plan_destroy.argtypes = [vp]
    L.ibim_plan_destroy.restype = None
    L.ibim_plan_sync.argtypes = [vp]
    L.ibim_plan_set_stream.argtypes = [vp, vp]
    L.ibim_plan_set_precision.argtypes = [vp, C.c_int]
    L.ibim_plan_footprint.argtypes = [vp]
    L.ibim_plan_footprint.restype = i64
    L.ibim_matvec.argtypes = [vp, d, d]
    L.ibim_matvec_device.argtypes = [vp, vp, vp]
    L.ibim_kernel_fmm.argtypes = [vp, C.c_int, d, d]
    L.ibim_rhs.argtypes = [vp, i64, d, d, d]
    L.ibim_rhs_device.argtypes = [vp, i64, vp, vp, vp]
    L.ibim_polarize.argtypes = [vp, i64, d, d, d]
    L.ibim_polarize_device.argtypes = [vp, i64, vp, vp, vp]
    L.ibim_polarize_device_noalloc.argtypes = [vp, i64, vp, vp, vp]
    L.ibim_gmres.argtypes = [vp, d, d, C.c_int, C.c_int, C.c_double, C.POINTER(SolveInfo)]
    L.ibim_gmres_device.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_double, C.POINTER(SolveInfo)]
    L.ibim_tree_stats_get.argtypes = [vp, C.POINTER(TreeStats)]
    L.ibim_tree_export.argtypes = [vp, i, i, i, i, i, i, d, d, i, i, i, i]
    L.ibim_tree_lists.argtypes = [vp, C.c_int, i, i, i64]
    L.ibim_tree_points.argtypes = [vp, C.c_int, i, i64]
    L.ibim_plan_profile.argtypes = [vp, C.c_int]
    L.ibim_plan_set_trace.argtypes = [vp, C.c_int]
    L.ibim_plan_profile_stages.argtypes = [vp, C.c_uint]
    L.ibim_plan_stage_times.argtypes = [vp, d, C.POINTER(C.c_int64), C.c_int]
    L.ibim_plan_stage_report.argtypes = [vp, C.POINTER(StageReport), C.c_int]
    L.ibim_server_profile_stages.argtypes = [vp, C.c_uint]
    L.ibim_server_stage_report.argtypes = [vp, C.POINTER(StageReport), C.c_int]
    L.ibim_surface_from_atoms.argtypes = [C.POINTER(GeomParams), i64, d, d, pp]
    L.ibim_surface_from_grid.argtypes = [C.POINTER(GeomParams), d, pp]
    L.ibim_surface_get_info.argtypes = [vp, C.POINTER(SurfaceInfo)]
    L.ibim_surface_copy.argtypes = [vp, d, d, d]
    L.ibim_surface_device.argtypes = [vp, pp, pp, pp]
    L.ibim_surface_destroy.argtypes = [vp]
    L.ibim_surface_destroy.restype = None
    L.ibim_synthetic_globule.argtypes = [i64, C.c_uint64, C.c_double, d, d, d]
    L.ibim_server_create.argtypes = [C.c_int, i64, pp]
    L.ibim_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
    L.ibim_server_wait.argtypes = [vp, C.c_int64, C.POINTER(Result)]
    L.ibim_server_now.argtypes = [vp]
    L.ibim_server_now.restype = C.c_double
    L.ibim_server_destroy.argtypes = [vp]
    L.ibim_server_destroy.restype = None
    L.ibim_kernel_fmm_targets.argtypes = [vp, C.c_int, i64, d, d, d]
    L.ibim_group_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, pp]
    L.ibim_group_matvec_device.argtypes = [vp, C.POINTER(vp), C.POINTER(vp)]
    L.ibim_group_stage_report.argtypes = [vp, C.POINTER(StageReport), C.c_int]
    L.ibim_group_destroy.argtypes = [vp]
    L.ibim_group_destroy.restype = None
    L.ibim_server_set_fusion.argtypes = [vp, i64, C.c_int]
    L.ibim_server_fusion_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.ibim_last_error.restype = C.c_char_p
    L.ibim_version.restype = C.c_char_p
    _lib = L
    return L


def declared_symbols():
    """Function names declared in include/ibim.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(ibim_[a-z_0-9]+)\s*\(", txt)))


def _check(rc, what):
    if rc < 0:
        raise IbimError(f"{what} failed ({rc}): {lib().ibim_last_error().decode()}")
    return rc


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def params(p=4, max_points=64, max_level=10, kappa=0.0, eps_in=1.0, eps_out=80.0, vacant_radius=0.0,
           precision=PREC_FP64, kernel_mode=KERNEL_PB, inner_order=0):
    """inner_order (mixed forms): the Chebyshev order of the inner solves'
    operator, 2..p-1, or 0 for p; the FP64 residual keeps order p."""
    return Params(p, max_points, max_level, kappa, eps_in, eps_out, vacant_radius, precision, kernel_mode,
                  inner_order)


class Plan:
    """A device-resident IBIM plan: octree, lists and FMM buffers for one
    surface point set (xyz, normals: (n,3); weights: (n,))."""

    def __init__(self, xyz=None, normals=None, weights=None, device_ptrs=None, **kw):
        self.prm = params(**kw)
        self._h = C.c_void_p()
        if device_ptrs is not None:
            n, dxyz, dnrm, dw = device_ptrs
            _check(lib().ibim_plan_create_device(C.byref(self.prm), n, dxyz, dnrm, dw, C.byref(self._h)),
                   "ibim_plan_create_device")
            self.n = int(n)
        else:
            xyz = _f64(xyz, (-1, 3))
            n = len(xyz)
            nrm = _f64(np.zeros((n, 3)) if normals is None else normals, (-1, 3))
            w = _f64(np.ones(n) if weights is None else weights)
            _check(lib().ibim_plan_create(C.byref(self.prm), n, _d(xyz), _d(nrm), _d(w), C.byref(self._h)),
                   "ibim_plan_create")
            self.n = n

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().ibim_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------- hot path --
    def matvec(self, x):
        x = _f64(x)
        assert x.shape == (2 * self.n,)
        y = np.zeros(2 * self.n)
        _check(lib().ibim_matvec(self._h, _d(x), _d(y)), "ibim_matvec")
        return y

    def matvec_device(self, d_x, d_y):
        _check(lib().ibim_matvec_device(self._h, d_x, d_y), "ibim_matvec_device")

    def kernel_fmm(self, kernel, charge):
        kid = KID[kernel] if isinstance(kernel, str) else int(kernel)
        q = _f64(charge)
        out = np.zeros(self.n)
        _check(lib().ibim_kernel_fmm(self._h, kid, _d(q), _d(out)), "ibim_kernel_fmm")
        return out

    def rhs(self, atom_xyz, atom_q):
        axyz = _f64(atom_xyz, (-1, 3))
        aq = _f64(atom_q)
        b = np.zeros(2 * self.n)
        _check(lib().ibim_rhs(self._h, len(axyz), _d(axyz), _d(aq), _d(b)), "ibim_rhs")
        return b

    def polarize(self, atom_xyz, x):
        axyz = _f64(atom_xyz, (-1, 3))
        out = np.zeros(len(axyz))
        _check(lib().ibim_polarize(self._h, len(axyz), _d(axyz), _d(_f64(x)), _d(out)), "ibim_polarize")
        return out

    def gmres(self, b, x0=None, restart=20, maxit=100, tol=1e-8):
        b = _f64(b)
        x = np.zeros(2 * self.n) if x0 is None else _f64(x0).copy()
        info = SolveInfo()
        rc = _check(lib().ibim_gmres(self._h, _d(b), _d(x), restart, maxit, tol, C.byref(info)), "ibim_gmres")
        return x, rc, info.as_dict()

    def sync(self):
        _check(lib().ibim_plan_sync(self._h), "ibim_plan_sync")

    def set_trace(self, on=True):
        """Print GMRES's iteration table (iteration, relative residual, seconds) to stdout."""
        _check(lib().ibim_plan_set_trace(self._h, int(on)), "ibim_plan_set_trace")

    def set_stream(self, stream):
        """Order the plan's work on a caller's stream (a hipStream_t as int, e.g.
        torch.cuda.current_stream().cuda_stream); None: the plan's own stream."""
        _check(lib().ibim_plan_set_stream(self._h, C.c_void_p(stream) if stream else None), "ibim_plan_set_stream")

    # ---------------------------------------------- device-pointer variants --
    def rhs_device(self, m, d_axyz, d_aq, d_b):
        _check(lib().ibim_rhs_device(self._h, m, d_axyz, d_aq, d_b), "ibim_rhs_device")

    def gmres_device(self, d_b, d_x, restart=20, maxit=100, tol=1e-8):
        info = SolveInfo()
        rc = _check(lib().ibim_gmres_device(self._h, d_b, d_x, restart, maxit, tol, C.byref(info)),
                    "ibim_gmres_device")
        return rc, info.as_dict()

    def polarize_device(self, m, d_axyz, d_x, d_phi):
        _check(lib().ibim_polarize_device(self._h, m, d_axyz, d_x, d_phi), "ibim_polarize_device")

    def polarize_device_noalloc(self, m, d_axyz, d_x, d_phi):
        _check(lib().ibim_polarize_device_noalloc(self._h, m, d_axyz, d_x, d_phi), "ibim_polarize_device_noalloc")

    # ------------------------------------------------------------ timing --
    def profile(self, enable=True, stages=None):
        """Stage timers on (every stage), off, or only for the named stages
        (a list of stage names such as ["M2L_MIXED"])."""
        if stages is None:
            _check(lib().ibim_plan_profile(self._h, int(enable)), "ibim_plan_profile")
            return
        mask = stage_mask(stages)
        _check(lib().ibim_plan_profile_stages(self._h, mask if enable else 0), "ibim_plan_profile_stages")

    def stage_times(self, reset=True):
        """{stage: (milliseconds, launches)} from HIP events on the plan's stream."""
        ms = np.zeros(16)
        cnt = np.zeros(16, np.int64)
        _check(lib().ibim_plan_stage_times(self._h, _d(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64)), int(reset)),
               "ibim_plan_stage_times")
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(STAGES)}

    def stage_report(self, reset=True):
        """Stage intervals with the algorithmic work done inside them."""
        r = StageReport()
        _check(lib().ibim_plan_stage_report(self._h, C.byref(r), int(reset)), "ibim_plan_stage_report")
        return r.as_dict()

    def set_precision(self, precision):
        _check(lib().ibim_plan_set_precision(self._h, precision), "ibim_plan_set_precision")

    def footprint(self):
        return int(lib().ibim_plan_footprint(self._h))

    # -------------------------------------------------------- inspection --
    def stats(self):
        s = TreeStats()
        _check(lib().ibim_tree_stats_get(self._h, C.byref(s)), "ibim_tree_stats_get")
        return s.as_dict()

    def tree(self):
        nn = self.stats()["n_nodes"]
        ints = {k: np.zeros(nn, np.int32) for k in
                ["level", "parent", "slot", "leaf", "empty", "nsrc", "nu", "nv", "nw", "nx"]}
        centre = np.zeros((nn, 3))
        radius = np.zeros((nn, 3))
        _check(lib().ibim_tree_export(self._h, _i(ints["level"]), _i(ints["parent"]), _i(ints["slot"]),
                                      _i(ints["leaf"]), _i(ints["empty"]), _i(ints["nsrc"]), _d(centre),
                                      _d(radius), _i(ints["nu"]), _i(ints["nv"]), _i(ints["nw"]),
                                      _i(ints["nx"])), "ibim_tree_export")
        ints["centre"] = centre
        ints["radius"] = radius
        return ints

    def lists(self, kind):
        k = {"U": 0, "V": 1, "W": 2, "X": 3}[kind]
        nn = self.stats()["n_nodes"]
        off = np.zeros(nn + 1, np.int32)
        _check(lib().ibim_tree_lists(self._h, k, _i(off), None, 0), "ibim_tree_lists")
        items = np.zeros(max(int(off[-1]), 1), np.int32)
        _check(lib().ibim_tree_lists(self._h, k, _i(off), _i(items), len(items)), "ibim_tree_lists")
        return off, items[:off[-1]]

    def points(self, node):
        cnt = _check(lib().ibim_tree_points(self._h, node, None, 0), "ibim_tree_points")
        out = np.zeros(max(cnt, 1), np.int32)
        _check(lib().ibim_tree_points(self._h, node, _i(out), len(out)), "ibim_tree_points")
        return out[:cnt]


# ------------------------------------------------------------ geometry --
def geom_params(h, lo=(0.0, 0.0, 0.0), dims=(0, 0, 0), pad=None, probe=1.4, band=8, thickness=2, thres=1e-4,
                reinit_time=None, scale=1.0, init=0.0):
    """Grid and level-set options.  reinit_time defaults to 26 h, the final time
    the reference's input.cfg (reinit_step = 60) gives at its usual grid of
    256 cells over 600 scaled units (main.cpp:38-40).  init is the grid's value
    before expand (main.cpp:43 uses -2 * size); 0 picks -2 max(dims) h."""
    if pad is None:
        pad = (band + 4) * h
    if reinit_time is None:
        reinit_time = 26.0 * h
    return GeomParams(h, (C.c_double * 3)(*lo), (C.c_int * 3)(*dims), pad, probe, band, thickness, thres,
                      reinit_time, scale, init)


class Surface:
    """Tube quadrature nodes of a level set, resident on the GPU."""

    def __init__(self, handle):
        self._h = handle
        info = SurfaceInfo()
        _check(lib().ibim_surface_get_info(self._h, C.byref(info)), "ibim_surface_get_info")
        self.info = info.as_dict()
        self.n = int(info.n_nodes)

    @classmethod
    def from_atoms(cls, atom_xyz, radii, **kw):
        gp = geom_params(**kw)
        axyz = _f64(atom_xyz, (-1, 3))
        h = C.c_void_p()
        _check(lib().ibim_surface_from_atoms(C.byref(gp), len(axyz), _d(axyz), _d(_f64(radii)), C.byref(h)),
               "ibim_surface_from_atoms")
        return cls(h)

    @classmethod
    def from_grid(cls, phi, h, lo, thickness=2, scale=1.0):
        phi = _f64(phi)
        gp = geom_params(h, lo=lo, dims=phi.shape, thickness=thickness, scale=scale)
        hd = C.c_void_p()
        _check(lib().ibim_surface_from_grid(C.byref(gp), _d(phi.ravel()), C.byref(hd)), "ibim_surface_from_grid")
        return cls(hd)

    def arrays(self):
        xyz = np.zeros((self.n, 3))
        nrm = np.zeros((self.n, 3))
        w = np.zeros(self.n)
        _check(lib().ibim_surface_copy(self._h, _d(xyz), _d(nrm), _d(w)), "ibim_surface_copy")
        return xyz, nrm, w

    def device_ptrs(self):
        a, b, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().ibim_surface_device(self._h, C.byref(a), C.byref(b), C.byref(c)), "ibim_surface_device")
        return self.n, a, b, c

    def close(self):
        if self._h:
            lib().ibim_surface_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synthetic_globule(m, seed, confine_radius):
    """Seeded synthetic molecule: (xyz (m,3), radii (m,), charges (m,))."""
    xyz = np.zeros((m, 3))
    r = np.zeros(m)
    q = np.zeros(m)
    _check(lib().ibim_synthetic_globule(m, seed, confine_radius, _d(xyz), _d(r), _d(q)), "ibim_synthetic_globule")
    return xyz, r, q


# ------------------------------------------------------------- serving --
class Group:
    """A fused matvec over several plans (ibim_group_*): one launch per FMM
    stage for all members.  mixed: each member's mixed-precision matvec,
    else FP64."""

    def __init__(self, plans, mixed=False):
        self._plans = list(plans)  # (kept alive: the group must go first)
        hs = (C.c_void_p * len(self._plans))(*[pl.handle for pl in self._plans])
        self._h = C.c_void_p()
        _check(lib().ibim_group_create(hs, len(self._plans), int(bool(mixed)), C.byref(self._h)),
               "ibim_group_create")

    def matvec_device(self, d_x, d_y):
        """d_x, d_y: one device pointer (int or c_void_p) per member."""
        k = len(self._plans)
        assert len(d_x) == k and len(d_y) == k
        xs = (C.c_void_p * k)(*[C.c_void_p(int(getattr(v, "value", v))) for v in d_x])
        ys = (C.c_void_p * k)(*[C.c_void_p(int(getattr(v, "value", v))) for v in d_y])
        _check(lib().ibim_group_matvec_device(self._h, xs, ys), "ibim_group_matvec_device")

    def stage_report(self, reset=True):
        r = StageReport()
        _check(lib().ibim_group_stage_report(self._h, C.byref(r), int(reset)), "ibim_group_stage_report")
        return r.as_dict()

    def close(self):
        if self._h:
            lib().ibim_group_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Continuous batching of PB solves: `slots` concurrent requests, admission
    by estimated device footprint within `mem_budget` bytes (0 = 90% of free).
    fuse_points > 0: solves of at most that many tube points share fused
    matvec launches (ibim_server_set_fusion), waiting up to fuse_wait_us for
    partners."""

    def __init__(self, slots=4, mem_budget=0, fuse_points=0, fuse_wait_us=2000):
        self._h = C.c_void_p()
        _check(lib().ibim_server_create(slots, mem_budget, C.byref(self._h)), "ibim_server_create")
        self._keep = {}
        if fuse_points:
            self.set_fusion(fuse_points, fuse_wait_us)

    def set_fusion(self, max_points, wait_us=2000):
        _check(lib().ibim_server_set_fusion(self._h, int(max_points), int(wait_us)), "ibim_server_set_fusion")

    def fusion_stats(self):
        """(fused group matvecs, member matvecs they ran) so far."""
        a, b = C.c_int64(), C.c_int64()
        _check(lib().ibim_server_fusion_stats(self._h, C.byref(a), C.byref(b)), "ibim_server_fusion_stats")
        return int(a.value), int(b.value)

    def submit(self, atom_xyz, radii, charges, h=0.35, thickness=2, probe=1.4, restart=40, maxit=2000, tol=1e-8,
               **fmm):
        axyz = _f64(atom_xyz, (-1, 3))
        rad = _f64(radii)
        q = _f64(charges)
        req = Request(len(axyz), _d(axyz), _d(rad), _d(q), geom_params(h, thickness=thickness, probe=probe),
                      params(**fmm), restart, maxit, tol)
        rid = C.c_int64()
        _check(lib().ibim_server_submit(self._h, C.byref(req), C.byref(rid)), "ibim_server_submit")
        return int(rid.value)

    def wait(self, rid):
        res = Result()
        _check(lib().ibim_server_wait(self._h, rid, C.byref(res)), "ibim_server_wait")
        return res.as_dict()

    def now(self):
        return float(lib().ibim_server_now(self._h))

    def profile(self, stages):
        """Record the named stages in the plan of every request started from now on."""
        _check(lib().ibim_server_profile_stages(self._h, stage_mask(stages)), "ibim_server_profile_stages")

    def stage_report(self, reset=True):
        """Stage intervals and their algorithmic work, summed over finished requests."""
        r = StageReport()
        _check(lib().ibim_server_stage_report(self._h, C.byref(r), int(reset)), "ibim_server_stage_report")
        return r.as_dict()

    def close(self):
        if self._h:
            lib().ibim_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
