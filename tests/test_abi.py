"""The C-ABI library loads and exports every function include/ibim.h declares
(no compute: these run without a GPU)."""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_library_exports_every_declared_symbol(ibim):
    L = ibim.lib()
    names = ibim.declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), f"{name} not exported"


def test_version_and_error_strings(ibim):
    L = ibim.lib()
    assert b"gfx950" in L.ibim_version()
    assert isinstance(L.ibim_last_error(), bytes)


def test_struct_layouts_match_header(ibim, tmp_path):
    """Every ctypes mirror has the C struct's size and field offsets (gcc on
    include/ibim.h), so a field added to the header but not to the binding,
    or in another place, fails here."""
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "native", "abi_layout.c"), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    c = dict(zip(lines[0::2], (int(v) for v in lines[1::2])))
    mirrors = {"ibim_params": ibim.Params, "ibim_tree_stats": ibim.TreeStats, "ibim_solve_info": ibim.SolveInfo,
               "ibim_stage_report": ibim.StageReport, "ibim_geom_params": ibim.GeomParams,
               "ibim_surface_info": ibim.SurfaceInfo, "ibim_request": ibim.Request, "ibim_result": ibim.Result}
    for name, cls in mirrors.items():
        assert ctypes.sizeof(cls) == c[name], name
        fields = [f for f, _ in cls._fields_]
        assert fields == [k.split(".", 1)[1] for k in c if k.startswith(name + ".")], name
        for f in fields:
            assert getattr(cls, f).offset == c[f"{name}.{f}"], f"{name}.{f}"


def test_plan_create_rejects_bad_arguments_before_any_device_call(ibim):
    """Argument checks come first, so they run without a GPU: each bad value
    returns IBIM_ERR_ARG (-1), leaves the output handle NULL and names the
    problem in ibim_last_error()."""
    L = ibim.lib()
    one = (ctypes.c_double * 3)(0.0, 0.0, 0.0)

    def params(**kw):
        base = dict(p=5, max_points=64, max_level=10, kappa=0.1257, eps_in=2.0, eps_out=80.0, vacant_radius=0.1,
                    precision=ibim.PREC_FP64, kernel_mode=ibim.KERNEL_PB)
        base.update(kw)
        return ibim.Params(**base)

    cases = [
        (params(p=1), 1, b"Chebyshev order"),
        (params(p=9), 1, b"Chebyshev order"),
        (params(), 0, b"number of points"),
        (params(max_points=0), 1, b"max_points"),
        (params(max_level=21), 1, b"max_level"),
        (params(eps_in=0.0), 1, b"eps_in"),
        (params(kappa=-1.0), 1, b"kappa"),
        (params(vacant_radius=0.0), 1, b"vacant_radius"),
        (params(precision=7), 1, b"precision"),
        (params(kernel_mode=5), 1, b"kernel_mode"),
        (params(inner_order=1), 1, b"inner_order"),
        (params(inner_order=5), 1, b"inner_order"),
    ]
    for prm, n, msg in cases:
        plan = ctypes.c_void_p(123)
        rc = L.ibim_plan_create(ctypes.byref(prm), n, one, one, one, ctypes.byref(plan))
        assert rc == -1, (msg, rc)
        assert plan.value is None
        assert msg in L.ibim_last_error(), (msg, L.ibim_last_error())
    plan = ctypes.c_void_p(123)
    assert L.ibim_plan_create(ctypes.byref(params()), 1, None, one, one, ctypes.byref(plan)) == -1
    assert b"required" in L.ibim_last_error()
