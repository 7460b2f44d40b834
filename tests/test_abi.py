"""The C-ABI library loads and exports every function include/ibim.h declares
(no compute: these run without a GPU)."""
import ctypes


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


def test_struct_layouts_match_header(ibim):
    # ibim_params: 3 ints, 4 doubles, 2 ints (C layout with padding)
    assert ctypes.sizeof(ibim.Params) == 56
    assert ctypes.sizeof(ibim.SolveInfo) == 24


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
