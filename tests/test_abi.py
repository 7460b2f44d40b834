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
