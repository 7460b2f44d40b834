"""GPU level set and tube (ibim_surface_from_grid, ibim_surface_from_atoms)
against the oracle's restatement of levelset.cpp:91-586 and main.cpp:43-62
on the same grids.

The GPU's WENO5 weights use refined reciprocals where the oracle divides,
so values agree to rounding rather than bit for bit: nodes and normals to
1e-9 absolute, weights to 1e-9 relative.  The discrete outcomes (tube
membership, flood fill, inclusions, step counts of evolve and of each
reinitialisation, which stop on an exact count) must be identical.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-9


def compare(xyz, nrm, w, oxyz, onrm, ow, what):
    assert len(w) == len(ow), f"{what}: {len(w)} GPU nodes vs {len(ow)} oracle nodes"
    assert np.abs(xyz - oxyz).max() < TOL, what
    assert np.abs(nrm - onrm).max() < TOL, what
    assert np.abs(w - ow).max() < TOL * np.abs(ow).max(), what


@pytest.mark.parametrize("thickness", [2, 4])
def test_surface_from_grid_sphere_and_cavity(ibim, gpu, oracle, thickness):
    h = 0.25
    lp = oracle.ls_params((49, 47, 45), (-6.1, -5.9, -5.6), h, thickness=thickness)
    P = oracle.grid_points(lp)
    r = np.linalg.norm(P - [0.05, -0.1, 0.02], axis=-1)
    for what, phi in [("sphere", r - 4.0), ("shell with a cavity", np.maximum(r - 5.0, 2.0 - r))]:
        oxyz, onrm, ow, oinc = oracle.surface_from_grid(lp, phi)
        s = ibim.Surface.from_grid(phi, h=h, lo=tuple(lp.lo), thickness=thickness)
        xyz, nrm, w = s.arrays()
        info = s.info
        s.close()
        print(f"{what}: thickness {thickness}, {len(w)} nodes, {info['inclusions']} inclusions", flush=True)
        assert info["inclusions"] == oinc, what
        compare(xyz, nrm, w, oxyz, onrm, ow, what)


def test_surface_from_grid_flood_start(ibim, gpu, oracle):
    """Grid point 0 just above / just below -tube: the fill reaches the tube
    or nothing (the GPU then reports that the level set has no tube points)."""
    h, T = 0.25, 4
    lp = oracle.ls_params((12, 12, 12), (0, 0, 0), h, thickness=T)
    s_ = oracle.grid_points(lp).sum(axis=-1)
    phi = s_ - (T * h - 0.5 * h)
    oxyz, onrm, ow, oinc = oracle.surface_from_grid(lp, phi)
    s = ibim.Surface.from_grid(phi, h=h, lo=(0, 0, 0), thickness=T)
    xyz, nrm, w = s.arrays()
    assert s.info["inclusions"] == oinc == 0
    s.close()
    compare(xyz, nrm, w, oxyz, onrm, ow, "reached")
    phi = s_ - (T * h + 0.5 * h)
    assert len(oracle.surface_from_grid(lp, phi)[2]) == 0
    with pytest.raises(ibim.IbimError, match="no tube points"):
        ibim.Surface.from_grid(phi, h=h, lo=(0, 0, 0), thickness=T)


def fitted_grid(xyz, rad, h, probe, band):
    pad = (band + 4) * h
    lo = (xyz - rad[:, None] - probe).min(axis=0) - pad
    hi = (xyz + rad[:, None] + probe).max(axis=0) + pad
    dims = tuple(int(math.ceil((hi[q] - lo[q]) / h)) + 1 for q in range(3))
    return tuple(float(v) for v in lo), dims


CASES = {
    # three overlapping atoms: the first reinitialisation stops early on the
    # gradient count, the second takes no step
    "three_atoms_early_stop": dict(atoms=([[0.0, 0.1, -0.05], [2.6, -0.2, 0.1], [1.1, 2.2, 0.3]], [1.8, 1.6, 1.7]),
                                   h=0.3, T=3, thres=0.1, rt=26 * 0.3, init="reference"),
    # a 24-atom globule whose kinks keep both passes running to their final time
    "globule24_full_reinit": dict(atoms=(24, 7, 7.0), h=0.4, T=3, thres=0.05, rt=26 * 0.4, init="reference"),
    # the library's default initial value (-2 max(dims) h) and the tight default tolerance
    "globule12_default_init": dict(atoms=(12, 11, 6.0), h=0.45, T=2, thres=1e-4, rt=4 * 0.45, init="default"),
}


@pytest.mark.parametrize("case", list(CASES))
def test_surface_from_atoms_matches_oracle(ibim, gpu, oracle, case):
    c = CASES[case]
    if isinstance(c["atoms"][0], int):
        xyz, rad, _ = ibim.synthetic_globule(*c["atoms"])
    else:
        xyz, rad = np.array(c["atoms"][0]), np.array(c["atoms"][1])
    h, T, thres, rt = c["h"], c["T"], c["thres"], c["rt"]
    probe, band = 1.4, 8
    lo, dims = fitted_grid(xyz, rad, h, probe, band)
    init = -2.0 * max(dims) if c["init"] == "reference" else -2.0 * max(dims) * h
    print(f"{case}: {len(xyz)} atoms, grid {dims}, init {init}", flush=True)
    lp = oracle.ls_params(dims, lo, h, probe=probe, band=band, thickness=T, thres=thres, reinit_time=rt)
    oxyz, onrm, ow, oinfo, _ = oracle.surface_from_atoms(lp, xyz, rad, init)
    kw = dict(h=h, lo=lo, dims=dims, probe=probe, band=band, thickness=T, thres=thres, reinit_time=rt)
    if c["init"] == "reference":
        kw["init"] = init
    s = ibim.Surface.from_atoms(xyz, rad, **kw)
    xyz_g, nrm_g, w_g = s.arrays()
    info = s.info
    s.close()
    print(f"  GPU steps {info['evolve_steps']} {info['reinit_steps']}, oracle {oinfo.evolve_steps} "
          f"{list(oinfo.reinit_steps)}; nodes {len(w_g)} / {len(ow)}", flush=True)
    assert info["dims"] == list(dims)
    assert info["evolve_steps"] == oinfo.evolve_steps
    assert info["reinit_steps"] == list(oinfo.reinit_steps)
    assert info["inclusions"] == oinfo.inclusions
    assert info["quality"][1] == oinfo.quality[1]
    assert info["quality"][0] == pytest.approx(oinfo.quality[0], rel=1e-9, abs=1e-15)
    compare(xyz_g, nrm_g, w_g, oxyz, onrm, ow, case)
    assert info["area"] == pytest.approx(ow.sum(), rel=1e-12)
