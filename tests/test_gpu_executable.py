"""The drop-in executable `levelset <input.cfg>` (main.cpp:14-79 of the
reference) end to end on the GPU: PQR -> level set -> tube nodes -> PB solve
-> polarized energy, for the Born ion (closed-form energy)."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ibim-levelset_amd", "bin", "levelset")
BORN = 1.0 / (8 * np.pi * 2.0) * (1 / 80.0 - 1.0)


def run(tmp_path, cfg_lines):
    cfg = tmp_path / "input.cfg"
    cfg.write_text("\n".join(cfg_lines) + "\n")
    out = subprocess.run([EXE, str(cfg)], capture_output=True, text=True, timeout=240, cwd=tmp_path)
    assert out.returncode == 0, out.stderr + out.stdout
    return out.stdout


def born_cfg(tmp_path, pqr, precision):
    return ["# Born ion, Angstrom grid", f"pqr_file = {pqr}", f"energy_file = {tmp_path / 'out.energy'}",
            f"potent_file = {tmp_path / 'out.potent'}", "grid_h = 0.125", "lvlset_thickness = 2",
            "lvlset_thres = 1e-4", "gmres_restart = 20", "gmres_maxiter = 200", "gmres_tol = 1e-8",
            "fmm_np = 4", "fmm_maxpoint = 64", "fmm_maxlevel = 10", "tau = 1.0", "solvent_kappa = 0.0",
            "solvent_dE = 80.0", "solvent_dI = 1.0", f"precision = {precision}"]


def test_born_ion_energy(tmp_path, gpu):
    pqr = tmp_path / "ion.pqr"
    pqr.write_text("ATOM      1  NA  ION     1       0.000   0.000   0.000  1.0000 2.0000\n")
    stdout = run(tmp_path, born_cfg(tmp_path, pqr, "fp64"))
    m = re.search(r"polarized energy:\s+(\S+)", stdout)
    assert m, stdout
    energy = float(m.group(1))
    assert energy == pytest.approx(BORN, rel=0.03), stdout
    assert "CONVERGED" in stdout
    n = int(re.search(r"POINTS NUM\s+(\d+)", stdout).group(1))
    assert len(open(tmp_path / "out.potent").read().split()) == 2 * n
    phi = float(open(tmp_path / "out.energy").read().split()[0])
    assert 0.5 * phi == pytest.approx(energy, rel=1e-5)  # files carry 6 significant digits, as the reference writes them


@pytest.mark.parametrize("precision", ["mixed", "mixed16"])
def test_born_ion_mixed_precision_matches_fp64(tmp_path, gpu, precision):
    """The executable's `precision` key: the mixed forms give the FP64 energy
    within 1e-6 relative; the energy is printed as the reference prints it
    (7 significant digits), so two printed values may differ by up to 1e-6
    more from rounding: bar 2e-6."""
    pqr = tmp_path / "ion.pqr"
    pqr.write_text("ATOM      1  NA  ION     1       0.000   0.000   0.000  1.0000 2.0000\n")
    e = {}
    for prec in ("fp64", precision):
        out = run(tmp_path, born_cfg(tmp_path, pqr, prec))
        assert "CONVERGED" in out, out
        e[prec] = float(re.search(r"polarized energy:\s+(\S+)", out).group(1))
    assert abs(e[precision] - e["fp64"]) <= 2e-6 * abs(e["fp64"]), e


def test_missing_pqr_fails_loudly(tmp_path, gpu):
    cfg = tmp_path / "bad.cfg"
    cfg.write_text("pqr_file = /nonexistent.pqr\ngrid_h = 0.25\n")
    out = subprocess.run([EXE, str(cfg)], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0 and "pqr_file" in out.stderr


def test_reference_scaled_grid_mode(tmp_path, gpu):
    """The reference's own grid mode (grid_size, no grid_h: main.cpp:33-45):
    the molecule is scaled so its half extent is 240 units of a [-300,300]^3
    grid.  An ion of radius 6 A keeps its probe-expanded sphere (240 + 1.4 *
    40 = 296 units) inside the grid; at grid_size 96 (h = 0.156 A) its
    energy is the closed form's within the discretisation error."""
    pqr = tmp_path / "ion6.pqr"
    pqr.write_text("ATOM      1  NA  ION     1       0.000   0.000   0.000  1.0000 6.0000\n")
    cfg = [f"pqr_file = {pqr}", f"energy_file = {tmp_path / 'out.energy'}", f"potent_file = {tmp_path / 'out.potent'}",
           "grid_size = 96", "lvlset_thickness = 2", "lvlset_thres = 1e-4", "reinit_step = 60",
           "gmres_restart = 20", "gmres_maxiter = 200", "gmres_tol = 1e-8", "fmm_np = 4", "fmm_maxpoint = 64",
           "fmm_maxlevel = 10", "tau = 1.0", "solvent_kappa = 0.0", "solvent_dE = 80.0", "solvent_dI = 1.0"]
    stdout = run(tmp_path, cfg)
    print(stdout[-600:])
    assert "CONVERGED" in stdout
    h = float(re.search(r"h\s+(\S+)\s+Angstroms", stdout).group(1))
    assert h == pytest.approx(600.0 / 96 / 40.0, rel=1e-6)
    energy = float(re.search(r"polarized energy:\s+(\S+)", stdout).group(1))
    born6 = 1.0 / (8 * np.pi * 6.0) * (1 / 80.0 - 1.0)
    assert energy == pytest.approx(born6, rel=0.03), (energy, born6)
    area = float(re.search(r"AREA APPROX\s+(\S+)", stdout).group(1))
    assert area == pytest.approx(4 * np.pi * 36.0, rel=0.03)
